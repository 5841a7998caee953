// inqx/ks_hamiltonian.hpp -- header-only C++ adapter over the inqx C ABI that
// mirrors the reference's Hamiltonian call surface, so duck-typed reference
// templates (operations::exponential, real_time::etrs, energy::calculate,
// ground_state::subspace_diagonalization) can be written against it.
//
//   inqx::ks_hamiltonian     hamiltonian::ks_hamiltonian (src/hamiltonian/ks_hamiltonian.hpp:35-275)
//   inqx::orbital_block      the G-sphere form of states::orbital_set (src/states/orbital_set.hpp:18-192)
//   inqx::operations::*      operations::overlap / orthogonalize / rotate / exponential /
//                            preconditioner (src/operations/*.hpp)
//   inqx::eigensolvers::*    eigensolvers::steepest_descent (src/eigensolvers/steepest_descent.hpp)
//
// Errors: every C status < 0 becomes inqx::error (std::runtime_error) with the
// library's message; nothing aborts.
#pragma once

#include <complex>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../inqx.h"

namespace inqx {

struct error : std::runtime_error {
  int status;
  error(int st, const std::string &m) : std::runtime_error(m), status(st) {}
};

inline void check(int st, inqx_ctx ctx, const char *what) {
  if (st < 0) throw error(st, std::string(what) + ": " + (ctx ? inqx_last_error(ctx) : "null context"));
}

class context {
  inqx_ctx h_ = nullptr;

 public:
  explicit context(int device = 0, void *stream = nullptr) { check(inqx_ctx_create(device, stream, &h_), nullptr, "ctx"); }
  context(const context &) = delete;
  context &operator=(const context &) = delete;
  ~context() { inqx_ctx_destroy(h_); }
  inqx_ctx get() const { return h_; }
  void sync() const { check(inqx_sync(h_), h_, "sync"); }
};

// tag of the orbital_block constructor that allocates the same shape without copying
struct uninitialized_t {};
inline constexpr uninitialized_t uninitialized{};

// device array of complex states [nst][npw] (band-major G sphere)
template <typename Complex> class orbital_block {
  inqx_ctx ctx_;
  inqx_basis basis_;
  int nst_;
  long npw_;
  void *ptr_ = nullptr;

 public:
  static constexpr int precision = sizeof(Complex) == 16 ? INQX_FP64 : INQX_FP32;
  orbital_block(const context &ctx, inqx_basis basis, int nst) : orbital_block(ctx.get(), basis, nst) {}
  orbital_block(inqx_ctx ctx, inqx_basis basis, int nst) : ctx_(ctx), basis_(basis), nst_(nst) {
    check(inqx_basis_size(basis, &npw_), ctx_, "basis size");
    check(inqx_malloc(ctx_, sizeof(Complex) * (size_t)nst * npw_, &ptr_), ctx_, "malloc");
  }
  // the same basis and number of states, contents undefined (no copy, no synchronisation)
  orbital_block(const orbital_block &o, uninitialized_t) : orbital_block(o.ctx_, o.basis_, o.nst_) {}
  // orbital_set copies are deep (the reference returns states by value); the
  // copy is ordered on the context's stream like every other call, so no host
  // synchronisation is needed
  orbital_block(const orbital_block &o) : orbital_block(o.ctx_, o.basis_, o.nst_) {
    check(inqx_memcpy_d2d(ctx_, ptr_, o.ptr_, bytes()), ctx_, "copy");
  }
  orbital_block(orbital_block &&o) noexcept
      : ctx_(o.ctx_), basis_(o.basis_), nst_(o.nst_), npw_(o.npw_), ptr_(o.ptr_) {
    o.ptr_ = nullptr;
  }
  orbital_block &operator=(const orbital_block &) = delete;
  ~orbital_block() { inqx_free(ctx_, ptr_); }
  size_t bytes() const { return sizeof(Complex) * (size_t)nst_ * npw_; }
  int set_size() const { return nst_; }  // orbital_set::set_size
  long basis_size() const { return npw_; }
  inqx_basis basis() const { return basis_; }
  inqx_ctx ctx() const { return ctx_; }
  void *data() { return ptr_; }
  const void *data() const { return ptr_; }
  void upload(const std::vector<Complex> &h) { check(inqx_memcpy_h2d(ctx_, ptr_, h.data(), bytes()), ctx_, "h2d"); }
  std::vector<Complex> download() const {
    std::vector<Complex> h((size_t)nst_ * npw_);
    check(inqx_memcpy_d2h(ctx_, h.data(), ptr_, bytes()), ctx_, "d2h");
    return h;
  }
};

// the terms of hamiltonian::energy::calculate (energy.hpp:61-113) for a set of
// states, occupation-weighted and reduced in FP64, with the electronic dipole
struct energy_terms {
  double kinetic = 0.0, eigenvalues = 0.0, non_local = 0.0, norm = 0.0;
  double dipole[3] = {0.0, 0.0, 0.0};
};

class ks_hamiltonian {
  inqx_ctx ctx_;
  inqx_ham h_ = nullptr;

 public:
  // ks_hamiltonian(basis, ...) (ks_hamiltonian.hpp:92-102): cell and grid
  ks_hamiltonian(const context &ctx, const double lattice[9], const int n[3]) : ctx_(ctx.get()) {
    check(inqx_ham_create(ctx_, lattice, n, &h_), ctx_, "ham create");
  }
  ks_hamiltonian(const ks_hamiltonian &) = delete;
  ~ks_hamiltonian() { inqx_ham_destroy(h_); }
  inqx_ham get() const { return h_; }

  // scalar_potential() (ks_hamiltonian.hpp:241-243): [N][ncomp] in Hartree
  void scalar_potential(const std::vector<double> &v, int ncomp = 1) {
    check(inqx_ham_set_potential(h_, v.data(), ncomp), ctx_, "potential");
  }
  // uniform_vector_potential() (ks_hamiltonian.hpp:257-262), covariant
  void uniform_vector_potential(const double a_cov[3]) { check(inqx_ham_set_vector_potential(h_, a_cov), ctx_, "A"); }
  // one projector of projector_all (projector_all.hpp:40-109)
  int add_projector(const double center[3], double radius, int nlm, const double *kb_coeff, double width,
                    const double *matrix = nullptr) {
    const int st = inqx_ham_add_projector(h_, center, radius, nlm, matrix, kb_coeff, width);
    check(st, ctx_, "projector");
    return st;
  }
  inqx_basis make_basis(const double kcov[3], double ecut, int spin = 0) const {
    inqx_basis b;
    check(inqx_basis_create(h_, kcov, ecut, spin, &b), ctx_, "basis");
    return b;
  }

  // operator()(orbital_set<fourier_space>) (ks_hamiltonian.hpp:160-180)
  template <typename Complex> orbital_block<Complex> operator()(const orbital_block<Complex> &phi) const {
    orbital_block<Complex> hphi(phi, uninitialized);  // every element is written by the apply
    check(inqx_ham_apply(h_, phi.basis(), phi.set_size(), orbital_block<Complex>::precision, phi.data(), hphi.data()),
          ctx_, "apply");
    return hphi;
  }
  template <typename Complex> void apply(const orbital_block<Complex> &phi, orbital_block<Complex> &hphi) const {
    check(inqx_ham_apply(h_, phi.basis(), phi.set_size(), orbital_block<Complex>::precision, phi.data(), hphi.data()),
          ctx_, "apply");
  }

  // energy::calculate's sums for one block (one H application); occ has set_size() entries
  template <typename Complex>
  energy_terms energy(const orbital_block<Complex> &phi, const std::vector<double> &occ) const {
    if ((int)occ.size() != phi.set_size()) throw error(INQX_ERR_ARG, "energy: one occupation per state");
    inqx_basis b = phi.basis();
    const int nst = phi.set_size();
    void *p = const_cast<void *>(phi.data());
    const double *o = occ.data();
    double out[8];
    check(inqx_energy(h_, 1, &b, &nst, orbital_block<Complex>::precision, &p, &o, out), ctx_, "energy");
    energy_terms e;
    e.kinetic = out[0];
    e.eigenvalues = out[1];
    e.non_local = out[2];
    e.norm = out[3];
    for (int d = 0; d < 3; ++d) e.dipole[d] = out[4 + d];
    return e;
  }
  // non_local_energy(phi, occupations) (ks_hamiltonian.hpp:124-133)
  template <typename Complex>
  double non_local_energy(const orbital_block<Complex> &phi, const std::vector<double> &occ) const {
    return energy(phi, occ).non_local;
  }
};

namespace operations {

// operations::exponential_in_place (exponential.hpp:21-34)
template <typename Complex>
void exponential_in_place(const ks_hamiltonian &ham, std::complex<double> factor, orbital_block<Complex> &phi,
                          int order = 4) {
  check(inqx_exponential(ham.get(), phi.basis(), phi.set_size(), orbital_block<Complex>::precision, factor.real(),
                         factor.imag(), order, phi.data()),
        phi.ctx(), "exponential");
}

// operations::orthogonalize (orthogonalize.hpp:25-32), vol_el = V/N^2 on the sphere
template <typename Complex> void orthogonalize(orbital_block<Complex> &phi, double vol_el) {
  check(inqx_orthogonalize(phi.ctx(), orbital_block<Complex>::precision, phi.set_size(), phi.basis_size(), vol_el,
                           phi.data()),
        phi.ctx(), "orthogonalize");
}

// operations::preconditioner (preconditioner.hpp:26-58), Teter-Payne-Allan, in
// place on Fourier-space states; it reads |G|^2 from the Hamiltonian's grid
class preconditioner {
  const ks_hamiltonian &ham_;

 public:
  explicit preconditioner(const ks_hamiltonian &ham) : ham_(ham) {}
  const ks_hamiltonian &hamiltonian() const { return ham_; }
  template <typename Complex> void operator()(orbital_block<Complex> &phi) const {
    check(inqx_precondition(ham_.get(), phi.basis(), phi.set_size(), orbital_block<Complex>::precision, phi.data()),
          phi.ctx(), "precondition");
  }
};

// operations::no_preconditioner (preconditioner.hpp:68-73)
struct no_preconditioner {
  template <typename Block> void operator()(Block &) const {}
};

}  // namespace operations

namespace eigensolvers {

// eigensolvers::steepest_descent(ham, prec, phi) (steepest_descent.hpp:23-86):
// five preconditioned line minimisations per state, then orthogonalize
template <typename Complex>
void steepest_descent(const ks_hamiltonian &ham, const operations::preconditioner &, orbital_block<Complex> &phi) {
  check(inqx_steepest_descent(ham.get(), phi.basis(), phi.set_size(), orbital_block<Complex>::precision, 5, 1,
                              phi.data()),
        phi.ctx(), "steepest_descent");
}
template <typename Complex>
void steepest_descent(const ks_hamiltonian &ham, const operations::no_preconditioner &, orbital_block<Complex> &phi) {
  check(inqx_steepest_descent(ham.get(), phi.basis(), phi.set_size(), orbital_block<Complex>::precision, 5, 0,
                              phi.data()),
        phi.ctx(), "steepest_descent");
}

}  // namespace eigensolvers
}  // namespace inqx
