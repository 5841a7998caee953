// propagate.cpp -- Taylor exponentials, the ETRS real-time step, per-step
// energies and dipole, and the hipGraph-captured propagation loop.
//
// exponential_in_place (src/operations/exponential.hpp:21-34): the Taylor
// terms (f H)^n/n! phi are accumulated into phi by the last pass of each H
// application (no separate shift/axpy kernels, src/operations/shift.hpp:33-44).
// exponential_2_for_1 (exponential.hpp:49-69): two accumulators share H^n phi.
// etrs (src/real_time/etrs.hpp:22-79) for a Hamiltonian that does not depend
// on the density: the half step exp_T4(-i dt/2 H) phi is the first output of
// the 2-for-1 call, the full-step prediction only feeds the density mixing,
// and every SCF pass of etrs.hpp:59-77 reproduces the same states, so the
// step is two in-place Taylor-4 half steps (8 H applications).
// The per-step observables (propagate.hpp:99 energy.calculate,
// energy.hpp:61-113, observables::dipole, dipole.hpp:24-59) come from a 9th H
// application whose x pass also accumulates the dipole of the real-space
// orbitals; everything is reduced in FP64 into a device ring of records.
#include "internal.hpp"

#include <map>
#include <memory>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>

using namespace inqx;

namespace inqx {

void launch_sphere_dots(Ctx &cx, int prec, const void *psi, const void *hpsi, const void *kin, int nst, int npw,
                        double *partial);
void launch_obs_combine(Ctx &cx, int prec, const void *obs_blocks, int nblocks, int natoms, int ntile, double vol_g,
                        double dip_scale, double *ring, int ring_cap, int *counter);
int dots_chunks(int npw);

struct ObsBlockHost {
  const double *dots;
  const void *proj;
  const void *dp;
  const double *dip;
  const double *occ;
  int nst, nch;
};

// grows a basis workspace unless a captured graph uses the basis
static void ensure_ws(Basis &B, DevMem &m, size_t bytes) {
  if (m.p && m.bytes >= bytes) return;
  INQX_CHECK(B.pinned == 0, INQX_ERR_STATE, "basis workspaces are pinned by a captured propagation graph");
  m.ensure(bytes);
}

template <typename T>
static void exp_in_place(inqx_ham_s *h, int nb, Basis *const *bases, const int *nst, void *const *phi,
                         std::complex<double> f, int order, void *const *acc = nullptr) {
  const int prec = std::is_same<T, double>::value ? INQX_FP64 : INQX_FP32;
  const size_t cs = sizeof(cplx<T>);
  std::vector<void *> hA(nb), hB(nb);
  for (int i = 0; i < nb; ++i) {
    Basis &B = *bases[i];
    ensure_ws(B, B.hn[prec][0], cs * (size_t)nst[i] * B.npw);
    ensure_ws(B, B.hn[prec][1], cs * (size_t)nst[i] * B.npw);
    hA[i] = B.hn[prec][0].p;
    hB[i] = B.hn[prec][1].p;
  }
  std::complex<double> coeff = 1.0;
  std::vector<const void *> in(nb);
  std::vector<void *> out(nb);
  for (int it = 1; it <= order; ++it) {
    coeff *= f / (double)it;
    for (int i = 0; i < nb; ++i) {
      in[i] = (it == 1) ? phi[i] : ((it % 2 == 0) ? hA[i] : hB[i]);
      out[i] = (it % 2 == 1) ? hA[i] : hB[i];
    }
    ApplyOpts o;
    // the Taylor terms accumulate into phi, or into the complex128 copy acc of
    // an FP32 state (phi then holds the FP32 input of the first application)
    o.acc1 = acc ? acc : phi;
    o.acc_wide = acc != nullptr;
    o.c1r = coeff.real();
    o.c1i = coeff.imag();
    o.write_h = it < order;
    launch_apply<T>(h, nb, bases, nst, in.data(), out.data(), &o);
  }
}

// complex64 <-> complex128 copies of a state block (n complex values)
__global__ void __launch_bounds__(256) k_widen(const float2 *a, double2 *b, long n) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) {
    const float2 v = a[i];
    b[i] = double2{(double)v.x, (double)v.y};
  }
}
__global__ void __launch_bounds__(256) k_narrow(const double2 *a, float2 *b, long n) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) {
    const double2 v = a[i];
    b[i] = float2{(float)v.x, (float)v.y};
  }
}
static void launch_widen(Ctx &cx, const void *a, void *b, long n) {
  hipLaunchKernelGGL(k_widen, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, cx.stream, (const float2 *)a, (double2 *)b, n);
  INQX_HIP(hipGetLastError());
}
static void launch_narrow(Ctx &cx, const void *a, void *b, long n) {
  ProfScope ps(cx, "state_narrow");
  hipLaunchKernelGGL(k_narrow, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, cx.stream, (const double2 *)a, (float2 *)b, n);
  INQX_HIP(hipGetLastError());
}

template <typename T>
static void exp_2_for_1(inqx_ham_s *h, Basis *B, int nst, void *phi, void *expphi, std::complex<double> f1,
                        std::complex<double> f2, int order) {
  const int prec = std::is_same<T, double>::value ? INQX_FP64 : INQX_FP32;
  const size_t cs = sizeof(cplx<T>);
  const size_t bytes = cs * (size_t)nst * B->npw;
  ensure_ws(*B, B->hn[prec][0], bytes);
  ensure_ws(*B, B->hn[prec][1], bytes);
  INQX_HIP(hipMemcpyAsync(expphi, phi, bytes, hipMemcpyDeviceToDevice, h->ctx->c.stream));
  std::complex<double> c1 = 1.0, c2 = 1.0;
  void *hA = B->hn[prec][0].p, *hB = B->hn[prec][1].p;
  for (int it = 1; it <= order; ++it) {
    c1 *= f1 / (double)it;
    c2 *= f2 / (double)it;
    const void *in = (it == 1) ? phi : ((it % 2 == 0) ? hA : hB);
    void *out = (it % 2 == 1) ? hA : hB;
    void *a1 = expphi, *a2 = phi;
    ApplyOpts o;
    o.acc1 = &a1;
    o.acc2 = &a2;
    o.c1r = c1.real();
    o.c1i = c1.imag();
    o.c2r = c2.real();
    o.c2i = c2.imag();
    o.write_h = it < order;
    launch_apply<T>(h, 1, &B, &nst, &in, &out, &o);
  }
}

// observables of the current states of every block into ring[counter % cap]
template <typename T>
static void energy_record(inqx_ham_s *h, int nb, Basis *const *bases, const int *nst, void *const *phi,
                          const double *const *occ, double *ring, int cap, int *counter, DevMem &obs_desc,
                          std::vector<char> &obs_host, bool capturing) {
  const int prec = std::is_same<T, double>::value ? INQX_FP64 : INQX_FP32;
  const size_t cs = sizeof(cplx<T>);
  std::vector<const void *> in(nb);
  std::vector<void *> out(nb);
  for (int i = 0; i < nb; ++i) {
    Basis &B = *bases[i];
    ensure_ws(B, B.hn[prec][0], cs * (size_t)nst[i] * B.npw);
    in[i] = phi[i];
    out[i] = B.hn[prec][0].p;
  }
  ApplyOpts o;
  o.observables = true;
  launch_apply<T>(h, nb, bases, nst, in.data(), out.data(), &o);
  std::vector<ObsBlockHost> ob(nb);
  for (int i = 0; i < nb; ++i) {
    Basis &B = *bases[i];
    const int nch = dots_chunks((int)B.npw);
    ensure_ws(B, B.dots, 32 * (size_t)nst[i] * nch);
    const void *kin = prec == INQX_FP64 ? B.d_kin64.p : B.d_kin32.p;
    launch_sphere_dots(h->ctx->c, prec, phi[i], out[i], kin, nst[i], (int)B.npw, B.dots.as<double>());
    ob[i].dots = B.dots.as<double>();
    ob[i].proj = h->kb.natoms > 0 ? B.proj[prec].p : nullptr;
    ob[i].dp = B.dp[prec].p;
    ob[i].dip = B.dip.as<double>();
    ob[i].occ = occ[i];
    ob[i].nst = nst[i];
    ob[i].nch = nch;
  }
  // the descriptor array is uploaded again only when a pointer changed
  const size_t ob_bytes = sizeof(ObsBlockHost) * nb;
  if (obs_host.size() != ob_bytes || std::memcmp(obs_host.data(), ob.data(), ob_bytes) != 0) {
    INQX_CHECK(!capturing, INQX_ERR_STATE, "observable buffers changed during graph capture");
    obs_desc.ensure(ob_bytes);
    INQX_HIP(hipMemcpyAsync(obs_desc.p, ob.data(), ob_bytes, hipMemcpyHostToDevice, h->ctx->c.stream));
    INQX_HIP(hipStreamSynchronize(h->ctx->c.stream));
    obs_host.assign(reinterpret_cast<const char *>(ob.data()), reinterpret_cast<const char *>(ob.data()) + ob_bytes);
  }
  const Grid &g = h->g;
  const double vol_g = g.vol / ((double)g.N * (double)g.N);
  // dipole = dV sum r |psi|^2 with psi = u / N
  const double dip_scale = (g.vol / (double)g.N) / ((double)g.N * (double)g.N);
  launch_obs_combine(h->ctx->c, prec, obs_desc.p, nb, h->kb.natoms, h->kb.ntile, vol_g, dip_scale, ring, cap, counter);
}

// ---------------------------------------------------------------- Kohn-Sham potential (Hartree level)
// update_hamiltonian (src/hamiltonian/self_consistency.hpp:100-125) without
// XC: V_KS = vion + V_H[total density] for every spin component, written in
// place into the Hamiltonian's device potential (V/N in FP64 and FP32), with
// the external and Hartree energies reduced in a fixed order.
constexpr int kScfChunk = 8192;

__global__ void __launch_bounds__(256) k_spin_total(const double *den, long N, int ncomp, double *tot) {
  const long p = (long)blockIdx.x * 256 + threadIdx.x;
  if (p >= N) return;
  double t = 0.0;
  for (int c = 0; c < ncomp; ++c) t += den[p * ncomp + c];
  tot[p] = t;
}

// v64/v32[c][p] = (vion + vh) / N; part[chunk] = {sum rho vion, sum rho vh}
__global__ void __launch_bounds__(256) k_vks_assemble(const double *vion, const double *vh, const double *rho, long N,
                                                     int ncomp, double invN, double *v64, float *v32, double *part) {
  __shared__ double red[2][256];
  const int tid = threadIdx.x;
  const long p0 = (long)blockIdx.x * kScfChunk;
  double ae = 0.0, ah = 0.0;
  for (long p = p0 + tid; p < p0 + kScfChunk && p < N; p += 256) {
    const double vi = vion[p], h = vh[p];
    const double v = (vi + h) * invN;
    for (int c = 0; c < ncomp; ++c) {
      v64[(size_t)c * N + p] = v;
      v32[(size_t)c * N + p] = (float)v;
    }
    ae += rho[p] * vi;
    ah += rho[p] * h;
  }
  red[0][tid] = ae;
  red[1][tid] = ah;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) {
      red[0][tid] += red[0][tid + w];
      red[1][tid] += red[1][tid + w];
    }
    __syncthreads();
  }
  if (tid == 0) {
    part[2 * blockIdx.x] = red[0][0];
    part[2 * blockIdx.x + 1] = red[1][0];
  }
}

// part[chunk] = sum |a - b| over the chunk (operations::integral_sum_absdiff before dV)
__global__ void __launch_bounds__(256) k_absdiff(const double *a, const double *b, long n, double *part) {
  __shared__ double red[256];
  const int tid = threadIdx.x;
  const long p0 = (long)blockIdx.x * kScfChunk;
  double acc = 0.0;
  for (long p = p0 + tid; p < p0 + kScfChunk && p < n; p += 256) acc += fabs(a[p] - b[p]);
  red[tid] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) red[tid] += red[tid + w];
    __syncthreads();
  }
  if (tid == 0) part[blockIdx.x] = red[0];
}

static int scf_chunks(long n) { return (int)((n + kScfChunk - 1) / kScfChunk); }

// workspaces of the potential update: total density, V_H, Hartree work, partial sums
struct KsWork {
  DevMem tot, vh, part;
};

// V_KS from the spin density den [N][ncomp] (FP64, device); energies (external,
// Hartree) on the host when requested (the call then synchronises)
static void ks_update(inqx_ham_s *h, const double *vion, const double *den, KsWork &w, double *energies) {
  Ctx &cx = h->ctx->c;
  const long N = h->g.N;
  const int ncomp = h->ncomp;
  INQX_CHECK(pair_grid_supported(h->g.n), INQX_ERR_UNSUPPORTED, "the Hartree solve needs built grid sizes");
  INQX_CHECK(h->v64.p && h->v32.p && h->v64.bytes >= sizeof(double) * (size_t)N * ncomp, INQX_ERR_STATE,
             "set a potential (inqx_ham_set_potential) before updating it");
  const double *rho = den;
  if (ncomp > 1) {
    w.tot.ensure(sizeof(double) * (size_t)N);
    hipLaunchKernelGGL(k_spin_total, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, cx.stream, den, N, ncomp,
                       w.tot.as<double>());
    INQX_HIP(hipGetLastError());
    rho = w.tot.as<double>();
  }
  w.vh.ensure(sizeof(double) * (size_t)N);
  h->hartree_work.ensure(16 * (size_t)N);
  h->hartree_part.ensure(sizeof(double) * (size_t)hartree_chunks(N));
  launch_hartree(h, rho, w.vh.as<double>(), h->hartree_work.p, h->hartree_part.as<double>());
  const int nch = scf_chunks(N);
  w.part.ensure(sizeof(double) * 2 * (size_t)nch);
  {
    ProfScope ps(cx, "vks_assemble");
    hipLaunchKernelGGL(k_vks_assemble, dim3(nch), dim3(256), 0, cx.stream, vion, w.vh.as<double>(), rho, N, ncomp,
                       1.0 / (double)N, h->v64.as<double>(), h->v32.as<float>(), w.part.as<double>());
  }
  INQX_HIP(hipGetLastError());
  if (energies) {
    std::vector<double> part(2 * (size_t)nch);
    INQX_HIP(hipMemcpyAsync(part.data(), w.part.p, sizeof(double) * part.size(), hipMemcpyDeviceToHost, cx.stream));
    INQX_HIP(hipStreamSynchronize(cx.stream));
    double e = 0.0, eh = 0.0;
    for (int c = 0; c < nch; ++c) {
      e += part[2 * c];
      eh += part[2 * c + 1];
    }
    const double dv = h->g.vol / (double)N;
    energies[0] = e * dv;         // energy.external (self_consistency.hpp:115)
    energies[1] = 0.5 * eh * dv;  // energy.hartree (self_consistency.hpp:120)
  }
}

// workspaces of the standalone potential update, one set per Hamiltonian
// (dropped with it: inqx_ham_destroy calls release_ks_work)
static std::map<inqx_ham_s *, std::unique_ptr<KsWork>> g_ks_work;
static KsWork &ks_work_of(inqx_ham_s *h) {
  auto &w = g_ks_work[h];
  if (!w) w.reset(new KsWork());
  return *w;
}
void release_ks_work(inqx_ham_s *h) { g_ks_work.erase(h); }

// spin density of every block's sphere states: den[N][ncomp] = sum occ |to_real(phi)|^2
template <typename T>
static void blocks_density(inqx_ham_s *h, int nb, Basis *const *bases, const int *nst, void *const *phi,
                           const double *const *occ, DevMem &rs, double *den) {
  Ctx &cx = h->ctx->c;
  const long N = h->g.N;
  const int prec = std::is_same<T, double>::value ? INQX_FP64 : INQX_FP32;
  launch_fill_zero(cx, den, sizeof(double) * (size_t)N * h->ncomp);
  for (int i = 0; i < nb; ++i) {
    INQX_CHECK(bases[i]->spin < h->ncomp, INQX_ERR_ARG, "a block's spin index has no density component");
    rs.ensure(sizeof(cplx<T>) * (size_t)nst[i] * N);
    ApplyOpts o;
    o.mode = 1;
    void *rsp = rs.p;
    o.rs = &rsp;
    o.rs_layout = 0;
    const void *pin = phi[i];
    void *pout = nullptr;
    Basis *b = bases[i];
    launch_apply<T>(h, 1, &b, &nst[i], &pin, &pout, &o);
    launch_density_add(cx, prec, rs.p, occ[i], nst[i], N, h->ncomp, bases[i]->spin, den);
  }
}

}  // namespace inqx

struct inqx_prop_s {
  inqx_ham ham;
  int prec;
  double dt;
  std::vector<Basis *> bases;
  std::vector<int> nst;
  std::vector<void *> phi;
  DevMem occ_dev;
  std::vector<const double *> occ;
  DevMem ring, counter, obs_desc;
  std::vector<char> obs_host;
  bool capturing = false;
  int cap = 0;
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int steps_done = 0;
  std::vector<std::shared_ptr<void>> keep;  // launch descriptors the graph uses
  // self-consistent steps (inqx_prop_set_scf): ionic potential, SCF limits,
  // half-step states, densities and workspaces, statistics of the last step
  const double *vion = nullptr;
  int nscf = 0;
  double scf_threshold = 5e-5;
  std::vector<DevMem> save;
  DevMem rho, rho_old, rs_work, absdiff_part;
  KsWork ks;
  int last_passes = 0;
  double last_delta = 0.0;
  // complex128 copies of FP32 states (inqx_prop_set_state_precision): the
  // propagated state; the caller's arrays receive it rounded after each half step
  bool wide = false;
  std::vector<DevMem> phi64;
  std::vector<void *> phi64p;
};

template <typename T> static void prop_step_eager(inqx_prop_s *p) {
  const std::complex<double> half(0.0, -p->dt / 2.0);
  const int nb = (int)p->bases.size();
  if (p->wide) {
    for (int hs = 0; hs < 2; ++hs) {
      exp_in_place<T>(p->ham, nb, p->bases.data(), p->nst.data(), p->phi.data(), half, 4, p->phi64p.data());
      for (int i = 0; i < nb; ++i)
        launch_narrow(p->ham->ctx->c, p->phi64p[i], p->phi[i], (long)p->nst[i] * p->bases[i]->npw);
    }
  } else {
    exp_in_place<T>(p->ham, nb, p->bases.data(), p->nst.data(), p->phi.data(), half, 4);
    exp_in_place<T>(p->ham, nb, p->bases.data(), p->nst.data(), p->phi.data(), half, 4);
  }
  energy_record<T>(p->ham, nb, p->bases.data(), p->nst.data(), p->phi.data(), p->occ.data(), p->ring.as<double>(),
                   p->cap, p->counter.as<int>(), p->obs_desc, p->obs_host, p->capturing);
}

// real_time::etrs with a density-dependent H (src/real_time/etrs.hpp:22-79):
// 2-for-1 half and full steps with H(t), density of the full-step prediction,
// update_hamiltonian, restore the half step, then up to nscf passes of the
// second half step with H(t + dt), each followed by the density, its change
// integral |drho| dV and update_hamiltonian, stopping when the change is below
// the threshold.  The pass count depends on the density change, which the host
// reads after each pass: these steps run eagerly.
template <typename T> static void prop_step_scf(inqx_prop_s *p) {
  inqx_ham_s *h = p->ham;
  Ctx &cx = h->ctx->c;
  const int nb = (int)p->bases.size();
  const std::complex<double> fhalf(0.0, -p->dt / 2.0), ffull(0.0, -p->dt);
  const long N = h->g.N;
  const size_t dbytes = sizeof(double) * (size_t)N * h->ncomp;
  p->rho.ensure(dbytes);
  p->rho_old.ensure(dbytes);
  std::vector<size_t> bytes(nb);
  for (int i = 0; i < nb; ++i) {
    bytes[i] = sizeof(cplx<T>) * (size_t)p->nst[i] * p->bases[i]->npw;
    p->save[i].ensure(bytes[i]);
    exp_2_for_1<T>(h, p->bases[i], p->nst[i], p->phi[i], p->save[i].p, fhalf, ffull, 4);
  }
  blocks_density<T>(h, nb, p->bases.data(), p->nst.data(), p->phi.data(), p->occ.data(), p->rs_work,
                    p->rho.as<double>());
  ks_update(h, p->vion, p->rho.as<double>(), p->ks, nullptr);
  const int nch = scf_chunks((long)N * h->ncomp);
  p->absdiff_part.ensure(sizeof(double) * (size_t)nch);
  std::vector<double> part(nch);
  int passes = 0;
  double delta = 0.0;
  for (int iscf = 0; iscf < p->nscf; ++iscf) {
    for (int i = 0; i < nb; ++i)
      INQX_HIP(hipMemcpyAsync(p->phi[i], p->save[i].p, bytes[i], hipMemcpyDeviceToDevice, cx.stream));
    exp_in_place<T>(h, nb, p->bases.data(), p->nst.data(), p->phi.data(), fhalf, 4);
    INQX_HIP(hipMemcpyAsync(p->rho_old.p, p->rho.p, dbytes, hipMemcpyDeviceToDevice, cx.stream));
    blocks_density<T>(h, nb, p->bases.data(), p->nst.data(), p->phi.data(), p->occ.data(), p->rs_work,
                      p->rho.as<double>());
    hipLaunchKernelGGL(k_absdiff, dim3(nch), dim3(256), 0, cx.stream, p->rho_old.as<double>(), p->rho.as<double>(),
                       (long)N * h->ncomp, p->absdiff_part.as<double>());
    INQX_HIP(hipGetLastError());
    INQX_HIP(hipMemcpyAsync(part.data(), p->absdiff_part.p, sizeof(double) * nch, hipMemcpyDeviceToHost, cx.stream));
    INQX_HIP(hipStreamSynchronize(cx.stream));
    delta = 0.0;
    for (int c = 0; c < nch; ++c) delta += part[c];
    delta *= h->g.vol / (double)N;
    const bool done = delta < p->scf_threshold || iscf == p->nscf - 1;
    ks_update(h, p->vion, p->rho.as<double>(), p->ks, nullptr);
    ++passes;
    if (done) break;
  }
  p->last_passes = passes;
  p->last_delta = delta;
  energy_record<T>(h, nb, p->bases.data(), p->nst.data(), p->phi.data(), p->occ.data(), p->ring.as<double>(), p->cap,
                   p->counter.as<int>(), p->obs_desc, p->obs_host, false);
}

extern "C" {

int inqx_ks_potential(inqx_ham h, const double *vion, const double *density, double *energies) {
  inqx_ctx ctx = h ? h->ctx : nullptr;
  API_BEGIN
  INQX_CHECK(h && vion && density, INQX_ERR_ARG, "null argument");
  INQX_CHECK(!ctx->c.host_only, INQX_ERR_STATE, "host-only context");
  INQX_CHECK(h->pinned == 0, INQX_ERR_STATE, "potential pinned by a captured propagation graph");
  ks_update(h, vion, density, ks_work_of(h), energies);
  if (!energies) INQX_HIP(hipStreamSynchronize(ctx->c.stream));
  API_END(ctx)
}

int inqx_prop_set_scf(inqx_prop p, const double *vion, int max_passes, double threshold) {
  inqx_ctx ctx = p ? p->ham->ctx : nullptr;
  API_BEGIN
  INQX_CHECK(p && max_passes >= 0 && std::isfinite(threshold) && threshold >= 0.0, INQX_ERR_ARG, "bad arguments");
  INQX_CHECK(!(p->wide && max_passes > 0), INQX_ERR_UNSUPPORTED, "FP64 states with FP32 applications: fixed-H steps only");
  INQX_CHECK(max_passes == 0 || vion, INQX_ERR_ARG, "the self-consistent step needs the ionic potential");
  INQX_CHECK(!p->exec, INQX_ERR_STATE, "the propagator already holds a captured graph of the fixed-H step");
  if (max_passes > 0) INQX_CHECK(pair_grid_supported(p->ham->g.n), INQX_ERR_UNSUPPORTED,
                                 "the Hartree solve needs built grid sizes");
  p->vion = max_passes > 0 ? vion : nullptr;
  p->nscf = max_passes;
  p->scf_threshold = threshold;
  p->save.clear();
  p->save = std::vector<DevMem>(p->bases.size());
  API_END(ctx)
}

int inqx_prop_set_state_precision(inqx_prop p, int precision) {
  inqx_ctx ctx = p ? p->ham->ctx : nullptr;
  API_BEGIN
  INQX_CHECK(p && (precision == INQX_FP64 || precision == INQX_FP32), INQX_ERR_ARG, "bad arguments");
  INQX_CHECK(!p->exec, INQX_ERR_STATE, "the propagator already holds a captured graph");
  const bool wide = precision == INQX_FP64 && p->prec == INQX_FP32;
  INQX_CHECK(!(wide && p->nscf > 0), INQX_ERR_UNSUPPORTED, "FP64 states with FP32 applications: fixed-H steps only");
  p->wide = wide;
  p->phi64.clear();
  p->phi64p.clear();
  if (wide) {
    p->phi64 = std::vector<DevMem>(p->bases.size());
    for (size_t i = 0; i < p->bases.size(); ++i) {
      const long n = (long)p->nst[i] * p->bases[i]->npw;
      p->phi64[i].ensure(16 * (size_t)n);
      launch_widen(ctx->c, p->phi[i], p->phi64[i].p, n);
      p->phi64p.push_back(p->phi64[i].p);
    }
    INQX_HIP(hipStreamSynchronize(ctx->c.stream));
  }
  API_END(ctx)
}

int inqx_prop_scf_info(inqx_prop p, int *passes, double *delta) {
  inqx_ctx ctx = p ? p->ham->ctx : nullptr;
  API_BEGIN
  INQX_CHECK(p, INQX_ERR_ARG, "null propagator");
  if (passes) *passes = p->last_passes;
  if (delta) *delta = p->last_delta;
  API_END(ctx)
}

int inqx_exponential(inqx_ham h, inqx_basis bs, int nst, int precision, double f_re, double f_im, int order,
                     void *phi) {
  inqx_ctx ctx = h ? h->ctx : nullptr;
  API_BEGIN
  INQX_CHECK(h && bs && phi && nst > 0 && order >= 1, INQX_ERR_ARG, "bad arguments");
  INQX_CHECK(!ctx->c.host_only, INQX_ERR_STATE, "host-only context");
  Basis *b = &bs->b;
  void *pp[1] = {phi};
  if (precision == INQX_FP64) exp_in_place<double>(h, 1, &b, &nst, pp, {f_re, f_im}, order);
  else if (precision == INQX_FP32) exp_in_place<float>(h, 1, &b, &nst, pp, {f_re, f_im}, order);
  else INQX_CHECK(false, INQX_ERR_ARG, "bad precision");
  API_END(ctx)
}

int inqx_exponential_2_for_1(inqx_ham h, inqx_basis bs, int nst, int precision, double f1_re, double f1_im,
                             double f2_re, double f2_im, int order, void *phi, void *expphi) {
  inqx_ctx ctx = h ? h->ctx : nullptr;
  API_BEGIN
  INQX_CHECK(h && bs && phi && expphi && phi != expphi && nst > 0 && order >= 1, INQX_ERR_ARG, "bad arguments");
  INQX_CHECK(!ctx->c.host_only, INQX_ERR_STATE, "host-only context");
  if (precision == INQX_FP64)
    exp_2_for_1<double>(h, &bs->b, nst, phi, expphi, {f1_re, f1_im}, {f2_re, f2_im}, order);
  else if (precision == INQX_FP32)
    exp_2_for_1<float>(h, &bs->b, nst, phi, expphi, {f1_re, f1_im}, {f2_re, f2_im}, order);
  else INQX_CHECK(false, INQX_ERR_ARG, "bad precision");
  API_END(ctx)
}

int inqx_prop_create(inqx_ham h, int nblocks, const inqx_basis *bases, const int *nst, int precision, void *const *phi,
                     const double *const *occ, double dt, int ring_capacity, inqx_prop *out) {
  inqx_ctx ctx = h ? h->ctx : nullptr;
  API_BEGIN
  INQX_CHECK(h && bases && nst && phi && occ && out && nblocks > 0 && ring_capacity > 0, INQX_ERR_ARG, "null argument");
  INQX_CHECK(!ctx->c.host_only, INQX_ERR_STATE, "host-only context");
  INQX_CHECK(precision == INQX_FP64 || precision == INQX_FP32, INQX_ERR_ARG, "bad precision");
  for (int i = 0; i < nblocks; ++i)
    INQX_CHECK(bases[i] && bases[i]->b.ham == h && nst[i] > 0 && phi[i] && occ[i], INQX_ERR_ARG, "bad block");
  std::unique_ptr<inqx_prop_s> p(new inqx_prop_s());
  p->ham = h;
  p->prec = precision;
  p->dt = dt;
  for (int i = 0; i < nblocks; ++i) {
    p->bases.push_back(&bases[i]->b);
    p->nst.push_back(nst[i]);
    p->phi.push_back(phi[i]);
  }
  std::vector<double> occ_all;
  for (int i = 0; i < nblocks; ++i) occ_all.insert(occ_all.end(), occ[i], occ[i] + nst[i]);
  upload(p->occ_dev, occ_all, ctx->c.stream);
  size_t off = 0;
  for (int i = 0; i < nblocks; ++i) {
    p->occ.push_back(p->occ_dev.as<double>() + off);
    off += nst[i];
  }
  p->cap = ring_capacity;
  p->ring.ensure(sizeof(double) * 8 * (size_t)ring_capacity);
  p->counter.ensure(sizeof(int));
  const int zero = 0;
  INQX_HIP(hipMemcpyAsync(p->counter.p, &zero, sizeof(int), hipMemcpyHostToDevice, ctx->c.stream));
  INQX_HIP(hipStreamSynchronize(ctx->c.stream));
  *out = p.release();
  API_END(ctx)
}

int inqx_prop_energy(inqx_prop p) {
  inqx_ctx ctx = p ? p->ham->ctx : nullptr;
  API_BEGIN
  INQX_CHECK(p, INQX_ERR_ARG, "null propagator");
  const int nb = (int)p->bases.size();
  if (p->prec == INQX_FP64)
    energy_record<double>(p->ham, nb, p->bases.data(), p->nst.data(), p->phi.data(), p->occ.data(),
                          p->ring.as<double>(), p->cap, p->counter.as<int>(), p->obs_desc, p->obs_host, false);
  else
    energy_record<float>(p->ham, nb, p->bases.data(), p->nst.data(), p->phi.data(), p->occ.data(),
                         p->ring.as<double>(), p->cap, p->counter.as<int>(), p->obs_desc, p->obs_host, false);
  API_END(ctx)
}

int inqx_prop_step(inqx_prop p, int nsteps, int use_graph) {
  inqx_ctx ctx = p ? p->ham->ctx : nullptr;
  API_BEGIN
  INQX_CHECK(p && nsteps >= 0, INQX_ERR_ARG, "bad arguments");
  Ctx &cx = ctx->c;
  INQX_CHECK(!(use_graph && p->nscf > 0), INQX_ERR_UNSUPPORTED,
             "self-consistent steps run eagerly (the SCF pass count depends on the density change)");
  for (int s = 0; s < nsteps; ++s) {
    if (p->nscf > 0) {
      if (p->prec == INQX_FP64) prop_step_scf<double>(p);
      else prop_step_scf<float>(p);
      ++p->steps_done;
      continue;
    }
    if (!use_graph || p->steps_done == 0) {
      // eager step; the first one also creates every workspace and launch descriptor
      if (p->prec == INQX_FP64) prop_step_eager<double>(p);
      else prop_step_eager<float>(p);
      ++p->steps_done;
      continue;
    }
    if (!p->exec) {
      INQX_HIP(hipStreamSynchronize(cx.stream));
      if (!p->stream) INQX_HIP(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
      hipStream_t saved = cx.stream;
      const bool prof = cx.prof;
      cx.prof = false;
      cx.stream = p->stream;
      INQX_HIP(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
      set_capture_keep(&p->keep);
      p->capturing = true;
      try {
        if (p->prec == INQX_FP64) prop_step_eager<double>(p);
        else prop_step_eager<float>(p);
      } catch (...) {
        p->capturing = false;
        set_capture_keep(nullptr);
        hipGraph_t g;
        (void)hipStreamEndCapture(p->stream, &g);
        if (g) (void)hipGraphDestroy(g);
        cx.stream = saved;
        cx.prof = prof;
        p->keep.clear();
        throw;
      }
      p->capturing = false;
      set_capture_keep(nullptr);
      INQX_HIP(hipStreamEndCapture(p->stream, &p->graph));
      cx.stream = saved;
      cx.prof = prof;
      INQX_HIP(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
      for (Basis *b : p->bases) ++b->pinned;
      ++p->ham->pinned;
    }
    {
      ProfScope ps(cx, "etrs_step_graph");
      INQX_HIP(hipGraphLaunch(p->exec, cx.stream));
    }
    ++p->steps_done;
  }
  API_END(ctx)
}

int inqx_energy(inqx_ham h, int nblocks, const inqx_basis *bases, const int *nst, int precision, void *const *phi,
                const double *const *occ, double *out) {
  inqx_ctx ctx = h ? h->ctx : nullptr;
  API_BEGIN
  INQX_CHECK(h && nblocks > 0 && bases && nst && phi && occ && out, INQX_ERR_ARG, "bad arguments");
  INQX_CHECK(!ctx->c.host_only, INQX_ERR_STATE, "host-only context");
  INQX_CHECK(precision == INQX_FP64 || precision == INQX_FP32, INQX_ERR_ARG, "bad precision");
  std::vector<Basis *> bb(nblocks);
  std::vector<double> occ_h;
  for (int i = 0; i < nblocks; ++i) {
    INQX_CHECK(bases[i] && bases[i]->b.ham == h, INQX_ERR_ARG, "basis belongs to another Hamiltonian");
    INQX_CHECK(nst[i] > 0 && phi[i] && occ[i], INQX_ERR_ARG, "bad block");
    bb[i] = &bases[i]->b;
    occ_h.insert(occ_h.end(), occ[i], occ[i] + nst[i]);
  }
  h->en_ring.ensure(8 * sizeof(double));
  h->en_counter.ensure(sizeof(int));
  h->en_occ.ensure(sizeof(double) * occ_h.size());
  const int zero = 0;
  INQX_HIP(hipMemcpyAsync(h->en_occ.p, occ_h.data(), sizeof(double) * occ_h.size(), hipMemcpyHostToDevice,
                          ctx->c.stream));
  INQX_HIP(hipMemcpyAsync(h->en_counter.p, &zero, sizeof(int), hipMemcpyHostToDevice, ctx->c.stream));
  std::vector<const double *> occ_d(nblocks);
  size_t off = 0;
  for (int i = 0; i < nblocks; ++i) {
    occ_d[i] = h->en_occ.as<double>() + off;
    off += nst[i];
  }
  if (precision == INQX_FP64)
    energy_record<double>(h, nblocks, bb.data(), nst, phi, occ_d.data(), h->en_ring.as<double>(), 1,
                          h->en_counter.as<int>(), h->en_desc, h->en_host, false);
  else
    energy_record<float>(h, nblocks, bb.data(), nst, phi, occ_d.data(), h->en_ring.as<double>(), 1,
                         h->en_counter.as<int>(), h->en_desc, h->en_host, false);
  INQX_HIP(hipMemcpyAsync(out, h->en_ring.p, 8 * sizeof(double), hipMemcpyDeviceToHost, ctx->c.stream));
  INQX_HIP(hipStreamSynchronize(ctx->c.stream));  // also keeps occ_h and zero alive until copied
  API_END(ctx)
}

int inqx_prop_records(inqx_prop p, int *count, double *out, int cap) {
  inqx_ctx ctx = p ? p->ham->ctx : nullptr;
  API_BEGIN
  INQX_CHECK(p && count, INQX_ERR_ARG, "null argument");
  int n = 0;
  INQX_HIP(hipMemcpyAsync(&n, p->counter.p, sizeof(int), hipMemcpyDeviceToHost, ctx->c.stream));
  INQX_HIP(hipStreamSynchronize(ctx->c.stream));
  *count = n;
  if (out && cap > 0) {
    const int m = std::min(std::min(n, cap), p->cap);
    std::vector<double> all((size_t)8 * p->cap);
    INQX_HIP(hipMemcpyAsync(all.data(), p->ring.p, all.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->c.stream));
    INQX_HIP(hipStreamSynchronize(ctx->c.stream));
    // the last m records in step order
    for (int k = 0; k < m; ++k) {
      const int step = n - m + k;
      std::memcpy(out + 8 * (size_t)k, all.data() + 8 * (size_t)(step % p->cap), 8 * sizeof(double));
    }
  }
  API_END(ctx)
}

int inqx_prop_destroy(inqx_prop p) {
  if (!p) return INQX_OK;
  if (p->exec) {
    for (Basis *b : p->bases) --b->pinned;
    --p->ham->pinned;
  }
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  if (p->stream) {
    (void)hipStreamSynchronize(p->stream);
    (void)hipStreamDestroy(p->stream);
  }
  delete p;
  return INQX_OK;
}

}  // extern "C"
