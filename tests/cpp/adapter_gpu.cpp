// GPU run of the header-only C++ adapter (include/inqx/ks_hamiltonian.hpp):
// reads a potential and sphere states written by tests/test_gpu_cpp_adapter.py,
// applies H through ks_hamiltonian::operator(), takes one Taylor-4 half step
// through a reference-shaped template and evaluates the energy terms, and
// writes the results back for the test to check against the oracle.
// Usage: adapter_gpu <dir>
#include <inqx/ks_hamiltonian.hpp>

#include <complex>
#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

template <typename T> std::vector<T> read_all(const std::string &path, size_t count) {
  std::vector<T> v(count);
  std::ifstream f(path, std::ios::binary);
  f.read(reinterpret_cast<char *>(v.data()), sizeof(T) * count);
  if ((size_t)f.gcount() != sizeof(T) * count) throw std::runtime_error("short read: " + path);
  return v;
}
template <typename T> void write_all(const std::string &path, const std::vector<T> &v) {
  std::ofstream f(path, std::ios::binary);
  f.write(reinterpret_cast<const char *>(v.data()), sizeof(T) * v.size());
}

// the shape of a reference-side caller (real_time::etrs's half step, etrs.hpp:41)
template <typename Ham, typename Block> void half_step(const Ham &ham, Block &phi, double dt) {
  inqx::operations::exponential_in_place(ham, std::complex<double>(0.0, -dt / 2.0), phi, 4);
}

int main(int argc, char **argv) {
  if (argc < 2) return 2;
  const std::string dir = argv[1];
  try {
    inqx::context ctx(0);
    const double lat[9] = {6, 0, 0, 0, 6, 0, 0, 0, 6};
    const int n[3] = {16, 16, 16};
    const int nst = 4;
    inqx::ks_hamiltonian ham(ctx, lat, n);
    ham.scalar_potential(read_all<double>(dir + "/v.bin", 16 * 16 * 16));
    const std::vector<double> proj = read_all<double>(dir + "/proj.bin", 3 + 16);
    ham.add_projector(proj.data(), 1.6, 4, proj.data() + 3, 0.7);
    const double k[3] = {0.0, 0.0, 0.0};
    inqx_basis b = ham.make_basis(k, 5.0);
    using cd = std::complex<double>;
    inqx::orbital_block<cd> phi(ctx, b, nst);
    phi.upload(read_all<cd>(dir + "/phi.bin", (size_t)nst * phi.basis_size()));
    auto hphi = ham(phi);
    write_all(dir + "/hphi.bin", hphi.download());
    std::vector<double> occ(nst, 2.0);
    inqx::energy_terms e = ham.energy(phi, occ);
    write_all(dir + "/energy.bin", std::vector<double>{e.kinetic, e.eigenvalues, e.non_local, e.norm});
    half_step(ham, phi, 0.05);
    write_all(dir + "/phi_half.bin", phi.download());
    std::printf("adapter_gpu: npw %ld ok\n", phi.basis_size());
  } catch (const std::exception &ex) {
    std::fprintf(stderr, "adapter_gpu: %s\n", ex.what());
    return 1;
  }
  return 0;
}
