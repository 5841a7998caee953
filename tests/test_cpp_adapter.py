"""The header-only C++ adapter (include/inqx/ks_hamiltonian.hpp) compiles
against the C ABI header and links against libinqx.so (CPU only: no call is
made)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = r'''
#include <inqx/ks_hamiltonian.hpp>
#include <complex>
// the shape of a reference-side caller: exponential_in_place over a block
template <typename Ham, typename Block> void half_step(const Ham &ham, Block &phi, double dt) {
  inqx::operations::exponential_in_place(ham, std::complex<double>(0.0, -dt / 2.0), phi, 4);
}
int main(int argc, char **) {
  if (argc > 5) {  // never taken: link-time check only
    inqx::context ctx(0);
    double lat[9] = {6, 0, 0, 0, 6, 0, 0, 0, 6};
    int n[3] = {16, 16, 16};
    inqx::ks_hamiltonian ham(ctx, lat, n);
    double k[3] = {0, 0, 0};
    inqx_basis b = ham.make_basis(k, 5.0);
    inqx::orbital_block<std::complex<float>> phi(ctx, b, 4);
    auto hphi = ham(phi);
    half_step(ham, phi, 0.05);
    inqx::operations::orthogonalize(phi, 1e-3);
    // the ground-state caller's shape (src/ground_state/calculator.hpp:160-168)
    inqx::operations::preconditioner prec(ham);
    inqx::eigensolvers::steepest_descent(ham, prec, phi);
    inqx::eigensolvers::steepest_descent(ham, inqx::operations::no_preconditioner{}, phi);
    std::vector<double> occ(4, 2.0);
    inqx::energy_terms e = ham.energy(phi, occ);
    double enl = ham.non_local_energy(phi, occ);
    if (e.kinetic + enl < -1e300) return 2;
  }
  return inqx_version() >= 1 ? 0 : 1;
}
'''


def test_adapter_compiles_links_and_runs_without_gpu(tmp_path):
    src = tmp_path / "adapter.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "adapter"
    lib = os.path.join(ROOT, "inq_amd")
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                           "-L", lib, "-linqx", "-Wl,-rpath," + lib])
    assert subprocess.call([str(exe)]) == 0
