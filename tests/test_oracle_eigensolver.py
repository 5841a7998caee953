"""The oracle's steepest descent and TPA preconditioner, pinned without the GPU.

* steepest_descent: the reference's own known answer
  (src/eigensolvers/steepest_descent.hpp:99-167): a 100x100 diagonal operator
  diag(1..100), 12 vectors exp(0.1 i ip ivec) on a trivial basis with volume
  element 1/100, orthogonalised, then 100 calls with the identity
  preconditioner; the printed eigenvalues to 12 decimals.
* preconditioner: the TPA function's closed values (K(0) = 1, K(1) = 65/81,
  K(x) -> 1/(2x)) and a single plane wave, whose mean kinetic energy is its own
  1/2|G|^2 so that it is scaled by K(1).
* a Kohn-Sham block: the ground-state loop (subspace diagonalisation, then
  steepest descent) converges to the four lowest eigenvalues of the dense H
  matrix (brute force on a 147-point sphere), keeps every Ritz value above its
  eigenvalue and leaves the states orthonormal.
"""
import numpy as np

from inq_amd import systems as S
from tests.helpers import oracle_ham


def test_sd_reference_known_answer():
    from oracle import oracle as O
    npoint, nvec = 100, 12
    diag = np.arange(1, npoint + 1, dtype=float)
    phi = np.exp(1j * 0.1 * np.outer(np.arange(nvec), np.arange(npoint)))
    vol = 1.0 / npoint
    phi, info = O.orthogonalize(phi, vol, True)
    assert info == 0

    def apply(x):
        return x * diag[None, :]

    for _ in range(100):
        phi, info = O.steepest_descent(apply, None, phi, vol)
        assert info == 0
    ev = np.abs(O.overlap_diagonal(phi, apply(phi), vol))
    gold = [1.000000001634, 2.000000003689, 3.000000001955, 4.000000001760, 5.000000002561, 6.000000003127,
            7.000000002312, 8.000000000292, 8.999999999033, 9.999999998497, 10.999999998768, 11.999999998422]
    # the reference prints 12 decimals; its own check is looser (Catch Approx)
    assert np.max(np.abs(ev - gold)) < 2e-12


def test_tpa_function_values():
    from oracle import oracle as O
    assert O.tpa_k(0.0) == 1.0
    assert abs(O.tpa_k(1.0) - 65.0 / 81.0) < 1e-16
    x = 1e6
    assert abs(O.tpa_k(x) * 2.0 * x - 1.0) < 1e-5


def test_tpa_plane_wave():
    from oracle import oracle as O
    g2h = np.array([0.0, 0.7, 2.5, 4.0])
    phi = np.zeros((2, 4), dtype=complex)
    phi[0, 2] = 0.3 - 0.4j
    phi[1, 1] = 2.0j
    out = O.tpa_precondition(phi, g2h)
    assert abs(out[0, 2] - phi[0, 2] * 65.0 / 81.0) < 1e-16
    assert abs(out[1, 1] - phi[1, 1] * 65.0 / 81.0) < 1e-15
    assert np.count_nonzero(out) == 2


def test_sd_ground_state_loop_reaches_dense_eigenvalues():
    """subspace diagonalisation + steepest descent, the ground-state loop order
    (src/ground_state/calculator.hpp:146-167), converges to the lowest
    eigenvalues of the dense H matrix built column by column on a 147-point sphere"""
    from oracle import oracle as O
    sysm = S.small_config(seed=5, n=(16, 16, 16), natoms=2, nst=6, ecut=6.0)
    oh = oracle_ham(sysm)
    g = oh.gsphere((0, 0, 0), sysm.ecut)
    npw = len(g)
    vol = S.sphere_volume_element(sysm)
    g2h = oh.gsphere_kinetic((0, 0, 0), g)
    dense = oh.apply_sphere(np.eye(npw, dtype=complex), g).reshape(npw, npw).T
    assert np.max(np.abs(dense - dense.conj().T)) < 1e-12
    w = np.linalg.eigvalsh(dense)[:sysm.nst]
    phi = S.random_states(npw, sysm.nst, 3, vol)

    def apply(x):
        return oh.apply_sphere(x, g).reshape(x.shape)

    for _ in range(60):
        phi, ev = oh.subspace_diag_sphere(phi, g)
        phi, info = O.steepest_descent(apply, lambda r: O.tpa_precondition(r, g2h), phi, vol)
        assert info == 0
    # the lowest four levels (0.04 Ha below the rest) converge; every Ritz value
    # bounds its eigenvalue from above (Rayleigh-Ritz)
    assert np.max(np.abs(ev[:4] - w[:4])) < 1e-8
    assert np.all(ev >= w - 1e-10)
    s = vol * (phi.conj() @ phi.T)
    assert np.max(np.abs(s - np.eye(sysm.nst))) < 1e-12
