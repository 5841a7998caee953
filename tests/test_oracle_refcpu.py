"""The CPU baseline of the benchmark (oracle/refcpu.c: the reference's CPU data
flow on the FFTW3 / BLAS / LAPACK interfaces) computes the same operators as
the FP64 oracle: H on the full cube masked to the sphere, the full-grid overlap
and its Cholesky factor (against numpy), and the exchange pair solves."""
import numpy as np
import pytest

from inq_amd import systems as S
from oracle import oracle as orc
from tests.helpers import oracle_ham


@pytest.mark.parametrize("threads", [1, 3])
def test_reference_path_apply_matches_oracle(threads):
    sysm = S.small_config(seed=7, n=(16, 18, 20), natoms=3, nst=5, ecut=6.0)
    oh = oracle_ham(sysm)
    g = oh.gsphere((0, 0, 0), sysm.ecut)
    c = S.random_states(len(g), 5, 5, S.sphere_volume_element(sysm))
    ref = oh.apply_sphere(c, g)
    rc = orc.RefCPU(oh, 5, threads)
    cube = np.zeros((oh.N, 5), np.complex128)
    cube[g] = c.T
    out = rc.apply(cube)
    assert np.max(np.abs(out[g].T - ref)) / np.max(np.abs(ref)) < 1e-12


def test_reference_path_overlap_cholesky():
    rng = np.random.default_rng(3)
    phi = rng.standard_normal((500, 6)) + 1j * rng.standard_normal((500, 6))
    S_, info = orc.ref_overlap_cholesky(phi, 0.1)
    assert info == 0
    s = 0.1 * phi.conj().T @ phi
    L = np.tril(S_.T)  # column-major lower factor, read row-major
    L = np.conj(L)
    assert np.allclose(L @ L.conj().T, s, rtol=1e-12, atol=1e-12)


def test_reference_path_exchange_pairs():
    sysm = S.small_config(seed=9, n=(16, 16, 16), natoms=1, nst=4, ecut=6.0)
    oh = orc.Ham(sysm.lattice, sysm.n)
    rng = np.random.default_rng(5)
    phi = rng.standard_normal((oh.N, 4)) + 1j * rng.standard_normal((oh.N, 4))
    hf = np.ascontiguousarray(phi[:, :2])
    occ = np.array([2.0, 0.5])
    ref = oh.exchange_direct(phi, hf, occ, exx_coeff=1.0, scale=1.0)
    got = orc.ref_exchange_pairs(sysm.lattice, sysm.n, phi, hf, -0.5 * occ, 2)
    assert np.max(np.abs(got - ref)) / np.max(np.abs(ref)) < 1e-12
