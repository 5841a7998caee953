"""The FP64 Hermitian eigensolver on the device (inqx_diagonalize, and the
subspace step built on it).

Checkers: the reference's own known answers (diagonalize.hpp:213-317: a 2x2
diagonal, a 3x3 real and a 3x3 complex matrix, eigenvalues to 1e-10), scipy's
eigh on random Hermitian matrices up to n = 432 (eigenvalues to 1e-12 of the
norm), and the defining properties of the output: A conj(r_k) = lambda_k
conj(r_k) for every row r_k (the reference's orientation) and orthonormal rows
(1e-11), also for matrices with 3-fold degenerate eigenvalues and for a
diagonal matrix.  The subspace step at 128 states agrees with the oracle's.
"""
import numpy as np
import pytest
import scipy.linalg as sla

from inq_amd import Context, DeviceArray
from inq_amd import operations as ops
from inq_amd import systems as S
from tests.helpers import gpu_ham, oracle_ham, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return Context(0)


def _diag(ctx, a):
    d = DeviceArray.from_numpy(ctx, np.ascontiguousarray(a, dtype=np.complex128))
    w = ops.diagonalize(ctx, d)
    return w, d.numpy()


def _check_vectors(a, w, rows, tol=1e-11):
    n = len(w)
    u = rows.conj().T  # columns: eigenvectors of a
    scale = max(1.0, np.max(np.abs(w)))
    assert np.max(np.abs(a @ u - u * w)) < tol * scale
    assert np.max(np.abs(u.conj().T @ u - np.eye(n))) < tol


def test_known_answers(ctx):
    say("case: diagonalize KATs")
    w, v = _diag(ctx, np.array([[4.0, 0.0], [0.0, 2.0]]))
    assert np.allclose(w, [2.0, 4.0], atol=1e-14)
    assert abs(abs(v[0, 1]) - 1.0) < 1e-14 and abs(v[0, 0]) < 1e-14
    a = np.array([[0.088958, 1.183407, 1.191946], [1.183407, 1.371884, 0.705297], [1.191946, 0.705297, 0.392459]])
    w, v = _diag(ctx, a)
    assert np.allclose(w, [-1.0626903983, 0.1733844724, 2.7426069258], atol=2e-10)
    _check_vectors(a, w, v)
    c = np.array([[0.088958, 1.183407 + 0.08285j, 1.191946 + 0.09413j],
                  [1.183407 - 0.08285j, 1.371884, 0.705297 + 0.12840j],
                  [1.191946 - 0.09413j, 0.705297 - 0.12840j, 0.392459]])
    w, v = _diag(ctx, c)
    assert np.allclose(w, [-1.0703967402, 0.1722879629, 2.7514097773], atol=2e-10)
    _check_vectors(c, w, v)
    w, v = _diag(ctx, np.array([[3.5]]))
    assert w[0] == 3.5 and abs(abs(v[0, 0]) - 1.0) < 1e-15


@pytest.mark.parametrize("n", [8, 40, 128, 432])
def test_random_hermitian_vs_scipy(ctx, n):
    say("case: diagonalize random", n)
    rng = np.random.default_rng(n)
    x = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    a = x + x.conj().T
    w, v = _diag(ctx, a)
    ref = sla.eigh(a, eigvals_only=True)
    assert np.max(np.abs(w - ref)) < 1e-12 * np.max(np.abs(ref))
    _check_vectors(a, w, v, 1e-11)


@pytest.mark.parametrize("n", [60, 432])
def test_degenerate_and_diagonal(ctx, n):
    say("case: diagonalize degenerate", n)
    rng = np.random.default_rng(7 + n)
    q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    lam = np.repeat(np.arange(n // 3, dtype=float), 3)
    a = (q * lam) @ q.conj().T
    a = 0.5 * (a + a.conj().T)
    w, v = _diag(ctx, a)
    assert np.max(np.abs(w - lam)) < 1e-12 * n
    _check_vectors(a, w, v, 1e-10)
    dg = np.diag(rng.standard_normal(n)).astype(np.complex128)
    w, v = _diag(ctx, dg)
    assert np.allclose(w, np.sort(np.diag(dg).real), atol=1e-14)
    _check_vectors(dg, w, v)


def test_subspace_diag_128_states_vs_oracle(ctx):
    say("case: subspace diag 128 states 24^3")
    sysm = S.small_config(seed=13, n=(24, 24, 24), a=10.0, natoms=4, nst=128, ecut=10.0)
    gh = gpu_ham(ctx, sysm)
    oh = oracle_ham(sysm)
    b = gh.basis((0, 0, 0), sysm.ecut)
    vol = S.sphere_volume_element(sysm)
    c = S.random_states(b.npw, sysm.nst, 5, vol)
    d = DeviceArray.from_numpy(ctx, c)
    ev = ops.subspace_diagonalization(gh, b, d)
    phi_ref, ev_ref = oh.subspace_diag_sphere(c, b.gindex())
    assert np.max(np.abs(ev - ev_ref)) < 1e-10 * max(1.0, np.max(np.abs(ev_ref)))
    out = d.numpy()
    gaps = np.diff(ev_ref)
    ok = np.ones(len(ev_ref), bool)  # states of well-separated eigenvalues agree up to a phase
    ok[:-1] &= gaps > 1e-6
    ok[1:] &= gaps > 1e-6
    ph = np.sum(phi_ref.conj() * out, 1) * vol
    assert np.max(np.abs(np.abs(ph[ok]) - 1.0)) < 1e-8
