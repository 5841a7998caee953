"""GPU parity of overlap / Cholesky / triangular inverse / rotate /
orthogonalize / subspace diagonalisation.

References: numpy/scipy for the dense algebra (textbook routines), the
reference's own 2x2 known answers (src/matrix/cholesky.hpp:125-145,
src/matrix/invert_triangular.hpp:72-125), and the FP64 oracle for the
Gram-Schmidt orientation and the subspace step.
Tolerances: FP64 states 1e-12 relative; FP32 states 2e-6 relative on
overlaps (FP32 MFMA products, FP64 accumulation across K chunks) and 1e-5 on
orthonormality.
"""
import numpy as np
import pytest
import scipy.linalg as sla

from inq_amd import Context, DeviceArray
from inq_amd import operations as ops
from inq_amd import systems as S
from tests.helpers import gpu_ham, oracle_ham, rel_err, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return Context(0)


def _rand(shape, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-13), (np.complex64, 2e-6)])
@pytest.mark.parametrize("n1,n2,npw", [(37, 70, 3001), (128, 128, 20000), (5, 3, 17)])
def test_overlap_vs_numpy(ctx, dtype, tol, n1, n2, npw):
    say("case: overlap", n1, n2, npw, np.dtype(dtype).name)
    a = _rand((n1, npw), 1).astype(dtype)
    b = _rand((n2, npw), 2).astype(dtype)
    S_d = ops.overlap(ctx, DeviceArray.from_numpy(ctx, a), DeviceArray.from_numpy(ctx, b), 0.37)
    ref = 0.37 * (a.astype(np.complex128).conj() @ b.astype(np.complex128).T)
    assert rel_err(S_d.numpy(), ref) < tol


def test_cholesky_kat_and_scipy(ctx):
    say("case: cholesky")
    m = np.array([[6432.12, 4502.48], [4502.48, 3151.74]], dtype=np.complex128)
    d = DeviceArray.from_numpy(ctx, m)
    L = ops.cholesky(ctx, d).numpy()
    assert L[0, 0].real == pytest.approx(80.2005, rel=1e-6)
    assert L[0, 1] == 0.0
    assert L[1, 0].real == pytest.approx(56.1402992511, rel=1e-9)
    assert L[1, 1].real == pytest.approx(0.0824620974, rel=1e-7)
    a = _rand((37, 37), 3)
    s = a @ a.conj().T + 37 * np.eye(37)
    L = ops.cholesky(ctx, DeviceArray.from_numpy(ctx, s)).numpy()
    assert np.max(np.abs(L - sla.cholesky(s, lower=True))) < 1e-12


def test_cholesky_not_positive_definite_is_a_status(ctx):
    from inq_amd import InqxError
    a = _rand((8, 8), 4)
    s = a @ a.conj().T + 8 * np.eye(8)
    s[3, 3] = -5.0
    with pytest.raises(InqxError) as e:
        ops.cholesky(ctx, DeviceArray.from_numpy(ctx, s))
    assert "status -4" in str(e.value)


@pytest.mark.parametrize("n", [32, 33, 100, 128, 129, 433])
def test_blocked_cholesky_and_inverse_vs_scipy(ctx, n):
    """Sizes on both Cholesky paths (one work-group up to 128, 32-wide panels
    above, a ragged last panel included) and across several inverse block rows;
    the strict upper triangle of the input holds garbage that must be ignored."""
    say("case: blocked cholesky / invert_triangular", n)
    a = _rand((n, n), 10 + n)
    s = a @ a.conj().T + n * np.eye(n)
    s_in = s.copy()
    s_in[np.triu_indices(n, 1)] = 7.0 - 3.0j
    L = ops.cholesky(ctx, DeviceArray.from_numpy(ctx, s_in)).numpy()
    L_ref = sla.cholesky(s, lower=True)
    assert np.max(np.abs(L - L_ref)) < 1e-12 * np.max(np.abs(L_ref))
    assert np.all(L[np.triu_indices(n, 1)] == 0.0)
    X = ops.invert_triangular(ctx, DeviceArray.from_numpy(ctx, L_ref)).numpy()
    assert np.max(np.abs(X @ L_ref - np.eye(n))) < 1e-12
    assert np.all(X[np.triu_indices(n, 1)] == 0.0)


@pytest.mark.parametrize("n,bad", [(100, 70), (200, 150)])
def test_cholesky_not_positive_definite_in_a_later_panel(ctx, n, bad):
    from inq_amd import InqxError
    a = _rand((n, n), 11)
    s = a @ a.conj().T + n * np.eye(n)
    s[bad, bad] = -5.0
    with pytest.raises(InqxError) as e:
        ops.cholesky(ctx, DeviceArray.from_numpy(ctx, s))
    assert f"minor {bad + 1} " in str(e.value)
    # the context stays usable, and the next factorisation starts clean
    s[bad, bad] = s[bad - 1, bad - 1]
    L = ops.cholesky(ctx, DeviceArray.from_numpy(ctx, s)).numpy()
    assert np.max(np.abs(L - sla.cholesky(s, lower=True))) < 1e-11


def test_invert_triangular(ctx):
    say("case: invert_triangular")
    m = np.array([[4.0, 0.0], [-1.0, 2.0]], dtype=np.complex128)
    X = ops.invert_triangular(ctx, DeviceArray.from_numpy(ctx, m)).numpy()
    assert np.allclose(X, [[0.25, 0.0], [0.125, 0.5]], rtol=0, atol=1e-15)
    nn = 15
    ii, jj = np.meshgrid(np.arange(nn), np.arange(nn), indexing="ij")
    arr = np.where(ii < jj, 0.0, np.cos(ii) * (jj + 0.1) + np.sin(jj - 0.25) * (ii + 1.0)).astype(np.complex128)
    arr *= np.exp(1j * (ii + 1.0) * (np.cos(jj) - 0.25))
    arr = np.tril(arr)
    X = ops.invert_triangular(ctx, DeviceArray.from_numpy(ctx, arr)).numpy()
    assert np.max(np.abs(X @ arr - np.eye(nn))) < 1e-12


@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-13), (np.complex64, 2e-6)])
def test_rotate_forms(ctx, dtype, tol):
    say("case: rotate", np.dtype(dtype).name)
    n, npw = 70, 5003
    phi = _rand((n, npw), 5).astype(dtype)
    rot = _rand((n, n), 6)
    d = DeviceArray.from_numpy(ctx, phi)
    ops.rotate(ctx, DeviceArray.from_numpy(ctx, rot), d)
    p64 = phi.astype(np.complex128)
    # Phi <- Phi rot^H with Phi = [points][states] = phi^T
    ref = (p64.T @ rot.conj().T).T
    assert rel_err(d.numpy(), ref) < tol
    psi = _rand((n, npw), 7).astype(dtype)
    dpsi = DeviceArray.from_numpy(ctx, psi)
    ops.rotate_add(ctx, DeviceArray.from_numpy(ctx, rot), DeviceArray.from_numpy(ctx, phi), dpsi, alpha=-1.0, beta=1.0)
    ref = (psi.astype(np.complex128).T - p64.T @ rot).T
    assert rel_err(dpsi.numpy(), ref) < tol * 10


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_orthogonalize_vs_oracle(ctx, dtype):
    from oracle import oracle as orc
    say("case: orthogonalize", np.dtype(dtype).name)
    n, npw, vol = 64, 9001, 0.013
    phi = _rand((n, npw), 8).astype(dtype)
    d = DeviceArray.from_numpy(ctx, phi)
    ops.orthogonalize(ctx, d, vol)
    out = d.numpy().astype(np.complex128)
    s = vol * (out.conj() @ out.T)
    assert np.max(np.abs(s - np.eye(n))) < (1e-12 if dtype == np.complex128 else 1e-5)
    ref, info = orc.orthogonalize(phi.astype(np.complex128), vol, True)
    assert info == 0
    assert rel_err(out, ref) < (1e-11 if dtype == np.complex128 else 5e-5)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_subspace_diagonalization_vs_oracle(ctx, dtype):
    sysm = S.small_config(seed=9, n=(20, 20, 20), natoms=3, nst=8, ecut=7.0)
    say("case: subspace diag", np.dtype(dtype).name)
    gh = gpu_ham(ctx, sysm)
    oh = oracle_ham(sysm)
    b = gh.basis((0, 0, 0), sysm.ecut)
    vol = S.sphere_volume_element(sysm)
    c = S.random_states(b.npw, sysm.nst, 21, vol)
    d = DeviceArray.from_numpy(ctx, c.astype(dtype))
    ev = ops.subspace_diagonalization(gh, b, d)
    phi_ref, ev_ref = oh.subspace_diag_sphere(c.astype(dtype).astype(np.complex128), b.gindex())
    tol = 1e-10 if dtype == np.complex128 else 2e-5
    assert np.max(np.abs(ev - ev_ref)) < tol * max(1.0, np.max(np.abs(ev_ref)))
    out = d.numpy().astype(np.complex128)
    # each state agrees with the oracle's up to a phase (eigenvalues are non-degenerate here)
    assert np.min(np.diff(ev_ref)) > 1e-3
    ph = np.sum(phi_ref.conj() * out, 1) * vol
    assert np.max(np.abs(np.abs(ph) - 1.0)) < (1e-9 if dtype == np.complex128 else 1e-4)
