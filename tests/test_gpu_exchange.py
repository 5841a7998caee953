"""GPU parity of the sphere <-> real-space transforms, the batched Poisson
solves and exact exchange (direct and ACE) against the FP64 oracle and
numpy's FFT.

Tolerances (max abs error / max abs reference):
  FP64: 1e-11 for transforms and Poisson, 1e-10 for exchange and ACE
        (sums over pairs of FP64 FFT round trips);
  FP32: 2e-5 for transforms and Poisson, 5e-5 for exchange and ACE.
Exchange energies: FP64 relative 1e-10, FP32 relative 1e-5.
The chunked pair loop (small pair workspace) must reproduce the single-chunk
result bit for bit: the accumulation order over j is the same.
"""
import numpy as np
import pytest

from inq_amd import Context, DeviceArray, KSHamiltonian
from inq_amd import systems as S
from inq_amd._lib import InqxError
from inq_amd.exchange import ExchangeOperator
from oracle import oracle as orc
from tests.helpers import rel_err, say

pytestmark = pytest.mark.gpu

LAT = np.eye(3) * 6.0
NG = (16, 16, 16)
N = int(np.prod(NG))
V = 216.0
TOL_T = {np.complex128: 1e-11, np.complex64: 2e-5}
TOL_X = {np.complex128: 1e-10, np.complex64: 5e-5}
TOL_E = {np.complex128: 1e-10, np.complex64: 1e-5}
DTYPES = [np.complex128, np.complex64]


@pytest.fixture(scope="module")
def ctx():
    return Context(0)


def _ham(ctx):
    return KSHamiltonian(ctx, LAT, NG)


def _cube(c, gidx):
    cube = np.zeros((c.shape[0], N), dtype=np.complex128)
    cube[:, gidx] = c
    return cube.reshape((c.shape[0],) + NG)


def _to_real(c, gidx):
    """transform::to_real of sphere states: [nst][N]"""
    return np.fft.ifftn(_cube(c, gidx), axes=(1, 2, 3)).reshape(c.shape[0], N)


def _to_sphere(rs, gidx):
    return np.fft.fftn(rs.reshape((rs.shape[0],) + NG), axes=(1, 2, 3)).reshape(rs.shape[0], N)[:, gidx]


def _states(b, nst, seed):
    return S.random_states(b.npw, nst, seed, V / float(N) ** 2)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("layout", [0, 1])
def test_sphere_real_transforms(ctx, dtype, layout):
    say("case: transforms 16^3", np.dtype(dtype).name, "layout", layout)
    gh = _ham(ctx)
    b = gh.basis((0, 0, 0), 6.0)
    g = b.gindex()
    c = _states(b, 5, 3).astype(dtype)
    d = DeviceArray.from_numpy(ctx, c)
    rs = gh.to_real(b, d, layout).numpy()
    ref = _to_real(c.astype(np.complex128), g)
    got = rs if layout == 0 else rs.T
    assert rel_err(got, ref) < TOL_T[dtype]
    # mask(FFT(.)) of a random full-grid field, and accumulation onto existing values
    rng = np.random.default_rng(4)
    f = (rng.standard_normal((5, N)) + 1j * rng.standard_normal((5, N))).astype(dtype)
    df = DeviceArray.from_numpy(ctx, f if layout == 0 else np.ascontiguousarray(f.T))
    out = DeviceArray.from_numpy(ctx, c)
    gh.to_sphere(b, df, 5, layout, out=out, accumulate=True)
    assert rel_err(out.numpy(), c + _to_sphere(f.astype(np.complex128), g)) < TOL_T[dtype]
    # round trip sphere -> real -> sphere is the identity
    back = gh.to_sphere(b, DeviceArray.from_numpy(ctx, rs), 5, layout).numpy()
    assert rel_err(back, c) < TOL_T[dtype]


@pytest.mark.parametrize("dtype", DTYPES)
def test_poisson_vs_oracle(ctx, dtype):
    say("case: poisson 16^3", np.dtype(dtype).name)
    gh = _ham(ctx)
    oh = orc.Ham(LAT, NG)
    rng = np.random.default_rng(8)
    rho = (rng.standard_normal((3, N)) + 1j * rng.standard_normal((3, N))).astype(dtype)
    for q, z in (((0.0, 0.0, 0.0), 0.0), ((0.0, 0.0, 0.0), 0.7), ((0.1, -0.2, 0.05), 0.0)):
        d = DeviceArray.from_numpy(ctx, rho)
        got = gh.poisson(d, q, z).numpy()
        ref = oh.poisson_in_place(np.ascontiguousarray(rho.astype(np.complex128).T), gshift=q, zeroterm=z).T
        assert rel_err(got, ref) < TOL_T[dtype], (q, z)
    # plane wave closed form (poisson.hpp:315-351): e^{ikx} -> 4 pi/k^2 e^{ikx}
    kk = 2 * np.pi / 6.0
    x = S.rgrid(LAT, NG)[:, 0]
    pw = (np.exp(1j * kk * x)[None, :] * (1.0 + np.arange(2))[:, None]).astype(dtype)
    got = gh.poisson(DeviceArray.from_numpy(ctx, pw)).numpy()
    assert rel_err(got, 4 * np.pi / kk ** 2 * pw.astype(np.complex128)) < TOL_T[dtype]


@pytest.mark.parametrize("dtype", DTYPES)
def test_exchange_direct_vs_oracle(ctx, dtype):
    say("case: exchange direct 16^3", np.dtype(dtype).name)
    gh = _ham(ctx)
    b = gh.basis((0, 0, 0), 6.0)
    g = b.gindex()
    phi = _states(b, 5, 21).astype(dtype)
    hf = _states(b, 4, 22).astype(dtype)
    occ = np.array([2.0, 0.0, 1.5, 0.3])  # the zero occupation is skipped
    init = _states(b, 5, 23).astype(dtype)
    ex = ExchangeOperator(gh, b, exchange_coefficient=0.25)
    out = DeviceArray.from_numpy(ctx, init)
    ex.direct(DeviceArray.from_numpy(ctx, phi), DeviceArray.from_numpy(ctx, hf), occ, scale=1.0, exxphi=out)
    oh = orc.Ham(LAT, NG)
    p_rs = _to_real(phi.astype(np.complex128), g).T
    h_rs = _to_real(hf.astype(np.complex128), g).T
    k_rs = oh.exchange_direct(np.ascontiguousarray(p_rs), np.ascontiguousarray(h_rs), occ, exx_coeff=0.25,
                              scale=1.0)
    ref = init.astype(np.complex128) + _to_sphere(np.ascontiguousarray(k_rs.T), g)
    assert rel_err(out.numpy() - init, ref - init) < TOL_X[dtype]


@pytest.mark.parametrize("dtype", DTYPES)
def test_ace_build_and_apply_vs_oracle(ctx, dtype):
    say("case: ACE 16^3", np.dtype(dtype).name)
    gh = _ham(ctx)
    b = gh.basis((0, 0, 0), 6.0)
    g = b.gindex()
    nst = 6
    phi = _states(b, nst, 31).astype(dtype)
    occ = np.array([2.0, 2.0, 2.0, 1.0, 0.5, 0.25])
    d = DeviceArray.from_numpy(ctx, phi)
    ex = ExchangeOperator(gh, b, exchange_coefficient=1.0)
    e_gpu = ex.update(d, occ)
    oh = orc.Ham(LAT, NG)
    p_rs = np.ascontiguousarray(_to_real(phi.astype(np.complex128), g).T)
    xi_rs, e_ref = oh.ace_build(p_rs, occ, exx_coeff=1.0)
    assert e_gpu == pytest.approx(e_ref, rel=TOL_E[dtype])
    xi_ref = _to_sphere(np.ascontiguousarray(xi_rs.T), g)
    assert rel_err(ex.xi.numpy(), xi_ref) < TOL_X[dtype]
    # ACE reproduces the direct exchange on the states it was built from
    kace = ex(d).numpy()
    kdir = ex.direct(d, d, occ, scale=1.0).numpy()
    assert rel_err(kace, kdir) < TOL_X[dtype]
    # and on other states it equals the oracle's ACE apply, sphere-projected
    psi = _states(b, 3, 33).astype(dtype)
    kpsi = ex(DeviceArray.from_numpy(ctx, psi)).numpy()
    ps_rs = np.ascontiguousarray(_to_real(psi.astype(np.complex128), g).T)
    ref = _to_sphere(np.ascontiguousarray(oh.ace_apply(xi_rs, ps_rs).T), g)
    assert rel_err(kpsi, ref) < TOL_X[dtype]


def _cart(kcov):
    """covariant k (the units of 2 pi m) in cartesian components"""
    return np.asarray(kcov, dtype=np.float64) @ np.linalg.inv(LAT).T


@pytest.mark.parametrize("dtype", DTYPES)
def test_exchange_kpoints_vs_oracle(ctx, dtype):
    """states at k against occupied states at k and at q (two hf blocks, each
    with its own singularity value): every pair solve of block q shifted by
    -k + q (exchange_operator.hpp:151), against the oracle; the ACE built from
    the same blocks reproduces the direct exchange on its own states and its
    energy equals -1/2 sum occ Re <phi_i|W_i> from the oracle's W"""
    say("case: exchange at two k-points 16^3", np.dtype(dtype).name)
    gh = _ham(ctx)
    k, q = (0.4, -0.3, 0.2), (-0.5, 0.6, 0.1)
    bk, bq = gh.basis(k, 6.0), gh.basis(q, 6.0)
    gk, gq = bk.gindex(), bq.gindex()
    phi = _states(bk, 5, 41).astype(dtype)
    hk = _states(bk, 3, 42).astype(dtype)
    hq = _states(bq, 4, 43).astype(dtype)
    ok, oq = np.array([2.0, 1.0, 0.5]), np.array([2.0, 0.0, 1.5, 0.3])
    zk, zq = 1.7, 0.9
    alpha = 0.25
    ex = ExchangeOperator(gh, bk, exchange_coefficient=alpha)
    init = _states(bk, 5, 44).astype(dtype)
    out = DeviceArray.from_numpy(ctx, init)
    dphi = DeviceArray.from_numpy(ctx, phi)
    blocks = [(bk, DeviceArray.from_numpy(ctx, hk), ok, zk), (bq, DeviceArray.from_numpy(ctx, hq), oq, zq)]
    ex.direct_kpts(dphi, blocks, scale=1.0, exxphi=out)
    oh = orc.Ham(LAT, NG)
    rs = lambda c, g: np.ascontiguousarray(_to_real(c.astype(np.complex128), g).T)
    p_rs = rs(phi, gk)
    w_rs = (oh.exchange_direct(p_rs, rs(hk, gk), ok, exx_coeff=alpha, gshift=(0.0, 0.0, 0.0), zeroterm=zk) +
            oh.exchange_direct(p_rs, rs(hq, gq), oq, exx_coeff=alpha, gshift=_cart(q) - _cart(k), zeroterm=zq))
    ref = init.astype(np.complex128) + _to_sphere(np.ascontiguousarray(w_rs.T), gk)
    assert rel_err(out.numpy() - init, ref - init) < TOL_X[dtype]
    # ACE of the block phi (occupations of its own states for the energy)
    occ_phi = np.array([2.0, 2.0, 1.0, 1.0, 0.5])
    e = ex.update_kpts(dphi, occ_phi, blocks)
    dv = V / N
    m_ii = dv * np.sum(np.conj(-w_rs) * p_rs, axis=0)  # W = direct(phi, -1) = -K phi
    assert e == pytest.approx(-0.5 * float(np.sum(occ_phi * m_ii.real)), rel=TOL_E[dtype])
    kace = ex(dphi).numpy()
    kdir = ex.direct_kpts(dphi, blocks, scale=1.0).numpy()
    assert rel_err(kace, kdir) < TOL_X[dtype]


def test_exchange_at_one_kpoint_equals_kpts_form(ctx):
    """hf states at the block's own k-point: inqx_exchange_direct and the
    k-point form with one block give the same result bit for bit"""
    say("case: exchange one k-point, both entry points 16^3")
    gh = _ham(ctx)
    b = gh.basis((0.25, 0.5, -0.75), 6.0)
    phi = DeviceArray.from_numpy(ctx, _states(b, 4, 51).astype(np.complex64))
    hf = DeviceArray.from_numpy(ctx, _states(b, 3, 52).astype(np.complex64))
    occ = np.array([2.0, 1.0, 0.25])
    ex = ExchangeOperator(gh, b, exchange_coefficient=1.0, zeroterm=0.6)
    a = ex.direct(phi, hf, occ).numpy()
    c = ex.direct_kpts(phi, [(b, hf, occ, 0.6)]).numpy()
    assert np.array_equal(a, c)


def test_two_plane_waves_closed_form(ctx):
    say("case: exchange two plane waves 16^3")
    gh = _ham(ctx)
    b = gh.basis((0, 0, 0), 6.0)
    g = list(b.gindex())
    n = NG[0]
    ia = (1 * n + 0) * n + 0          # G_a = 2 pi/6 (1, 0, 0)
    ib = (0 * n + (n - 1)) * n + 2    # G_b = 2 pi/6 (0, -1, 2)
    c = np.zeros((2, b.npw), dtype=np.complex128)
    c[0, g.index(ia)] = N / np.sqrt(V)
    c[1, g.index(ib)] = N / np.sqrt(V)
    occ = np.array([2.0, 0.7])
    alpha = 0.25
    ex = ExchangeOperator(gh, b, exchange_coefficient=alpha)
    d = DeviceArray.from_numpy(ctx, c)
    k = ex.direct(d, d, occ, scale=1.0).numpy()
    d2 = (2 * np.pi / 6.0) ** 2 * (1 + 1 + 4)
    lam = np.array([-2 * np.pi * alpha * occ[1] / (V * d2), -2 * np.pi * alpha * occ[0] / (V * d2)])
    assert rel_err(k, lam[:, None] * c) < 1e-12
    e = ex.update(d, occ)
    assert e == pytest.approx(-2 * np.pi * alpha * occ[0] * occ[1] / (V * d2), rel=1e-12)


def test_chunked_pairs_are_bitwise_identical(ctx):
    say("case: exchange chunking 16^3")
    gh = _ham(ctx)
    b = gh.basis((0, 0, 0), 6.0)
    phi = _states(b, 5, 41)
    occ = np.array([2.0, 1.0, 0.0, 0.5, 2.0])
    ex = ExchangeOperator(gh, b, exchange_coefficient=1.0)
    d = DeviceArray.from_numpy(ctx, phi)
    one = ex.direct(d, d, occ).numpy()
    gh.set_exchange_workspace(1)  # one j per chunk
    many = ex.direct(d, d, occ).numpy()
    gh.set_exchange_workspace(0)
    assert np.array_equal(one, many)


@pytest.mark.parametrize("dtype", DTYPES)
def test_symmetric_pairs_match_full_pairs(ctx, dtype):
    """hf = phi (the ACE build's case) solves each unordered pair of occupied
    states once and takes the other's potential as its conjugate; with hf a copy
    of phi (another array) every ordered pair is solved.  Both agree with each
    other and with the oracle, with unoccupied states among them, and the
    symmetric form does not depend on the chunking (bit for bit)"""
    say("case: symmetric pairs 16^3", np.dtype(dtype).name)
    gh = _ham(ctx)
    b = gh.basis((0, 0, 0), 6.0)
    g = b.gindex()
    phi = _states(b, 7, 71).astype(dtype)
    occ = np.array([2.0, 0.0, 1.0, 2.0, 0.0, 0.5, 2.0])
    ex = ExchangeOperator(gh, b, exchange_coefficient=0.7, zeroterm=0.3)
    d = DeviceArray.from_numpy(ctx, phi)
    sym = ex.direct(d, d, occ).numpy()
    full = ex.direct(d, DeviceArray.from_numpy(ctx, phi.copy()), occ).numpy()
    assert rel_err(sym, full) < TOL_X[dtype]
    oh = orc.Ham(LAT, NG)
    p_rs = np.ascontiguousarray(_to_real(phi.astype(np.complex128), g).T)
    k_rs = oh.exchange_direct(p_rs, p_rs, occ, exx_coeff=0.7, gshift=(0.0, 0.0, 0.0), zeroterm=0.3)
    assert rel_err(sym, _to_sphere(np.ascontiguousarray(k_rs.T), g)) < TOL_X[dtype]
    gh.set_exchange_workspace(1)  # one j per chunk
    sym1 = ex.direct(d, d, occ).numpy()
    gh.set_exchange_workspace(0)
    assert np.array_equal(sym, sym1)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("n", [72, 80, 96])
def test_large_grid_paths(ctx, dtype, n):
    """the largest line lengths (their kernels run close to the register limit):
    transforms and Poisson at n^3 against numpy / the oracle, and the direct
    exchange of 2 states with 1 occupied state"""
    say("case: large grid", n, np.dtype(dtype).name)
    lat = np.eye(3) * (n * 0.3)
    ng = (n, n, n)
    N = n ** 3
    gh = KSHamiltonian(ctx, lat, ng)
    b = gh.basis((0, 0, 0), 4.0)
    g = b.gindex()
    vol = abs(np.linalg.det(lat))
    c = S.random_states(b.npw, 2, 50 + n, vol / float(N) ** 2).astype(dtype)
    rs = gh.to_real(b, DeviceArray.from_numpy(ctx, c)).numpy()
    cube = np.zeros((2, N), dtype=np.complex128)
    cube[:, g] = c
    ref = np.fft.ifftn(cube.reshape((2,) + ng), axes=(1, 2, 3)).reshape(2, N)
    assert rel_err(rs, ref) < TOL_T[dtype], "to_real"
    back = gh.to_sphere(b, DeviceArray.from_numpy(ctx, rs), 2).numpy()
    assert rel_err(back, c) < TOL_T[dtype], "to_sphere"
    rng = np.random.default_rng(n)
    rho = (rng.standard_normal((2, N)) + 1j * rng.standard_normal((2, N))).astype(dtype)
    got = gh.poisson(DeviceArray.from_numpy(ctx, rho)).numpy()
    G = 2 * np.pi * np.fft.fftfreq(n, d=lat[0, 0] / n)
    g2 = G[:, None, None] ** 2 + G[None, :, None] ** 2 + G[None, None, :] ** 2
    kern = np.where(g2 < 1e-6, 0.0, 4 * np.pi / np.where(g2 < 1e-6, 1.0, g2))
    ref = np.fft.ifftn(np.fft.fftn(rho.astype(np.complex128).reshape((2,) + ng), axes=(1, 2, 3)) * kern,
                       axes=(1, 2, 3)).reshape(2, N)
    assert rel_err(got, ref) < TOL_T[dtype], "poisson"
    ex = ExchangeOperator(gh, b, exchange_coefficient=1.0)
    d = DeviceArray.from_numpy(ctx, c)
    hf = DeviceArray.from_numpy(ctx, np.ascontiguousarray(c[:1]))
    if dtype == np.complex128 and n > 64:
        # complex128 exchange is not offered above 64^3
        with pytest.raises(InqxError, match="status -3"):
            ex.direct(d, hf, np.array([2.0]))
        return
    k = ex.direct(d, hf, np.array([2.0])).numpy()
    rs64 = ref_rs = np.fft.ifftn(cube.reshape((2,) + ng), axes=(1, 2, 3)).reshape(2, N)
    v = np.fft.ifftn(np.fft.fftn((np.conj(rs64[0])[None, :] * rs64).reshape((2,) + ng), axes=(1, 2, 3)) * kern,
                     axes=(1, 2, 3)).reshape(2, N)
    kr = -0.5 * 2.0 * rs64[0][None, :] * v
    kref = np.fft.fftn(kr.reshape((2,) + ng), axes=(1, 2, 3)).reshape(2, N)[:, g]
    assert rel_err(k, kref) < TOL_X[dtype], "exchange"


@pytest.mark.parametrize("dtype", DTYPES)
def test_many_states_large_grid_transforms(ctx, dtype):
    """regression: 60 states on a 96^3 grid (thousands of line work-groups);
    a few states are checked against numpy, all against the round trip"""
    say("case: 60 states 96^3 transforms", np.dtype(dtype).name)
    n, nst = 96, 60
    lat = np.eye(3) * 28.8
    ng = (n, n, n)
    N = n ** 3
    gh = KSHamiltonian(ctx, lat, ng)
    b = gh.basis((0, 0, 0), 3.0)
    g = b.gindex()
    c = S.random_states(b.npw, nst, 77, 1.0).astype(dtype)
    rs = gh.to_real(b, DeviceArray.from_numpy(ctx, c))
    pick = [0, 1, 31, 59]
    cube = np.zeros((len(pick), N), dtype=np.complex128)
    cube[:, g] = c[pick]
    ref = np.fft.ifftn(cube.reshape((len(pick),) + ng), axes=(1, 2, 3)).reshape(len(pick), N)
    got = rs.numpy()
    assert rel_err(got[pick], ref) < TOL_T[dtype]
    back = gh.to_sphere(b, rs, nst).numpy()
    assert rel_err(back, c) < TOL_T[dtype]


@pytest.mark.skip(reason="complex128 exchange above 64^3 returns INQX_ERR_UNSUPPORTED: repeated calls on "
                         "identical inputs differed in 1-8 of 432 states (~1e-3 relative) with no cause shown")
def test_fp64_exchange_96_is_deterministic(ctx):
    """two complex128 direct exchanges at 96^3 with 4 occupied states must agree
    bit for bit, and 8 states must agree with the oracle"""
    n, nst, nhf = 96, 16, 4
    lat = np.eye(3) * 30.78
    gh = KSHamiltonian(ctx, lat, (n, n, n))
    b = gh.basis((0, 0, 0), 6.0)
    c = S.random_states(b.npw, nst, 11, 1.0)
    ex = ExchangeOperator(gh, b, exchange_coefficient=1.0)
    d = DeviceArray.from_numpy(ctx, c)
    hf = DeviceArray.from_numpy(ctx, np.ascontiguousarray(c[:nhf]))
    k1 = ex.direct(d, hf, np.full(nhf, 2.0)).numpy()
    k2 = ex.direct(d, hf, np.full(nhf, 2.0)).numpy()
    assert np.array_equal(k1, k2)


def test_c5_si216_exchange_subset(ctx):
    """the Si216 exchange workload (96^3, 432 complex64 states, L = 30.78):
    direct exchange of all 432 states with one occupied state, the first 8
    outputs against the FP64 oracle; then ACE built from 16 of the states
    reproduces the direct exchange on them"""
    say("case: C5 Si216 96^3 432 states exchange, 8 checked")
    sysm = S.config(5)
    gh = KSHamiltonian(ctx, sysm.lattice, sysm.n)
    b = gh.basis((0, 0, 0), sysm.ecut)
    g = b.gindex()
    ng = tuple(sysm.n)
    NN = sysm.N
    vol = S.sphere_volume_element(sysm)
    rng = np.random.default_rng(sysm.seed)
    c = (rng.standard_normal((sysm.nst, b.npw)) + 1j * rng.standard_normal((sysm.nst, b.npw)))
    c /= np.sqrt(vol * b.npw * 2.0)  # unit norm on average
    c = c.astype(np.complex64)
    d = DeviceArray.from_numpy(ctx, c)
    hf = DeviceArray.from_numpy(ctx, np.ascontiguousarray(c[:1]))
    occ = np.array([2.0])
    ex = ExchangeOperator(gh, b, exchange_coefficient=1.0)
    k = ex.direct(d, hf, occ, scale=1.0).numpy()
    ns = 8
    cube = np.zeros((ns, NN), dtype=np.complex128)
    cube[:, g] = c[:ns]
    rs = np.fft.ifftn(cube.reshape((ns,) + ng), axes=(1, 2, 3)).reshape(ns, NN)
    oh = orc.Ham(sysm.lattice, sysm.n)
    kr = oh.exchange_direct(np.ascontiguousarray(rs.T), np.ascontiguousarray(rs[:1].T), occ, exx_coeff=1.0, scale=1.0)
    kref = np.fft.fftn(np.ascontiguousarray(kr.T).reshape((ns,) + ng), axes=(1, 2, 3)).reshape(ns, NN)[:, g]
    assert rel_err(k[:ns], kref) < TOL_X[np.complex64]
    del d, hf
    nb = 16
    sub = DeviceArray.from_numpy(ctx, np.ascontiguousarray(c[:nb]))
    occb = np.full(nb, 2.0)
    ex.update(sub, occb)
    kace = ex(sub).numpy()
    kdir = ex.direct(sub, sub, occb, scale=1.0).numpy()
    assert rel_err(kace, kdir) < TOL_X[np.complex64]


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("ng2,lens", [((24, 16, 16), (9.0, 6.0, 6.0)), ((16, 20, 24), (6.0, 7.5, 9.0))])
def test_noncubic_grid_poisson_exchange_hartree(ctx, dtype, ng2, lens):
    """grids that are not cubic (24 x 16 x 16: the first axis longer; 16 x 20 x
    24: planes that are not square, column and row kernels): batched Poisson
    with shift and zeroterm, exchange of states at k with occupied states at k
    and q, and the Hartree potential, against the oracle"""
    say("case: non-cubic Poisson / exchange / Hartree", ng2, np.dtype(dtype).name)
    lat2 = np.diag(lens)
    vol = float(np.prod(lens))
    n2 = int(np.prod(ng2))
    gh = KSHamiltonian(ctx, lat2, ng2)
    oh = orc.Ham(lat2, ng2)
    rng = np.random.default_rng(12)
    rho = (rng.standard_normal((2, n2)) + 1j * rng.standard_normal((2, n2))).astype(dtype)
    for q, z in (((0.0, 0.0, 0.0), 0.4), ((0.1, -0.2, 0.05), 0.0)):
        got = gh.poisson(DeviceArray.from_numpy(ctx, rho), q, z).numpy()
        ref = oh.poisson_in_place(np.ascontiguousarray(rho.astype(np.complex128).T), gshift=q, zeroterm=z).T
        assert rel_err(got, ref) < TOL_T[dtype], (q, z)
    # exchange at two k-points
    k, q = (0.3, -0.2, 0.1), (-0.4, 0.5, 0.2)
    bk, bq = gh.basis(k, 5.0), gh.basis(q, 5.0)
    vel = vol / float(n2) ** 2
    phi = S.random_states(bk.npw, 3, 61, vel).astype(dtype)
    hq = S.random_states(bq.npw, 2, 62, vel).astype(dtype)
    ex = ExchangeOperator(gh, bk, exchange_coefficient=0.5)
    got = ex.direct_kpts(DeviceArray.from_numpy(ctx, phi),
                         [(bq, DeviceArray.from_numpy(ctx, hq), np.array([2.0, 1.0]), 0.0)]).numpy()

    def rs(c, g):
        cube = np.zeros((c.shape[0], n2), dtype=np.complex128)
        cube[:, g] = c
        return np.ascontiguousarray(np.fft.ifftn(cube.reshape((c.shape[0],) + ng2), axes=(1, 2, 3))
                                    .reshape(c.shape[0], n2).T)
    gk, gq = bk.gindex(), bq.gindex()
    cart = lambda v: np.asarray(v, dtype=np.float64) @ np.linalg.inv(lat2).T
    w = oh.exchange_direct(rs(phi.astype(np.complex128), gk), rs(hq.astype(np.complex128), gq),
                           np.array([2.0, 1.0]), exx_coeff=0.5, gshift=cart(q) - cart(k), zeroterm=0.0)
    ref = np.fft.fftn(np.ascontiguousarray(w.T).reshape((3,) + ng2), axes=(1, 2, 3)).reshape(3, n2)[:, gk]
    assert rel_err(got, ref) < TOL_X[dtype]
    # hf = phi: each unordered pair of occupied states solved once
    dp = DeviceArray.from_numpy(ctx, phi)
    occ3 = np.array([2.0, 1.0, 0.0])
    ek = ExchangeOperator(gh, bk, exchange_coefficient=0.5, zeroterm=0.2)
    full = ek.direct(dp, DeviceArray.from_numpy(ctx, phi.copy()), occ3).numpy()
    assert rel_err(ek.direct(dp, dp, occ3).numpy(), full) < TOL_X[dtype]
    # Hartree potential of a real density (complex128 path)
    dens = 0.01 + rng.random(n2) * 0.02
    vh, eh = gh.hartree(DeviceArray.from_numpy(ctx, dens))
    vref = oh.poisson_in_place(dens.astype(np.complex128)[:, None], zeroterm=0.0)[:, 0].real
    assert rel_err(vh.numpy(), vref) < 1e-11
    assert eh == pytest.approx(0.5 * float(np.sum(dens * vref)) * vol / n2, rel=1e-11)


@pytest.mark.parametrize("ng2,lens", [((128, 128, 128), (12.0, 12.0, 12.0)), ((100, 120, 108), (10.0, 12.0, 10.8))])
def test_grids_above_96_points_poisson_hartree(ctx, ng2, lens):
    """batched Poisson (both precisions) and the Hartree potential on grids
    with lines above 96 points: complex128 pair planes that do not fit the
    160 KB of LDS run the column and row kernels; complex64 exchange at k"""
    say("case: Poisson / Hartree / exchange above 96 points", ng2)
    lat2 = np.diag(lens)
    vol = float(np.prod(lens))
    n2 = int(np.prod(ng2))
    gh = KSHamiltonian(ctx, lat2, ng2)
    oh = orc.Ham(lat2, ng2)
    rng = np.random.default_rng(14)
    rho = rng.standard_normal((2, n2)) + 1j * rng.standard_normal((2, n2))
    q = (0.1, -0.2, 0.05)
    ref = oh.poisson_in_place(np.ascontiguousarray(rho.T), gshift=q, zeroterm=0.3).T
    for dtype in DTYPES:
        got = gh.poisson(DeviceArray.from_numpy(ctx, rho.astype(dtype)), q, 0.3).numpy()
        assert rel_err(got, ref) < TOL_T[dtype], dtype
        # the solve is in place: a fresh copy of rho; the line pass (four lanes
        # per line for complex128 here) repeats bit for bit
        again = gh.poisson(DeviceArray.from_numpy(ctx, rho.astype(dtype)), q, 0.3).numpy()
        assert np.array_equal(again, got), dtype
    dens = 0.01 + rng.random(n2) * 0.02
    vh, eh = gh.hartree(DeviceArray.from_numpy(ctx, dens))
    vref = oh.poisson_in_place(dens.astype(np.complex128)[:, None], zeroterm=0.0)[:, 0].real
    assert rel_err(vh.numpy(), vref) < 1e-11
    assert eh == pytest.approx(0.5 * float(np.sum(dens * vref)) * vol / n2, rel=1e-11)
    # complex64 exchange of states at k with occupied states at q
    k, kq = (0.3, -0.2, 0.1), (-0.4, 0.5, 0.2)
    bk, bq = gh.basis(k, 8.0), gh.basis(kq, 8.0)
    vel = vol / float(n2) ** 2
    phi = S.random_states(bk.npw, 2, 63, vel).astype(np.complex64)
    hq = S.random_states(bq.npw, 1, 64, vel).astype(np.complex64)
    ex = ExchangeOperator(gh, bk, exchange_coefficient=0.5)
    got = ex.direct_kpts(DeviceArray.from_numpy(ctx, phi),
                         [(bq, DeviceArray.from_numpy(ctx, hq), np.array([2.0]), 0.0)]).numpy()

    def rs(c, g):
        cube = np.zeros((c.shape[0], n2), dtype=np.complex128)
        cube[:, g] = c
        return np.ascontiguousarray(np.fft.ifftn(cube.reshape((c.shape[0],) + ng2), axes=(1, 2, 3))
                                    .reshape(c.shape[0], n2).T)
    gk, gq = bk.gindex(), bq.gindex()
    cart = lambda v: np.asarray(v, dtype=np.float64) @ np.linalg.inv(lat2).T
    w = oh.exchange_direct(rs(phi.astype(np.complex128), gk), rs(hq.astype(np.complex128), gq),
                           np.array([2.0]), exx_coeff=0.5, gshift=cart(kq) - cart(k), zeroterm=0.0)
    ref = np.fft.fftn(np.ascontiguousarray(w.T).reshape((2,) + ng2), axes=(1, 2, 3)).reshape(2, n2)[:, gk]
    assert rel_err(got, ref) < TOL_X[np.complex64]

