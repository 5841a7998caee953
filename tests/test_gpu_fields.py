"""GPU parity of the real-space and per-state entry points:
  * H on real-space states with a whole-cube basis against the oracle's
    real-space operator (a1) and full-cube Fourier operator (a2);
  * density::calculate_add against numpy on the oracle-free transform;
  * overlap_diagonal (plain and normalised) against numpy;
  * randomize: bit-exact against the oracle's pcg32 restatement, which
    reproduces the reference's own known answers (randomize.hpp:96-144);
  * the launch-shape autotuner changes no result.
Tolerances: FP64 1e-11, FP32 2e-5 (max abs error / max abs reference);
overlap_diagonal FP64 1e-13, FP32 1e-6 relative.
"""
import numpy as np
import pytest

from inq_amd import Context, DeviceArray
from inq_amd import operations as ops
from inq_amd import systems as S
from oracle import oracle as orc
from tests.helpers import gpu_ham, oracle_ham, rel_err, say

pytestmark = pytest.mark.gpu

TOL = {np.complex128: 1e-11, np.complex64: 2e-5}
DTYPES = [np.complex128, np.complex64]


@pytest.fixture(scope="module")
def ctx():
    return Context(0)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("k", [(0.0, 0.0, 0.0), (0.25, -0.1, 0.3)])
def test_real_space_apply_whole_cube(ctx, dtype, k):
    sysm = S.small_config(seed=3, n=(16, 16, 16), natoms=2, nst=3, ecut=6.0)
    say("case: H on real-space states 16^3", np.dtype(dtype).name, "k", k)
    gh = gpu_ham(ctx, sysm)
    oh = oracle_ham(sysm)
    b = gh.basis(k, 1e30)
    N = sysm.N
    assert b.npw == N
    rng = np.random.default_rng(5)
    psi = (rng.standard_normal((N, 3)) + 1j * rng.standard_normal((N, 3))).astype(dtype)  # not band limited
    for layout in (1, 0):
        arr = psi if layout == 1 else np.ascontiguousarray(psi.T)
        out = gh.apply_real(b, DeviceArray.from_numpy(ctx, arr), 3, layout).numpy()
        got = out if layout == 1 else out.T
        ref = oh.apply_rs(psi.astype(np.complex128), k)
        assert rel_err(got, ref) < TOL[dtype], layout
    # the whole-cube sphere form is the reference's Fourier-space operator a2
    cube = np.fft.fftn(psi.astype(np.complex128).reshape(16, 16, 16, 3), axes=(0, 1, 2)).reshape(N, 3)
    g = b.gindex()
    sph = np.ascontiguousarray(cube[g, :].T).astype(dtype)
    hs = DeviceArray(ctx, sph.shape, dtype)
    gh.apply(b, DeviceArray.from_numpy(ctx, sph), hs)
    ref_fs = oh.apply_fs(cube.astype(dtype).astype(np.complex128), k)
    assert rel_err(hs.numpy().T, ref_fs[g, :]) < TOL[dtype]


@pytest.mark.parametrize("dtype", DTYPES)
def test_density_add(ctx, dtype):
    sysm = S.small_config(seed=4, n=(16, 16, 16), natoms=2, nst=4, ecut=6.0)
    say("case: density 16^3", np.dtype(dtype).name)
    gh = gpu_ham(ctx, sysm)
    b = gh.basis((0, 0, 0), sysm.ecut)
    c = S.random_states(b.npw, 4, 9, S.sphere_volume_element(sysm)).astype(dtype)
    occ = np.array([2.0, 1.5, 0.5, 0.0])
    den = DeviceArray.from_numpy(ctx, np.full((sysm.N, 2), 0.25))
    gh.density_add(b, DeviceArray.from_numpy(ctx, c), occ, den, ncomp=2, ispin=1)
    cube = np.zeros((4, sysm.N), dtype=np.complex128)
    cube[:, b.gindex()] = c
    rs = np.fft.ifftn(cube.reshape(4, 16, 16, 16), axes=(1, 2, 3)).reshape(4, sysm.N)
    ref = np.full((sysm.N, 2), 0.25)
    ref[:, 1] += np.sum(occ[:, None] * np.abs(rs) ** 2, 0)
    got = den.numpy()
    assert np.array_equal(got[:, 0], ref[:, 0])
    assert rel_err(got[:, 1] - 0.25, ref[:, 1] - 0.25) < 10 * TOL[dtype]
    # the density integrates to sum occ (states orthonormal on the sphere)
    vol = abs(np.linalg.det(sysm.lattice))
    assert np.sum(got[:, 1] - 0.25) * vol / sysm.N == pytest.approx(4.0, rel=10 * TOL[dtype])


@pytest.mark.parametrize("dtype", DTYPES)
def test_overlap_diagonal(ctx, dtype):
    say("case: overlap_diagonal", np.dtype(dtype).name)
    rng = np.random.default_rng(2)
    a = (rng.standard_normal((7, 3001)) + 1j * rng.standard_normal((7, 3001))).astype(dtype)
    b = (rng.standard_normal((7, 3001)) + 1j * rng.standard_normal((7, 3001))).astype(dtype)
    da, db = DeviceArray.from_numpy(ctx, a), DeviceArray.from_numpy(ctx, b)
    a64, b64 = a.astype(np.complex128), b.astype(np.complex128)
    tol = 1e-13 if dtype == np.complex128 else 1e-6
    got = ops.overlap_diagonal(ctx, da, db, 0.37).numpy()
    ref = 0.37 * np.sum(a64.conj() * b64, 1)
    assert rel_err(got, ref) < tol
    got = ops.overlap_diagonal(ctx, da, db, 0.37, normalized=True).numpy()
    ref = np.sum(a64.conj() * b64, 1) / np.sum(np.abs(b64) ** 2, 1)
    assert rel_err(got, ref) < tol


def test_randomize_matches_reference_kats(ctx):
    say("case: randomize pcg32")
    # reference grid: cubic 10 bohr at spacing 0.49672941 -> 20^3 points, 12 states
    npts, nst = 8000, 12
    ref_r = orc.randomize(npts, nst, 0, complex_values=False)
    ref_c = orc.randomize(npts, nst, 0, complex_values=True)
    got_r = ops.randomize(ctx, npts, nst, 0, np.float64).numpy()
    got_c = ops.randomize(ctx, npts, nst, 0, np.complex128).numpy()
    assert np.array_equal(got_r, ref_r)
    assert np.array_equal(got_c, ref_c)
    got_s = ops.randomize(ctx, npts, nst, 0, np.complex64, layout=0).numpy()
    assert np.array_equal(got_s, ref_c.T.astype(np.complex64))
    # the reference's known answers: norms vol_el sum |u|^2 (randomize.hpp:96-144)
    kat_r = [330.1381395023, 330.5444105287, 331.5435469092, 340.0201855451, 332.959828166, 333.1967703871,
             328.9429576951, 336.8659386972, 331.6228884743, 334.5374311686, 334.6847273315, 334.1846706334]
    kat_c = [669.1459385544, 664.8544595319, 669.6590362331, 668.4278672543, 668.0965017232, 665.1942561788,
             665.6688584485, 672.3987040517, 671.7988968555, 672.6486737858, 664.3526206348, 674.5678452953]
    d = DeviceArray.from_numpy(ctx, np.ascontiguousarray(got_c.T))
    norms = ops.overlap_diagonal(ctx, d, d, 1000.0 / npts).numpy().real
    assert np.allclose(norms, kat_c, rtol=1e-10)
    assert np.allclose(0.125 * np.sum(got_r ** 2, 0), kat_r, rtol=1e-10)
    # a different index changes the stream
    assert not np.array_equal(ops.randomize(ctx, npts, nst, 1, np.float64).numpy(), got_r)


def test_autotune_keeps_results(ctx):
    sysm = S.config(1)
    say("case: autotune C1")
    gh = gpu_ham(ctx, sysm)
    b = gh.basis((0, 0, 0), sysm.ecut)
    c = S.random_states(b.npw, sysm.nst, sysm.seed, S.sphere_volume_element(sysm)).astype(np.complex64)
    d = DeviceArray.from_numpy(ctx, c)
    outs = []
    for w in (8, 4, 2, 1):
        gh.set_launch_shape(np.complex64, w)
        o = DeviceArray(ctx, c.shape, np.complex64)
        gh.apply(b, d, o)
        outs.append(o.numpy())
    assert all(np.array_equal(outs[0], o) for o in outs[1:])
    w, ms = gh.autotune(b, sysm.nst, np.complex64)
    assert w in (1, 2, 4, 8) and ms > 0
    o = DeviceArray(ctx, c.shape, np.complex64)
    gh.apply(b, d, o)
    assert np.array_equal(o.numpy(), outs[0])


def test_autotune_batch_cache_and_results(ctx, tmp_path):
    """the batch autotuner searches x-pass and plane-pass shapes on uneven Al4
    k-point blocks, stores the winner in its cache file, applies it from the
    cache on a second call (same shape, no search), and no shape changes a
    result bit"""
    sysm = S.config(3)
    say("case: autotune batch C3 (8 blocks)")
    gh = gpu_ham(ctx, sysm)
    blocks = list(sysm.blocks())[:8]
    bases = [gh.basis(k, ec, 0) for (k, spin, ec) in blocks]
    vol = S.sphere_volume_element(sysm)
    ins = [DeviceArray.from_numpy(ctx, S.random_states(b.npw, sysm.nst, 40 + i, vol).astype(np.complex64))
           for i, b in enumerate(bases)]

    def run():
        outs = [DeviceArray(ctx, d.shape, np.complex64) for d in ins]
        gh.apply_batch(bases, ins, outs)
        return [o.numpy() for o in outs]

    ref = run()
    cache = tmp_path / "tune.txt"
    shape, ms = gh.autotune_batch(bases, [sysm.nst] * len(bases), np.complex64, cache)
    assert shape[0] in (1, 2, 4, 8) and shape[1] >= 1 and ms > 0
    lines = cache.read_text().split()
    assert len(lines) == 4 and int(lines[1]) == shape[0] and int(lines[2]) == shape[1]
    shape2, ms2 = gh.autotune_batch(bases, [sysm.nst] * len(bases), np.complex64, cache)
    assert shape2 == shape and ms2 == pytest.approx(ms, rel=1e-5)
    assert len(cache.read_text().split()) == 4  # read from the cache, nothing appended
    got = run()
    assert all(np.array_equal(a, b) for a, b in zip(got, ref))
    # every plane batch gives the same bits
    for pb in (1, 2, 3, 0):
        gh.set_plane_batch(np.complex64, pb)
        assert all(np.array_equal(a, b) for a, b in zip(run(), ref)), pb
