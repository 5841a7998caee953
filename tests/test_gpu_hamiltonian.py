"""GPU parity of H|psi> on G-sphere blocks against the FP64 oracle.

Tolerances (written here, per precision):
  FP64 path: max |H psi - oracle| / max |oracle| < 1e-11 (the oracle itself
             agrees with the reference's own KATs to better than 1e-9);
  FP32 path: same ratio < 2e-5 (complex64 FFTs and pointwise work).
Full-size configurations are also checked through size-independent
properties: hermiticity <a|H b> = <H a|b> and the Rayleigh quotients.
"""
import numpy as np
import pytest

from inq_amd import Context, DeviceArray
from inq_amd import systems as S
from tests.helpers import gpu_ham, oracle_ham, rel_err, say

pytestmark = pytest.mark.gpu

TOL = {np.complex128: 1e-11, np.complex64: 2e-5}


@pytest.fixture(scope="module")
def ctx():
    return Context(0)


def _run(ctx, gh, sysm, k, ecut, nst, dtype, seed, spin=0):
    b = gh.basis(k, ecut, spin)
    vol = S.sphere_volume_element(sysm)
    c = S.random_states(b.npw, nst, seed, vol)
    d_in = DeviceArray.from_numpy(ctx, c.astype(dtype))
    d_out = DeviceArray(ctx, c.shape, dtype)
    gh.apply(b, d_in, d_out)
    return b, c, d_out.numpy()


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("k", [(0.0, 0.0, 0.0), (0.3, -0.2, 0.45)])
def test_small_random_system(ctx, dtype, k):
    sysm = S.small_config(seed=7, n=(16, 16, 16), natoms=3, nst=5, ecut=6.0)
    say("case: small16", np.dtype(dtype).name, "k =", k)
    gh = gpu_ham(ctx, sysm)
    oh = oracle_ham(sysm)
    b, c, out = _run(ctx, gh, sysm, k, sysm.ecut, sysm.nst, dtype, 11)
    ref = oh.apply_sphere(c.astype(dtype).astype(np.complex128), b.gindex(), k)
    assert rel_err(out, ref) < TOL[dtype]


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("n,lat", [((32, 24, 24), (8.0, 6.0, 6.0)), ((20, 30, 30), (5.0, 7.5, 7.5))])
def test_grid_n0_differs_from_n1(ctx, dtype, n, lat):
    """a grid whose first axis differs from the other two (orthorhombic cell,
    k != 0, three atoms): x-pass line length and plane size are independent"""
    rng = np.random.default_rng(17)
    L = np.diag(lat)
    sysm = S.System(f"ortho{n[0]}x{n[1]}", 190, L, n, rng.uniform(-0.5, 0.5, (3, 3)) @ L, 8.0, 5,
                    v_sigma=0.6, kb_width=0.7, kb_radius=1.6)
    k = (0.35, -0.25, 0.15)
    say("case: grid", n, np.dtype(dtype).name)
    gh = gpu_ham(ctx, sysm)
    oh = oracle_ham(sysm)
    b, c, out = _run(ctx, gh, sysm, k, sysm.ecut, sysm.nst, dtype, 19)
    ref = oh.apply_sphere(c.astype(dtype).astype(np.complex128), b.gindex(), k)
    assert rel_err(out, ref) < TOL[dtype]


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("n,lat", [((24, 20, 30), (6.0, 5.0, 7.5)), ((20, 32, 16), (5.0, 8.0, 4.0))])
def test_plane_n1_differs_from_n2(ctx, dtype, n, lat):
    """planes that are not square (n1 != n2 runs the row and column kernels)
    at k != 0 with three atoms: H on the sphere and H on a whole-cube basis
    (every plane row holds sphere points) against the oracle, and the sphere
    <-> real-space transforms against numpy's FFT"""
    rng = np.random.default_rng(23)
    L = np.diag(lat)
    sysm = S.System(f"rect{n[1]}x{n[2]}", 191, L, n, rng.uniform(-0.5, 0.5, (3, 3)) @ L, 8.0, 5,
                    v_sigma=0.6, kb_width=0.7, kb_radius=1.6)
    k = (-0.3, 0.2, 0.35)
    say("case: plane", n, np.dtype(dtype).name)
    gh = gpu_ham(ctx, sysm)
    oh = oracle_ham(sysm)
    b, c, out = _run(ctx, gh, sysm, k, sysm.ecut, sysm.nst, dtype, 29)
    ref = oh.apply_sphere(c.astype(dtype).astype(np.complex128), b.gindex(), k)
    assert rel_err(out, ref) < TOL[dtype]
    # the launch-shape autotuner on these planes (row and column kernels at
    # every plane batch it tries) changes no result bit
    shape, ms = gh.autotune_batch([b], [sysm.nst], dtype, None)
    assert shape[1] >= 1 and ms > 0
    d_in = DeviceArray.from_numpy(ctx, c.astype(dtype))
    assert np.array_equal(gh(b, d_in).numpy(), out)
    bw, cw, outw = _run(ctx, gh, sysm, k, 1e6, 3, dtype, 31)
    assert bw.npw == sysm.N
    refw = oh.apply_sphere(cw.astype(dtype).astype(np.complex128), bw.gindex(), k)
    assert rel_err(outw, refw) < TOL[dtype]
    # sphere -> real space (x 1/N) and back (unnormalised), both layouts
    g = b.gindex()
    cube = np.zeros((sysm.nst, sysm.N), dtype=np.complex128)
    cube[:, g] = c.astype(dtype)
    want = np.fft.ifftn(cube.reshape((sysm.nst,) + tuple(n)), axes=(1, 2, 3)).reshape(sysm.nst, sysm.N)
    d = DeviceArray.from_numpy(ctx, c.astype(dtype))
    for layout in (0, 1):
        rs = gh.to_real(b, d, layout).numpy()
        assert rel_err(rs if layout == 0 else rs.T, want) < TOL[dtype]
        back = gh.to_sphere(b, DeviceArray.from_numpy(ctx, rs), sysm.nst, layout).numpy()
        assert rel_err(back, c.astype(dtype)) < TOL[dtype]


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("n,lat", [((128, 128, 128), (12.0, 12.0, 12.0)), ((100, 120, 108), (10.0, 12.0, 10.8))])
def test_grids_above_96_points(ctx, dtype, n, lat):
    """line lengths above 96 (lines over 96 points use 128-bit point masks; the
    complex64 x passes take 4 states per work-group there): H on the sphere at
    k != 0 with three atoms against the oracle, and the sphere <-> real-space
    transforms against numpy's FFT"""
    rng = np.random.default_rng(29)
    L = np.diag(lat)
    sysm = S.System(f"big{n[0]}", 192, L, n, rng.uniform(-0.5, 0.5, (3, 3)) @ L, 20.0, 3,
                    v_sigma=0.6, kb_width=0.7, kb_radius=1.6)
    k = (0.25, -0.15, 0.3)
    say("case: grid", n, np.dtype(dtype).name)
    gh = gpu_ham(ctx, sysm)
    oh = oracle_ham(sysm)
    b, c, out = _run(ctx, gh, sysm, k, sysm.ecut, sysm.nst, dtype, 37)
    ref = oh.apply_sphere(c.astype(dtype).astype(np.complex128), b.gindex(), k)
    assert rel_err(out, ref) < TOL[dtype]
    g = b.gindex()
    cube = np.zeros((sysm.nst, sysm.N), dtype=np.complex128)
    cube[:, g] = c.astype(dtype)
    want = np.fft.ifftn(cube.reshape((sysm.nst,) + tuple(n)), axes=(1, 2, 3)).reshape(sysm.nst, sysm.N)
    rs = gh.to_real(b, DeviceArray.from_numpy(ctx, c.astype(dtype)), 0).numpy()
    assert rel_err(rs, want) < TOL[dtype]


def test_complex128_plane_window_beyond_lds_is_a_status(ctx):
    """a complex128 sphere whose plane row window does not fit the 160 KB of
    LDS (a whole-cube basis at 128^3) is refused with INQX_ERR_UNSUPPORTED; the
    complex64 apply of the same basis runs"""
    from inq_amd import InqxError
    sysm = S.small_config(seed=4, n=(128, 128, 128), natoms=0, nst=1, ecut=6.0)
    say("case: complex128 plane window beyond LDS at 128^3")
    from inq_amd import KSHamiltonian
    gh = KSHamiltonian(ctx, sysm.lattice, sysm.n)
    gh.set_scalar_potential(np.zeros(sysm.N))
    b = gh.basis((0, 0, 0), 1e6)
    c = np.zeros((1, b.npw), dtype=np.complex128)
    c[0, 5] = 1.0
    with pytest.raises(InqxError) as e:
        gh(b, DeviceArray.from_numpy(ctx, c))
    assert "status -3" in str(e.value)
    out = gh(b, DeviceArray.from_numpy(ctx, c.astype(np.complex64))).numpy()
    assert out.shape == c.shape


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_c1_diamond(ctx, dtype):
    sysm = S.config(1)
    say("case: C1 diamond 32^3", np.dtype(dtype).name)
    gh = gpu_ham(ctx, sysm)
    oh = oracle_ham(sysm)
    b, c, out = _run(ctx, gh, sysm, (0, 0, 0), sysm.ecut, sysm.nst, dtype, sysm.seed)
    ref = oh.apply_sphere(c.astype(dtype).astype(np.complex128), b.gindex(), (0, 0, 0))
    assert rel_err(out, ref) < TOL[dtype]


def test_no_projectors_and_plane_waves(ctx):
    """H on single plane waves with constant V: (|G+k|^2/2 + V0) exactly."""
    sysm = S.small_config(seed=3, n=(20, 20, 20), natoms=0, nst=4, ecut=8.0)
    say("case: plane waves 20^3")
    from inq_amd import KSHamiltonian
    gh = KSHamiltonian(ctx, sysm.lattice, sysm.n)
    gh.set_scalar_potential(np.full(sysm.N, -0.3))
    k = (0.1, 0.2, -0.3)
    b = gh.basis(k, sysm.ecut)
    kin = b.kinetic()
    sel = [0, 3, b.npw // 2, b.npw - 1]
    c = np.zeros((4, b.npw), dtype=np.complex128)
    for s, i in enumerate(sel):
        c[s, i] = 1.0
    d_in = DeviceArray.from_numpy(ctx, c)
    out = gh(b, d_in).numpy()
    exp = np.zeros_like(c)
    for s, i in enumerate(sel):
        exp[s, i] = kin[i] - 0.3
    assert np.max(np.abs(out - exp)) < 1e-12


def test_c3_kpoint_batch(ctx):
    """several uneven (k, spin) blocks of both spins in one batched launch: a
    two-component potential [N][2], each block's spin index selecting its
    component, against the oracle with that component"""
    sysm = S.config(3)
    say("case: C3 Al4 batch of 8 blocks, both spins, fp64 and fp32")
    v = S.local_potential(sysm)
    v2 = np.stack([v, v + 0.3 * np.random.default_rng(5).standard_normal(v.shape)], axis=1)
    gh = gpu_ham(ctx, sysm)
    gh.set_scalar_potential(v2, ncomp=2)
    oh = [oracle_ham(sysm, np.ascontiguousarray(v2[:, s])) for s in (0, 1)]
    blocks = sysm.blocks()[:: 15][:8]
    assert {spin for _, spin, _ in blocks} == {0, 1}
    vol = S.sphere_volume_element(sysm)
    for dtype in (np.complex128, np.complex64):
        bases, ins, outs, cs = [], [], [], []
        for i, (k, spin, ec) in enumerate(blocks):
            b = gh.basis(k, ec, spin)
            c = S.random_states(b.npw, sysm.nst, 100 + i, vol)
            bases.append(b)
            cs.append(c)
            ins.append(DeviceArray.from_numpy(ctx, c.astype(dtype)))
            outs.append(DeviceArray(ctx, c.shape, dtype))
        gh.apply_batch(bases, ins, outs)
        for (k, spin, ec), b, c, o in zip(blocks, bases, cs, outs):
            ref = oh[spin].apply_sphere(c.astype(dtype).astype(np.complex128), b.gindex(), k)
            assert rel_err(o.numpy(), ref) < TOL[dtype], (k, spin, dtype)
            if spin == 1:  # the other component gives a different result
                assert rel_err(o.numpy(), oh[0].apply_sphere(c.astype(dtype).astype(np.complex128),
                                                             b.gindex(), k)) > 1e-3


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_c2_si64_subset_and_hermiticity(ctx, dtype):
    """Si64 72^3: 8 states against the oracle; all 128 states through
    hermiticity of the subspace matrix."""
    sysm = S.config(2)
    say("case: C2 Si64 72^3", np.dtype(dtype).name)
    gh = gpu_ham(ctx, sysm)
    b = gh.basis((0, 0, 0), sysm.ecut)
    vol = S.sphere_volume_element(sysm)
    c = S.random_states(b.npw, sysm.nst, sysm.seed, vol)
    d_in = DeviceArray.from_numpy(ctx, c.astype(dtype))
    out = gh(b, d_in).numpy().astype(np.complex128)
    ci = c.astype(dtype).astype(np.complex128)
    hsub = vol * (ci.conj() @ out.T)
    herm = np.max(np.abs(hsub - hsub.conj().T)) / np.max(np.abs(hsub))
    assert herm < (1e-12 if dtype == np.complex128 else 1e-5)
    oh = oracle_ham(sysm)
    ref = oh.apply_sphere(ci[:8], b.gindex(), (0, 0, 0))
    assert rel_err(out[:8], ref) < TOL[dtype]


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_reference_layouts(ctx, dtype):
    """the reference's Fourier-space layouts: point-major sphere [npw][nst] and
    the full cube [Gx][Gy][Gz][ist].  Conversions move values bit for bit (the
    cube gets zeros off the sphere); H in the point-major layouts equals the
    band-major H bit for bit; on a basis holding the whole cube, H on the
    cube layout is the reference's full-cube operator (oracle a2)"""
    sysm = S.small_config(seed=9, n=(16, 16, 16), natoms=2, nst=5, ecut=6.0)
    say("case: layouts 16^3", np.dtype(dtype).name)
    gh = gpu_ham(ctx, sysm)
    k = (0.2, -0.4, 0.1)
    b = gh.basis(k, sysm.ecut)
    g = b.gindex()
    c = S.random_states(b.npw, sysm.nst, 12, S.sphere_volume_element(sysm)).astype(dtype)
    N = sysm.N
    cube = np.zeros((N, sysm.nst), dtype=dtype)
    cube[g, :] = c.T
    rng = np.random.default_rng(3)
    off = np.ones(N, bool)
    off[g] = False
    noisy = cube.copy()
    noisy[off, :] = (rng.standard_normal((int(off.sum()), sysm.nst)) * 7).astype(dtype)  # dropped on input
    sph = gh.states_to_sphere(b, DeviceArray.from_numpy(ctx, noisy), gh.CUBE_POINT).numpy()
    assert np.array_equal(sph, c)
    pm = np.ascontiguousarray(c.T)
    assert np.array_equal(gh.states_to_sphere(b, DeviceArray.from_numpy(ctx, pm), gh.SPHERE_POINT).numpy(), c)
    d = DeviceArray.from_numpy(ctx, c)
    assert np.array_equal(gh.states_from_sphere(b, d, gh.CUBE_POINT).numpy(), cube)
    assert np.array_equal(gh.states_from_sphere(b, d, gh.SPHERE_POINT).numpy(), pm)
    # H in the point-major layouts equals the band-major H
    hb = gh(b, d).numpy()
    out_pm = DeviceArray(ctx, pm.shape, dtype)
    gh.apply_layout(b, DeviceArray.from_numpy(ctx, pm), out_pm, gh.SPHERE_POINT)
    assert np.array_equal(out_pm.numpy(), np.ascontiguousarray(hb.T))
    out_cube = DeviceArray(ctx, cube.shape, dtype)
    gh.apply_layout(b, DeviceArray.from_numpy(ctx, noisy), out_cube, gh.CUBE_POINT)
    want = np.zeros_like(cube)
    want[g, :] = hb.T
    assert np.array_equal(out_cube.numpy(), want)
    # whole-cube basis: the reference's Fourier-space operator on [Gx][Gy][Gz][ist]
    bw = gh.basis(k, 1e6)
    assert bw.npw == N
    full = (rng.standard_normal((N, 3)) + 1j * rng.standard_normal((N, 3))).astype(dtype)
    out_full = DeviceArray(ctx, full.shape, dtype)
    gh.apply_layout(bw, DeviceArray.from_numpy(ctx, full), out_full, gh.CUBE_POINT)
    oh = oracle_ham(sysm)
    ref = oh.apply_sphere(np.ascontiguousarray(full.T).astype(np.complex128), bw.gindex(), k)
    assert rel_err(out_full.numpy(), ref.T) < TOL[dtype]


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_d_channel_projectors(ctx, dtype):
    """projectors with 9 (s + p + d), 4 and 1 channels and full coefficient
    matrices coupling every channel, at k != 0: H|psi> and the non-local
    energy against the oracle (a 9-channel projector runs as three KB groups
    coupled in kb_reduce)"""
    from inq_amd import KSHamiltonian
    from oracle import oracle as orc
    sysm = S.small_config(seed=11, n=(20, 20, 20), natoms=3, nst=5, ecut=7.0)
    say("case: d-channel projectors 20^3", np.dtype(dtype).name)
    rng = np.random.default_rng(23)
    v = S.local_potential(sysm)
    gh = KSHamiltonian(ctx, sysm.lattice, sysm.n)
    gh.set_scalar_potential(v)
    oh = orc.Ham(sysm.lattice, sysm.n)
    oh.set_potential(v)
    for p, nlm in zip(sysm.positions, [9, 4, 1]):
        A = rng.standard_normal((nlm, nlm))
        D = 0.5 * (A + A.T)
        gh.add_projector(p, sysm.kb_radius, nlm, D, width=sysm.kb_width)
        oh.add_projector_gauss(p, sysm.kb_radius, nlm, sysm.kb_width, D)
    k = (0.3, -0.15, 0.2)
    b, c, out = _run(ctx, gh, sysm, k, sysm.ecut, sysm.nst, dtype, 29)
    ci = c.astype(dtype).astype(np.complex128)
    ref = oh.apply_sphere(ci, b.gindex(), k)
    assert rel_err(out, ref) < TOL[dtype]
    occ = np.linspace(2.0, 0.5, sysm.nst)
    want = oh.energy_sphere(ci, b.gindex(), occ, k=b.kpoint)
    got = gh.energy([b], [DeviceArray.from_numpy(ctx, c.astype(dtype))], [occ])
    etol = 1e-10 if dtype == np.complex128 else 1e-5
    assert abs(got["non_local"] - want[2]) < etol * max(1.0, abs(want[2])), (got, want)
