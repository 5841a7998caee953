"""Pins the FP64 CPU oracle against the reference's own known-answer tests,
closed forms and textbook library routines (numpy.fft, scipy.linalg).

Every number quoted below comes from the reference's unit tests (file:line
given next to it) or from mathematics; the oracle is never checked against a
retyped copy of its own formula.
"""
import numpy as np
import pytest
import scipy.linalg as sla

from oracle import oracle as orc


def spacing_grid(lattice, spacing):
    """real_space(cell, spacing) sizes: good_size(round(|a_d|/h))
    (src/basis/real_space.hpp:244-258)."""
    lattice = np.asarray(lattice, dtype=float).reshape(3, 3)
    n = []
    for d in range(3):
        m = int(round(np.linalg.norm(lattice[d]) / spacing))
        n.append(orc.lib().orc_good_size(m))
    return tuple(n)


def sym(i, n):
    return np.where(i >= (n + 1) // 2, i - n, i)


def rgrid(lattice, n):
    """cartesian r of every grid point in the symmetric range, [N,3]."""
    lattice = np.asarray(lattice, dtype=float).reshape(3, 3)
    i0, i1, i2 = np.meshgrid(np.arange(n[0]), np.arange(n[1]), np.arange(n[2]), indexing="ij")
    f = np.stack([sym(i0, n[0]) / n[0], sym(i1, n[1]) / n[1], sym(i2, n[2]) / n[2]], -1).reshape(-1, 3)
    return f @ lattice


def gcart(lattice, n):
    lattice = np.asarray(lattice, dtype=float).reshape(3, 3)
    b = 2 * np.pi * np.linalg.inv(lattice).T
    i0, i1, i2 = np.meshgrid(np.arange(n[0]), np.arange(n[1]), np.arange(n[2]), indexing="ij")
    m = np.stack([sym(i0, n[0]), sym(i1, n[1]), sym(i2, n[2])], -1).reshape(-1, 3)
    return m @ b


# --------------------------------------------------------------------------- FFT

@pytest.mark.parametrize("n,nst", [((8, 6, 10), 3), ((24, 24, 24), 2), ((20, 15, 7), 1), ((72, 8, 9), 2)])
def test_fft_matches_numpy(n, nst):
    rng = np.random.default_rng(1)
    a = rng.standard_normal(n + (nst,)) + 1j * rng.standard_normal(n + (nst,))
    f = orc.fft3d(n, a.copy(), -1).reshape(n + (nst,))
    ref = np.fft.fftn(a, axes=(0, 1, 2))
    assert np.max(np.abs(f - ref)) < 1e-12 * np.max(np.abs(ref))
    b = orc.fft3d(n, a.copy(), +1).reshape(n + (nst,))
    refb = np.fft.ifftn(a, axes=(0, 1, 2)) * np.prod(n)
    assert np.max(np.abs(b - refb)) < 1e-12 * np.max(np.abs(refb))


def test_fft_gaussian_closed_form():
    """transform.hpp:289-350: exp(-s r^2) -> (pi/s)^{3/2}/fs_vol exp(-G^2/4s),
    mean error < 1e-3 and round trip < 1e-15."""
    lat = np.eye(3) * 16.66
    n = spacing_grid(lat, 0.306320257)
    h = orc.Ham(lat, n)
    N = h.N
    r = rgrid(lat, n)
    r2 = np.sum(r * r, 1)
    nst = 7
    sig = 0.5 * (np.arange(nst) + 1)
    phi = np.exp(-np.outer(r2, sig)).astype(np.complex128)
    fphi = h.to_fourier(phi)
    fs_vol = N * h.volume / (N * N)  # fourier_space volume_element = V/N^2
    assert fs_vol == pytest.approx(0.0293659268, rel=1e-6)
    g = gcart(lat, n)
    g2 = np.sum(g * g, 1)
    ana = (np.pi / sig) ** 1.5 / fs_vol * np.exp(-0.25 * np.outer(g2, 1.0 / sig))
    assert np.sum(np.abs(fphi - ana)) / N < 1e-3
    phi2 = h.to_real(fphi)
    assert np.sum(np.abs(phi - phi2)) / phi.size < 1e-15


def test_zero_outside_sphere_fraction():
    """transform.hpp:273-287: fraction of the Fourier cube inside the
    inscribed ellipsoid sum (m_d/n_d)^2 <= 1/4 is 0.5240308896."""
    lat = np.eye(3) * 16.66
    n = spacing_grid(lat, 0.306320257)
    i0, i1, i2 = np.meshgrid(*[np.arange(k) for k in n], indexing="ij")
    xx = sym(i0, n[0]) / n[0]
    yy = sym(i1, n[1]) / n[1]
    zz = sym(i2, n[2]) / n[2]
    inside = (xx * xx + yy * yy + zz * zz <= 0.25).sum()
    assert inside / np.prod(n) == pytest.approx(0.5240308896, rel=1e-9)


# --------------------------------------------------------------------------- spheres

@pytest.mark.parametrize("center", [(0.0, 0.0, 0.0), (-5.0, 0.0, 0.0), (5.0, 0.0, 0.0), (-5.0, -5.0, -5.0),
                                    (5.0, 5.0, 5.0), (2.5, 2.5, 2.5)])
def test_sphere_count_cubic(center):
    """spherical_grid.hpp:272-355: 257 points, 0.9586598525 of the ball."""
    lat = np.eye(3) * 10.0
    n = spacing_grid(lat, 0.49672941)
    assert n == (20, 20, 20)
    h = orc.Ham(lat, n)
    coords, dist, rel = h.sphere(center, 2.0)
    assert len(coords) == 257
    theo = 4.0 / 3.0 * np.pi * 2.0 ** 3
    assert len(coords) * h.volume / h.N / theo == pytest.approx(0.9586598525, rel=1e-9)
    # brute force over the grid with minimum images
    r = rgrid(lat, n)
    d = r - np.asarray(center)
    d -= 10.0 * np.round(d / 10.0)
    assert np.sum(np.sum(d * d, 1) <= 4.0) == 257


def test_sphere_count_noncubic():
    """spherical_grid.hpp:357-372: 15963 points (with periodic images)."""
    lat = np.array([[7.57325, 0.0, 0.0], [-3.78662, 6.55862, 0.0], [5.78917e-16, 1.00271e-15, 9.45443]])
    n = spacing_grid(lat, 0.405578)
    h = orc.Ham(lat, n)
    coords, dist, rel = h.sphere((3.78659, -2.18619, 1.65434), 5.71)
    assert len(coords) == 15963
    theo = 4.0 / 3.0 * np.pi * 5.71 ** 3
    assert len(coords) * h.volume / h.N / theo == pytest.approx(1.0013282291, rel=1e-9)
    # every point is within the radius of some image, distances agree
    r = rgrid(lat, n)
    idx = (coords[:, 0] * n[1] + coords[:, 1]) * n[2] + coords[:, 2]
    rel_cart = rel.astype(np.float64) @ lat
    assert np.all(np.abs(np.linalg.norm(rel_cart, axis=1) - dist) < 1e-4)
    assert np.all(dist <= 5.71 + 1e-6)
    # the point minus its relative position is a lattice image of the center
    img = r[idx] - rel_cart - np.array([3.78659, -2.18619, 1.65434])
    frac = img @ np.linalg.inv(lat)
    assert np.max(np.abs(frac - np.round(frac))) < 1e-5


# --------------------------------------------------------------------------- H known answers

def _ho_setup():
    lat = np.eye(3) * 10.0
    n = spacing_grid(lat, 0.49672941)
    h = orc.Ham(lat, n)
    return lat, n, h


def test_h_constant_function():
    """ks_hamiltonian.hpp:324-357: H 1 = 0 with V = 0."""
    lat, n, h = _ho_setup()
    h.set_potential(np.zeros(h.N))
    phi = np.ones((h.N, 3), dtype=np.complex128)
    assert np.mean(np.abs(h.apply_rs(phi))) < 1e-14
    fs = h.to_fourier(phi)
    assert np.mean(np.abs(h.to_real(h.apply_fs(fs)))) < 1e-14


def test_h_plane_waves():
    """ks_hamiltonian.hpp:359-402,448-493: e^{i j k x} -> (j k)^2/2."""
    lat, n, h = _ho_setup()
    h.set_potential(np.zeros(h.N))
    kk = 2 * np.pi / 10.0
    x = rgrid(lat, n)[:, 0]
    nst = 6
    j = np.arange(nst)
    phi = np.exp(1j * np.outer(x, j * kk))
    hphi = h.apply_rs(phi)
    assert np.mean(np.abs(hphi - 0.5 * (j * kk) ** 2 * phi)) < 1e-14
    hphi2 = h.to_real(h.apply_fs(h.to_fourier(phi)))
    assert np.mean(np.abs(hphi2 - 0.5 * (j * kk) ** 2 * phi)) < 1e-14


def test_h_harmonic_oscillator():
    """ks_hamiltonian.hpp:404-445,495-536: mean |H phi - 1.5 w phi| = 0.0051420503."""
    lat, n, h = _ho_setup()
    ww = 2.0
    r2 = np.sum(rgrid(lat, n) ** 2, 1)
    h.set_potential(0.5 * ww * ww * r2)
    phi = np.repeat(np.exp(-ww * r2)[:, None], 6, 1).astype(np.complex128)
    d = np.mean(np.abs(h.apply_rs(phi) - 1.5 * ww * phi))
    assert d == pytest.approx(0.0051420503, rel=1e-7)
    d2 = np.mean(np.abs(h.to_real(h.apply_fs(h.to_fourier(phi))) - 1.5 * ww * phi))
    assert d2 == pytest.approx(0.0051420503, rel=1e-7)


def test_plane_wave_with_kpoint():
    """H e^{i(G0+k).r} = (|G0+k|^2/2 + V0) e^{i(G0+k).r} for constant V0."""
    lat = np.array([[6.0, 0.3, 0.0], [0.0, 5.5, 0.4], [0.2, 0.0, 6.3]])
    n = (12, 12, 15)
    h = orc.Ham(lat, n)
    h.set_potential(np.full(h.N, -0.7))
    b = 2 * np.pi * np.linalg.inv(lat).T
    kcov = np.array([0.3, -0.2, 0.45])
    kcart = kcov @ (b / (2 * np.pi))
    r = rgrid(lat, n)
    g0 = np.array([1, -2, 0]) @ b
    # periodic part e^{i G0 r} times Bloch phase handled through k: psi stores the periodic part
    phi = np.exp(1j * r @ g0)[:, None].astype(np.complex128)
    hphi = h.apply_rs(phi, kcov)
    e = 0.5 * np.sum((g0 + kcart) ** 2) - 0.7
    assert np.max(np.abs(hphi - e * phi)) < 1e-12


# --------------------------------------------------------------------------- KB projectors

def _kb_setup(seed=3):
    lat = np.array([[7.0, 0.0, 0.2], [0.3, 6.5, 0.0], [0.0, -0.4, 7.5]])
    n = (14, 13, 15)
    h = orc.Ham(lat, n)
    rng = np.random.default_rng(seed)
    h.set_potential(rng.standard_normal(h.N) * 0.3)
    atoms = [(0.1, 0.2, 0.3), (3.0, 3.1, 2.9), (-3.3, 0.5, 3.6)]
    h.coeffs = []
    for c in atoms:
        D = np.diag(rng.uniform(-2, 2, 4))
        D[1, 2] = D[2, 1] = 0.3  # exercise off-diagonal couplings
        h.add_projector_gauss(c, 2.6, 4, 0.9, D)
        h.coeffs.append(D)
    return h, rng


def test_kb_dense_bruteforce():
    """V_nl = sum_a B_a D_a B_a^H dV, B_a[x,l] = beta e^{-ik.pos} on its sphere
    points (projector_all.hpp:113-288), built densely and compared with the
    oracle's gather/contract/scatter."""
    h, rng = _kb_setup()
    kcov = np.array([0.4, -0.25, 0.1])
    nst = 3
    phi = rng.standard_normal((h.N, nst)) + 1j * rng.standard_normal((h.N, nst))
    dv = h.volume / h.N
    vnl = np.zeros((h.N, h.N), dtype=np.complex128)
    for a in range(3):
        coords, dist, rel, mat = h.projector_points(a, 4)
        idx = (coords[:, 0] * h.n[1] + coords[:, 1]) * h.n[2] + coords[:, 2]
        phase = np.exp(-1j * (rel.astype(np.float64) @ kcov))
        B = np.zeros((h.N, 4), dtype=np.complex128)
        for l in range(4):
            np.add.at(B[:, l], idx, mat[l] * phase)
        vnl += B @ h.coeffs[a] @ B.conj().T * dv
    full = h.apply_rs(phi, kcov)
    h0 = orc.Ham(h.lattice, h.n)  # same grid and potential, no projectors
    h0.set_potential(h.v)
    loc = h0.apply_rs(phi, kcov)
    assert np.max(np.abs((full - loc) - vnl @ phi)) < 1e-11 * np.max(np.abs(vnl @ phi))
    assert np.max(np.abs(vnl - vnl.conj().T)) < 1e-14
    # non-local energy = sum occ <phi|V_nl|phi>
    occ = np.array([1.0, 0.5, 2.0])
    # with the reference's inner product <a|b> = dV sum conj(a) b
    e_ref = dv * np.real(np.sum(occ * np.einsum("ps,pq,qs->s", phi.conj(), vnl, phi)))
    assert h.nonlocal_energy(phi, occ, kcov) == pytest.approx(e_ref, rel=1e-12)


def test_kb_projector_values():
    """beta = exp(-r^2/(2w^2)) Y_lm with Y00 = 1/(2 sqrt pi), Y1m = sqrt(3/4pi)(y,z,x)/r,
    evaluated on the float-rounded sphere tables (projector.hpp:63,
    spherical_grid.hpp:40-43)."""
    h, _ = _kb_setup()
    coords, dist, rel, mat = h.projector_points(1, 4)
    cart = rel.astype(np.float64) @ h.lattice
    r = np.linalg.norm(cart, axis=1)
    rad = np.exp(-dist.astype(np.float64) ** 2 / (2 * 0.9 ** 2))
    c = np.sqrt(3 / (4 * np.pi))
    with np.errstate(invalid="ignore", divide="ignore"):
        ref = np.stack([rad * 0.5 / np.sqrt(np.pi), rad * c * np.where(r > 0, cart[:, 1] / r, 0),
                        rad * c * np.where(r > 0, cart[:, 2] / r, 0), rad * c * np.where(r > 0, cart[:, 0] / r, 0)])
    assert np.max(np.abs(mat - ref)) < 1e-15


def test_sphere_form_equals_masked_a2():
    """mask(a2(scatter(psi))) built from numpy pieces equals the oracle's sphere form,
    and the oracle's a2 equals FFT(a1(IFFT)) for band-limited input."""
    h, rng = _kb_setup()
    kcov = np.array([0.2, 0.1, -0.3])
    g = h.gsphere(kcov, 6.0)
    nst = 2
    c = rng.standard_normal((nst, len(g))) + 1j * rng.standard_normal((nst, len(g)))
    hs = h.apply_sphere(c, g, kcov)
    cube = h.sphere_to_cube(c, g)
    ref = h.cube_to_sphere(h.apply_fs(cube, kcov), g)
    assert np.max(np.abs(hs - ref)) == 0.0
    # a2 vs a1 conjugated by FFTs
    rs = h.to_real(cube)
    alt = h.to_fourier(h.apply_rs(rs, kcov))
    ref2 = h.apply_fs(cube, kcov)
    assert np.max(np.abs(alt - ref2)) < 1e-11 * np.max(np.abs(ref2))
    # hermiticity on the sphere: <a|H b> = <H a|b>
    a = c[0:1]
    bb = c[1:2]
    lhs = np.vdot(a, h.apply_sphere(bb, g, kcov))
    rhs = np.vdot(h.apply_sphere(a, g, kcov), bb)
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)


def test_gsphere_matches_bruteforce():
    lat = np.array([[7.0, 0.0, 0.2], [0.3, 6.5, 0.0], [0.0, -0.4, 7.5]])
    n = (14, 13, 15)
    h = orc.Ham(lat, n)
    kcov = np.array([0.5, -0.1, 0.2])
    g = h.gsphere(kcov, 7.0)
    gc = gcart(lat, n) + kcov @ (2 * np.pi * np.linalg.inv(lat).T / (2 * np.pi))
    ref = np.nonzero(0.5 * np.sum(gc * gc, 1) < 7.0)[0]
    assert np.array_equal(g, ref)
    kin = h.gsphere_kinetic(kcov, g)
    assert np.max(np.abs(kin - 0.5 * np.sum(gc[ref] ** 2, 1))) < 1e-12


# --------------------------------------------------------------------------- dense linear algebra

def test_cholesky_kat():
    """cholesky.hpp:125-145."""
    m = np.array([[6432.12, 4502.48], [4502.48, 3151.74]], dtype=np.complex128)
    L, info = orc.cholesky(m)
    assert info == 0
    assert L[0, 0].real == pytest.approx(80.2005, rel=1e-6)
    assert L[0, 1] == 0.0
    assert L[1, 0].real == pytest.approx(56.1402992511, rel=1e-9)
    assert L[1, 1].real == pytest.approx(0.0824620974, rel=1e-7)


def test_cholesky_vs_scipy():
    rng = np.random.default_rng(7)
    a = rng.standard_normal((37, 37)) + 1j * rng.standard_normal((37, 37))
    s = a @ a.conj().T + 37 * np.eye(37)
    L, info = orc.cholesky(s)
    assert info == 0
    assert np.max(np.abs(L - sla.cholesky(s, lower=True))) < 1e-12
    bad = s.copy()
    bad[5, 5] = -1.0
    assert orc.cholesky(bad)[1] > 0


def test_invert_triangular_kat():
    """invert_triangular.hpp:72-125."""
    m = np.array([[4.0, 0.0], [-1.0, 2.0]], dtype=np.complex128)
    inv = orc.invert_lower(m)
    assert np.allclose(inv, [[0.25, 0.0], [0.125, 0.5]], atol=0, rtol=0)
    nn = 15
    ii, jj = np.meshgrid(np.arange(nn), np.arange(nn), indexing="ij")
    arr = np.where(ii < jj, 0.0, np.cos(ii) * (jj + 0.1) + np.sin(jj - 0.25) * (ii + 1.0)).astype(np.complex128)
    arr *= np.exp(1j * (ii + 1.0) * (np.cos(jj) - 0.25))
    arr = np.tril(arr)
    inv = orc.invert_lower(arr)
    mul = inv @ arr
    assert np.max(np.abs(mul - np.eye(nn))) < 1e-12


def test_diagonalize_kats():
    """diagonalize.hpp:213-317."""
    w, v = orc.eigh(np.array([[4.0, 0.0], [0.0, 2.0]]))
    assert np.allclose(w, [2.0, 4.0])
    assert abs(abs(v[0, 1]) - 1.0) < 1e-14 and abs(v[0, 0]) < 1e-14
    a = np.array([[0.088958, 1.183407, 1.191946], [1.183407, 1.371884, 0.705297], [1.191946, 0.705297, 0.392459]])
    w, _ = orc.eigh(a)
    assert np.allclose(w, [-1.0626903983, 0.1733844724, 2.7426069258], atol=2e-10)
    c = np.array([[0.088958, 1.183407 + 0.08285j, 1.191946 + 0.09413j],
                  [1.183407 - 0.08285j, 1.371884, 0.705297 + 0.12840j],
                  [1.191946 - 0.09413j, 0.705297 - 0.12840j, 0.392459]])
    w, _ = orc.eigh(c)
    assert np.allclose(w, [-1.0703967402, 0.1722879629, 2.7514097773], atol=2e-10)


def test_eigh_vs_scipy_and_orientation():
    rng = np.random.default_rng(11)
    a = rng.standard_normal((40, 40)) + 1j * rng.standard_normal((40, 40))
    hm = a + a.conj().T
    w, rows = orc.eigh(hm)
    assert np.max(np.abs(w - sla.eigh(hm, eigvals_only=True))) < 1e-11
    u = rows.conj().T  # columns are eigenvectors of hm
    assert np.max(np.abs(hm @ u - u * w)) < 1e-10
    assert np.max(np.abs(u.conj().T @ u - np.eye(40))) < 1e-12


def test_orthogonalize_gram_schmidt_structure():
    """orthogonalize.hpp:25-32: S -> I, and state i only mixes states j <= i."""
    rng = np.random.default_rng(5)
    for nst, npnt, bm in [(3, 50, False), (100, 300, True), (37, 200, False)]:
        shape = (nst, npnt) if bm else (npnt, nst)
        phi = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
        out, info = orc.orthogonalize(phi, 0.37, bm)
        assert info == 0
        s = orc.overlap(out, out, 0.37, bm)
        assert np.max(np.abs(s - np.eye(nst))) < 1e-12
        A = phi.T if bm else phi
        B = out.T if bm else out
        # B = A R with R upper triangular (Gram-Schmidt), i.e. A = B R^{-1}
        R = np.linalg.lstsq(A, B, rcond=None)[0]
        assert np.max(np.abs(np.tril(R, -1))) < 1e-9 * np.max(np.abs(R))
        # classical Gram-Schmidt on the columns gives the same states up to phase (1 here: positive diag)
        q, r = np.linalg.qr(A * np.sqrt(0.37))
        q = q * np.sign(np.real(np.diag(r)))[None, :] / np.sqrt(0.37) if np.all(np.isreal(np.diag(r))) else q
        ph = np.sum(np.conj(q) * B, 0)
        assert np.max(np.abs(np.abs(ph) * 0.37 - 1.0)) < 1e-9


def test_rotate5():
    rng = np.random.default_rng(2)
    phi = rng.standard_normal((30, 4)) + 1j * rng.standard_normal((30, 4))
    rot = rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))
    y = rng.standard_normal((30, 4)) + 1j * rng.standard_normal((30, 4))
    out = orc.rotate5(rot, phi, y, -1.0, 1.0, False)
    assert np.max(np.abs(out - (y - phi @ rot))) < 1e-13


# --------------------------------------------------------------------------- Poisson / singularity

@pytest.mark.slow
def test_poisson_point_charge():
    """poisson.hpp:266-312."""
    lat = np.diag([10.0, 10.0, 13.7])
    n = spacing_grid(lat, 0.1)
    assert n == (100, 100, 140)
    h = orc.Ham(lat, n)
    nst = 5
    dens = np.zeros((h.N, 1), dtype=np.complex128)
    dens[0, 0] = -1.0
    dset = np.zeros((h.N, nst), dtype=np.complex128)
    dset[0, :] = -(1.0 + np.arange(nst))
    pot = h.poisson_in_place(dens)
    pset = h.poisson_in_place(dset)
    s0 = np.sum(np.abs(pot.real)) + np.sum(np.abs(pset.real) / (1.0 + np.arange(nst)))
    s1 = np.sum(np.abs(pot.imag)) + np.sum(np.abs(pset.imag) / (1.0 + np.arange(nst)))
    # the reference checks these with Catch's default Approx (relative 100 float
    # epsilon = 1.19e-5); the recorded constants are good to about 1e-6
    assert s0 / (nst + 1.0) == pytest.approx(82.9383793318, rel=1.19e-5)
    assert s1 / (nst + 1.0) <= 5e-12
    assert pot[0, 0].real == pytest.approx(-0.0238339212, rel=1.19e-5)


def test_poisson_plane_wave():
    """poisson.hpp:315-351 (smaller grid): plane wave -> 4pi/k^2."""
    lat = np.diag([10.0, 10.0, 13.7])
    n = (20, 20, 28)
    h = orc.Ham(lat, n)
    kk = 2 * np.pi / 10.0
    x = rgrid(lat, n)[:, 0]
    nst = 5
    d = np.exp(1j * kk * x)[:, None] * (1.0 + np.arange(nst))[None, :]
    p = h.poisson_in_place(d)
    assert np.mean(np.abs(p / (1.0 + np.arange(nst)) - 4 * np.pi / kk ** 2 * np.exp(1j * kk * x)[:, None])) < 1e-13


def test_singularity_kats():
    """singularity_correction.hpp:132-145."""
    aa = 10.18
    lat = np.eye(3) * aa
    assert orc.sing_auxiliary(lat, 2 * np.pi * np.array([0.0, -0.5, 0.0])) == pytest.approx(25.9081, rel=1e-5)
    assert orc.sing_auxiliary(lat, 2 * np.pi * np.array([8.3333333333333332E-003, 7.4999999999999997E-002,
                                                         0.26666666666666666])) == pytest.approx(42.650855183181122,
                                                                                               rel=1e-12)
    assert orc.sing_auxiliary(lat, 2 * np.pi * np.array([0.11666666666666667, 0.20000000000000001,
                                                         0.21666666666666667])) == pytest.approx(29.780683447124286,
                                                                                               rel=1e-12)
    # 2x2x2 unshifted grid: coordinates 0 or 1/2 -> wrapped to -1/2 (brillouin.hpp:78-84)
    ks = np.array([[a, b, c] for a in (0, 1) for b in (0, 1) for c in (0, 1)], dtype=float) * 0.5
    ks = np.where(ks >= 0.5, ks - 1.0, ks) * 2 * np.pi
    fk, fz = orc.sing_correction(lat, ks, np.full(8, 1 / 8))
    assert fz == pytest.approx(0.30983869660201141, rel=1e-10)
    assert np.allclose(fk, 0.18644848345224296, rtol=1e-10, atol=0)
    assert -8 * aa ** 3 * (fk[0] - fz) == pytest.approx(1041.3915164701, rel=1e-9)


# --------------------------------------------------------------------------- propagation

def test_exponential_diagonal_closed_form():
    """exponential.hpp:94-150 idea: for an eigenstate, exp(f H) phi = e^{f e} phi up
    to the Taylor truncation; with order 16 and |f e| < 1 this is below 1e-12."""
    lat = np.eye(3) * 6.0
    n = (12, 12, 12)
    h = orc.Ham(lat, n)
    h.set_potential(np.full(h.N, 0.25))
    g = h.gsphere((0, 0, 0), 3.0)
    kin = h.gsphere_kinetic((0, 0, 0), g)
    nst = 4
    c = np.zeros((nst, len(g)), dtype=np.complex128)
    sel = [0, 5, 11, 20]
    for s, i in enumerate(sel):
        c[s, i] = 1.0
    e = kin[sel] + 0.25
    out = h.exponential_in_place(c, g, -0.1, order=16)
    assert np.max(np.abs(out[np.arange(nst), sel] - np.exp(-0.1 * e))) < 1e-12
    phi2, half = h.exponential_2_for_1(c, g, -0.1, -0.2, order=16)
    assert np.max(np.abs(half[np.arange(nst), sel] - np.exp(-0.1 * e))) < 1e-12
    assert np.max(np.abs(phi2[np.arange(nst), sel] - np.exp(-0.2 * e))) < 1e-12


def test_etrs_step_unitary_and_phase():
    """one ETRS step = exp_T4(-i dt/2 H) applied twice; for an eigenstate the
    result is the Taylor-4 approximant squared."""
    lat = np.eye(3) * 6.0
    n = (12, 12, 12)
    h = orc.Ham(lat, n)
    h.set_potential(np.full(h.N, -0.3))
    g = h.gsphere((0, 0, 0), 4.0)
    kin = h.gsphere_kinetic((0, 0, 0), g)
    c = np.zeros((1, len(g)), dtype=np.complex128)
    c[0, 7] = 1.0
    e = kin[7] - 0.3
    dt = 0.05
    z = -1j * dt / 2 * e
    t4 = 1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24
    out = h.etrs_step(c, g, dt)
    assert abs(out[0, 7] - t4 ** 2) < 1e-14


def test_randomize_kats():
    """randomize.hpp:96-144: norms of pcg32 U[-1,1) fields on the 20^3 grid of a
    10 bohr cube (spacing 0.49672941), 12 states, volume element 1000/8000."""
    npts, nst = 8000, 12
    r = orc.randomize(npts, nst, 0, complex_values=False)
    c = orc.randomize(npts, nst, 0, complex_values=True)
    kat_r = [330.1381395023, 330.5444105287, 331.5435469092, 340.0201855451, 332.959828166, 333.1967703871,
             328.9429576951, 336.8659386972, 331.6228884743, 334.5374311686, 334.6847273315, 334.1846706334]
    kat_c = [669.1459385544, 664.8544595319, 669.6590362331, 668.4278672543, 668.0965017232, 665.1942561788,
             665.6688584485, 672.3987040517, 671.7988968555, 672.6486737858, 664.3526206348, 674.5678452953]
    assert np.allclose(0.125 * np.sum(r * r, 0), kat_r, rtol=1e-10, atol=0)
    assert np.allclose(0.125 * np.sum(np.abs(c) ** 2, 0), kat_c, rtol=1e-10, atol=0)
    assert np.all(np.abs(c.real) <= 1.0) and np.all(np.abs(c.imag) <= 1.0)
