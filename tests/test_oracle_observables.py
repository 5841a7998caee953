"""Pins the oracle's energy and dipole reductions (energy::calculate,
src/hamiltonian/energy.hpp:61-113; observables::dipole,
src/observables/dipole.hpp:24-41) with closed forms, so that the GPU records
checked against them are anchored to mathematics and not to the oracle alone.

* single plane waves e^{i(G+k).r} under a constant potential V0 and no
  projectors: kinetic = sum occ |G+k|^2/2, eigenvalue sum = sum occ
  (|G+k|^2/2 + V0), norm = 1 per state, non-local = 0;  G+k is converted to
  cartesian here from the lattice (skewed cell, k != 0);
* for normalised states with projectors, the eigenvalue sum splits exactly
  into kinetic + sum occ <phi|V|phi> + non-local, with <phi|V|phi> summed in
  numpy over the real-space density (pins the scale of the non-local term);
* occupation weighting: every record entry except the norm is linear in occ;
* dipole of densities with compact support on a whole-cube basis: charge
  times the centre in the symmetric range, for a centre inside, one next to
  the cell edge, one past the middle (negative coordinates) and one whose
  support straddles the i = n/2 boundary (sym(n/2) = -n/2), on a skewed cell.
"""
import numpy as np

from oracle import oracle as orc

LAT = np.array([[6.0, 0.0, 0.0], [1.0, 7.0, 0.0], [0.5, 0.3, 6.5]])
NG = (16, 18, 15)
N = int(np.prod(NG))
VOL = abs(np.linalg.det(LAT))


def sym(i, n):
    return i - n if i >= (n + 1) // 2 else i


def gk_cart(gidx, kcov):
    """cartesian G+k of flattened FFT indices: covariant 2 pi m + k mapped through
    the dual basis (rows of inv(LAT)^T)."""
    i2 = gidx % NG[2]
    i1 = (gidx // NG[2]) % NG[1]
    i0 = gidx // (NG[1] * NG[2])
    m = np.stack([[sym(a, NG[0]) for a in i0], [sym(a, NG[1]) for a in i1], [sym(a, NG[2]) for a in i2]], -1)
    return (2 * np.pi * m + np.asarray(kcov)) @ np.linalg.inv(LAT).T


def test_plane_waves_closed_form():
    h = orc.Ham(LAT, NG)
    v0 = -0.37
    h.set_potential(np.full(N, v0))
    k = (0.3, -0.2, 0.1)
    g = h.gsphere(k, 9.0)
    vol_el = VOL / float(N) ** 2
    pick = [0, 7, len(g) // 3, len(g) // 2, len(g) - 1]
    c = np.zeros((len(pick), len(g)), dtype=np.complex128)
    for s, p in enumerate(pick):
        c[s, p] = np.exp(0.7j * s) / np.sqrt(vol_el)
    occ = np.array([2.0, 1.0, 0.5, 1.5, 0.25])
    kin = 0.5 * np.sum(gk_cart(g[pick], k) ** 2, 1)
    out = h.energy_sphere(c, g, occ, k)
    assert abs(out[0] - np.sum(occ * kin)) < 1e-11 * np.sum(occ * kin)
    assert abs(out[1] - np.sum(occ * (kin + v0))) < 1e-11 * np.sum(occ * kin)
    assert abs(out[2]) < 1e-14
    assert abs(out[3] - len(pick)) < 1e-12


def _system_with_projectors(seed):
    rng = np.random.default_rng(seed)
    h = orc.Ham(LAT, NG)
    r = np.stack(np.meshgrid(*[np.arange(n) for n in NG], indexing="ij"), -1).reshape(-1, 3)
    v = -1.5 * np.exp(-np.sum(((r / np.array(NG) - 0.4) @ LAT) ** 2, 1) / 2.0) + 0.2
    h.set_potential(v)
    for a in range(2):
        pos = rng.uniform(-0.4, 0.4, 3) @ LAT
        d = rng.uniform(-2, 2, (4, 4))
        h.add_projector_gauss(pos, 1.6, 4, 0.7, (d + d.T) / 2)
    return h, v


def _random_normalised(npw, nst, seed, vol_el):
    rng = np.random.default_rng(seed)
    c = rng.standard_normal((nst, npw)) + 1j * rng.standard_normal((nst, npw))
    return c / np.sqrt(vol_el * np.sum(np.abs(c) ** 2, 1))[:, None]


def test_eigenvalue_sum_splits_into_terms():
    h, v = _system_with_projectors(3)
    k = (0.25, 0.1, -0.3)
    g = h.gsphere(k, 7.0)
    vol_el = VOL / float(N) ** 2
    c = _random_normalised(len(g), 4, 5, vol_el)
    occ = np.array([2.0, 1.0, 0.5, 1.5])
    out = h.energy_sphere(c, g, occ, k)
    cube = np.zeros((4, N), dtype=np.complex128)
    cube[:, g] = c
    phi = np.fft.ifftn(cube.reshape((4,) + NG), axes=(1, 2, 3)).reshape(4, N)
    dv = VOL / N
    vexp = np.sum(np.abs(phi) ** 2 * v[None, :], 1) * dv
    assert abs(np.sum(np.abs(phi) ** 2) * dv - 4.0) < 1e-12  # the states are normalised
    assert abs(out[2]) > 1e-3  # the projectors contribute
    assert abs(out[1] - (out[0] + np.sum(occ * vexp) + out[2])) < 1e-10 * abs(out[1])


def test_occupation_weighting_and_phase():
    h, _ = _system_with_projectors(4)
    k = (0.0, 0.2, 0.0)
    g = h.gsphere(k, 7.0)
    vol_el = VOL / float(N) ** 2
    c = _random_normalised(len(g), 3, 6, vol_el)
    occ = np.array([2.0, 0.75, 1.25])
    tot = h.energy_sphere(c, g, occ, k)
    parts = np.array([h.energy_sphere(c, g, np.eye(3)[s], k) for s in range(3)])
    assert np.allclose(tot[:3], occ @ parts[:, :3], rtol=1e-12, atol=1e-13)
    assert abs(tot[3] - 3.0) < 1e-12 and np.allclose(parts[:, 3], 3.0, atol=1e-12)
    ph = c * np.exp(1j * np.array([0.3, -1.1, 2.0]))[:, None]
    assert np.allclose(h.energy_sphere(ph, g, occ, k), tot, rtol=1e-12, atol=1e-13)


def _blob(center, w):
    """density weights w[d] (-1..1, symmetric) around grid index `center`,
    as a wavefunction sqrt(weights) on the grid"""
    phi = np.zeros(NG)
    for dx in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dz in (-1, 0, 1):
                idx = tuple((center[d] + o) % NG[d] for d, o in enumerate((dx, dy, dz)))
                phi[idx] = np.sqrt(w[abs(dx)] * w[abs(dy)] * w[abs(dz)])
    return phi


def test_dipole_closed_forms():
    h = orc.Ham(LAT, NG)
    h.set_potential(np.zeros(N))
    g = h.gsphere((0, 0, 0), 1e6)
    assert len(g) == N  # whole cube
    dv = VOL / N
    w = (0.5, 0.25)  # weights of the centre and of its neighbours along each axis
    charge1 = (w[0] + 2 * w[1]) ** 3 * dv
    # centre inside the cell, next to the +edge (x index 6 of 16: support 5..7;
    # y index 7 of 18: support 6..8, sym(9) would be -9),
    # past the middle (x index 11 -> -5, y index 12 of 18 -> -6)
    cases = [((2, 3, 1), (2 / 16, 3 / 18, 1 / 15)),
             ((6, 7, 6), (6 / 16, 7 / 18, 6 / 15)),
             ((11, 12, 9), (-5 / 16, -6 / 18, -6 / 15))]
    for centre, frac in cases:
        phi = _blob(centre, w)
        c = np.fft.fftn(phi).reshape(1, N)[:, g]
        _, dip = h.density_dipole_sphere(c, g, np.array([2.0]))
        want = 2.0 * charge1 * (np.array(frac) @ LAT)
        assert np.allclose(dip, want, rtol=1e-12, atol=1e-12), (centre, dip, want)
    # support straddling i0 = 8 of 16: indices 7, 8, 9 map to 7, -8, -7 (sym(8) = -8)
    phi = _blob((8, 2, 3), w)
    c = np.fft.fftn(phi).reshape(1, N)[:, g]
    _, dip = h.density_dipole_sphere(c, g, np.array([1.0]))
    yz = (w[0] + 2 * w[1]) ** 2
    fx = (7 * w[1] - 8 * w[0] - 7 * w[1]) / 16 * yz * dv
    fy = 2 / 18 * charge1
    fz = 3 / 15 * charge1
    assert np.allclose(dip, np.array([fx, fy, fz]) @ LAT, rtol=1e-12, atol=1e-12), dip
    # occupation weighting of two states
    p1, p2 = _blob((2, 3, 1), w), _blob((11, 12, 9), w)
    c = np.stack([np.fft.fftn(p1).ravel()[g], np.fft.fftn(p2).ravel()[g]])
    rho, dip = h.density_dipole_sphere(c, g, np.array([1.5, 0.5]))
    want = charge1 * ((1.5 * np.array(cases[0][1]) + 0.5 * np.array(cases[2][1])) @ LAT)
    assert np.allclose(dip, want, rtol=1e-12, atol=1e-12)
    assert abs(rho.sum() * dv - 2.0 * charge1) < 1e-12
