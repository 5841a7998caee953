"""The oracle's atomic-sphere sum (atomic_potential::local_potential /
ionic_density, src/hamiltonian/atomic_potential.hpp:116-200) pinned by closed
forms, CPU only:

* one atom, analytic radial function: the sum equals the function at the
  minimum-image distance wherever only one image is inside the sphere;
* the periodic sum of a normalised Gaussian over all images equals its
  reciprocal-space (Poisson summation) series (1/V) sum_G exp(-sigma^2 G^2 / 2)
  e^{iG.r}, which pins the image enumeration on cubic and fcc cells;
* the declared ionic convention integrates to -Z (and its table is the
  Gaussian of width sigma).
The tolerances cover the linear interpolation of the radial tables
(max |f''| dr^2 / 8 relative to the peak).
"""
import math

import numpy as np

from oracle import oracle as orc


def _grid_points(lat, n):
    i0, i1, i2 = np.meshgrid(np.arange(n[0]) / n[0], np.arange(n[1]) / n[1], np.arange(n[2]) / n[2], indexing="ij")
    frac = np.stack([i0.ravel(), i1.ravel(), i2.ravel()], axis=1)
    return frac @ lat


def _gauss(r, s):
    return np.exp(-0.5 * (r / s) ** 2) / (2 * math.pi * s * s) ** 1.5


def test_single_atom_matches_the_analytic_function():
    L = 8.0
    lat = np.eye(3) * L
    n = (16, 16, 16)
    pos = np.array([[1.3, 2.9, 6.1]])
    s, rc, dr = 0.9, 3.0, 1e-3
    r = np.arange(int(rc / dr) + 2) * dr
    f = orc.atomic_field(lat, n, pos, rc, 0, _gauss(r, s)[None, :], dr, images=1)
    pts = _grid_points(lat, n)
    d = pts - pos[0]
    d -= L * np.round(d / L)  # minimum image (rc < L / 2: one image at most)
    dist = np.linalg.norm(d, axis=1)
    ref = np.where(dist <= rc, _gauss(dist, s), 0.0)
    assert np.max(np.abs(f - ref)) < 1e-6 * _gauss(0.0, s)
    assert np.count_nonzero(f) == np.count_nonzero(dist <= rc)


def _poisson_sum(lat, pts, center, s):
    """(1/V) sum_G exp(-s^2 |G|^2 / 2) cos(G.(r - R)), G = 2 pi m . inv(lat)^T"""
    V = abs(np.linalg.det(lat))
    b = 2 * math.pi * np.linalg.inv(lat).T
    m = np.arange(-12, 13)
    M = np.array(np.meshgrid(m, m, m, indexing="ij")).reshape(3, -1).T
    G = M @ b
    w = np.exp(-0.5 * s * s * np.sum(G * G, axis=1))
    keep = w > 1e-18
    G, w = G[keep], w[keep]
    return (np.cos((pts - center) @ G.T) @ w) / V


def test_periodic_gaussian_equals_its_reciprocal_series_cubic():
    L = 3.0  # radius 4.2 > L: several images of the atom reach each point
    lat = np.eye(3) * L
    n = (8, 8, 8)
    pos = np.array([[0.2, 2.7, 1.1]])
    s, dr = 0.6, 2e-4
    rc = s * math.sqrt(-2 * math.log(1e-14))
    r = np.arange(int(rc / dr) + 2) * dr
    f = orc.atomic_field(lat, n, pos, rc, 0, _gauss(r, s)[None, :], dr, images=3)
    ref = _poisson_sum(lat, _grid_points(lat, n), pos[0], s)
    assert np.max(np.abs(f - ref)) < 1e-6 * _gauss(0.0, s)


def test_periodic_gaussian_equals_its_reciprocal_series_fcc():
    a = 4.0
    lat = 0.5 * a * np.array([[0.0, 1.0, 1.0], [1.0, 0.0, 1.0], [1.0, 1.0, 0.0]])
    n = (8, 8, 8)
    pos = np.array([[0.3, -0.4, 1.7]])
    s, dr = 0.55, 2e-4
    rc = s * math.sqrt(-2 * math.log(1e-14))
    r = np.arange(int(rc / dr) + 2) * dr
    f = orc.atomic_field(lat, n, pos, rc, 0, _gauss(r, s)[None, :], dr, images=4)
    ref = _poisson_sum(lat, _grid_points(lat, n), pos[0], s)
    assert np.max(np.abs(f - ref)) < 1e-6 * _gauss(0.0, s)


def test_declared_ionic_convention_integrates_to_minus_z():
    from inq_amd import potential as pot
    s = 0.625
    r = np.linspace(0.0, 6.0, 7)
    assert np.allclose(pot.long_range_density(r, s), -_gauss(r, s), rtol=1e-15, atol=0)
    # radial integral of the unit density: -1
    rr = np.linspace(0.0, pot.gaussian_radius(s), 200001)
    q = np.trapezoid(4 * math.pi * rr * rr * pot.long_range_density(rr, s), rr)
    assert abs(q + 1.0) < 1e-9
    # and its potential, from the radial Poisson integral V(r) = 4 pi [ (1/r) int_0^r rho r'^2 + int_r^inf rho r' ]
    for r0 in (0.3, 1.0, 2.5):
        inner = np.linspace(0.0, r0, 20001)
        outer = np.linspace(r0, pot.gaussian_radius(s), 200001)
        v = 4 * math.pi * (np.trapezoid(pot.long_range_density(inner, s) * inner ** 2, inner) / r0
                           + np.trapezoid(pot.long_range_density(outer, s) * outer, outer))
        assert abs(v - pot.long_range_potential(np.array([r0]), s)[0]) < 1e-8
