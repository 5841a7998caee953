"""Pins the oracle's projector channels up to l = 2 (projector.hpp:38-93,
projector_all.hpp:113-288 restated in oracle/oracle.c).

  * the real spherical harmonics of the channels (l = 0, 1, 2; m = -l..l in
    the order of projector.hpp:63) equal the standard real harmonics built
    from scipy's complex ones: Y_lm = sqrt(2) (-1)^m Im Y_l^|m| (m < 0),
    Y_l0, sqrt(2) (-1)^m Re Y_l^m (m > 0).  A swapped m or a dropped
    normalisation fails it; they are orthonormal on the sphere (a
    Gauss-Legendre x uniform-angle product rule, exact for these degrees).
  * the non-local term of projectors with nine channels (s + p + d) and a
    full coefficient matrix coupling every channel equals the dense
    sum_a B_a D_a B_a^H dV built from the oracle's projector tables.
The reference takes the harmonics from the un-vendored sharmonic library; the
convention here is the standard real one (parity against sharmonic unpinned).
"""
import numpy as np
import pytest
import scipy.special as ss

from oracle import oracle as orc


def _scipy_real_ylm(l, m, v):
    r = np.linalg.norm(v)
    theta = np.arccos(np.clip(v[2] / r, -1, 1))
    phi = np.arctan2(v[1], v[0])
    y = ss.sph_harm_y(l, abs(m), theta, phi)
    if m < 0:
        return np.sqrt(2) * (-1) ** m * y.imag
    if m == 0:
        return y.real
    return np.sqrt(2) * (-1) ** m * y.real


def test_real_harmonics_against_scipy():
    rng = np.random.default_rng(5)
    for v in rng.standard_normal((40, 3)):
        for l in range(3):
            for m in range(-l, l + 1):
                assert orc.ylm_real(l, m, v) == pytest.approx(_scipy_real_ylm(l, m, v), abs=1e-13)
    assert orc.ylm_real(2, 0, np.zeros(3)) == 0.0
    assert orc.ylm_real(0, 0, np.zeros(3)) == pytest.approx(0.5 / np.sqrt(np.pi), rel=1e-15)


def test_real_harmonics_orthonormal():
    x, w = np.polynomial.legendre.leggauss(8)
    nphi = 16
    phis = 2 * np.pi * np.arange(nphi) / nphi
    chans = [(l, m) for l in range(3) for m in range(-l, l + 1)]
    gram = np.zeros((9, 9))
    for ct, wt in zip(x, w):
        st = np.sqrt(1 - ct * ct)
        for ph in phis:
            v = np.array([st * np.cos(ph), st * np.sin(ph), ct])
            y = np.array([orc.ylm_real(l, m, v) for l, m in chans])
            gram += wt * (2 * np.pi / nphi) * np.outer(y, y)
    assert np.max(np.abs(gram - np.eye(9))) < 1e-13


def test_kb_nine_channels_dense_bruteforce():
    lat = np.array([[7.0, 0.0, 0.2], [0.3, 6.5, 0.0], [0.0, -0.4, 7.5]])
    n = (14, 13, 15)
    h = orc.Ham(lat, n)
    rng = np.random.default_rng(8)
    h.set_potential(rng.standard_normal(h.N) * 0.3)
    nlms = [9, 4, 9]
    coeffs = []
    for c, nlm in zip([(0.1, 0.2, 0.3), (3.0, 3.1, 2.9), (-3.3, 0.5, 3.6)], nlms):
        A = rng.standard_normal((nlm, nlm))
        D = 0.5 * (A + A.T)  # every channel coupled to every other
        h.add_projector_gauss(c, 2.6, nlm, 0.9, D)
        coeffs.append(D)
    kcov = np.array([0.4, -0.25, 0.1])
    nst = 3
    phi = rng.standard_normal((h.N, nst)) + 1j * rng.standard_normal((h.N, nst))
    dv = h.volume / h.N
    vnl = np.zeros((h.N, h.N), dtype=np.complex128)
    for a, nlm in enumerate(nlms):
        coords, dist, rel, mat = h.projector_points(a, nlm)
        idx = (coords[:, 0] * h.n[1] + coords[:, 1]) * h.n[2] + coords[:, 2]
        phase = np.exp(-1j * (rel.astype(np.float64) @ kcov))
        B = np.zeros((h.N, nlm), dtype=np.complex128)
        for l in range(nlm):
            np.add.at(B[:, l], idx, mat[l] * phase)
        vnl += B @ coeffs[a] @ B.conj().T * dv
    h0 = orc.Ham(lat, n)
    h0.set_potential(h.v)
    got = h.apply_rs(phi, kcov) - h0.apply_rs(phi, kcov)
    assert np.max(np.abs(got - vnl @ phi)) < 1e-11 * np.max(np.abs(vnl @ phi))
    # the d-channel values are the Gaussian times the l = 2 harmonics
    coords, dist, rel, mat = h.projector_points(0, 9)
    cart = rel.astype(np.float64) @ lat
    rad = np.exp(-dist.astype(np.float64) ** 2 / (2 * 0.9 ** 2))
    for ip in range(0, len(dist), 97):
        for ch in range(4, 9):
            assert mat[ch, ip] == pytest.approx(rad[ip] * _scipy_real_ylm(2, ch - 6, cart[ip]) if
                                                np.linalg.norm(cart[ip]) > 0 else 0.0, abs=1e-13)
