"""Pins the oracle's exact exchange (exchange_operator.hpp:52-231 restated in
oracle/oracle.c) by closed forms and by the defining property of ACE.

  * two plane waves phi_a = e^{i G_a r}/sqrt(V), phi_b = e^{i G_b r}/sqrt(V):
    rho_ab = e^{i (G_a - G_b) r}/V, so Poisson gives 4 pi/(V |dG|^2) rho_ab and
    K phi_a = -(2 pi alpha occ_b / (V |dG|^2)) phi_a (rho_aa has only G = 0,
    which is excluded); the ACE energy is -2 pi alpha occ_a occ_b/(V |dG|^2).
    A dropped 1/2, 4 pi, occupation or volume factor fails this.
  * the same with the hf state at another k-point: the pair solve is
    shifted by s = -k + q (exchange_operator.hpp:151), so K phi_a =
    -(2 pi alpha occ_b / (V |G_a - G_b + s|^2)) phi_a; with G_b = G_a the
    density is uniform and the shift alone sets the value; with s = 0 the
    uniform density takes the zeroterm.  A shift of the wrong sign fails it.
  * ACE is exact on the states it is built from: -xi xi^H phi dV == K phi.
  * K is Hermitian: <a|K b> = conj(<b|K a>).
"""
import numpy as np
import pytest

from inq_amd.systems import rgrid
from oracle import oracle as orc

LAT = np.diag([6.0, 6.0, 7.0])
NG = (12, 12, 14)


def _plane_waves(G):
    r = rgrid(LAT, NG)
    V = abs(np.linalg.det(LAT))
    return np.stack([np.exp(1j * (r @ g)) for g in G], axis=1) / np.sqrt(V), V


def _random_bandlimited(nst, seed, gmax=2):
    """states with a few low Fourier components: real-space [N][nst]"""
    rng = np.random.default_rng(seed)
    cube = np.zeros(NG + (nst,), dtype=np.complex128)
    for i0 in range(-gmax, gmax + 1):
        for i1 in range(-gmax, gmax + 1):
            for i2 in range(-gmax, gmax + 1):
                cube[i0, i1, i2, :] = rng.standard_normal(nst) + 1j * rng.standard_normal(nst)
    phi = np.fft.ifftn(cube, axes=(0, 1, 2)).reshape(-1, nst)
    V = abs(np.linalg.det(LAT))
    dv = V / phi.shape[0]
    # orthonormalise with the real-space volume element
    s = dv * phi.conj().T @ phi
    L = np.linalg.cholesky(s)
    return np.linalg.solve(L.conj(), phi.T).T, dv


def test_two_plane_waves_closed_form():
    b = 2 * np.pi / np.diag(LAT)
    G = [np.array([b[0], 0, 0]), np.array([0, -b[1], 2 * b[2]])]
    phi, V = _plane_waves(G)
    occ = np.array([2.0, 0.7])
    alpha = 0.25
    h = orc.Ham(LAT, NG)
    kphi = h.exchange_direct(phi, phi, occ, exx_coeff=alpha, scale=1.0)
    d2 = float(np.sum((G[0] - G[1]) ** 2))
    lam_a = -2 * np.pi * alpha * occ[1] / (V * d2)
    lam_b = -2 * np.pi * alpha * occ[0] / (V * d2)
    assert np.max(np.abs(kphi[:, 0] - lam_a * phi[:, 0])) < 1e-13 * np.max(np.abs(phi))
    assert np.max(np.abs(kphi[:, 1] - lam_b * phi[:, 1])) < 1e-13 * np.max(np.abs(phi))
    xi, ex = h.ace_build(phi, occ, exx_coeff=alpha)
    assert ex == pytest.approx(-2 * np.pi * alpha * occ[0] * occ[1] / (V * d2), rel=1e-12)


def test_zeroterm_constant_density():
    """a uniform density has only G = 0: Poisson returns zeroterm / V times N/N."""
    h = orc.Ham(LAT, NG)
    V = abs(np.linalg.det(LAT))
    rho = np.full((h.N, 2), 1.0 / V, dtype=np.complex128)
    out = h.poisson_in_place(rho, zeroterm=3.5)
    assert np.max(np.abs(out - 3.5 / V)) < 1e-13


def test_ace_exact_on_building_set_and_hermitian():
    h = orc.Ham(LAT, NG)
    phi, dv = _random_bandlimited(4, 5)
    occ = np.array([2.0, 2.0, 1.0, 0.5])
    alpha = 1.0
    kphi = h.exchange_direct(phi, phi, occ, exx_coeff=alpha, scale=1.0)
    xi, ex = h.ace_build(phi, occ, exx_coeff=alpha)
    ace = h.ace_apply(xi, phi)
    assert np.max(np.abs(ace - kphi)) < 1e-11 * np.max(np.abs(kphi))
    # applying to a subset of the states (fewer columns than xi) gives the same columns
    sub = h.ace_apply(xi, np.ascontiguousarray(phi[:, 1:3]))
    assert np.max(np.abs(sub - ace[:, 1:3])) < 1e-13 * np.max(np.abs(ace))
    # Hermiticity of the direct operator on another random set
    psi, _ = _random_bandlimited(3, 9)
    kpsi = h.exchange_direct(psi, phi, occ, exx_coeff=alpha, scale=1.0)
    m = dv * psi.conj().T @ kpsi
    assert np.max(np.abs(m - m.conj().T)) < 1e-12 * np.max(np.abs(m))
    # exchange energy = 1/2 sum occ <phi|K phi> (negative definite K)
    e2 = 0.5 * sum(occ[i] * dv * np.vdot(phi[:, i], kphi[:, i]).real for i in range(4))
    assert ex == pytest.approx(e2, rel=1e-12)
    assert ex < 0


def test_shifted_pair_solves_closed_form():
    b = 2 * np.pi / np.diag(LAT)
    Ga = np.array([b[0], 0, 0])
    Gb = np.array([0, -b[1], 2 * b[2]])
    phi, V = _plane_waves([Ga, Gb])
    alpha, occ_b = 0.25, 0.7
    s = np.array([0.31, -0.22, 0.17])  # cartesian -k + q, not orthogonal to G_a - G_b
    h = orc.Ham(LAT, NG)
    ka = h.exchange_direct(phi[:, :1], phi[:, 1:], [occ_b], exx_coeff=alpha, gshift=s)
    lam = -2 * np.pi * alpha * occ_b / (V * float(np.sum((Ga - Gb + s) ** 2)))
    assert np.max(np.abs(ka[:, 0] - lam * phi[:, 0])) < 1e-13 * np.max(np.abs(phi))
    wrong = -2 * np.pi * alpha * occ_b / (V * float(np.sum((Ga - Gb - s) ** 2)))
    assert abs(wrong - lam) > 1e-3 * abs(lam)
    # the same plane wave at another k-point: uniform density, shifted
    kaa = h.exchange_direct(phi[:, :1], phi[:, :1], [occ_b], exx_coeff=alpha, gshift=s)
    lam0 = -2 * np.pi * alpha * occ_b / (V * float(np.sum(s * s)))
    assert np.max(np.abs(kaa[:, 0] - lam0 * phi[:, 0])) < 1e-13 * np.max(np.abs(phi))
    # no shift: the uniform density takes the zeroterm, -alpha/2 occ zeroterm / V
    kz = h.exchange_direct(phi[:, :1], phi[:, :1], [occ_b], exx_coeff=alpha, gshift=(0, 0, 0), zeroterm=3.5)
    assert np.max(np.abs(kz[:, 0] + 0.5 * alpha * occ_b * 3.5 / V * phi[:, 0])) < 1e-13 * np.max(np.abs(phi))
