"""Pins the oracle's self-consistent ETRS step (etrs.hpp:22-79 with the Hartree
potential of self_consistency.hpp:100-125) against a dense calculation that
shares none of its code: the Hamiltonian of a small sphere basis as an explicit
matrix (numpy FFTs), V_H from numpy FFTs, and exact exponentials
(scipy.linalg.expm) in place of Taylor-4.  At dt = 0.01 the two differ only by
the Taylor truncation (~1e-10 here).  The test also shows that two plausible
mistakes move the result far outside the tolerance: feeding the half-step
density instead of the full-step prediction into H(t + dt) (one SCF pass), and
SCF passes that do not restart from the half-step states (five passes).
"""
import numpy as np
import scipy.linalg as sla

from oracle import oracle as orc

LAT = np.eye(3) * 5.0
NG = (12, 12, 12)
N = int(np.prod(NG))
VOL = 125.0
DV = VOL / N


def gvec2(n=NG, lat=LAT):
    b = 2 * np.pi * np.linalg.inv(lat).T
    m = [np.fft.fftfreq(k, 1.0 / k) for k in n]
    M = np.stack(np.meshgrid(*m, indexing="ij"), -1).reshape(-1, 3)
    g = M @ b
    return np.sum(g * g, 1)


G2 = gvec2()


def hartree(rho):
    k = np.where(G2 < 1e-6, 0.0, 4 * np.pi / np.where(G2 < 1e-6, 1.0, G2))
    return np.real(np.fft.ifftn((np.fft.fftn(rho.reshape(NG)).ravel() * k).reshape(NG))).ravel()


def to_real(c, g):
    cube = np.zeros((c.shape[0], N), dtype=np.complex128)
    cube[:, g] = c
    return np.fft.ifftn(cube.reshape((c.shape[0],) + NG), axes=(1, 2, 3)).reshape(c.shape[0], N)


def density(c, g, occ):
    return np.sum(occ[:, None] * np.abs(to_real(c, g)) ** 2, 0)


def dense_h(v, g):
    """H on the sphere: mask(FFT(V IFFT(scatter c))) + |G|^2/2 c, as a matrix"""
    npw = len(g)
    e = np.eye(npw, dtype=np.complex128)
    vr = to_real(e, g) * v[None, :]
    hv = np.fft.fftn(vr.reshape((npw,) + NG), axes=(1, 2, 3)).reshape(npw, N)[:, g]
    return hv.T + np.diag(0.5 * G2[g])


def dense_etrs_sc(c, g, dt, vion, occ, nscf, thr, half_step_density=False, no_restore=False):
    rho0 = density(c, g, occ)
    h0 = dense_h(vion + hartree(rho0), g)
    half = (sla.expm(-0.5j * dt * h0) @ c.T).T
    full = (sla.expm(-1j * dt * h0) @ c.T).T
    rho = density(half if half_step_density else full, g, occ)
    h = dense_h(vion + hartree(rho), g)
    passes = 0
    phi = half
    for iscf in range(nscf):
        phi = (sla.expm(-0.5j * dt * h) @ (phi if no_restore else half).T).T
        old, rho = rho, density(phi, g, occ)
        d = np.sum(np.abs(old - rho)) * DV
        h = dense_h(vion + hartree(rho), g)
        passes += 1
        if d < thr or iscf == nscf - 1:
            break
    return phi, rho, passes


def _system():
    oh = orc.Ham(LAT, NG)
    g = oh.gsphere((0, 0, 0), 7.0)
    rng = np.random.default_rng(4)
    r = np.stack(np.meshgrid(*[np.arange(k) / k for k in NG], indexing="ij"), -1).reshape(-1, 3) @ LAT
    vion = -2.0 * np.exp(-np.sum((r - 2.5) ** 2, 1) / 1.5)
    c = rng.standard_normal((2, len(g))) + 1j * rng.standard_normal((2, len(g)))
    c *= np.exp(-0.5 * G2[g] / 4.0)[None, :]  # smooth states: low kinetic energy
    c /= np.sqrt(VOL / N ** 2 * np.sum(np.abs(c) ** 2, 1))[:, None]
    occ = np.array([2.0, 1.5])
    return oh, g, vion, c, occ


def test_sc_etrs_matches_dense_exact_exponentials():
    oh, g, vion, c, occ = _system()
    dt = 0.01
    for nscf, thr, want_passes in ((5, 0.0, 5), (5, 1e9, 1)):
        oh.set_potential(vion + hartree(density(c, g, occ)))
        phi, rho, passes, delta = oh.etrs_sc_step(c, g, dt, vion, occ, nscf=nscf, threshold=thr)
        ref_phi, ref_rho, ref_passes = dense_etrs_sc(c, g, dt, vion, occ, nscf, thr)
        assert passes == ref_passes == want_passes
        err = np.max(np.abs(phi - ref_phi)) / np.max(np.abs(ref_phi))
        assert err < 1e-7, err
        assert np.max(np.abs(rho - ref_rho)) / np.max(rho) < 1e-7
        # plausible mistakes are different algorithms at this tolerance: with one
        # pass, the half-step density in H(t + dt); with five, passes that do not
        # restart from the half-step states
        if passes == 1:
            wrong, _, _ = dense_etrs_sc(c, g, dt, vion, occ, nscf, thr, half_step_density=True)
        else:
            wrong, _, _ = dense_etrs_sc(c, g, dt, vion, occ, nscf, thr, no_restore=True)
        assert np.max(np.abs(wrong - ref_phi)) / np.max(np.abs(ref_phi)) > 100 * max(err, 1e-9)
