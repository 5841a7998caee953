"""The Hartree-level Kohn-Sham potential update (update_hamiltonian,
self_consistency.hpp:100-125) and the self-consistent ETRS step
(etrs.hpp:22-79) on the device.

* ks_potential: the Hamiltonian with V = vion + V_H[density] set on the device
  applies exactly as one whose potential was set from the host with V_H from
  numpy FFTs (FP64 1e-12); the energies equal integral rho vion and 1/2
  integral rho V_H (closed form for a real plane-wave density: V_H =
  4 pi/|G|^2 rho); a two-component spin density acts through its total.
* self-consistent steps against the oracle's restatement (itself pinned by a
  dense exact-exponential calculation): states 1e-10 (FP64) / 2e-5 (FP32),
  records 1e-10 / 1e-5 relative, and the same number of SCF passes.
"""
import numpy as np
import pytest

from inq_amd import Context, DeviceArray
from inq_amd import real_time as rt
from inq_amd import systems as S
from tests.helpers import gpu_ham, oracle_ham, rel_err, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return Context(0)


def _hartree_np(rho, lat, n):
    b = 2 * np.pi * np.linalg.inv(lat).T
    m = [np.fft.fftfreq(k, 1.0 / k) for k in n]
    M = np.stack(np.meshgrid(*m, indexing="ij"), -1).reshape(-1, 3)
    g2 = np.sum((M @ b) ** 2, 1)
    k = np.where(g2 < 1e-6, 0.0, 4 * np.pi / np.where(g2 < 1e-6, 1.0, g2))
    return np.real(np.fft.ifftn((np.fft.fftn(rho.reshape(n)).ravel() * k).reshape(n))).ravel()


def test_ks_potential_matches_host_potential(ctx):
    say("case: ks_potential 20^3")
    sysm = S.small_config(seed=11, n=(20, 20, 20), natoms=2, nst=3, ecut=6.0)
    lat, n, N = sysm.lattice, sysm.n, sysm.N
    vion = S.local_potential(sysm)
    r = S.rgrid(lat, n)
    b = 2 * np.pi * np.linalg.inv(lat).T
    G = 2 * b[0] - b[1] + b[2]
    rho = 0.3 + 0.1 * np.cos(r @ G)
    dv = abs(np.linalg.det(lat)) / N
    g1 = gpu_ham(ctx, sysm, vion)
    e_ext, e_h = g1.ks_potential(DeviceArray.from_numpy(ctx, vion), DeviceArray.from_numpy(ctx, rho))
    vh = 4 * np.pi / (G @ G) * 0.1 * np.cos(r @ G)  # the constant part is dropped (G = 0)
    assert e_ext == pytest.approx(np.sum(rho * vion) * dv, rel=1e-12)
    assert e_h == pytest.approx(0.5 * np.sum(rho * vh) * dv, rel=1e-12)
    g2 = gpu_ham(ctx, sysm, vion + _hartree_np(rho, lat, n))
    bs1, bs2 = g1.basis((0.1, 0.0, -0.2), sysm.ecut), g2.basis((0.1, 0.0, -0.2), sysm.ecut)
    c = S.random_states(bs1.npw, 3, 4, S.sphere_volume_element(sysm))
    h1 = g1(bs1, DeviceArray.from_numpy(ctx, c)).numpy()
    h2 = g2(bs2, DeviceArray.from_numpy(ctx, c)).numpy()
    assert rel_err(h1, h2) < 1e-12
    # spin density: both components feel the total density's Hartree potential
    g3 = gpu_ham(ctx, sysm, np.repeat(vion[:, None], 2, 1).ravel())
    g3.set_scalar_potential(np.repeat(vion[:, None], 2, 1), ncomp=2)
    spin = np.stack([0.25 * rho, 0.75 * rho], 1)
    g3.ks_potential(DeviceArray.from_numpy(ctx, vion), DeviceArray.from_numpy(ctx, spin))
    for s in (0, 1):
        b3 = g3.basis((0.1, 0.0, -0.2), sysm.ecut, spin=s)
        h3 = g3(b3, DeviceArray.from_numpy(ctx, c)).numpy()
        assert rel_err(h3, h2) < 1e-12


@pytest.mark.parametrize("dtype,tol,etol,thr", [(np.complex128, 1e-10, 1e-10, 5e-5), (np.complex64, 2e-5, 1e-5, 0.0)])
def test_self_consistent_steps_vs_oracle(ctx, dtype, tol, etol, thr):
    say("case: self-consistent ETRS 20^3", np.dtype(dtype).name)
    sysm = S.small_config(seed=12, n=(20, 20, 20), natoms=2, nst=4, ecut=6.0)
    vion = S.local_potential(sysm)
    occ = np.array([2.0, 2.0, 1.0, 0.5])
    gh = gpu_ham(ctx, sysm, vion)
    oh = oracle_ham(sysm, vion)
    b = gh.basis((0.0, 0.0, 0.0), sysm.ecut)
    g = b.gindex()
    c = S.random_states(b.npw, sysm.nst, 21, S.sphere_volume_element(sysm))
    c = c.astype(dtype).astype(np.complex128)
    # potential of the initial density on both sides
    rho0, _ = oh.density_dipole_sphere(c, g, occ)
    d_vion = DeviceArray.from_numpy(ctx, vion)
    gh.ks_potential(d_vion, DeviceArray.from_numpy(ctx, rho0))
    oh.set_potential(vion + _hartree_np(rho0, sysm.lattice, sysm.n))
    d = DeviceArray.from_numpy(ctx, c.astype(dtype))
    prop = rt.Propagator(gh, [b], [d], [occ], dt=0.05)
    prop.set_scf(d_vion, 5, thr)
    with pytest.raises(Exception):
        prop.step(1, graph=True)  # self-consistent steps run eagerly
    ref = c
    for step in range(2):
        prop.step(1, graph=False)
        ref, _, passes, delta = oh.etrs_sc_step(ref, g, 0.05, vion, occ, nscf=5, threshold=thr)
        got_passes, got_delta = prop.scf_info()
        assert got_passes == passes, (step, got_passes, passes, got_delta, delta)
        assert rel_err(d.numpy(), ref) < tol, step
        e = oh.energy_sphere(ref, g, occ)
        _, dip = oh.density_dipole_sphere(ref, g, occ)
        want = np.concatenate([e, dip])
        got = prop.records()[-1, :7]
        assert np.max(np.abs(got - want) / np.maximum(np.abs(want), 1.0)) < etol, (step, got, want)
    prop.close()
