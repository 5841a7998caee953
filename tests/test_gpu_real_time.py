"""GPU parity of the Taylor exponentials, the ETRS step, the per-step
energies/dipole and the hipGraph replay, against the FP64 oracle.

Tolerances: FP64 1e-11 relative on states, 1e-10 on energies; FP32 states
2e-5, energies 1e-5 relative.  The graph test checks every replay against an
eager run of the same steps, with host synchronisation between replays.
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


def _setup(ctx, seed=9, k=(0.0, 0.0, 0.0)):
    sysm = S.small_config(seed=seed, n=(20, 20, 20), natoms=3, nst=6, ecut=6.0)
    gh = gpu_ham(ctx, sysm)
    oh = oracle_ham(sysm)
    b = gh.basis(k, sysm.ecut)
    c = S.random_states(b.npw, sysm.nst, 31, S.sphere_volume_element(sysm))
    return sysm, gh, oh, b, c


@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-11), (np.complex64, 2e-5)])
def test_exponential_and_2_for_1(ctx, dtype, tol):
    sysm, gh, oh, b, c = _setup(ctx)
    say("case: exponential", np.dtype(dtype).name)
    ci = c.astype(dtype).astype(np.complex128)
    d = DeviceArray.from_numpy(ctx, c.astype(dtype))
    rt.exponential(gh, b, d, -0.05j, 4)
    ref = oh.exponential_in_place(ci, b.gindex(), -0.05j, 4)
    assert rel_err(d.numpy(), ref) < tol
    d = DeviceArray.from_numpy(ctx, c.astype(dtype))
    half = rt.exponential_2_for_1(gh, b, d, -0.025j, -0.05j, 4)
    full_ref, half_ref = oh.exponential_2_for_1(ci, b.gindex(), -0.025j, -0.05j, 4)
    assert rel_err(half.numpy(), half_ref) < tol
    assert rel_err(d.numpy(), full_ref) < tol


@pytest.mark.parametrize("dtype,tol,etol", [(np.complex128, 1e-11, 1e-10), (np.complex64, 2e-5, 1e-5)])
def test_etrs_steps_and_records(ctx, dtype, tol, etol):
    sysm, gh, oh, b, c = _setup(ctx)
    say("case: etrs", np.dtype(dtype).name)
    occ = np.full(sysm.nst, 2.0)
    d = DeviceArray.from_numpy(ctx, c.astype(dtype))
    prop = rt.Propagator(gh, [b], [d], [occ], dt=0.05)
    prop.energy()
    prop.step(3, graph=False)
    recs = prop.records()
    assert recs.shape == (4, 8) and list(recs[:, 7]) == [0, 1, 2, 3]
    ref = c.astype(dtype).astype(np.complex128)
    g = b.gindex()
    for step in range(4):
        if step > 0:
            ref = oh.etrs_step(ref, g, 0.05)
        e = oh.energy_sphere(ref, g, occ)
        _, dip = oh.density_dipole_sphere(ref, g, occ)
        want = np.concatenate([e, dip])
        got = recs[step, :7]
        assert np.max(np.abs(got - want) / np.maximum(np.abs(want), 1.0)) < etol, (step, got, want)
    assert rel_err(d.numpy(), ref) < tol


def test_graph_replays_match_eager(ctx):
    sysm, gh, oh, b, c = _setup(ctx, seed=4)
    say("case: graph replay")
    occ = np.full(sysm.nst, 1.0)
    d_eager = DeviceArray.from_numpy(ctx, c)
    d_graph = DeviceArray.from_numpy(ctx, c)
    pe = rt.Propagator(gh, [b], [d_eager], [occ], dt=0.05)
    pg = rt.Propagator(gh, [b], [d_graph], [occ], dt=0.05)
    for rep in range(4):
        pe.step(1, graph=False)
        pg.step(1, graph=True)
        ctx.sync()
        assert np.array_equal(d_eager.numpy(), d_graph.numpy()), rep
    re, rg = pe.records(), pg.records()
    assert np.array_equal(re, rg)


def test_velocity_gauge_kick(ctx):
    """uniform vector potential A (kick.hpp:36-48) enters as k + A in the
    kinetic factors and the KB phases"""
    sysm, gh, oh, b, c = _setup(ctx, seed=5)
    say("case: kick")
    A_cart = np.array([-0.01, 0.0, 0.0])
    A_cov = sysm.lattice @ A_cart
    gh.uniform_vector_potential = A_cov
    d = DeviceArray.from_numpy(ctx, c)
    out = gh(b, d).numpy()
    ref = oh.apply_sphere(c, b.gindex(), A_cov)
    assert rel_err(out, ref) < 1e-11
    gh.uniform_vector_potential = np.zeros(3)
    out0 = gh(b, d).numpy()
    assert rel_err(out0, oh.apply_sphere(c, b.gindex(), (0, 0, 0))) < 1e-11


_C4_REF = {}


def _c4_reference():
    """the Si64 TDDFT workload (72^3, E_cut 30 Ha, 64 atoms, velocity-gauge kick
    A = -0.01 x): the first 8 seeded states propagated 2 ETRS steps by the
    oracle, with the energy/dipole record of those 8 states before and after
    every step.  H does not depend on the density, so every state evolves on
    its own and 8 states of the 128-state block are a valid subset."""
    if not _C4_REF:
        sysm = S.config(4)
        v = S.local_potential(sysm)
        A = sysm.lattice @ np.array([-0.01, 0.0, 0.0])
        from inq_amd import KSHamiltonian
        gh0 = KSHamiltonian(Context(-1), sysm.lattice, sysm.n)
        g = gh0.basis((0, 0, 0), sysm.ecut).gindex()
        c = S.random_states(len(g), sysm.nst, sysm.seed, S.sphere_volume_element(sysm))
        oh = oracle_ham(sysm, v)
        occ = np.full(8, 2.0)
        ref = c[:8].copy()
        recs = []
        for step in range(3):
            if step > 0:
                ref = oh.etrs_step(ref, g, 0.05, k=A)
            e = oh.energy_sphere(ref, g, occ, k=A)
            _, dip = oh.density_dipole_sphere(ref, g, occ)
            recs.append(np.concatenate([e, dip]))
        _C4_REF.update(sysm=sysm, v=v, A=A, c=c, states=ref, recs=np.array(recs))
    return _C4_REF


@pytest.mark.parametrize("dtype,tol,etol", [(np.complex64, 2e-5, 1e-5), (np.complex128, 1e-11, 1e-10)])
def test_c4_si64_block_two_steps(ctx, dtype, tol, etol):
    """two graph-captured ETRS steps of the whole 128-state Si64 block with the
    kick; states 0..7 against the oracle, and the records of an 8-state block
    against the oracle's energy and dipole of the same states"""
    say("case: C4 Si64 72^3 128 states, 2 ETRS steps", np.dtype(dtype).name)
    R = _c4_reference()
    sysm = R["sysm"]
    gh = gpu_ham(ctx, sysm, R["v"])
    gh.uniform_vector_potential = R["A"]
    b = gh.basis((0, 0, 0), sysm.ecut)
    d = DeviceArray.from_numpy(ctx, R["c"].astype(dtype))
    prop = rt.Propagator(gh, [b], [d], [np.full(sysm.nst, 2.0)], dt=0.05)
    prop.step(1, graph=False)
    prop.step(1, graph=True)
    ctx.sync()
    assert rel_err(d.numpy()[:8], R["states"]) < tol
    prop.close()
    d8 = DeviceArray.from_numpy(ctx, np.ascontiguousarray(R["c"][:8]).astype(dtype))
    p8 = rt.Propagator(gh, [b], [d8], [np.full(8, 2.0)], dt=0.05)
    p8.energy()
    p8.step(2, graph=False)
    recs = p8.records()
    assert recs.shape == (3, 8)
    want = R["recs"]
    err = np.max(np.abs(recs[:, :7] - want) / np.maximum(np.abs(want), 1.0))
    assert err < etol, (recs[:, :7], want)
    assert rel_err(d8.numpy(), R["states"]) < tol
    p8.close()


def test_fp64_state_accumulation_reduces_fp32_drift(ctx):
    """FP32 H applications accumulating into complex128 states
    (inqx_prop_set_state_precision): over 40 graph steps the states stay closer
    to the FP64 propagation than with complex64 states, and match the FP64
    propagation to the FP32 application error (2e-6)"""
    sysm, gh, oh, b, c = _setup(ctx, seed=6)
    say("case: FP64 state accumulation, 40 steps")
    occ = np.full(sysm.nst, 2.0)
    runs = {}
    for name, dtype, wide in (("fp64", np.complex128, False), ("fp32", np.complex64, False),
                              ("wide", np.complex64, True)):
        d = DeviceArray.from_numpy(ctx, c.astype(dtype))
        prop = rt.Propagator(gh, [b], [d], [occ], dt=0.05)
        if wide:
            prop.set_state_precision(np.complex128)
        prop.step(40, graph=True)
        ctx.sync()
        runs[name] = d.numpy().astype(np.complex128)
        prop.close()
    e32 = rel_err(runs["fp32"], runs["fp64"])
    ew = rel_err(runs["wide"], runs["fp64"])
    assert ew < e32, (ew, e32)
    assert ew < 2e-6, ew
