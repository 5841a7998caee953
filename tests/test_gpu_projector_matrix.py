"""Projectors given as values on their sphere points (the matrix argument of
inqx_ham_add_projector: a pseudopotential's radial functions evaluated by the
caller, projector.hpp:38-113 / projector_all.hpp:40-109) against the oracle
built from the same matrix.  The radial parts are not Gaussian
((1 + r) e^{-r^2} for s, r e^{-r} for p) with the declared real harmonics
Y_00 = 1/(2 sqrt pi), Y_1m = sqrt(3/4pi) (y, z, x)/r.  Tolerances as in
tests/test_gpu_hamiltonian.py: FP64 1e-11, FP32 2e-5 relative.
"""
import math

import numpy as np
import pytest

from inq_amd import DeviceArray, KSHamiltonian
from inq_amd import systems as S
from tests.helpers import rel_err, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from inq_amd import Context
    return Context(0)


def _matrix(lattice, dist, rel, nlm):
    r = dist.astype(np.float64)
    cart = rel.astype(np.float64) @ lattice  # contravariant -> cartesian
    m = np.zeros((nlm, len(r)))
    m[0] = (1.0 + r) * np.exp(-r * r) / (2.0 * math.sqrt(math.pi))
    if nlm > 1:
        safe = np.where(r > 0, r, 1.0)
        rad = r * np.exp(-r) * math.sqrt(3.0 / (4.0 * math.pi)) / safe
        for row, comp in zip((1, 2, 3), (1, 2, 0)):  # m = -1, 0, 1 -> y, z, x
            m[row] = np.where(r > 0, rad * cart[:, comp], 0.0)
    return m


@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-11), (np.complex64, 2e-5)])
def test_tabulated_projectors_vs_oracle(ctx, dtype, tol):
    from oracle import oracle as orc
    sysm = S.small_config(seed=7, n=(16, 16, 16), natoms=3, nst=5, ecut=6.0)
    say("case: projector matrix", np.dtype(dtype).name)
    v = S.local_potential(sysm)
    probe = KSHamiltonian(ctx, sysm.lattice, sysm.n)  # to learn the sphere points
    gh = KSHamiltonian(ctx, sysm.lattice, sysm.n)
    gh.set_scalar_potential(v)
    oh = orc.Ham(sysm.lattice, sysm.n)
    oh.set_potential(v)
    for a, (p, D) in enumerate(zip(sysm.positions, sysm.kb_coeffs())):
        probe.add_projector(p, sysm.kb_radius, sysm.nlm, D, width=sysm.kb_width)
        coords, dist, rel = probe.projector_points(a)
        M = _matrix(sysm.lattice, dist, rel, sysm.nlm)
        gh.add_projector(p, sysm.kb_radius, sysm.nlm, D, matrix=M)
        i = oh.add_projector_matrix(p, sysm.kb_radius, sysm.nlm, M, D)
        oc, _, _, om = oh.projector_points(i, sysm.nlm)
        assert np.array_equal(oc, coords)  # same point order on both sides
        assert np.array_equal(om, M)
    b = gh.basis((0.1, -0.2, 0.3), sysm.ecut)
    c = S.random_states(b.npw, sysm.nst, 13, S.sphere_volume_element(sysm))
    d_in = DeviceArray.from_numpy(ctx, c.astype(dtype))
    d_out = DeviceArray(ctx, c.shape, dtype)
    gh.apply(b, d_in, d_out)
    ref = oh.apply_sphere(c.astype(dtype).astype(np.complex128), b.gindex(), (0.1, -0.2, 0.3))
    assert rel_err(d_out.numpy(), ref) < tol
