"""Hartree potential and energy through the C ABI (poisson::solve of a real
field, src/solvers/poisson.hpp:217-223, and the Hartree term of
self_consistency::update_hamiltonian, src/hamiltonian/self_consistency.hpp:115-120).

Checkers: the closed form of a real plane wave (the reference's own "Real plane
wave" case, poisson.hpp:355-390: V = 4 pi / |G|^2 rho, E_H = 1/2 V/N sum rho V),
a constant density (its G = 0 term is dropped: V = 0), and the FP64 oracle's
Poisson solve on a random density.  Tolerance 1e-12 relative (FP64).
"""
import numpy as np
import pytest

from inq_amd import Context, DeviceArray, KSHamiltonian
from tests.helpers import rel_err, say

pytestmark = pytest.mark.gpu

LAT = np.array([[7.0, 0.0, 0.0], [0.3, 6.5, 0.0], [-0.2, 0.4, 8.1]])
N = (24, 24, 24)


@pytest.fixture(scope="module")
def ham():
    ctx = Context(0)
    return KSHamiltonian(ctx, LAT, N)


def _points():
    i = np.indices(N).reshape(3, -1).T.astype(float)
    return (i / np.array(N)) @ LAT  # cartesian r of every point, order (i0, i1, i2)


def test_real_plane_wave_closed_form(ham):
    say("case: hartree plane wave")
    b = 2.0 * np.pi * np.linalg.inv(LAT).T  # rows: reciprocal vectors
    g = 2 * b[0] - b[1] + 3 * b[2]
    rho = np.cos(_points() @ g)
    d = DeviceArray.from_numpy(ham.ctx, rho)
    vh, e = ham.hartree(d)
    ref = 4.0 * np.pi / (g @ g) * rho
    assert rel_err(vh.numpy(), ref) < 1e-12
    vol = abs(np.linalg.det(LAT))
    e_ref = 0.5 * vol / rho.size * np.sum(rho * ref)
    assert abs(e - e_ref) < 1e-12 * abs(e_ref)


def test_constant_density_has_no_potential(ham):
    say("case: hartree constant")
    d = DeviceArray.from_numpy(ham.ctx, np.full(int(np.prod(N)), 0.37))
    vh, e = ham.hartree(d)
    assert np.max(np.abs(vh.numpy())) < 1e-14 and abs(e) < 1e-12


def test_random_density_vs_oracle(ham):
    from oracle import oracle as orc
    say("case: hartree random vs oracle")
    rho = np.random.default_rng(7).random(int(np.prod(N)))
    d = DeviceArray.from_numpy(ham.ctx, rho)
    vh, e = ham.hartree(d)
    oh = orc.Ham(LAT, N)
    ref = oh.poisson_in_place(rho.astype(np.complex128).reshape(-1, 1)).real.ravel()
    assert rel_err(vh.numpy(), ref) < 1e-12
    vol = abs(np.linalg.det(LAT))
    assert abs(e - 0.5 * vol / rho.size * np.sum(rho * ref)) < 1e-12 * abs(e)
