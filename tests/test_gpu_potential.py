"""GPU parity of inqx_atomic_field_add (the sphere sums of
atomic_potential::local_potential / ionic_density,
src/hamiltonian/atomic_potential.hpp:116-200) against the oracle's brute-force
sum, which tests/test_oracle_potential.py pins by closed forms.  FP64: 1e-12
relative to the largest value (the same linear interpolation, a different
summation order).
"""
import math

import numpy as np
import pytest

from inq_amd import DeviceArray, InqxError, KSHamiltonian
from inq_amd import potential as pot
from oracle import oracle as orc
from tests.helpers import say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from inq_amd import Context
    return Context(0)


def _zeros(ctx, n):
    d = DeviceArray(ctx, (n,), np.float64)
    d.upload(np.zeros(n))
    return d


def _tables(nspecies, nr, dr, seed):
    rng = np.random.default_rng(seed)
    r = np.arange(nr) * dr
    return np.stack([rng.uniform(-3, 3) * np.exp(-r * r / rng.uniform(0.5, 2.0)) + 0.01 * rng.standard_normal(nr)
                     for _ in range(nspecies)])


def test_diamond_two_species_vs_oracle(ctx):
    say("case: atomic_field_add diamond 32^3, 8 atoms, 2 species")
    a = 6.74
    lat = np.eye(3) * a
    n = (32, 32, 32)
    frac = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0],
                     [.25, .25, .25], [.25, .75, .75], [.75, .25, .75], [.75, .75, .25]])
    pos = frac * a + 0.013
    species = np.array([0, 1, 0, 1, 0, 1, 0, 1], dtype=np.int32)
    radius = np.where(species == 0, 2.0, 2.7)
    dr, nr = 0.01, 300
    tab = _tables(2, nr, dr, 7)
    h = KSHamiltonian(ctx, lat, n)
    f = h.atomic_field_add(pos, radius, species, tab, dr, _zeros(ctx, int(np.prod(n)))).numpy()
    ref = orc.atomic_field(lat, n, pos, radius, species, tab, dr, images=1)
    assert np.max(np.abs(f - ref)) < 1e-12 * np.max(np.abs(ref))
    assert np.count_nonzero(f) == np.count_nonzero(ref)


def test_fcc_cell_radius_beyond_the_cell_vs_oracle(ctx):
    say("case: atomic_field_add fcc 16^3, radius > cell")
    a = 4.0
    lat = 0.5 * a * np.array([[0.0, 1.0, 1.0], [1.0, 0.0, 1.0], [1.0, 1.0, 0.0]])
    n = (16, 16, 16)
    pos = np.array([[0.3, -0.4, 1.7], [-2.9, 5.1, 0.2]])  # outside the home cell on purpose
    dr, nr = 0.005, 1001
    tab = _tables(1, nr, dr, 8)
    radius = [4.6, 3.1]
    h = KSHamiltonian(ctx, lat, n)
    f = h.atomic_field_add(pos, radius, 0, tab, dr, _zeros(ctx, int(np.prod(n)))).numpy()
    ref = orc.atomic_field(lat, n, pos, radius, 0, tab, dr, images=4)
    assert np.max(np.abs(f - ref)) < 1e-12 * np.max(np.abs(ref))


def test_ionic_density_charge_and_accumulation(ctx):
    say("case: ionic_density C8 32^3")
    a = 6.74
    lat = np.eye(3) * a
    n = (32, 32, 32)
    frac = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0],
                     [.25, .25, .25], [.25, .75, .75], [.75, .25, .75], [.75, .75, .25]])
    pos = frac * a
    h = KSHamiltonian(ctx, lat, n)
    z = np.full(8, 4.0)
    rho = pot.ionic_density(h, pos, z).numpy()
    dv = a ** 3 / np.prod(n)
    assert abs(rho.sum() * dv + 32.0) < 1e-5  # -sum Z, up to the table interpolation
    # accumulating twice doubles the field exactly, and the sum is reproducible bit for bit
    d = _zeros(ctx, int(np.prod(n)))
    pot.ionic_density(h, pos, z, field=d)
    pot.ionic_density(h, pos, z, field=d)
    assert np.array_equal(d.numpy(), 2.0 * rho)


def test_no_atoms_and_bad_arguments(ctx):
    lat = np.eye(3) * 5.0
    n = (12, 12, 12)
    h = KSHamiltonian(ctx, lat, n)
    d = _zeros(ctx, int(np.prod(n)))
    h.atomic_field_add(np.zeros((0, 3)), [], [], np.zeros((1, 4)), 0.1, d)
    assert not np.any(d.numpy())
    tab = np.ones((1, 11))
    with pytest.raises(InqxError, match="radial table"):
        h.atomic_field_add([[0, 0, 0]], 1.5, 0, tab, 0.1, d)  # radius > (nr - 1) dr
    with pytest.raises(InqxError, match="species"):
        h.atomic_field_add([[0, 0, 0]], 0.5, 1, tab, 0.1, d)
    with pytest.raises(InqxError, match="images"):
        h.atomic_field_add([[0, 0, 0]], 100.0, 0, np.ones((1, 2002)), 0.05, d)
    assert not np.any(d.numpy())
