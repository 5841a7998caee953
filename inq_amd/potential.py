"""The local-potential step before H (hamiltonian::atomic_potential):
atomic-sphere sums on the device through inqx_atomic_field_add.

local_potential  sum over atoms of short-range radial functions, each on its
                 atom's sphere (atomic_potential.hpp:116-165)
ionic_density    sum over atoms of charge x long-range density on the sphere of
                 the range separation (atomic_potential.hpp:170-200)

The reference takes both radial functions from its pseudopotential library
(pseudo::math::erf_range_separation with sigma = 0.625, atomic_potential.hpp:58,
and the short-range splines of each pseudopotential); that library is not part
of the reference tree, so its constants are not reproduced here.  This module
uses a declared convention instead: the long-range charge of an atom of
valence charge Z is a normalised Gaussian of width sigma,
    rho(r) = -Z exp(-r^2 / (2 sigma^2)) / (2 pi sigma^2)^(3/2),
whose potential is -Z erf(r / (sqrt(2) sigma)) / r, and it is truncated where
the Gaussian falls below 1e-14 of its peak.  Callers with the pseudopotential's
own tables pass them to local_potential or straight to
KSHamiltonian.atomic_field_add.
"""
import math

import numpy as np

from .hamiltonian import DeviceArray

GAUSS_CUT = 1e-14


def _zeros(ham):
    npts = int(np.prod(ham.n))
    field = DeviceArray(ham.ctx, (npts,), np.float64)
    field.upload(np.zeros(npts))
    return field


def gaussian_radius(sigma):
    """radius where exp(-r^2 / 2 sigma^2) = GAUSS_CUT"""
    return sigma * math.sqrt(-2.0 * math.log(GAUSS_CUT))


def long_range_density(r, sigma):
    """unit-charge Gaussian density of the declared convention (negative: electrons' sign)"""
    r = np.asarray(r, dtype=np.float64)
    return -np.exp(-0.5 * (r / sigma) ** 2) / (2.0 * math.pi * sigma * sigma) ** 1.5


def long_range_potential(r, sigma):
    """potential of long_range_density: -erf(r / (sqrt 2 sigma)) / r, -sqrt(2/pi)/sigma at r = 0"""
    from scipy.special import erf
    r = np.asarray(r, dtype=np.float64)
    out = np.empty_like(r)
    small = r < 1e-12
    out[small] = -math.sqrt(2.0 / math.pi) / sigma
    out[~small] = -erf(r[~small] / (math.sqrt(2.0) * sigma)) / r[~small]
    return out


def ionic_density(ham, positions, charges, sigma=0.625, field=None, dr=None):
    """sum_a Z_a rho_sigma(|r - R_a|) on the grid (atomic_potential.hpp:170-200), a
    device float64 array [N]; added to field when given."""
    pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
    z = np.ascontiguousarray(np.broadcast_to(np.asarray(charges, dtype=np.float64), (pos.shape[0],)))
    rc = gaussian_radius(sigma)
    # linear interpolation shifts the integrated charge by about (dr^2 / 12) int f'' dV:
    # 4e-8 relative at sigma / 2000
    dr = dr if dr is not None else sigma / 2000.0
    nr = int(math.ceil(rc / dr)) + 2
    r = np.arange(nr) * dr
    species, inv = np.unique(z, return_inverse=True)
    tables = np.stack([zz * long_range_density(r, sigma) for zz in species])
    if field is None:
        field = _zeros(ham)
    return ham.atomic_field_add(pos, rc, inv.astype(np.int32), tables, dr, field)


def local_potential(ham, positions, species, tables, dr, radius, field=None):
    """sum_a f_{species[a]}(|r - R_a|) within radius[a] (atomic_potential.hpp:116-165):
    tables [nspecies][nr] of the short-range radial potentials at r = j dr."""
    if field is None:
        field = _zeros(ham)
    return ham.atomic_field_add(positions, radius, species, tables, dr, field)
