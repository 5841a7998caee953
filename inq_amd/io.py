"""operations::io over the inqx C ABI: the reference's restart files
(src/operations/io.hpp:41-211, src/systems/electrons.hpp:323-381).

States are band-major DeviceArrays [nst][npw] on the G sphere; fields and
arrays are host numpy float64 arrays.
"""
import os

import numpy as np

from ._lib import DP, check, lib
from .hamiltonian import precision_of


def _path(dirname):
    return os.fsencode(os.fspath(dirname))


def save_states(ham, basis, phi, dirname, first_state=0):
    """io::save(dirname, orbital_set) (io.hpp:168-183): one raw complex128
    real-space file per state, dirname/<ist %010d>_0000000000."""
    check(lib().inqx_states_save(ham.h, basis.h, phi.shape[0], precision_of(phi.dtype), phi.ptr, _path(dirname),
                                 int(first_state)), ham.ctx.h, "inqx_states_save")


def load_states(ham, basis, phi, dirname, first_state=0):
    """io::load(dirname, orbital_set) (io.hpp:187-211), projected on the sphere."""
    check(lib().inqx_states_load(ham.h, basis.h, phi.shape[0], precision_of(phi.dtype), _path(dirname),
                                 int(first_state), phi.ptr), ham.ctx.h, "inqx_states_load")
    return phi


def save_field(ctx, dirname, v):
    """io::save of a real field_set (io.hpp:124-140): v [npoints] or [npoints][ncomp]."""
    a = np.ascontiguousarray(v, dtype=np.float64)
    a2 = a.reshape(a.shape[0], -1)
    check(lib().inqx_field_save(ctx.h, _path(dirname), a2.shape[0], a2.shape[1], a2.ctypes.data_as(DP)), ctx.h,
          "inqx_field_save")


def load_field(ctx, dirname, npoints, ncomp=1):
    out = np.zeros((int(npoints), int(ncomp)))
    check(lib().inqx_field_load(ctx.h, _path(dirname), int(npoints), int(ncomp), out.ctypes.data_as(DP)), ctx.h,
          "inqx_field_load")
    return out


def save_array(ctx, dirname, v):
    """io::save of an array (io.hpp:104-112), e.g. occupations: dirname/array."""
    a = np.ascontiguousarray(v, dtype=np.float64).ravel()
    check(lib().inqx_array_save(ctx.h, _path(dirname), a.size, a.ctypes.data_as(DP)), ctx.h, "inqx_array_save")


def load_array(ctx, dirname, n):
    out = np.zeros(int(n))
    check(lib().inqx_array_load(ctx.h, _path(dirname), int(n), out.ctypes.data_as(DP)), ctx.h, "inqx_array_load")
    return out
