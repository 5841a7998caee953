"""Real-time propagation over the inqx C ABI: Taylor exponentials
(src/operations/exponential.hpp), the ETRS step (src/real_time/etrs.hpp) and
the propagate loop with per-step energies and dipole
(src/real_time/propagate.hpp:87-119), captured in a hipGraph."""
import ctypes as C

import numpy as np

from ._lib import check, lib
from .hamiltonian import DeviceArray, precision_of

RECORD_FIELDS = ("kinetic", "eigenvalues", "non_local", "norm", "dipole_x", "dipole_y", "dipole_z", "step")


def exponential(ham, basis, phi, factor, order=4):
    """exponential_in_place: phi <- sum_n (factor H)^n / n! phi."""
    f = complex(factor)
    check(lib().inqx_exponential(ham.h, basis.h, phi.shape[0], precision_of(phi.dtype), f.real, f.imag, int(order),
                                 phi.ptr), ham.ctx.h, "inqx_exponential")
    return phi


def exponential_2_for_1(ham, basis, phi, factor1, factor2, order=4):
    """returns exp(factor1 H) phi and replaces phi by exp(factor2 H) phi."""
    f1, f2 = complex(factor1), complex(factor2)
    out = DeviceArray(ham.ctx, phi.shape, phi.dtype)
    check(lib().inqx_exponential_2_for_1(ham.h, basis.h, phi.shape[0], precision_of(phi.dtype), f1.real, f1.imag,
                                         f2.real, f2.imag, int(order), phi.ptr, out.ptr), ham.ctx.h,
          "inqx_exponential_2_for_1")
    return out


class Propagator:
    """ETRS propagation of several blocks with FP64 per-step records."""

    def __init__(self, ham, bases, phis, occs, dt, ring_capacity=4096):
        self.ham = ham
        self.h = C.c_void_p()
        nb = len(bases)
        self._occ = [np.ascontiguousarray(o, dtype=np.float64) for o in occs]
        bb = (C.c_void_p * nb)(*[b.h.value for b in bases])
        nst = (C.c_int * nb)(*[p.shape[0] for p in phis])
        pp = (C.c_void_p * nb)(*[p.ptr.value for p in phis])
        oc = (C.c_void_p * nb)(*[o.ctypes.data for o in self._occ])
        check(lib().inqx_prop_create(ham.h, nb, bb, nst, precision_of(phis[0].dtype), pp, oc, float(dt),
                                     int(ring_capacity), C.byref(self.h)), ham.ctx.h, "inqx_prop_create")
        self.phis = phis
        self.bases = list(bases)  # captured graphs use the bases' workspaces: keep them alive

    def step(self, nsteps=1, graph=True):
        check(lib().inqx_prop_step(self.h, int(nsteps), int(bool(graph))), self.ham.ctx.h, "inqx_prop_step")

    def energy(self):
        check(lib().inqx_prop_energy(self.h), self.ham.ctx.h, "inqx_prop_energy")

    def records(self, cap=4096):
        n = C.c_int()
        out = np.zeros((cap, 8))
        check(lib().inqx_prop_records(self.h, C.byref(n), out.ctypes.data_as(C.POINTER(C.c_double)), cap),
              self.ham.ctx.h, "inqx_prop_records")
        m = min(n.value, cap)
        return out[:m]

    def set_scf(self, vion, max_passes=5, threshold=5e-5):
        """self-consistent steps (etrs.hpp:22-79) with V = vion + V_H[density]:
        vion is a device float64 array [N] kept alive by the propagator; the
        Hamiltonian's potential must already be that of the initial density
        (KSHamiltonian.ks_potential).  max_passes = 0 returns to fixed-H steps."""
        self._vion = vion
        check(lib().inqx_prop_set_scf(self.h, vion.ptr if vion is not None else None, int(max_passes),
                                      float(threshold)), self.ham.ctx.h, "inqx_prop_set_scf")

    def set_state_precision(self, dtype):
        """np.complex128 on an FP32 propagator: FP64 states accumulate the FP32 H
        applications (the caller's complex64 array gets the rounded state)"""
        check(lib().inqx_prop_set_state_precision(self.h, precision_of(dtype)), self.ham.ctx.h,
              "inqx_prop_set_state_precision")

    def scf_info(self):
        """(SCF passes, integral |drho| dV of the last pass) of the most recent step"""
        n = C.c_int()
        d = C.c_double()
        check(lib().inqx_prop_scf_info(self.h, C.byref(n), C.byref(d)), self.ham.ctx.h, "inqx_prop_scf_info")
        return n.value, d.value

    def close(self):
        if self.h.value:
            lib().inqx_prop_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
