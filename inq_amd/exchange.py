"""Exact exchange over the inqx C ABI: hamiltonian::exchange_operator
(src/hamiltonian/exchange_operator.hpp) with states on the G sphere of a
(k-point, spin) block.  ``update`` builds the ACE projectors xi
(exchange_operator.hpp:52-120), ``ace`` applies K psi = -xi (xi^H psi)
(:217-231) and ``direct`` runs the pair-density Poisson solves (:124-186).
The ``*_kpts`` forms take the occupied states of several k-points, each pair
solve shifted by -k + q (exchange_operator.hpp:151)."""
import ctypes as C

import numpy as np

from ._lib import check, lib
from .hamiltonian import DeviceArray, precision_of


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def singularity_correction(ctx, lattice, kpoints=((0.0, 0.0, 0.0),), weights=(1.0,)):
    """hamiltonian::singularity_correction (singularity_correction.hpp:19-103):
    returns (fk [nk], fzero) for covariant k-points and their weights."""
    lat = np.ascontiguousarray(lattice, dtype=np.float64).reshape(9)
    k = np.ascontiguousarray(kpoints, dtype=np.float64).reshape(-1, 3)
    w = np.ascontiguousarray(weights, dtype=np.float64).reshape(-1)
    fk = np.zeros(len(k))
    fz = C.c_double()
    check(lib().inqx_singularity_correction(ctx.h, _dp(lat), len(k), _dp(k), _dp(w), _dp(fk), C.byref(fz)), ctx.h,
          "inqx_singularity_correction")
    return fk, fz.value


def singularity_zeroterm(ctx, lattice, kpoints=((0.0, 0.0, 0.0),), weights=(1.0,), ik=0):
    """the G + q = 0 value of the exchange Poisson solves for the hf states of
    k-point ik: -nk V (fk[ik] - fzero) (singularity_correction.hpp:100-102)"""
    fk, fz = singularity_correction(ctx, lattice, kpoints, weights)
    vol = abs(np.linalg.det(np.asarray(lattice, dtype=np.float64).reshape(3, 3)))
    return -len(fk) * vol * (fk[ik] - fz)


class ExchangeOperator:
    def __init__(self, ham, basis, exchange_coefficient=1.0, zeroterm=0.0):
        """zeroterm: the G = 0 value of the pair Poisson solves, or "singularity"
        for the reference's Carrier et al. correction of a Gamma-only k-point set
        (exchange_operator.hpp:151 passes sing_(idx) to poisson::in_place)"""
        self.ham = ham
        self.basis = basis
        self.coefficient = float(exchange_coefficient)
        if isinstance(zeroterm, str):
            if zeroterm != "singularity":
                raise ValueError(f"unknown zeroterm {zeroterm!r}")
            zeroterm = singularity_zeroterm(ham.ctx, ham.lattice)
        self.zeroterm = float(zeroterm)
        self.xi = None
        self.energy = None

    @property
    def vol_el(self):
        N = float(np.prod(self.ham.n))
        return abs(np.linalg.det(self.ham.lattice)) / (N * N)

    def direct(self, phi, hf, occ, scale=1.0, exxphi=None):
        """exxphi += mask(FFT(sum_j -scale alpha/2 occ_j hf_j Poisson(conj(hf_j) phi)))."""
        occ = np.ascontiguousarray(occ, dtype=np.float64)
        if exxphi is None:
            exxphi = DeviceArray(self.ham.ctx, phi.shape, phi.dtype)
            exxphi.upload(np.zeros(phi.shape, phi.dtype))
        check(lib().inqx_exchange_direct(self.ham.h, self.basis.h, phi.shape[0], precision_of(phi.dtype), phi.ptr,
                                          hf.shape[0], hf.ptr, _dp(occ), self.coefficient, float(scale),
                                          self.zeroterm, exxphi.ptr), self.ham.ctx.h, "inqx_exchange_direct")
        return exxphi

    @staticmethod
    def _hf_arrays(hf_blocks):
        """hf_blocks: [(basis, hf DeviceArray, occupations, zeroterm)] -> ctypes arrays"""
        n = len(hf_blocks)
        bases = (C.c_void_p * n)(*[b.h for b, _, _, _ in hf_blocks])
        nhf = np.ascontiguousarray([h.shape[0] for _, h, _, _ in hf_blocks], dtype=np.int32)
        ptrs = (C.c_void_p * n)(*[h.ptr for _, h, _, _ in hf_blocks])
        occs = [np.ascontiguousarray(o, dtype=np.float64) for _, _, o, _ in hf_blocks]
        optr = (C.c_void_p * n)(*[o.ctypes.data for o in occs])
        zt = np.ascontiguousarray([z for _, _, _, z in hf_blocks], dtype=np.float64)
        return n, bases, nhf, ptrs, occs, optr, zt

    def direct_kpts(self, phi, hf_blocks, scale=1.0, exxphi=None):
        """exchange with the occupied states of several k-points:
        hf_blocks = [(basis, hf, occ, zeroterm)], each pair solve shifted by
        -k + q of its block"""
        if exxphi is None:
            exxphi = DeviceArray(self.ham.ctx, phi.shape, phi.dtype)
            exxphi.upload(np.zeros(phi.shape, phi.dtype))
        n, bases, nhf, ptrs, occs, optr, zt = self._hf_arrays(hf_blocks)
        check(lib().inqx_exchange_direct_kpts(self.ham.h, self.basis.h, phi.shape[0], precision_of(phi.dtype),
                                              phi.ptr, n, bases, nhf.ctypes.data_as(C.POINTER(C.c_int)), ptrs, optr,
                                              _dp(zt), self.coefficient, float(scale), exxphi.ptr), self.ham.ctx.h,
              "inqx_exchange_direct_kpts")
        return exxphi

    def update_kpts(self, phi, occ, hf_blocks):
        """ACE build of this block against the occupied states of every k-point
        (hf_blocks as in direct_kpts); returns this block's exchange energy"""
        occ = np.ascontiguousarray(occ, dtype=np.float64)
        if self.xi is None or self.xi.shape != phi.shape or self.xi.dtype != phi.dtype:
            self.xi = DeviceArray(self.ham.ctx, phi.shape, phi.dtype)
        n, bases, nhf, ptrs, occs, optr, zt = self._hf_arrays(hf_blocks)
        en = C.c_double()
        check(lib().inqx_ace_build_kpts(self.ham.h, self.basis.h, phi.shape[0], precision_of(phi.dtype), phi.ptr,
                                        _dp(occ), n, bases, nhf.ctypes.data_as(C.POINTER(C.c_int)), ptrs, optr,
                                        _dp(zt), self.coefficient, self.xi.ptr, C.byref(en)), self.ham.ctx.h,
              "inqx_ace_build_kpts")
        self.energy = en.value
        return self.energy

    def update(self, phi, occ):
        """ACE build from the states phi; returns the exchange energy."""
        occ = np.ascontiguousarray(occ, dtype=np.float64)
        if self.xi is None or self.xi.shape != phi.shape or self.xi.dtype != phi.dtype:
            self.xi = DeviceArray(self.ham.ctx, phi.shape, phi.dtype)
        en = C.c_double()
        check(lib().inqx_ace_build(self.ham.h, self.basis.h, phi.shape[0], precision_of(phi.dtype), phi.ptr, _dp(occ),
                                   self.coefficient, self.zeroterm, self.xi.ptr, C.byref(en)), self.ham.ctx.h,
              "inqx_ace_build")
        self.energy = en.value
        return self.energy

    def ace(self, psi, exxphi):
        """exxphi <- exxphi - xi (vol_el xi^H psi)."""
        assert self.xi is not None, "update() first"
        check(lib().inqx_ace_apply(self.ham.ctx.h, precision_of(psi.dtype), self.xi.shape[0], psi.shape[0],
                                   psi.shape[1], self.vol_el, self.xi.ptr, psi.ptr, exxphi.ptr), self.ham.ctx.h,
              "inqx_ace_apply")
        return exxphi

    def __call__(self, psi):
        out = DeviceArray(self.ham.ctx, psi.shape, psi.dtype)
        out.upload(np.zeros(psi.shape, psi.dtype))
        return self.ace(psi, out)
