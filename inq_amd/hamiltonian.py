"""Host-side mirror of the reference's Hamiltonian call surface over the inqx
C ABI.

Names follow the reference: ``KSHamiltonian`` is hamiltonian::ks_hamiltonian
(src/hamiltonian/ks_hamiltonian.hpp:35-275), ``Basis`` is the G sphere of one
orbital_set block (src/states/orbital_set.hpp:18-192), and calling the
Hamiltonian on a block of states applies H (ks_hamiltonian.hpp:160-180).
States live in device memory (``DeviceArray``); numpy arrays are copied in and
out explicitly.
"""
import ctypes as C

import numpy as np

from ._lib import INQX_FP32, INQX_FP64, check, lib

_DT = {INQX_FP64: np.complex128, INQX_FP32: np.complex64}


def precision_of(dtype):
    dtype = np.dtype(dtype)
    if dtype == np.complex128:
        return INQX_FP64
    if dtype == np.complex64:
        return INQX_FP32
    raise ValueError(f"unsupported state dtype {dtype}")


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Context:
    """inqx_ctx: one device and one stream (the default stream unless set)."""

    def __init__(self, device=0, stream=None):
        self.h = C.c_void_p()
        check(lib().inqx_ctx_create(int(device), C.c_void_p(stream or 0), C.byref(self.h)), None, "inqx_ctx_create")

    def sync(self):
        check(lib().inqx_sync(self.h), self.h, "inqx_sync")

    def profile(self, on=True):
        check(lib().inqx_prof_enable(self.h, int(on)), self.h, "inqx_prof_enable")

    def profile_reset(self):
        check(lib().inqx_prof_reset(self.h), self.h, "inqx_prof_reset")

    def kernel_times(self):
        """{kernel name: (launches, total ms)} accumulated with HIP events."""
        n = C.c_int()
        check(lib().inqx_prof_count(self.h, C.byref(n)), self.h, "inqx_prof_count")
        out = {}
        buf = C.create_string_buffer(128)
        for i in range(n.value):
            cnt = C.c_long()
            ms = C.c_double()
            check(lib().inqx_prof_get(self.h, i, buf, 128, C.byref(cnt), C.byref(ms)), self.h, "inqx_prof_get")
            out[buf.value.decode()] = (cnt.value, ms.value)
        return out

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib().inqx_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceArray:
    """A device allocation with a numpy shape/dtype."""

    def __init__(self, ctx, shape, dtype):
        self.ctx = ctx
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = C.c_void_p()
        check(lib().inqx_malloc(ctx.h, self.nbytes, C.byref(self.ptr)), ctx.h, "inqx_malloc")

    @classmethod
    def from_numpy(cls, ctx, a):
        a = np.ascontiguousarray(a)
        d = cls(ctx, a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        check(lib().inqx_memcpy_h2d(self.ctx.h, self.ptr, a.ctypes.data_as(C.c_void_p), self.nbytes), self.ctx.h,
              "inqx_memcpy_h2d")

    def numpy(self):
        out = np.empty(self.shape, dtype=self.dtype)
        check(lib().inqx_memcpy_d2h(self.ctx.h, out.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes), self.ctx.h,
              "inqx_memcpy_d2h")
        return out

    def copy_from(self, other):
        assert other.nbytes == self.nbytes
        check(lib().inqx_memcpy_d2d(self.ctx.h, self.ptr, other.ptr, self.nbytes), self.ctx.h, "inqx_memcpy_d2d")

    def free(self):
        if self.ptr.value:
            lib().inqx_free(self.ctx.h, self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Basis:
    """G sphere 1/2|G+k|^2 < ecut of one (k-point, spin) block."""

    def __init__(self, ham, kpoint, ecut, spin=0):
        self.ham = ham
        self.kpoint = np.asarray(kpoint, dtype=np.float64)
        self.ecut = float(ecut)
        self.spin = int(spin)
        self.h = C.c_void_p()
        check(lib().inqx_basis_create(ham.h, _dp(self.kpoint), self.ecut, self.spin, C.byref(self.h)), ham.ctx.h,
              "inqx_basis_create")
        n = C.c_long()
        check(lib().inqx_basis_size(self.h, C.byref(n)), ham.ctx.h, "inqx_basis_size")
        self.npw = n.value

    def gindex(self):
        g = np.zeros(self.npw, dtype=np.int32)
        check(lib().inqx_basis_gindex(self.h, g.ctypes.data_as(C.POINTER(C.c_int))), self.ham.ctx.h, "gindex")
        return g

    def kinetic(self):
        k = np.zeros(self.npw)
        check(lib().inqx_basis_kinetic(self.h, _dp(k)), self.ham.ctx.h, "kinetic")
        return k

    def __del__(self):
        try:
            if self.h.value:
                lib().inqx_basis_destroy(self.h)
                self.h = C.c_void_p()
        except Exception:
            pass


class KSHamiltonian:
    """hamiltonian::ks_hamiltonian on a grid with a local potential and KB projectors."""

    def __init__(self, ctx, lattice, n):
        self.ctx = ctx
        self.lattice = np.ascontiguousarray(lattice, dtype=np.float64).reshape(3, 3)
        self.n = tuple(int(x) for x in n)
        self.h = C.c_void_p()
        nn = (C.c_int * 3)(*self.n)
        check(lib().inqx_ham_create(ctx.h, _dp(self.lattice.ravel()), nn, C.byref(self.h)), ctx.h, "inqx_ham_create")
        self._A = np.zeros(3)

    def set_scalar_potential(self, v, ncomp=1):
        v = np.ascontiguousarray(v, dtype=np.float64)
        check(lib().inqx_ham_set_potential(self.h, _dp(v), int(ncomp)), self.ctx.h, "set_potential")

    @property
    def uniform_vector_potential(self):
        return self._A.copy()

    @uniform_vector_potential.setter
    def uniform_vector_potential(self, a):
        self._A = np.asarray(a, dtype=np.float64).copy()
        check(lib().inqx_ham_set_vector_potential(self.h, _dp(self._A)), self.ctx.h, "set_vector_potential")

    def add_projector(self, center, radius, nlm, kb_coeff, width=None, matrix=None):
        c = np.asarray(center, dtype=np.float64)
        d = np.ascontiguousarray(kb_coeff, dtype=np.float64)
        m = None if matrix is None else np.ascontiguousarray(matrix, dtype=np.float64)
        st = lib().inqx_ham_add_projector(self.h, _dp(c), float(radius), int(nlm),
                                          None if m is None else _dp(m), _dp(d), float(width or 0.0))
        return check(st, self.ctx.h, "add_projector")

    def projector_points(self, i):
        n = C.c_long()
        check(lib().inqx_ham_projector_size(self.h, int(i), C.byref(n)), self.ctx.h, "projector_size")
        coords = np.zeros((n.value, 3), dtype=np.int32)
        dist = np.zeros(n.value, dtype=np.float32)
        rel = np.zeros((n.value, 3), dtype=np.float32)
        check(lib().inqx_ham_projector_points(self.h, int(i), coords.ctypes.data_as(C.POINTER(C.c_int)),
                                              dist.ctypes.data_as(C.POINTER(C.c_float)),
                                              rel.ctypes.data_as(C.POINTER(C.c_float))), self.ctx.h,
              "projector_points")
        return coords, dist, rel

    def basis(self, kpoint, ecut, spin=0):
        return Basis(self, kpoint, ecut, spin)

    def prepare(self, basis, nst, dtype):
        check(lib().inqx_ham_prepare(self.h, basis.h, int(nst), precision_of(dtype)), self.ctx.h, "prepare")

    def apply(self, basis, psi, hpsi):
        """hpsi = H psi for one block (device arrays [nst][npw])."""
        nst = psi.shape[0]
        assert psi.shape == (nst, basis.npw) and hpsi.shape == psi.shape and hpsi.dtype == psi.dtype
        check(lib().inqx_ham_apply(self.h, basis.h, nst, precision_of(psi.dtype), psi.ptr, hpsi.ptr), self.ctx.h,
              "inqx_ham_apply")

    # Fourier-space layouts (inqx.h): band-major sphere [nst][npw], the
    # reference's point-major sphere [npw][nst], the full cube [N][nst]
    SPHERE_BAND, SPHERE_POINT, CUBE_POINT = 0, 1, 2

    def _rows(self, basis, layout):
        return {0: basis.npw, 1: basis.npw, 2: int(np.prod(self.n))}[layout]

    def apply_layout(self, basis, psi, hpsi, layout):
        """hpsi = H psi with both in `layout`: [nst][npw], [npw][nst] or the
        full Fourier cube [N][nst] ([Gx][Gy][Gz][ist])"""
        nst = psi.shape[0] if layout == 0 else psi.shape[1]
        check(lib().inqx_ham_apply_layout(self.h, basis.h, nst, precision_of(psi.dtype), layout, psi.ptr, hpsi.ptr),
              self.ctx.h, "inqx_ham_apply_layout")

    def states_to_sphere(self, basis, src, layout, out=None):
        """states in `layout` -> band-major sphere [nst][npw]"""
        nst = src.shape[0] if layout == 0 else src.shape[1]
        if out is None:
            out = DeviceArray(self.ctx, (nst, basis.npw), src.dtype)
        check(lib().inqx_states_to_sphere(self.h, basis.h, nst, precision_of(src.dtype), layout, src.ptr, out.ptr),
              self.ctx.h, "inqx_states_to_sphere")
        return out

    def states_from_sphere(self, basis, sph, layout, out=None):
        """band-major sphere [nst][npw] -> `layout`"""
        nst = sph.shape[0]
        if out is None:
            shape = (nst, basis.npw) if layout == 0 else (self._rows(basis, layout), nst)
            out = DeviceArray(self.ctx, shape, sph.dtype)
        check(lib().inqx_states_from_sphere(self.h, basis.h, nst, precision_of(sph.dtype), layout, sph.ptr,
                                            out.ptr), self.ctx.h, "inqx_states_from_sphere")
        return out

    def apply_batch(self, bases, psis, hpsis):
        nb = len(bases)
        prec = precision_of(psis[0].dtype)
        bb = (C.c_void_p * nb)(*[b.h.value for b in bases])
        nst = (C.c_int * nb)(*[p.shape[0] for p in psis])
        pp = (C.c_void_p * nb)(*[p.ptr.value for p in psis])
        hp = (C.c_void_p * nb)(*[p.ptr.value for p in hpsis])
        check(lib().inqx_ham_apply_batch(self.h, nb, bb, nst, prec, pp, hp), self.ctx.h, "inqx_ham_apply_batch")

    def to_real(self, basis, psi, layout=0, out=None):
        """transform::to_real of sphere states (transform.hpp:113-182): [nst][N]
        (layout 0) or the reference's [N][nst] (layout 1), normalised by 1/N."""
        nst = psi.shape[0]
        N = int(np.prod(self.n))
        shape = (nst, N) if layout == 0 else (N, nst)
        out = out if out is not None else DeviceArray(self.ctx, shape, psi.dtype)
        check(lib().inqx_sphere_to_real(self.h, basis.h, nst, precision_of(psi.dtype), psi.ptr, int(layout), out.ptr),
              self.ctx.h, "inqx_sphere_to_real")
        return out

    def to_sphere(self, basis, psi_rs, nst, layout=0, out=None, accumulate=False):
        """transform::to_fourier restricted to the sphere (transform.hpp:52-109), unnormalised."""
        out = out if out is not None else DeviceArray(self.ctx, (nst, basis.npw), psi_rs.dtype)
        check(lib().inqx_real_to_sphere(self.h, basis.h, int(nst), precision_of(psi_rs.dtype), psi_rs.ptr, int(layout),
                                        int(bool(accumulate)), out.ptr), self.ctx.h, "inqx_real_to_sphere")
        return out

    def apply_real(self, basis, psi_rs, nst, layout=0, out=None):
        """ks_hamiltonian::operator() on real-space states (ks_hamiltonian.hpp:137-156);
        basis must hold the whole cube."""
        out = out if out is not None else DeviceArray(self.ctx, psi_rs.shape, psi_rs.dtype)
        check(lib().inqx_ham_apply_real(self.h, basis.h, int(nst), precision_of(psi_rs.dtype), psi_rs.ptr, int(layout),
                                        out.ptr), self.ctx.h, "inqx_ham_apply_real")
        return out

    def density_add(self, basis, phi, occ, density=None, ncomp=1, ispin=0):
        """density::calculate_add (density.hpp:24-38): density [N][ncomp] += sum occ |phi(r)|^2."""
        N = int(np.prod(self.n))
        if density is None:
            density = DeviceArray(self.ctx, (N, ncomp), np.float64)
            density.upload(np.zeros((N, ncomp)))
        occ = np.ascontiguousarray(occ, dtype=np.float64)
        check(lib().inqx_density_add(self.h, basis.h, phi.shape[0], precision_of(phi.dtype), phi.ptr, _dp(occ),
                                     int(ncomp), int(ispin), density.ptr), self.ctx.h, "inqx_density_add")
        return density

    def poisson(self, rho, gshift=(0.0, 0.0, 0.0), zeroterm=0.0):
        """solvers::poisson::in_place on densities [n][N] (poisson.hpp:71-110,199-213), in place."""
        g = np.ascontiguousarray(gshift, dtype=np.float64)
        check(lib().inqx_poisson(self.h, rho.shape[0], precision_of(rho.dtype), rho.ptr, _dp(g), float(zeroterm)),
              self.ctx.h, "inqx_poisson")
        return rho

    def hartree(self, density, vh=None, energy=True):
        """Hartree potential and energy of a real density (self_consistency.hpp:115-120):
        density, vh device float64 arrays [N]; returns (vh, E_H or None)."""
        if vh is None:
            vh = DeviceArray(self.ctx, density.shape, np.float64)
        e = C.c_double()
        check(lib().inqx_hartree(self.h, density.ptr, vh.ptr, C.byref(e) if energy else None), self.ctx.h,
              "inqx_hartree")
        return vh, (e.value if energy else None)

    def ks_potential(self, vion, density):
        """update_hamiltonian at the Hartree level (self_consistency.hpp:100-125):
        the potential becomes vion + V_H[total density] on the device; vion [N]
        and density [N][ncomp] are device float64 arrays.  Returns the energies
        (external, Hartree)."""
        e = (C.c_double * 2)()
        check(lib().inqx_ks_potential(self.h, vion.ptr, density.ptr, e), self.ctx.h, "inqx_ks_potential")
        return e[0], e[1]

    def energy(self, bases, phis, occs):
        """energy::calculate (energy.hpp:61-113) of the states of every block, with
        the non-local energy (ks_hamiltonian.hpp:124-133) and the dipole
        (dipole.hpp:24-59): one H application, FP64 reductions.  Returns
        {kinetic, eigenvalues (sum occ eps), non_local, norm, dipole}."""
        nb = len(bases)
        occ = [np.ascontiguousarray(o, dtype=np.float64) for o in occs]
        bb = (C.c_void_p * nb)(*[b.h.value for b in bases])
        nst = (C.c_int * nb)(*[p.shape[0] for p in phis])
        pp = (C.c_void_p * nb)(*[p.ptr.value for p in phis])
        oc = (C.c_void_p * nb)(*[o.ctypes.data for o in occ])
        out = np.zeros(8)
        check(lib().inqx_energy(self.h, nb, bb, nst, precision_of(phis[0].dtype), pp, oc, _dp(out)), self.ctx.h,
              "inqx_energy")
        return {"kinetic": out[0], "eigenvalues": out[1], "non_local": out[2], "norm": out[3],
                "dipole": out[4:7].copy()}

    def atomic_field_add(self, positions, radius, species, tables, dr, field):
        """field[p] += sum_a sum_T f_{species[a]}(|r_p - R_a - T|) within radius[a]
        (the sphere sums of atomic_potential::local_potential / ionic_density,
        atomic_potential.hpp:116-200); tables [nspecies][nr] at r = j dr,
        interpolated linearly; field a device float64 array [N]."""
        pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
        nat = pos.shape[0]
        rad = np.ascontiguousarray(np.broadcast_to(np.asarray(radius, dtype=np.float64), (nat,)))
        spc = np.ascontiguousarray(np.broadcast_to(np.asarray(species, dtype=np.int32), (nat,)))
        tab = np.ascontiguousarray(np.atleast_2d(tables), dtype=np.float64)
        check(lib().inqx_atomic_field_add(self.h, nat, _dp(pos), _dp(rad), spc.ctypes.data_as(C.POINTER(C.c_int)),
                                          tab.shape[0], tab.shape[1], float(dr), _dp(tab), field.ptr), self.ctx.h,
              "inqx_atomic_field_add")
        return field

    def autotune(self, basis, nst, dtype):
        """times the x-pass launch shapes on this block size and keeps the fastest;
        returns (states per work-group, ms per apply)."""
        w = C.c_int()
        ms = C.c_double()
        check(lib().inqx_ham_autotune(self.h, basis.h, int(nst), precision_of(dtype), C.byref(w), C.byref(ms)),
              self.ctx.h, "inqx_ham_autotune")
        return w.value, ms.value

    def autotune_batch(self, bases, nsts, dtype, cache=None):
        """launch-shape search over (x-pass states per work-group, plane-pass planes
        per work-group) for these blocks, through the on-disk cache file `cache`
        (None: no cache); returns ((xwaves, plane batch), ms per launch sequence)."""
        nb = len(bases)
        bb = (C.c_void_p * nb)(*[b.h.value for b in bases])
        ns = (C.c_int * nb)(*[int(n) for n in nsts])
        shape = (C.c_int * 2)()
        ms = C.c_double()
        check(lib().inqx_ham_autotune_batch(self.h, nb, bb, ns, precision_of(dtype),
                                            None if cache is None else str(cache).encode(), shape, C.byref(ms)),
              self.ctx.h, "inqx_ham_autotune_batch")
        return (shape[0], shape[1]), ms.value

    def set_plane_batch(self, dtype, planes):
        check(lib().inqx_ham_set_plane_batch(self.h, precision_of(dtype), int(planes)), self.ctx.h,
              "set_plane_batch")

    def set_launch_shape(self, dtype, xwaves):
        check(lib().inqx_ham_set_launch_shape(self.h, precision_of(dtype), int(xwaves)), self.ctx.h,
              "set_launch_shape")

    def set_exchange_workspace(self, nbytes):
        check(lib().inqx_ham_set_exchange_workspace(self.h, int(nbytes)), self.ctx.h, "set_exchange_workspace")

    def __call__(self, basis, psi):
        """ks_hamiltonian::operator() on a block of sphere states: returns a new DeviceArray."""
        out = DeviceArray(self.ctx, psi.shape, psi.dtype)
        self.apply(basis, psi, out)
        return out

    def __del__(self):
        try:
            if self.h.value:
                lib().inqx_ham_destroy(self.h)
                self.h = C.c_void_p()
        except Exception:
            pass
