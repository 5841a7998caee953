"""ctypes wrapper of the FP64 CPU oracle (oracle/liborc.so).

TEST INFRASTRUCTURE ONLY.  Only tests/, ``__graft_entry__.smoke()`` and
bench.py's ``cpu_baseline`` leg may import this module, and only to check
results (or, for the baseline, as the timed CPU port).  The product package
``inq_amd`` never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE], stdout=subprocess.DEVNULL)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liborc.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        P = C.c_void_p
        dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
        cp = np.ctypeslib.ndpointer(dtype=np.complex128, flags="C_CONTIGUOUS")
        ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
        fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
        sig = {
            "orc_version": (C.c_int, []),
            "orc_set_threads": (None, [C.c_int]),
            "orc_good_size": (C.c_int, [C.c_int]),
            "orc_fft3d": (None, [ip, C.c_int, cp, C.c_int]),
            "orc_ham_create": (P, [dp, ip]),
            "orc_ham_destroy": (None, [P]),
            "orc_ham_volume": (C.c_double, [P]),
            "orc_ham_set_potential": (None, [P, dp]),
            "orc_sphere_build": (C.c_long, [P, dp, C.c_double, P, P, P, C.c_long]),
            "orc_ham_add_projector_gauss": (C.c_int, [P, dp, C.c_double, C.c_int, C.c_double, dp]),
            "orc_ham_add_projector_matrix": (C.c_int, [P, dp, C.c_double, C.c_int, dp, dp]),
            "orc_ham_num_projectors": (C.c_int, [P]),
            "orc_ham_projector_size": (C.c_long, [P, C.c_int]),
            "orc_ham_projector_points": (None, [P, C.c_int, ip, fp, fp, dp]),
            "orc_ham_apply_rs": (None, [P, C.c_int, dp, cp, cp]),
            "orc_ham_apply_fs": (None, [P, C.c_int, dp, cp, cp]),
            "orc_ham_apply_sphere": (None, [P, C.c_int, dp, C.c_long, ip, cp, cp]),
            "orc_ham_projections": (None, [P, C.c_int, dp, cp, C.c_int, cp]),
            "orc_ham_nonlocal_energy": (C.c_double, [P, C.c_int, dp, cp, dp]),
            "orc_gsphere": (C.c_long, [P, dp, C.c_double, P, C.c_long]),
            "orc_gsphere_kinetic": (None, [P, dp, C.c_long, ip, dp]),
            "orc_sphere_to_cube": (None, [P, C.c_int, C.c_long, ip, cp, cp]),
            "orc_cube_to_sphere": (None, [P, C.c_int, C.c_long, ip, cp, cp]),
            "orc_to_real": (None, [P, C.c_int, cp, cp, C.c_int]),
            "orc_to_fourier": (None, [P, C.c_int, cp, cp]),
            "orc_overlap": (None, [C.c_long, C.c_int, C.c_int, C.c_double, C.c_int, cp, cp, cp]),
            "orc_cholesky": (C.c_int, [C.c_int, cp]),
            "orc_invert_lower": (None, [C.c_int, cp]),
            "orc_rotate_h": (None, [C.c_long, C.c_int, C.c_int, cp, cp]),
            "orc_orthogonalize": (C.c_int, [C.c_long, C.c_int, C.c_double, C.c_int, cp]),
            "orc_eigh": (C.c_int, [C.c_int, cp, dp]),
            "orc_subspace_diag_sphere": (C.c_int, [P, C.c_int, dp, C.c_long, ip, cp, dp]),
            "orc_etrs_step": (None, [P, C.c_int, dp, C.c_long, ip, C.c_double, cp]),
            "orc_energy_sphere": (None, [P, C.c_int, dp, C.c_long, ip, cp, dp, dp]),
            "orc_etrs_sc_step": (C.c_int, [P, C.c_int, dp, C.c_long, ip, C.c_double, dp, dp, C.c_int, C.c_double,
                                           cp, dp, dp]),
            "orc_density_dipole_sphere": (None, [P, C.c_int, C.c_long, ip, cp, dp, P, dp]),
            "orc_poisson_in_place": (None, [P, C.c_int, cp, dp, C.c_double]),
            "orc_sing_auxiliary": (C.c_double, [dp, dp]),
            "orc_ylm_real": (C.c_double, [C.c_int, C.c_int, dp]),
            "orc_randomize": (None, [C.c_long, C.c_int, C.c_int, C.c_int, dp]),
            "orc_sing_correction": (None, [dp, C.c_int, dp, dp, dp, dp]),
            "orc_ace_build": (C.c_double, [P, C.c_int, cp, dp, C.c_double, cp]),
            "orc_exchange_direct": (None, [P, C.c_int, cp, C.c_int, cp, dp, C.c_double, C.c_double, cp]),
            "orc_exchange_direct_shift": (None, [P, C.c_int, cp, C.c_int, cp, dp, C.c_double, C.c_double, dp,
                                                 C.c_double, cp]),
            "orc_ace_apply": (None, [P, C.c_int, cp, C.c_int, cp, cp]),
            # refcpu.c: the reference CPU data flow on FFTW3/BLAS/LAPACK (CPU baseline)
            "rc_init": (C.c_int, []),
            "rc_backend": (C.c_int, []),
            "rc_ham_create": (P, [dp, ip, C.c_int, dp, C.c_int, ip, ip, C.c_int, dp, dp, C.c_int]),
            "rc_ham_destroy": (None, [P]),
            "rc_ham_apply": (None, [P, cp, cp]),
            "rc_overlap_cholesky": (C.c_int, [C.c_long, C.c_int, C.c_double, cp, cp]),
            "rc_exchange_pairs": (None, [dp, ip, C.c_int, cp, C.c_int, cp, dp, cp, C.c_int]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        # complex-by-value signatures
        class _C(C.Structure):
            _fields_ = [("re", C.c_double), ("im", C.c_double)]
        L._cplx = _C
        for name, args in {
            "orc_exponential_in_place": [P, C.c_int, dp, C.c_long, ip, _C, C.c_int, cp],
            "orc_exponential_2_for_1": [P, C.c_int, dp, C.c_long, ip, _C, _C, C.c_int, cp, cp],
            "orc_rotate5": [C.c_long, C.c_int, C.c_int, cp, cp, cp, _C, _C],
        }.items():
            f = getattr(L, name)
            f.restype = None
            f.argtypes = args
        _LIB = L
    return _LIB


def _c(z):
    z = complex(z)
    return lib()._cplx(z.real, z.imag)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def set_threads(n):
    lib().orc_set_threads(int(n))


def fft3d(n, data, sign):
    """In-place batched 3-D FFT of a [n0,n1,n2,nst] complex128 array."""
    data = _c128(data)
    nst = data.size // (n[0] * n[1] * n[2])
    lib().orc_fft3d(np.asarray(n, dtype=np.int32), int(nst), data, int(sign))
    return data


class Ham:
    """Oracle Hamiltonian: grid + local potential + KB projectors."""

    def __init__(self, lattice, n):
        self.lattice = _f64(lattice).reshape(3, 3)
        self.n = tuple(int(x) for x in n)
        self.N = self.n[0] * self.n[1] * self.n[2]
        self.h = lib().orc_ham_create(self.lattice.ravel(), np.asarray(self.n, dtype=np.int32))
        self.projs = []
        self._proj_meta = []

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.orc_ham_destroy(self.h)
            self.h = None

    @property
    def volume(self):
        return lib().orc_ham_volume(self.h)

    def set_potential(self, v):
        self.v = _f64(v).ravel()
        lib().orc_ham_set_potential(self.h, self.v)

    def sphere(self, center, radius):
        L = lib()
        c = _f64(center)
        cnt = L.orc_sphere_build(self.h, c, float(radius), None, None, None, 0)
        coords = np.zeros((cnt, 3), dtype=np.int32)
        dist = np.zeros(cnt, dtype=np.float32)
        rel = np.zeros((cnt, 3), dtype=np.float32)
        L.orc_sphere_build(self.h, c, float(radius), coords.ctypes.data, dist.ctypes.data, rel.ctypes.data, cnt)
        return coords, dist, rel

    def add_projector_gauss(self, center, radius, nlm, width, coeff):
        i = lib().orc_ham_add_projector_gauss(self.h, _f64(center), float(radius), int(nlm), float(width),
                                              _f64(coeff).ravel())
        self.projs.append(i)
        self._proj_meta.append((int(nlm), _f64(coeff).ravel().copy()))
        return i

    def add_projector_matrix(self, center, radius, nlm, matrix, coeff):
        i = lib().orc_ham_add_projector_matrix(self.h, _f64(center), float(radius), int(nlm),
                                               _f64(matrix).ravel(), _f64(coeff).ravel())
        self.projs.append(i)
        self._proj_meta.append((int(nlm), _f64(coeff).ravel().copy()))
        return i

    def projector_points(self, i, nlm):
        L = lib()
        n = L.orc_ham_projector_size(self.h, i)
        coords = np.zeros((n, 3), dtype=np.int32)
        dist = np.zeros(n, dtype=np.float32)
        rel = np.zeros((n, 3), dtype=np.float32)
        mat = np.zeros((nlm, n), dtype=np.float64)
        L.orc_ham_projector_points(self.h, i, coords, dist, rel, mat)
        return coords, dist, rel, mat

    # ---- H ----
    def apply_rs(self, phi, k=(0, 0, 0)):
        phi = _c128(phi)
        nst = phi.size // self.N
        out = np.zeros_like(phi)
        lib().orc_ham_apply_rs(self.h, nst, _f64(k), phi, out)
        return out

    def apply_fs(self, phi, k=(0, 0, 0)):
        phi = _c128(phi)
        nst = phi.size // self.N
        out = np.zeros_like(phi)
        lib().orc_ham_apply_fs(self.h, nst, _f64(k), phi, out)
        return out

    def gsphere(self, k, ecut):
        L = lib()
        kk = _f64(k)
        cnt = L.orc_gsphere(self.h, kk, float(ecut), None, 0)
        g = np.zeros(cnt, dtype=np.int32)
        L.orc_gsphere(self.h, kk, float(ecut), g.ctypes.data, cnt)
        return g

    def gsphere_kinetic(self, k, gidx):
        out = np.zeros(len(gidx), dtype=np.float64)
        lib().orc_gsphere_kinetic(self.h, _f64(k), len(gidx), np.ascontiguousarray(gidx, np.int32), out)
        return out

    def apply_sphere(self, phi, gidx, k=(0, 0, 0)):
        phi = _c128(phi)
        npw = len(gidx)
        nst = phi.size // npw
        out = np.zeros_like(phi)
        lib().orc_ham_apply_sphere(self.h, nst, _f64(k), npw, np.ascontiguousarray(gidx, np.int32), phi, out)
        return out

    def projections(self, phi_rs, k=(0, 0, 0), apply_coeff=True, maxnlm=4):
        phi_rs = _c128(phi_rs)
        nst = phi_rs.size // self.N
        out = np.zeros((len(self.projs), maxnlm, nst), dtype=np.complex128)
        lib().orc_ham_projections(self.h, nst, _f64(k), phi_rs, int(apply_coeff), out)
        return out

    def nonlocal_energy(self, phi_rs, occ, k=(0, 0, 0)):
        phi_rs = _c128(phi_rs)
        nst = phi_rs.size // self.N
        return lib().orc_ham_nonlocal_energy(self.h, nst, _f64(k), phi_rs, _f64(occ))

    def to_real(self, fs, normalize=True):
        fs = _c128(fs)
        out = np.zeros_like(fs)
        lib().orc_to_real(self.h, fs.size // self.N, fs, out, int(normalize))
        return out

    def to_fourier(self, rs):
        rs = _c128(rs)
        out = np.zeros_like(rs)
        lib().orc_to_fourier(self.h, rs.size // self.N, rs, out)
        return out

    def sphere_to_cube(self, phi, gidx):
        phi = _c128(phi)
        npw = len(gidx)
        nst = phi.size // npw
        out = np.zeros((self.N, nst), dtype=np.complex128)
        lib().orc_sphere_to_cube(self.h, nst, npw, np.ascontiguousarray(gidx, np.int32), phi, out)
        return out

    def cube_to_sphere(self, cube, gidx):
        cube = _c128(cube)
        nst = cube.size // self.N
        out = np.zeros((nst, len(gidx)), dtype=np.complex128)
        lib().orc_cube_to_sphere(self.h, nst, len(gidx), np.ascontiguousarray(gidx, np.int32), cube, out)
        return out

    # ---- propagation / energies on the sphere ----
    def subspace_diag_sphere(self, phi, gidx, k=(0, 0, 0)):
        phi = _c128(phi).copy()
        npw = len(gidx)
        nst = phi.size // npw
        ev = np.zeros(nst)
        lib().orc_subspace_diag_sphere(self.h, nst, _f64(k), npw, np.ascontiguousarray(gidx, np.int32), phi, ev)
        return phi, ev

    def exponential_in_place(self, phi, gidx, factor, order=4, k=(0, 0, 0)):
        phi = _c128(phi).copy()
        npw = len(gidx)
        lib().orc_exponential_in_place(self.h, phi.size // npw, _f64(k), npw, np.ascontiguousarray(gidx, np.int32),
                                       _c(factor), int(order), phi)
        return phi

    def exponential_2_for_1(self, phi, gidx, f1, f2, order=4, k=(0, 0, 0)):
        phi = _c128(phi).copy()
        npw = len(gidx)
        exp = np.zeros_like(phi)
        lib().orc_exponential_2_for_1(self.h, phi.size // npw, _f64(k), npw, np.ascontiguousarray(gidx, np.int32),
                                      _c(f1), _c(f2), int(order), phi, exp)
        return phi, exp

    def etrs_step(self, phi, gidx, dt, k=(0, 0, 0)):
        phi = _c128(phi).copy()
        npw = len(gidx)
        lib().orc_etrs_step(self.h, phi.size // npw, _f64(k), npw, np.ascontiguousarray(gidx, np.int32),
                            float(dt), phi)
        return phi

    def etrs_sc_step(self, phi, gidx, dt, vion, occ, nscf=5, threshold=5e-5, k=(0, 0, 0)):
        """self-consistent ETRS step with V = vion + V_H[rho] (the potential held
        by this oracle Hamiltonian is replaced); returns (phi, rho, passes, delta)"""
        phi = _c128(phi).copy()
        npw = len(gidx)
        rho = np.zeros(self.N)
        delta = np.zeros(1)
        vion = _f64(vion).ravel()
        n = lib().orc_etrs_sc_step(self.h, phi.size // npw, _f64(k), npw, np.ascontiguousarray(gidx, np.int32),
                                   float(dt), vion, _f64(occ), int(nscf), float(threshold), phi, rho, delta)
        return phi, rho, n, float(delta[0])

    def energy_sphere(self, phi, gidx, occ, k=(0, 0, 0)):
        phi = _c128(phi)
        npw = len(gidx)
        out = np.zeros(4)
        lib().orc_energy_sphere(self.h, phi.size // npw, _f64(k), npw, np.ascontiguousarray(gidx, np.int32), phi,
                                _f64(occ), out)
        return out

    def density_dipole_sphere(self, phi, gidx, occ):
        phi = _c128(phi)
        npw = len(gidx)
        rho = np.zeros(self.N)
        dip = np.zeros(3)
        lib().orc_density_dipole_sphere(self.h, phi.size // npw, npw, np.ascontiguousarray(gidx, np.int32), phi,
                                        _f64(occ), rho.ctypes.data, dip)
        return rho, dip

    def poisson_in_place(self, rho, gshift=(0, 0, 0), zeroterm=0.0):
        rho = _c128(rho).copy()
        lib().orc_poisson_in_place(self.h, rho.size // self.N, rho, _f64(gshift), float(zeroterm))
        return rho

    def ace_build(self, phi_rs, occ, exx_coeff=1.0):
        phi_rs = _c128(phi_rs)
        nst = phi_rs.size // self.N
        xi = np.zeros_like(phi_rs)
        ex = lib().orc_ace_build(self.h, nst, phi_rs, _f64(occ), float(exx_coeff), xi)
        return xi, ex

    def exchange_direct(self, phi_rs, hf_rs, occ, exx_coeff=1.0, scale=1.0, gshift=None, zeroterm=0.0):
        """exchange of phi (real space [N][nst]) with the hf states of one
        k-point; gshift: cartesian -k + q of the pair Poisson solves (None: Gamma)"""
        phi_rs = _c128(phi_rs)
        hf_rs = _c128(hf_rs)
        nst = phi_rs.size // self.N
        nhf = hf_rs.size // self.N
        out = np.zeros_like(phi_rs)
        if gshift is None and zeroterm == 0.0:
            lib().orc_exchange_direct(self.h, nst, phi_rs, nhf, hf_rs, _f64(occ), float(exx_coeff), float(scale),
                                      out)
        else:
            gs = _f64(gshift if gshift is not None else (0.0, 0.0, 0.0))
            lib().orc_exchange_direct_shift(self.h, nst, phi_rs, nhf, hf_rs, _f64(occ), float(exx_coeff),
                                            float(scale), gs, float(zeroterm), out)
        return out

    def ace_apply(self, xi, phi_rs):
        xi = _c128(xi)
        phi_rs = _c128(phi_rs)
        out = np.zeros_like(phi_rs)
        lib().orc_ace_apply(self.h, xi.size // self.N, xi, phi_rs.size // self.N, phi_rs, out)
        return out


# ---- dense linear algebra ----
def overlap(phi1, phi2, vol_el, band_major):
    phi1 = _c128(phi1)
    phi2 = _c128(phi2)
    if band_major:
        n1, npnt = phi1.shape
        n2 = phi2.shape[0]
    else:
        npnt, n1 = phi1.shape
        n2 = phi2.shape[1]
    out = np.zeros((n1, n2), dtype=np.complex128)
    lib().orc_overlap(npnt, n1, n2, float(vol_el), int(band_major), phi1, phi2, out)
    return out


def cholesky(m):
    m = _c128(m).copy()
    info = lib().orc_cholesky(m.shape[0], m)
    return m, info


def invert_lower(m):
    m = _c128(m).copy()
    lib().orc_invert_lower(m.shape[0], m)
    return m


def eigh(m):
    m = _c128(m).copy()
    w = np.zeros(m.shape[0])
    lib().orc_eigh(m.shape[0], m, w)
    return w, m


def orthogonalize(phi, vol_el, band_major):
    phi = _c128(phi).copy()
    if band_major:
        nst, npnt = phi.shape
    else:
        npnt, nst = phi.shape
    info = lib().orc_orthogonalize(npnt, nst, float(vol_el), int(band_major), phi)
    return phi, info


def rotate_h(rot, phi, band_major):
    phi = _c128(phi).copy()
    nst = rot.shape[0]
    npnt = phi.size // nst
    lib().orc_rotate_h(npnt, nst, int(band_major), _c128(rot), phi)
    return phi


def rotate5(rot, phi, rotphi, alpha, beta, band_major):
    rotphi = _c128(rotphi).copy()
    nst = rot.shape[0]
    npnt = phi.size // nst
    lib().orc_rotate5(npnt, nst, int(band_major), _c128(rot), _c128(phi), rotphi, _c(alpha), _c(beta))
    return rotphi


def randomize(npoints, nst, index=0, complex_values=True):
    """operations::randomize values [npoints][nst] (complex128 or float64)."""
    out = np.zeros((npoints, nst, 2 if complex_values else 1))
    lib().orc_randomize(int(npoints), int(nst), int(index), int(bool(complex_values)), out)
    return out[..., 0] + 1j * out[..., 1] if complex_values else out[..., 0]


def sing_auxiliary(lattice, q):
    return lib().orc_sing_auxiliary(_f64(lattice).ravel(), _f64(q))


def sing_correction(lattice, kcov, kweight):
    kcov = _f64(kcov).reshape(-1, 3)
    nk = kcov.shape[0]
    fk = np.zeros(nk)
    fz = np.zeros(1)
    lib().orc_sing_correction(_f64(lattice).ravel(), nk, kcov.ravel(), _f64(kweight), fk, fz)
    return fk, fz[0]


# ---- eigensolver ----
def overlap_diagonal(phi1, phi2, vol_el):
    """operations::overlap_diagonal (overlap_diagonal.hpp:25-60), band-major [nst][npnt]."""
    return vol_el * np.einsum("ip,ip->i", np.conj(phi1), phi2)


def tpa_k(x):
    """Teter-Payne-Allan preconditioning function (preconditioner.hpp:31-38)."""
    num = ((8.0 * x + 12.0) * x + 18.0) * x + 27.0
    den = (((16.0 * x + 8.0) * x + 12.0) * x + 18.0) * x + 27.0
    return num / den


def tpa_precondition(phi, g2half):
    """operations::preconditioner on Fourier-space states (preconditioner.hpp:40-58):
    <T>_i = Re(<(1/2)|G|^2 phi_i|phi_i> / <phi_i|phi_i>), phi_i(G) *= K((1/2)|G|^2 / <T>_i).
    g2half holds (1/2)|G|^2 per point (no k shift: point_op.g2)."""
    phi = np.asarray(phi, dtype=np.complex128)
    w = np.abs(phi) ** 2
    kin = (w @ g2half) / w.sum(axis=1)
    return phi * tpa_k(g2half[None, :] / kin[:, None])


def steepest_descent(apply, precond, phi, vol_el, num_steps=5):
    """eigensolvers::steepest_descent (steepest_descent.hpp:23-86) on band-major
    states: num_steps line-minimisation steps along the preconditioned residual,
    then operations::orthogonalize.  apply(phi) -> H phi; precond(res) -> res or None."""
    phi = np.array(phi, dtype=np.complex128)
    hphi = np.asarray(apply(phi), dtype=np.complex128).copy()
    for istep in range(num_steps):
        ev = overlap_diagonal(hphi, phi, vol_el)
        norm = overlap_diagonal(phi, phi, vol_el)
        res = hphi - (ev / norm.real)[:, None] * phi
        if precond is not None:
            res = precond(res)
        hres = np.asarray(apply(res), dtype=np.complex128)
        m0 = overlap_diagonal(res, res, vol_el)
        m1 = overlap_diagonal(phi, res, vol_el)
        m2 = overlap_diagonal(res, hres, vol_el)
        m3 = overlap_diagonal(phi, hres, vol_el)
        m4, m5 = ev, norm
        ca = (m0 * m3 - m2 * m1).real
        cb = (m5 * m2 - m4 * m0).real
        cc = (m4 * m1 - m3 * m5).real
        with np.errstate(invalid="ignore", divide="ignore"):
            den = cb + np.sqrt(cb * cb - 4.0 * ca * cc)
            lam = np.where(np.abs(den) < 1e-15, 0.0, 2.0 * cc / den)
        phi += lam[:, None] * res
        if istep != num_steps - 1:
            hphi += lam[:, None] * hres
    phi, info = orthogonalize(phi, vol_el, True)
    return phi, info


def atomic_field(lattice, n, positions, radius, species, tables, dr, images=2):
    """Brute-force atomic-sphere sum (atomic_potential::local_potential,
    src/hamiltonian/atomic_potential.hpp:116-165, and ::ionic_density, :170-200):
    field[p] = sum_a sum_T f_{species[a]}(|r_p - R_a - T|) for |...| <= radius[a],
    over every lattice image T = s . lattice with s in [-images, images]^3 (the
    caller makes the range wide enough), f interpolated linearly on r = j dr.
    Points in (i0, i1, i2) order, r_p = sum_d (i_d / n_d) a_d.  FP64 numpy."""
    lat = np.asarray(lattice, dtype=np.float64).reshape(3, 3)
    n = [int(x) for x in n]
    pos = np.asarray(positions, dtype=np.float64).reshape(-1, 3)
    nat = pos.shape[0]
    rad = np.broadcast_to(np.asarray(radius, dtype=np.float64), (nat,))
    spc = np.broadcast_to(np.asarray(species, dtype=np.int64), (nat,))
    tab = np.atleast_2d(np.asarray(tables, dtype=np.float64))
    rgrid = np.arange(tab.shape[1]) * dr
    i0, i1, i2 = np.meshgrid(np.arange(n[0]) / n[0], np.arange(n[1]) / n[1], np.arange(n[2]) / n[2], indexing="ij")
    frac = np.stack([i0.ravel(), i1.ravel(), i2.ravel()], axis=1)
    pts = frac @ lat
    out = np.zeros(pts.shape[0])
    rng = np.arange(-images, images + 1)
    shifts = np.array(np.meshgrid(rng, rng, rng, indexing="ij")).reshape(3, -1).T @ lat
    for a in range(nat):
        for t in shifts:
            d = np.linalg.norm(pts - pos[a] - t, axis=1)
            m = d <= rad[a]
            out[m] += np.interp(d[m], rgrid, tab[spc[a]])
    return out


class RefCPU:
    """The reference's CPU data flow for H on full-cube fourier_space orbitals
    [N][nst] (refcpu.c): batched FFTW3 transforms over the state axis, V psi,
    KB projectors as gather / DGEMM / coefficients / DGEMM / scatter, to_fourier,
    laplacian_add; OpenMP threads own chunks of states, the FFTW/BLAS library
    runs sequentially in each.  Gamma point only.  ``backend`` is "mkl-fftw3"
    when MKL's FFTW3/BLAS/LAPACK interfaces were loaded, else "own-fft"."""

    def __init__(self, ham, nst, threads):
        L = lib()
        L.rc_init()
        self.nst = int(nst)
        self.N = ham.N
        npts, pts, beta, coeff = [], [], [], []
        self.nlm = 4
        for i, (nlm, D) in enumerate(ham._proj_meta):
            co, _, _, mat = ham.projector_points(ham.projs[i], self.nlm)
            n = ham.n
            pts.append(((co[:, 0] % n[0]) * n[1] + (co[:, 1] % n[1])) * n[2] + (co[:, 2] % n[2]))
            npts.append(len(co))
            beta.append(mat)
            dd = np.zeros((self.nlm, self.nlm))
            dd[:nlm, :nlm] = np.asarray(D).reshape(nlm, nlm)
            coeff.append(dd)
        self._keep = (np.asarray(npts, np.int32), np.ascontiguousarray(np.concatenate(pts) if pts else [0], np.int32),
                      np.ascontiguousarray(np.concatenate([b.ravel() for b in beta]) if beta else [0.0]),
                      np.ascontiguousarray(np.concatenate([c.ravel() for c in coeff]) if coeff else [0.0]))
        self.h = L.rc_ham_create(ham.lattice.ravel(), np.asarray(ham.n, np.int32), self.nst, ham.v, len(npts),
                                 self._keep[0] if npts else np.zeros(1, np.int32), self._keep[1], self.nlm,
                                 self._keep[2], self._keep[3], int(threads))
        self.backend = "mkl-fftw3" if L.rc_backend() == 1 else "own-fft"

    def apply(self, psi_fs, out=None):
        """H on full-cube Fourier coefficients [N][nst] (complex128)"""
        psi_fs = _c128(psi_fs)
        if out is None:
            out = np.zeros_like(psi_fs)
        lib().rc_ham_apply(self.h, psi_fs, out)
        return out

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.rc_ham_destroy(self.h)
            self.h = None


def ref_overlap_cholesky(phi, vol_el):
    """S = vol_el Phi^H Phi (ZGEMM over the full grid, Phi [N][nst]) and ZPOTRF."""
    phi = _c128(phi)
    N, nst = phi.shape
    S = np.zeros((nst, nst), dtype=np.complex128)
    info = lib().rc_overlap_cholesky(N, nst, float(vol_el), phi, S)
    return S, info


def ref_exchange_pairs(lattice, n, phi_rs, hf_rs, coef, threads):
    """Gamma-point pair solves: kphi_i = sum_j coef_j hf_j Poisson(conj(hf_j) phi_i),
    phi_rs [N][nst], hf_rs [N][nhf] (G = 0 dropped)."""
    phi_rs, hf_rs = _c128(phi_rs), _c128(hf_rs)
    out = np.zeros_like(phi_rs)
    lib().rc_exchange_pairs(_f64(lattice).ravel(), np.asarray(n, np.int32), phi_rs.shape[1], phi_rs, hf_rs.shape[1],
                            hf_rs, _f64(coef), out, int(threads))
    return out


def ylm_real(l, m, v):
    """the oracle's real spherical harmonic Y_lm(v/|v|) (projector channels)"""
    return float(lib().orc_ylm_real(int(l), int(m), _f64(v)))
