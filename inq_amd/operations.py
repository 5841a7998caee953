"""operations:: and matrix:: functions of the reference over the inqx C ABI.

States are band-major DeviceArrays [n][npw] (complex64 or complex128);
matrices are device complex128 [n][n] row-major.  Names follow
src/operations/{overlap,orthogonalize,rotate}.hpp, src/matrix/{cholesky,
invert_triangular}.hpp and src/ground_state/subspace_diagonalization.hpp.
"""
import ctypes as C

import numpy as np

from ._lib import check, lib
from .hamiltonian import DeviceArray, precision_of


def overlap(ctx, phi1, phi2, vol_el):
    """S = vol_el phi1^H phi2 (overlap.hpp:26-95)."""
    n1, npw = phi1.shape
    n2 = phi2.shape[0]
    assert phi2.shape[1] == npw and phi1.dtype == phi2.dtype
    S = DeviceArray(ctx, (n1, n2), np.complex128)
    check(lib().inqx_overlap(ctx.h, precision_of(phi1.dtype), n1, n2, npw, float(vol_el), phi1.ptr, phi2.ptr, S.ptr),
          ctx.h, "inqx_overlap")
    return S


def overlap_diagonal(ctx, phi1, phi2, vol_el, normalized=False):
    """per-state vol_el <phi1_i|phi2_i> (overlap_diagonal.hpp:25-60), or with
    normalized <phi1_i|phi2_i>/<phi2_i|phi2_i> (:100-153); device complex128 [n]."""
    n, npw = phi1.shape
    assert phi2.shape == phi1.shape and phi1.dtype == phi2.dtype
    out = DeviceArray(ctx, (n,), np.complex128)
    check(lib().inqx_overlap_diagonal(ctx.h, precision_of(phi1.dtype), n, npw, float(vol_el), phi1.ptr, phi2.ptr,
                                      int(bool(normalized)), out.ptr), ctx.h, "inqx_overlap_diagonal")
    return out


def randomize(ctx, npoints, nst, index=0, dtype=np.complex128, layout=1):
    """operations::randomize (randomize.hpp:48-70): pcg32 U[-1,1) values, [npoints][nst] (layout 1)
    or [nst][npoints] (layout 0); dtype complex128, complex64 or float64."""
    dt = np.dtype(dtype)
    kind = {np.dtype(np.complex128): 0, np.dtype(np.complex64): 1, np.dtype(np.float64): 2}[dt]
    shape = (npoints, nst) if layout == 1 else (nst, npoints)
    out = DeviceArray(ctx, shape, dt)
    check(lib().inqx_randomize(ctx.h, kind, int(npoints), int(nst), int(index), int(layout), out.ptr), ctx.h,
          "inqx_randomize")
    return out


def cholesky(ctx, S):
    """in place S = L L^H, L lower (cholesky.hpp:28-100)."""
    n = S.shape[0]
    check(lib().inqx_cholesky(ctx.h, n, S.ptr), ctx.h, "inqx_cholesky")
    return S


def invert_triangular(ctx, L):
    n = L.shape[0]
    X = DeviceArray(ctx, (n, n), np.complex128)
    check(lib().inqx_invert_triangular(ctx.h, n, L.ptr, X.ptr), ctx.h, "inqx_invert_triangular")
    return X


def rotate(ctx, rot, phi):
    """phi <- phi rot^H (rotate.hpp:26-80)."""
    n, npw = phi.shape
    check(lib().inqx_rotate(ctx.h, precision_of(phi.dtype), n, npw, rot.ptr, phi.ptr), ctx.h, "inqx_rotate")
    return phi


def rotate_add(ctx, rot, phi, rotphi, alpha=1.0, beta=0.0):
    """rotphi <- alpha phi rot + beta rotphi (rotate.hpp:84-140)."""
    n, npw = phi.shape
    a, b = complex(alpha), complex(beta)
    check(lib().inqx_rotate_add(ctx.h, precision_of(phi.dtype), n, npw, rot.ptr, phi.ptr, rotphi.ptr, a.real, a.imag,
                                b.real, b.imag), ctx.h, "inqx_rotate_add")
    return rotphi


def orthogonalize(ctx, phi, vol_el):
    """overlap -> Cholesky -> rotate_trs, in place (orthogonalize.hpp:25-32)."""
    n, npw = phi.shape
    check(lib().inqx_orthogonalize(ctx.h, precision_of(phi.dtype), n, npw, float(vol_el), phi.ptr), ctx.h,
          "inqx_orthogonalize")
    return phi


def subspace_diagonalization(ham, basis, phi):
    """diagonalise phi^H H phi in FP64 and rotate phi onto the eigenvectors;
    returns the eigenvalues, ascending (subspace_diagonalization.hpp:21-29)."""
    nst = phi.shape[0]
    ev = np.zeros(nst)
    check(lib().inqx_subspace_diag(ham.h, basis.h, nst, precision_of(phi.dtype), phi.ptr,
                                   ev.ctypes.data_as(C.POINTER(C.c_double))), ham.ctx.h, "inqx_subspace_diag")
    return ev


def diagonalize(ctx, matrix):
    """matrix::diagonalize (diagonalize.hpp:167-189) of a device complex128 [n][n]
    Hermitian matrix, in place: rows become the conjugated eigenvectors (the
    reference's orientation); returns the eigenvalues, ascending."""
    n = matrix.shape[0]
    ev = np.zeros(n)
    check(lib().inqx_diagonalize(ctx.h, n, matrix.ptr, ev.ctypes.data_as(C.POINTER(C.c_double))), ctx.h,
          "inqx_diagonalize")
    return ev
