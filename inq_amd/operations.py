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
