"""eigensolvers:: and the preconditioner of operations:: over the inqx C ABI.

States are band-major DeviceArrays [nst][npw] on a basis of the Hamiltonian.
Names follow src/eigensolvers/steepest_descent.hpp and
src/operations/preconditioner.hpp.
"""
from ._lib import check, lib
from .hamiltonian import precision_of


def preconditioner(ham, basis, phi):
    """Teter-Payne-Allan preconditioner on Fourier-space states, in place
    (preconditioner.hpp:26-58)."""
    nst = phi.shape[0]
    check(lib().inqx_precondition(ham.h, basis.h, nst, precision_of(phi.dtype), phi.ptr), ham.ctx.h,
          "inqx_precondition")
    return phi


def steepest_descent(ham, basis, phi, precondition=True, num_steps=5):
    """steepest_descent(ham, prec, phi) (steepest_descent.hpp:23-86): num_steps
    preconditioned line minimisations per state, then orthogonalize; in place."""
    nst = phi.shape[0]
    check(lib().inqx_steepest_descent(ham.h, basis.h, nst, precision_of(phi.dtype), int(num_steps),
                                      int(bool(precondition)), phi.ptr), ham.ctx.h, "inqx_steepest_descent")
    return phi
