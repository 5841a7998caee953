"""inq_amd: MI355X-native Kohn-Sham H|psi> engine (C ABI in include/inqx.h,
HIP kernels in inq_amd/csrc, this package is the Python host mirror)."""
from ._lib import INQX_FP32, INQX_FP64, InqxError, lib  # noqa: F401
from .hamiltonian import Basis, Context, DeviceArray, KSHamiltonian, precision_of  # noqa: F401
