"""Loads the HIP library libinqx.so through its C ABI (include/inqx.h).

The product path has no CPU fallback: if the library is missing or fails to
load, every call raises.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libinqx.so")
_LIB = None

INQX_OK = 0
INQX_FP64 = 0
INQX_FP32 = 1
# status codes (inqx_status in include/inqx.h)
INQX_ERR_ARG = -1
INQX_ERR_HIP = -2
INQX_ERR_UNSUPPORTED = -3
INQX_ERR_NOT_PD = -4
INQX_ERR_STATE = -5

# name -> (restype, argtypes); every symbol include/inqx.h declares
P = C.c_void_p
PP = C.POINTER(C.c_void_p)
DP = C.POINTER(C.c_double)
IP = C.POINTER(C.c_int)
FP = C.POINTER(C.c_float)
LP = C.POINTER(C.c_long)
SIGNATURES = {
    "inqx_version": (C.c_int, []),
    "inqx_ctx_create": (C.c_int, [C.c_int, P, PP]),
    "inqx_ctx_destroy": (C.c_int, [P]),
    "inqx_ctx_set_stream": (C.c_int, [P, P]),
    "inqx_sync": (C.c_int, [P]),
    "inqx_last_error": (C.c_char_p, [P]),
    "inqx_prof_enable": (C.c_int, [P, C.c_int]),
    "inqx_prof_reset": (C.c_int, [P]),
    "inqx_prof_count": (C.c_int, [P, IP]),
    "inqx_prof_get": (C.c_int, [P, C.c_int, C.c_char_p, C.c_int, LP, DP]),
    "inqx_malloc": (C.c_int, [P, C.c_size_t, PP]),
    "inqx_free": (C.c_int, [P, P]),
    "inqx_memcpy_h2d": (C.c_int, [P, P, P, C.c_size_t]),
    "inqx_memcpy_d2h": (C.c_int, [P, P, P, C.c_size_t]),
    "inqx_memcpy_d2d": (C.c_int, [P, P, P, C.c_size_t]),
    "inqx_stream_copy": (C.c_int, [P, P, P, C.c_size_t]),
    "inqx_ham_create": (C.c_int, [P, DP, IP, PP]),
    "inqx_ham_destroy": (C.c_int, [P]),
    "inqx_ham_set_potential": (C.c_int, [P, DP, C.c_int]),
    "inqx_ham_set_vector_potential": (C.c_int, [P, DP]),
    "inqx_ham_add_projector": (C.c_int, [P, DP, C.c_double, C.c_int, DP, DP, C.c_double]),
    "inqx_ham_projector_size": (C.c_int, [P, C.c_int, LP]),
    "inqx_ham_projector_points": (C.c_int, [P, C.c_int, IP, FP, FP]),
    "inqx_basis_create": (C.c_int, [P, DP, C.c_double, C.c_int, PP]),
    "inqx_basis_destroy": (C.c_int, [P]),
    "inqx_basis_size": (C.c_int, [P, LP]),
    "inqx_basis_gindex": (C.c_int, [P, IP]),
    "inqx_basis_kinetic": (C.c_int, [P, DP]),
    "inqx_ham_prepare": (C.c_int, [P, P, C.c_int, C.c_int]),
    "inqx_ham_apply": (C.c_int, [P, P, C.c_int, C.c_int, P, P]),
    "inqx_ham_apply_batch": (C.c_int, [P, C.c_int, PP, IP, C.c_int, PP, PP]),
    "inqx_ham_apply_layout": (C.c_int, [P, P, C.c_int, C.c_int, C.c_int, P, P]),
    "inqx_states_to_sphere": (C.c_int, [P, P, C.c_int, C.c_int, C.c_int, P, P]),
    "inqx_states_from_sphere": (C.c_int, [P, P, C.c_int, C.c_int, C.c_int, P, P]),
    "inqx_sphere_to_real": (C.c_int, [P, P, C.c_int, C.c_int, P, C.c_int, P]),
    "inqx_real_to_sphere": (C.c_int, [P, P, C.c_int, C.c_int, P, C.c_int, C.c_int, P]),
    "inqx_poisson": (C.c_int, [P, C.c_int, C.c_int, P, DP, C.c_double]),
    "inqx_exchange_direct": (C.c_int, [P, P, C.c_int, C.c_int, P, C.c_int, P, DP, C.c_double, C.c_double,
                                       C.c_double, P]),
    "inqx_ace_build": (C.c_int, [P, P, C.c_int, C.c_int, P, DP, C.c_double, C.c_double, P, DP]),
    "inqx_exchange_direct_kpts": (C.c_int, [P, P, C.c_int, C.c_int, P, C.c_int, PP, IP, PP, PP, DP, C.c_double,
                                            C.c_double, P]),
    "inqx_ace_build_kpts": (C.c_int, [P, P, C.c_int, C.c_int, P, DP, C.c_int, PP, IP, PP, PP, DP, C.c_double, P,
                                      DP]),
    "inqx_ace_apply": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_long, C.c_double, P, P, P]),
    "inqx_ham_set_exchange_workspace": (C.c_int, [P, C.c_size_t]),
    "inqx_ham_autotune": (C.c_int, [P, P, C.c_int, C.c_int, IP, DP]),
    "inqx_ham_autotune_batch": (C.c_int, [P, C.c_int, P, IP, C.c_int, C.c_char_p, IP, DP]),
    "inqx_ham_set_plane_batch": (C.c_int, [P, C.c_int, C.c_int]),
    "inqx_randomize": (C.c_int, [P, C.c_int, C.c_long, C.c_int, C.c_int, C.c_int, P]),
    "inqx_overlap_diagonal": (C.c_int, [P, C.c_int, C.c_int, C.c_long, C.c_double, P, P, C.c_int, P]),
    "inqx_ham_apply_real": (C.c_int, [P, P, C.c_int, C.c_int, P, C.c_int, P]),
    "inqx_density_add": (C.c_int, [P, P, C.c_int, C.c_int, P, DP, C.c_int, C.c_int, P]),
    "inqx_ham_set_launch_shape": (C.c_int, [P, C.c_int, C.c_int]),
    "inqx_overlap": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_long, C.c_double, P, P, P]),
    "inqx_cholesky": (C.c_int, [P, C.c_int, P]),
    "inqx_invert_triangular": (C.c_int, [P, C.c_int, P, P]),
    "inqx_rotate": (C.c_int, [P, C.c_int, C.c_int, C.c_long, P, P]),
    "inqx_rotate_add": (C.c_int, [P, C.c_int, C.c_int, C.c_long, P, P, P, C.c_double, C.c_double, C.c_double,
                                  C.c_double]),
    "inqx_orthogonalize": (C.c_int, [P, C.c_int, C.c_int, C.c_long, C.c_double, P]),
    "inqx_subspace_diag": (C.c_int, [P, P, C.c_int, C.c_int, P, DP]),
    "inqx_diagonalize": (C.c_int, [P, C.c_int, P, DP]),
    "inqx_precondition": (C.c_int, [P, P, C.c_int, C.c_int, P]),
    "inqx_hartree": (C.c_int, [P, P, P, DP]),
    "inqx_hartree_noalloc": (C.c_int, [P, P, P, DP]),
    "inqx_atomic_field_add": (C.c_int, [P, C.c_int, DP, DP, IP, C.c_int, C.c_int, C.c_double, DP, P]),
    "inqx_states_save": (C.c_int, [P, P, C.c_int, C.c_int, P, C.c_char_p, C.c_int]),
    "inqx_states_load": (C.c_int, [P, P, C.c_int, C.c_int, C.c_char_p, C.c_int, P]),
    "inqx_field_save": (C.c_int, [P, C.c_char_p, C.c_long, C.c_int, DP]),
    "inqx_field_load": (C.c_int, [P, C.c_char_p, C.c_long, C.c_int, DP]),
    "inqx_array_save": (C.c_int, [P, C.c_char_p, C.c_long, DP]),
    "inqx_array_load": (C.c_int, [P, C.c_char_p, C.c_long, DP]),
    "inqx_steepest_descent": (C.c_int, [P, P, C.c_int, C.c_int, C.c_int, C.c_int, P]),
    "inqx_exponential": (C.c_int, [P, P, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, P]),
    "inqx_exponential_2_for_1": (C.c_int, [P, P, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                           C.c_int, P, P]),
    "inqx_prop_create": (C.c_int, [P, C.c_int, PP, IP, C.c_int, PP, PP, C.c_double, C.c_int, PP]),
    "inqx_prop_step": (C.c_int, [P, C.c_int, C.c_int]),
    "inqx_prop_energy": (C.c_int, [P]),
    "inqx_prop_records": (C.c_int, [P, IP, DP, C.c_int]),
    "inqx_prop_set_scf": (C.c_int, [P, P, C.c_int, C.c_double]),
    "inqx_prop_scf_info": (C.c_int, [P, IP, DP]),
    "inqx_prop_set_state_precision": (C.c_int, [P, C.c_int]),
    "inqx_ks_potential": (C.c_int, [P, P, P, DP]),
    "inqx_singularity_correction": (C.c_int, [P, DP, C.c_int, DP, DP, DP, DP]),
    "inqx_energy": (C.c_int, [P, C.c_int, PP, IP, C.c_int, PP, PP, DP]),
    "inqx_prop_destroy": (C.c_int, [P]),
}


class InqxError(RuntimeError):
    pass


def lib():
    """The loaded library; raises if it is not built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise InqxError(f"{LIB_PATH} is not built: run __graft_entry__.build() first")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def check(status, ctx=None, what=""):
    if status < 0:
        msg = ""
        if ctx is not None:
            m = lib().inqx_last_error(ctx)
            msg = m.decode() if m else ""
        raise InqxError(f"{what} failed with status {status}: {msg}")
    return status
