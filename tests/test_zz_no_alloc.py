"""The allocation-free Hartree entry point (inqx_hartree_noalloc) through the
C ABI: the same potential and energy as inqx_hartree, bit for bit, on two
calls in a row (96^3, then a smaller 48^3 grid on another Hamiltonian, so the
second call reuses the scratch left by the first), and a clear INQX_ERR_ARG
for a grid beyond its scratch (no kernel is launched for it).
"""
import ctypes as C

import numpy as np
import pytest

from inq_amd import Context, DeviceArray, KSHamiltonian
from inq_amd import _lib
from tests.helpers import say

pytestmark = pytest.mark.gpu


def _noalloc(ham, dens, vh):
    e = C.c_double()
    st = _lib.lib().inqx_hartree_noalloc(ham.h, dens.ptr, vh.ptr, C.byref(e))
    return st, e.value


@pytest.fixture(scope="module")
def ctx():
    return Context(0)


def test_noalloc_hartree_matches_hartree(ctx):
    rng = np.random.default_rng(12)
    for n, a in ((96, 30.78), (48, 11.0)):
        say("case: hartree_noalloc", n)
        lat = np.array([[a, 0.0, 0.0], [0.4, a, 0.0], [-0.3, 0.2, a]])
        ham = KSHamiltonian(ctx, lat, (n, n, n))
        rho = rng.standard_normal(n ** 3) + 0.5
        d = DeviceArray.from_numpy(ctx, rho)
        vh_ref, e_ref = ham.hartree(d)
        vh = DeviceArray(ctx, (n ** 3,), np.float64)
        st, e = _noalloc(ham, d, vh)
        assert st == 0, _lib.lib().inqx_last_error(ctx.h)
        assert np.array_equal(vh.numpy(), vh_ref.numpy())
        assert e == e_ref
        # energy only, then potential only
        e2 = C.c_double()
        assert _lib.lib().inqx_hartree_noalloc(ham.h, d.ptr, None, C.byref(e2)) == 0
        assert e2.value == e_ref
        vh.upload(np.zeros(n ** 3))
        assert _lib.lib().inqx_hartree_noalloc(ham.h, d.ptr, vh.ptr, None) == 0
        assert np.array_equal(vh.numpy(), vh_ref.numpy())


def test_noalloc_hartree_rejects_large_grids(ctx):
    say("case: hartree_noalloc 128^3 rejected")
    ham = KSHamiltonian(ctx, np.eye(3) * 20.0, (128, 128, 128))
    d = DeviceArray(ctx, (16,), np.float64)  # never read: the size check comes first
    st = _lib.lib().inqx_hartree_noalloc(ham.h, d.ptr, d.ptr, None)
    assert st == _lib.INQX_ERR_ARG
    assert b"exceeds the allocation-free Hartree limit" in _lib.lib().inqx_last_error(ctx.h)
