"""CPU checks of the C ABI: the library loads, exports every symbol
include/inqx.h declares, and its host-side set-up (projector spheres,
projector values, G spheres) matches the oracle bit for bit.  No compute call
is made: the context is host-only (device = -1)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from inq_amd import _lib, systems as S
from inq_amd.hamiltonian import Context, KSHamiltonian
from tests.helpers import oracle_ham

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "inqx.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(inqx_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared_symbols()
    assert len(names) > 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the Python binding covers them all
    assert not [n for n in names if n not in _lib.SIGNATURES]


def test_version_and_host_context():
    assert _lib.lib().inqx_version() >= 1
    ctx = Context(device=-1)
    assert ctx.h.value


def test_error_codes_not_exceptions():
    ctx = Context(device=-1)
    h = C.c_void_p()
    bad = (C.c_double * 9)(*([0.0] * 9))
    n = (C.c_int * 3)(8, 8, 8)
    st = _lib.lib().inqx_ham_create(ctx.h, bad, n, C.byref(h))
    assert st == -1
    assert b"volume" in _lib.lib().inqx_last_error(ctx.h)


@pytest.mark.parametrize("cid", [1, 3])
def test_projector_spheres_match_oracle(cid):
    sysm = S.config(cid)
    ctx = Context(device=-1)
    gh = KSHamiltonian(ctx, sysm.lattice, sysm.n)
    oh = oracle_ham(sysm, v=np.zeros(sysm.N))
    for a, (p, D) in enumerate(zip(sysm.positions, sysm.kb_coeffs())):
        gh.add_projector(p, sysm.kb_radius, sysm.nlm, D, width=sysm.kb_width)
        c1, d1, r1 = gh.projector_points(a)
        c2, d2, r2, _ = oh.projector_points(a, sysm.nlm)
        assert np.array_equal(c1, c2)
        assert np.array_equal(d1, d2)
        assert np.array_equal(r1, r2)


def test_noncubic_sphere_kat_through_abi():
    """15963 points on the hexagonal cell (spherical_grid.hpp:357-372)."""
    lat = np.array([[7.57325, 0.0, 0.0], [-3.78662, 6.55862, 0.0], [5.78917e-16, 1.00271e-15, 9.45443]])
    ctx = Context(device=-1)
    gh = KSHamiltonian(ctx, lat, (20, 20, 24))
    gh.add_projector((3.78659, -2.18619, 1.65434), 5.71, 1, np.eye(1), width=1.0)
    c, d, r = gh.projector_points(0)
    assert len(c) == 15963


@pytest.mark.parametrize("cid", [1, 2, 3])
def test_gsphere_matches_oracle(cid):
    sysm = S.config(cid)
    ctx = Context(device=-1)
    gh = KSHamiltonian(ctx, sysm.lattice, sysm.n)
    oh = oracle_ham(sysm, v=np.zeros(sysm.N)) if cid != 2 else None
    from oracle import oracle as orc
    if oh is None:
        oh = orc.Ham(sysm.lattice, sysm.n)
    for k, spin, ec in sysm.blocks()[:6]:
        b = gh.basis(k, ec, spin)
        g = b.gindex()
        assert np.array_equal(g, oh.gsphere(k, ec))
        assert np.max(np.abs(b.kinetic() - oh.gsphere_kinetic(k, g))) < 1e-12


@pytest.mark.parametrize("n,lat", [((128, 128, 128), (12.0, 12.0, 12.0)), ((100, 120, 108), (10.0, 12.0, 10.8)),
                                   ((24, 20, 30), (6.0, 5.0, 7.5))])
def test_gsphere_on_large_and_noncubic_grids(n, lat):
    """the G sphere and its kinetic factors on grids with lines above 96 points
    and with planes that are not square, bit-exact against the oracle"""
    from oracle import oracle as orc
    L = np.diag(lat)
    ctx = Context(device=-1)
    gh = KSHamiltonian(ctx, L, n)
    oh = orc.Ham(L, n)
    for k, ec in (((0.0, 0.0, 0.0), 20.0), ((0.25, -0.15, 0.3), 12.0)):
        b = gh.basis(k, ec)
        g = b.gindex()
        assert np.array_equal(g, oh.gsphere(k, ec))
        assert np.max(np.abs(b.kinetic() - oh.gsphere_kinetic(k, g))) < 1e-12


def test_potential_components_cover_every_basis_spin():
    """A basis with spin index s reads column s of the potential, so a potential
    with ncomp <= s is rejected (INQX_ERR_ARG), and so is the apply on such a
    basis; ncomp = s + 1 is accepted."""
    ctx = Context(device=-1)
    n = (12, 12, 12)
    h = KSHamiltonian(ctx, np.eye(3) * 6.0, n)
    b1 = h.basis((0.0, 0.0, 0.0), 4.0, spin=1)
    N = int(np.prod(n))
    v1 = np.zeros(N)
    st = _lib.lib().inqx_ham_set_potential(h.h, v1.ctypes.data_as(C.POINTER(C.c_double)), 1)
    assert st == _lib.INQX_ERR_ARG
    assert b"spin index 1" in _lib.lib().inqx_last_error(ctx.h)
    v2 = np.zeros((N, 2))
    h.set_scalar_potential(v2, ncomp=2)
    del b1
