"""Shared builders: one synthetic system set up identically in the HIP
library (through the C ABI) and in the FP64 oracle that checks it."""
import sys

import numpy as np

from inq_amd import systems as S


def say(*a):
    """names the case before a GPU step (flushed, so a crash still shows it)"""
    print(*a, flush=True)
    sys.stdout.flush()


def oracle_ham(sysm, v=None):
    from oracle import oracle as orc
    h = orc.Ham(sysm.lattice, sysm.n)
    h.set_potential(S.local_potential(sysm) if v is None else v)
    for p, D in zip(sysm.positions, sysm.kb_coeffs()):
        h.add_projector_gauss(p, sysm.kb_radius, sysm.nlm, sysm.kb_width, D)
    return h


def gpu_ham(ctx, sysm, v=None):
    from inq_amd import KSHamiltonian
    h = KSHamiltonian(ctx, sysm.lattice, sysm.n)
    h.set_scalar_potential(S.local_potential(sysm) if v is None else v)
    for p, D in zip(sysm.positions, sysm.kb_coeffs()):
        h.add_projector(p, sysm.kb_radius, sysm.nlm, D, width=sysm.kb_width)
    return h


def rel_err(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
