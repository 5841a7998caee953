"""Host emulation of the index maps of the four pair-Poisson kernels
(inq_amd/csrc/exx_kernels.hip: k_pair_plane_fwd, k_pair_xline, k_pair_plane_inv,
k_pair_reduce) at the 96^3 grid, with the work-group -> element arithmetic of
each kernel restated here: every element of the pair workspace t (and of the
accumulator w) is written exactly once per stage, each stage reads only
elements the previous stage wrote, and no work-group of a stage writes an
element another work-group of the same stage touches.  The maps are affine in
the state count, so a few states and occupied states cover the Si216 sizes;
the largest offsets at 432 states x 11 occupied states are checked against
the 32-bit buffer-offset range and the int range of the grid sizes.
"""
import numpy as np

N0 = N1 = N2 = 96
NP = N1 * N2
N = N0 * NP


def plane_pass(nst, nj):
    """(work-group, element) -> flat t index of the plane passes: wg = (jj, i0, i),
    elements e = y*N2 + z of the plane (pair_plane_fwd stores, pair_plane_inv
    reads and stores)"""
    wg = np.arange(nj * N0 * nst)
    i = wg % nst
    r = wg // nst
    i0 = r % N0
    jj = r // N0
    base = (jj * nst + i) * N + i0 * NP
    return wg[:, None], (base[:, None] + np.arange(NP)[None, :])


def xline_pass(nst, nj):
    """pair_xline: wg = (pair, group of 4 waves), wave -> tile of 32 lines, lane
    pair -> line L; the lower lane holds i0 = 2n, the upper 2n + 1 before the
    transform and i0 = k, k + 48 after it (stores)"""
    ntile = (NP + 31) // 32
    ngrp = (ntile + 3) // 4
    tot = nj * nst * ngrp
    wg = np.arange(tot)
    pair = wg // ngrp
    out_wg, out_idx = [], []
    for wave in range(4):
        tile = (wg - pair * ngrp) * 4 + wave
        ok = tile < ntile
        for lane in range(32):
            L = tile * 32 + lane
            valid = ok & (L < NP)
            for k in range(N0 // 2):
                for half in (0, 1):
                    i0 = k + half * (N0 // 2)
                    out_wg.append(wg[valid])
                    out_idx.append(pair[valid] * N + i0 * NP + L[valid])
    return np.concatenate(out_wg), np.concatenate(out_idx)


def reduce_pass(nst):
    """pair_reduce (complex128: one element per 16-byte vector): wg = (point
    range, i), 256 lanes x 4 vectors"""
    nvec = N
    per = (nvec + 1023) // 1024
    wg = np.arange(per * nst)
    i = wg % nst
    v0 = (wg // nst) * 1024
    v = v0[:, None] + (np.arange(256)[:, None] + 256 * np.arange(4)[None, :]).ravel()[None, :]
    valid = v < nvec
    return np.broadcast_to(wg[:, None], v.shape)[valid], (i[:, None] * nvec + v)[valid]


def _exactly_once(idx, size):
    counts = np.bincount(idx, minlength=size)
    return counts.min() == 1 and counts.max() == 1 and counts.size == size


def test_plane_passes_write_each_element_once():
    nst, nj = 5, 3
    wg, idx = plane_pass(nst, nj)
    assert _exactly_once(idx.ravel(), nj * nst * N)
    # a work-group touches one plane only: no two work-groups share an element
    owner = np.full(nj * nst * N, -1)
    owner[idx.ravel()] = np.repeat(wg.ravel(), NP)
    assert np.array_equal(owner[idx], np.broadcast_to(wg, idx.shape))


def test_xline_writes_each_element_once_within_its_pair():
    nst, nj = 3, 2
    wg, idx = xline_pass(nst, nj)
    assert _exactly_once(idx, nj * nst * N)
    # reads are the same elements (in place), and every element of a pair lies
    # in that pair's 96^3 slab
    ngrp = ((NP + 31) // 32 + 3) // 4
    assert np.array_equal(idx // N, wg // ngrp)


def test_reduce_writes_each_w_element_once():
    nst = 4
    _, idx = reduce_pass(nst)
    assert _exactly_once(idx, nst * N)


def test_offsets_fit_at_si216_sizes():
    nst, nj = 432, 11
    # buffer offsets of pair_xline (bytes, complex128): lane offset + sgpr offset
    voff_max = (NP - 1 + NP) * 16
    soff_max = 2 * (N0 // 2 - 1) * NP * 16
    assert voff_max + soff_max < N * 16 < 2 ** 31
    # work counts are ints
    assert nj * N0 * nst < 2 ** 31
    assert nj * nst * (((NP + 31) // 32 + 3) // 4) < 2 ** 31
    assert ((N + 1023) // 1024) * nst < 2 ** 31
    # element offsets use 64-bit arithmetic (size_t) in the kernels; they exceed 2^31
    assert (nj * nst) * N > 2 ** 31
