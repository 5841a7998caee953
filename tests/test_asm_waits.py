"""The wait-count check of tests/asm_waits.py: on small hand-written gfx950
snippets it flags a register read before its load has returned and accepts
the same code with the wait in place (straight-line, across a branch, with
stores sharing the counter, loads into LDS); with INQX_ASM_CHECK=1 it compiles
every HIP source of the library to assembly and requires no finding in any
kernel."""
import os
import subprocess
import pytest

from tests import asm_waits as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def k(body):
    return "_Zk:\n" + body + "\n\ts_endpgm\n"


def test_read_before_wait_is_found():
    f = W.check_kernel(k("\tglobal_load_dwordx4 v[4:7], v[0:1], off\n\tv_add_f64 v[8:9], v[4:5], v[6:7]"))
    assert f and "v4" in f[0]


def test_read_after_wait_passes():
    assert not W.check_kernel(k("\tglobal_load_dwordx4 v[4:7], v[0:1], off\n\ts_waitcnt vmcnt(0)\n"
                                "\tv_add_f64 v[8:9], v[4:5], v[6:7]"))


def test_in_order_retirement():
    # vmcnt(1) leaves only the newest load in flight
    body = ("\tglobal_load_dword v4, v[0:1], off\n\tglobal_load_dword v5, v[2:3], off\n\ts_waitcnt vmcnt(1)\n"
            "\tv_mov_b32 v6, v4")
    assert not W.check_kernel(k(body))
    assert W.check_kernel(k(body.replace("v_mov_b32 v6, v4", "v_mov_b32 v6, v5")))


def test_stores_share_the_counter():
    # the store issued after the load makes vmcnt(1) retire the load
    body = ("\tglobal_load_dword v4, v[0:1], off\n\tglobal_store_dword v[2:3], v9, off\n\ts_waitcnt vmcnt(1)\n"
            "\tv_mov_b32 v6, v4")
    assert not W.check_kernel(k(body))


def test_a_path_without_the_wait_is_found():
    body = ("\tglobal_load_dword v4, v[0:1], off\n\ts_cbranch_scc1 .LBB0_2\n\ts_waitcnt vmcnt(0)\n"
            ".LBB0_2:\n\tv_mov_b32 v6, v4")
    assert W.check_kernel(k(body))
    assert not W.check_kernel(k(body.replace("s_cbranch_scc1 .LBB0_2\n\ts_waitcnt vmcnt(0)",
                                             "s_waitcnt vmcnt(0)\n\ts_cbranch_scc1 .LBB0_2")))


def test_lds_reads_and_loads_into_lds():
    assert W.check_kernel(k("\tds_read_b128 v[4:7], v1\n\tv_mov_b32 v8, v5"))
    assert not W.check_kernel(k("\tds_read_b128 v[4:7], v1\n\ts_waitcnt lgkmcnt(0)\n\tv_mov_b32 v8, v5"))
    # operand 0 of a load into LDS is its address: reusing it at once is fine
    assert not W.check_kernel(k("\tglobal_load_lds_dwordx4 v[4:5], off\n\tv_mov_b32 v4, 0"))


@pytest.mark.skipif(os.environ.get("INQX_ASM_CHECK") != "1", reason="compiles every source (minutes); set INQX_ASM_CHECK=1")
@pytest.mark.parametrize("src", ["ham_kernels", "plane_rect", "exx_kernels", "linalg", "eigh", "eigensolver",
                                 "potential", "singularity", "layout"])
def test_library_kernels(src, tmp_path):
    csrc = os.path.join(ROOT, "inq_amd", "csrc")
    out = tmp_path / (src + ".s")
    subprocess.check_call(["/opt/rocm/bin/hipcc", "-O3", "-std=c++17", "--offload-arch=gfx950", "-Xclang",
                           "-target-feature", "-Xclang", "-packed-fp32-ops", "-fno-gpu-rdc",
                           "-I" + os.path.join(ROOT, "include"), "--cuda-device-only", "-S", "-x", "hip",
                           os.path.join(csrc, src + ".hip"), "-o", str(out)], stderr=subprocess.DEVNULL)
    bad, n = W.check_file(out.read_text())
    assert n > 0
    assert not bad, {name: f[:3] for name, f in list(bad.items())[:3]}
