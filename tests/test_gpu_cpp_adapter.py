"""The header-only C++ adapter on the GPU: tests/cpp/adapter_gpu (built by
__graft_entry__.build()) applies H through ks_hamiltonian::operator(), takes a
Taylor-4 half step through a reference-shaped template and evaluates the
energy terms on a 16^3 system; the results are checked against the FP64
oracle (1e-11 on states, 1e-10 on energies)."""
import os
import subprocess

import numpy as np
import pytest

from inq_amd import systems as S
from oracle import oracle as orc
from tests.helpers import rel_err, say

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "adapter_gpu")


def test_cpp_adapter_apply_half_step_energy(tmp_path):
    say("case: C++ adapter H, half step, energy")
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built: run __graft_entry__.build() first")
    lat = np.eye(3) * 6.0
    n = (16, 16, 16)
    sysm = S.small_config(seed=3, n=n, natoms=1, nst=4, ecut=5.0)
    v = S.local_potential(sysm)
    rng = np.random.default_rng(2)
    d = rng.uniform(-2, 2, (4, 4))
    d = (d + d.T) / 2
    center = np.array([0.5, -0.3, 0.2])
    oh = orc.Ham(lat, n)
    oh.set_potential(v)
    oh.add_projector_gauss(center, 1.6, 4, 0.7, d)
    g = oh.gsphere((0, 0, 0), 5.0)
    c = S.random_states(len(g), 4, 8, abs(np.linalg.det(lat)) / 4096.0 ** 2)
    v.astype(np.float64).tofile(tmp_path / "v.bin")
    np.concatenate([center, d.ravel()]).tofile(tmp_path / "proj.bin")
    c.astype(np.complex128).tofile(tmp_path / "phi.bin")
    out = subprocess.run([EXE, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    hphi = np.fromfile(tmp_path / "hphi.bin", dtype=np.complex128).reshape(4, len(g))
    assert rel_err(hphi, oh.apply_sphere(c, g)) < 1e-11
    half = np.fromfile(tmp_path / "phi_half.bin", dtype=np.complex128).reshape(4, len(g))
    assert rel_err(half, oh.exponential_in_place(c, g, -0.025j, 4)) < 1e-11
    e = np.fromfile(tmp_path / "energy.bin")
    want = oh.energy_sphere(c, g, np.full(4, 2.0))
    assert np.max(np.abs(e - want) / np.maximum(np.abs(want), 1.0)) < 1e-10
