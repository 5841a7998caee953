"""The exchange singularity correction on the device against the reference's
known answers (singularity_correction.hpp:132-145: a = 10.18 bohr cubic cell,
2x2x2 k-point grid) and against the oracle on a skewed cell with 3 k-points;
the Gamma-only zeroterm of ExchangeOperator(zeroterm="singularity") is V fzero."""
import numpy as np
import pytest

from inq_amd import Context
from inq_amd import exchange as exx
from oracle import oracle as orc
from tests.helpers import say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return Context(0)


def test_singularity_known_answers(ctx):
    say("case: singularity correction KATs")
    aa = 10.18
    lat = np.eye(3) * aa
    ks = np.array([[a, b, c] for a in (0, 1) for b in (0, 1) for c in (0, 1)], dtype=float) * 0.5
    ks = np.where(ks >= 0.5, ks - 1.0, ks) * 2 * np.pi
    fk, fz = exx.singularity_correction(ctx, lat, ks, np.full(8, 1 / 8))
    assert fz == pytest.approx(0.30983869660201141, rel=1e-10)
    assert np.allclose(fk, 0.18644848345224296, rtol=1e-10, atol=0)
    z = exx.singularity_zeroterm(ctx, lat, ks, np.full(8, 1 / 8), ik=0)
    assert z == pytest.approx(1041.3915164701, rel=1e-9)


def test_singularity_matches_oracle_skewed(ctx):
    say("case: singularity correction skewed cell")
    lat = np.array([[7.0, 0.5, 0.0], [0.3, 8.0, 0.2], [0.0, -0.4, 6.5]])
    ks = np.array([[0.0, 0.0, 0.0], [np.pi, 0.0, 0.0], [0.0, -np.pi, np.pi]])
    w = np.array([0.5, 0.25, 0.25])
    fk, fz = exx.singularity_correction(ctx, lat, ks, w)
    fk_o, fz_o = orc.sing_correction(lat, ks, w)
    assert fz == pytest.approx(fz_o, rel=1e-12)
    assert np.allclose(fk, fk_o, rtol=1e-13, atol=0)
    # Gamma only: fk = 0 and the zeroterm is V fzero
    fk1, fz1 = exx.singularity_correction(ctx, lat)
    assert fk1[0] == 0.0
    vol = abs(np.linalg.det(lat))
    assert exx.singularity_zeroterm(ctx, lat) == pytest.approx(vol * fz1, rel=1e-15)
