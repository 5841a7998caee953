"""Restart files of fields and arrays (host side, no GPU): the reference's raw
format (src/operations/io.hpp:104-162) -- one headerless native-endian file
per component named with "%010ld" (utils::num_to_str, whose own known answers
are src/utils/num_str.hpp:75-77), and dirname/array for plain arrays."""
import os

import numpy as np
import pytest

from inq_amd import Context, InqxError
from inq_amd import io as inqio


@pytest.fixture(scope="module")
def hctx():
    return Context(-1)  # host-only context: no device is touched


def test_field_round_trip_and_layout(hctx, tmp_path):
    rng = np.random.default_rng(3)
    v = rng.standard_normal((1000, 2))
    d = tmp_path / "spin_density"
    inqio.save_field(hctx, d, v)
    assert sorted(os.listdir(d)) == ["0000000000", "0000000001"]
    for c in range(2):
        raw = np.fromfile(d / f"{c:010d}", dtype=np.float64)
        assert raw.shape == (1000,) and np.array_equal(raw, v[:, c])
    assert np.array_equal(inqio.load_field(hctx, d, 1000, 2), v)


def test_array_round_trip(hctx, tmp_path):
    occ = np.array([2.0, 2.0, 1.5, 0.0])
    d = tmp_path / "kpin0000000000" / "occupations"
    inqio.save_array(hctx, d, occ)
    assert os.listdir(d) == ["array"]
    assert np.array_equal(np.fromfile(d / "array", dtype=np.float64), occ)
    assert np.array_equal(inqio.load_array(hctx, d, 4), occ)


def test_missing_or_short_files_are_statuses(hctx, tmp_path):
    with pytest.raises(InqxError):
        inqio.load_array(hctx, tmp_path / "nothing", 4)
    d = tmp_path / "short"
    inqio.save_array(hctx, d, np.ones(3))
    with pytest.raises(InqxError):
        inqio.load_array(hctx, d, 4)
