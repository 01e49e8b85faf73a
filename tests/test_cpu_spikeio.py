"""Spike file formats (brian_amd/spikeio.py) on the host."""
import numpy as np
import pytest


def test_raster_round_trip(tmp_path):
    from brian_amd.spikeio import save_raster, load_raster
    rng = np.random.default_rng(1)
    steps = rng.integers(0, 1000, 500).astype(np.int32)
    idx = rng.integers(0, 77, 500).astype(np.int32)
    p = tmp_path / "r.bmr"
    save_raster(p, steps, idx, 1e-4, 77)
    s, i, dt, N = load_raster(p)
    o = np.lexsort((idx, steps))
    np.testing.assert_array_equal(s, steps[o])
    np.testing.assert_array_equal(i, idx[o])
    assert dt == 1e-4 and N == 77
    raw = p.read_bytes()
    assert raw[:8] == b"BMRS0001" and len(raw) == 32 + 8 * 500


def test_aer_round_trip_and_layout(tmp_path):
    """AER-DAT 2.0 as tools/io.py:58-185 reads it: header lines, then
    big-endian (address, tick) pairs; times = ticks * tick."""
    from brian_amd.spikeio import save_aer, load_aer
    p = str(tmp_path / "x.aedat")
    save_aer(p, [3, 1, 4], [10, 11, 12], 1e-4)
    a, t = load_aer(p)
    np.testing.assert_array_equal(a, [3, 1, 4])
    np.testing.assert_allclose(t, [10e-4, 11e-4, 12e-4], rtol=1e-15)
    raw = open(p, "rb").read()
    body = raw.split(b"\n", 5)[5]                   # after the five header lines
    assert body[:8] == bytes([0, 0, 0, 3, 0, 0, 0, 10])
    with pytest.raises(ValueError):
        load_aer(str(tmp_path / "x.txt"))


def test_py2_float_str():
    """FileSpikeMonitor writes t with Python 2's str(float)."""
    from brian_amd.spikeio import py2_float_str
    assert py2_float_str(0.1) == "0.1"
    assert py2_float_str(1.0) == "1.0"
    assert py2_float_str(0.0) == "0.0"
    assert py2_float_str(3 * 0.1 * 0.001) == "0.0003"
    assert py2_float_str(1e-05) == "1e-05"
    assert py2_float_str(0.30000000000000004) == "0.3"
