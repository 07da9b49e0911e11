"""The reference's own input sequences (data/*.mat: disparity_sim,
disparity_sim_noclutter, fov_test, spiral; cells Z1 and Z2, the first 100
frames, the configuration's max_steps) as committed fixtures
(tests/golden/reference_sequences.npz, written by
tests/golden/make_reference_sequences.py), run through the reference main
loop (src/main.cpp:57-95): predict(0.5); update with Z1[k] and the fixed
camera on even k, with Z2[k] and the moving camera on odd k; checkStuff;
resample.  Configuration: the reference's config/calibration.cfg.

CPU: the fixtures equal what the library's MAT reader reads from the
reference's files (when present), and the oracle runs every sequence.  The
GPU comparison is tests/test_reference_sequences_gpu.py.
"""
import os

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
REF_DATA = "/root/reference/data"
FILES = ["disparity_sim", "disparity_sim_noclutter", "fov_test", "spiral"]


def sequence(name):
    """Frames of one file: list of (u, v, fixed) in main.cpp's order."""
    z = np.load(os.path.join(GOLDEN, "reference_sequences.npz"))
    cells = {}
    for var in ("Z1", "Z2"):
        counts = z[f"{name}_{var}_counts"]
        uv = z[f"{name}_{var}_uv"]
        off = np.concatenate([[0], np.cumsum(counts)])
        cells[var] = [(uv[0, off[k]:off[k + 1]].copy(), uv[1, off[k]:off[k + 1]].copy()) for k in range(len(counts))]
    n = min(len(cells["Z1"]), len(cells["Z2"]))
    return [(*cells["Z1"][k], True) if k % 2 == 0 else (*cells["Z2"][k], False) for k in range(n)]


def reference_params():
    import scphd_amd

    return scphd_amd.load_params(os.path.join(GOLDEN, "calibration.cfg"))


@pytest.mark.skipif(not os.path.isdir(REF_DATA), reason="the reference's data files are not present")
@pytest.mark.parametrize("name", FILES)
def test_fixture_equals_the_library_reader(name):
    import scphd_amd

    z = np.load(os.path.join(GOLDEN, "reference_sequences.npz"))
    for var in ("Z1", "Z2"):
        frames = scphd_amd.read_measurements(os.path.join(REF_DATA, name + ".mat"), var)[:100]
        counts = z[f"{name}_{var}_counts"]
        uv = z[f"{name}_{var}_uv"]
        assert len(frames) == len(counts)
        np.testing.assert_array_equal([len(u) for u, _ in frames], counts)
        np.testing.assert_array_equal(np.concatenate([u for u, _ in frames]), uv[0])
        np.testing.assert_array_equal(np.concatenate([v for _, v in frames]), uv[1])


@pytest.mark.parametrize("name", FILES)
def test_oracle_runs_the_reference_main_loop(name):
    p = reference_params()
    p["n_particles"] = 32
    o = oracle.Oracle(p, threads=8)
    for u, v, fixed in sequence(name)[:6]:
        o.predict(0.5)
        o.update(u, v, fixed)
        assert o.check() == 0
        o.resample()
    st, w = o.parents()
    assert np.isfinite(st).all() and abs(w.sum() - 1.0) < 1e-12
    assert o.counts()["n_gaussians"] > 0
