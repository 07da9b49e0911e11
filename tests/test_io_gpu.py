"""MAT output of the filter (writeMat) and the command-line driver, on the GPU.

* scphd_write_mat: the variables the reference's writeMat produces
  (particles, weights, map_weights, map_particles, map_particle_weights)
  read back with scipy and compared with the state read through the ABI.
* bin/scphd_camera_calibration cfg data.mat outdir: the reference main loop
  over a measurement sequence stored as MAT cells Z1 / Z2; prior.mat and
  k.mat are written for every frame.
"""
import os
import subprocess

import numpy as np
import pytest

import scene
import workloads

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(REPO, "scphd-camera-calibration_amd", "bin", "scphd_camera_calibration")


def _small():
    p, s = workloads.workload("c0", n_particles=32, particles_per_feature=40, n_frames=4)
    p["seed"] = 21
    return p, s


def test_write_mat_holds_parents_and_the_best_parents_map(tmp_path):
    import scipy.io

    import scphd_amd

    p, s = _small()
    sc = scene.generate(seed=p["seed"], **s)
    f = scphd_amd.Filter(p)
    for k in range(3):
        u, v, fixed = scene.frame_measurements(sc, k)
        f.predict(s["dt"])
        f.update(u, v, fixed)
        f.resample()
    f.predict(s["dt"])  # a deferred prediction is materialised before writing
    path = str(tmp_path / "k.mat")
    f.write_mat(path)
    m = scipy.io.loadmat(path)
    st, w = f.parents()
    gc, wg, smp = f.maps()
    np.testing.assert_array_equal(m["particles"], st.T)
    np.testing.assert_array_equal(m["weights"].ravel(), w)
    best = int(np.argmax(w))
    off = np.concatenate([[0], np.cumsum(gc)])
    S = p["particles_per_feature"]
    np.testing.assert_array_equal(m["map_weights"].ravel(), wg[off[best]:off[best + 1]])
    lo, hi = off[best] * S, off[best + 1] * S
    np.testing.assert_array_equal(m["map_particles"].reshape(6, -1), smp[:6, lo:hi])
    np.testing.assert_array_equal(m["map_particle_weights"].ravel(), smp[6, lo:hi])
    f.close()


def test_command_line_driver_runs_the_reference_main_loop(tmp_path):
    import scipy.io

    import scphd_amd

    p, s = _small()
    p["max_steps"] = 4
    cfg = tmp_path / "run.cfg"
    cfg.write_text(workloads.to_cfg(p))
    sc = scene.generate(seed=p["seed"], **s)
    z1 = np.empty((1, 4), dtype=object)
    z2 = np.empty((1, 4), dtype=object)
    for k in range(4):
        # frame k of each camera; the loop uses Z1 on even k, Z2 on odd k
        u1, v1, _ = scene.frame_measurements(sc, k - k % 2)
        u2, v2, _ = scene.frame_measurements(sc, k - k % 2 + 1)
        z1[0, k] = np.vstack([u1, v1])
        z2[0, k] = np.vstack([u2, v2])
    data = tmp_path / "seq.mat"
    scipy.io.savemat(str(data), {"Z1": z1, "Z2": z2})
    assert len(scphd_amd.read_measurements(str(data), "Z1")) == 4
    out = tmp_path / "out"
    out.mkdir()
    r = subprocess.run([CLI, str(cfg), str(data), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    for name in ["prior.mat"] + [f"{k}.mat" for k in range(4)]:
        m = scipy.io.loadmat(str(out / name))
        assert m["particles"].shape == (12, p["n_particles"])
        assert abs(m["weights"].sum() - 1.0) < 1e-9
    assert scipy.io.loadmat(str(out / "3.mat"))["map_weights"].size > 0


def test_monte_carlo_batch_runner_matches_runs_done_alone(tmp_path):
    import scipy.io

    import batch_run
    import scphd_amd

    p, s = _small()
    p["max_steps"] = 4
    cfg = tmp_path / "run.cfg"
    cfg.write_text(workloads.to_cfg(p))
    root = tmp_path / "data"
    seqs = {}
    for n in (1, 2):
        sc = scene.generate(seed=100 + n, **dict(s, n_frames=4))
        z = [scene.frame_measurements(sc, k)[:2] for k in range(4)]
        cells = np.empty((1, 4), dtype=object)
        for k in range(4):
            cells[0, k] = np.vstack(z[k])
        (root / str(n)).mkdir(parents=True)
        scipy.io.savemat(str(root / str(n) / "data.mat"), {"Z1": cells, "Z2": cells})
        seqs[n] = z
    res = tmp_path / "results"
    assert batch_run.main([str(cfg), str(root), "2", str(res), "--workers", "2"]) == 0
    for n in (1, 2):
        files = sorted(os.listdir(res / str(n)))
        assert files == ["0.mat", "1.mat", "2.mat", "3.mat", "prior.mat"]
        # 3.mat is written after the update of frame 3, before its resampling
        m = scipy.io.loadmat(str(res / str(n) / "3.mat"))
        g = scphd_amd.Filter(dict(p, seed=p["seed"] + n))
        for k in range(4):
            g.predict(0.5)
            g.update(*seqs[n][k], k % 2 == 0)
            if k < 3:
                g.resample()
        st, w = g.parents()
        np.testing.assert_array_equal(m["particles"], st.T)
        np.testing.assert_array_equal(m["weights"].ravel(), w)
        g.close()
