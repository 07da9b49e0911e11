"""The plain filter step at its full benchmark size -- 4096 parents, S = 200,
200 landmarks, clutter lambda 50, FP32 storage -- checked through a property
that holds at every size.

Each parent's map update depends only on that parent's own state and on
noise counters that carry its parent index (and the Gaussian index within
the run's layout, which is the same for the first parents of any run). The
first K parents of the 4096-parent GPU run therefore equal a K-parent oracle
run from the same seed, as long as no resampling mixes the parents. Frame 0
has a fixed camera and births only, so the weights stay uniform and nothing
is resampled. Frame 1 moves the camera and brings detection, non-detection,
merge and birth terms. For parents 0 .. K-1 the test compares:
  * bit-exact: survivor, merged and kept non-detection counts, and the
    resampling decision of frame 0;
  * within tolerance: predicted parent states (1e-13), normalisers eta
    (1e-11 relative), parent log-likelihoods (1e-11 relative, 1e-9
    absolute), and the parent weights as ratios to parent 0 (1e-8
    relative). Ratios, because the GPU normalises over 4096 parents and the
    oracle over K.
Every launch, scan and compaction runs at its full-size shape (4096
workgroups per per-parent kernel, ~10^6 Gaussians after frame 1).
"""
import numpy as np
import pytest

import oracle
import scene
import workloads

pytestmark = pytest.mark.gpu

K = 24


def _close(a, b, rtol, atol, what):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    assert a.shape == b.shape, f"{what}: shape {a.shape} vs {b.shape}"
    err = np.abs(a - b)
    lim = atol + rtol * np.maximum(np.abs(a), np.abs(b))
    bad = err > lim
    assert not bad.any(), (f"{what}: {bad.sum()} of {a.size} differ; worst {err.max():.3e} at "
                           f"{int(np.argmax(err))} ({a.flat[np.argmax(err)]!r} vs {b.flat[np.argmax(err)]!r})")


def test_first_parents_of_the_full_size_step_equal_a_small_oracle_run():
    import scphd_amd

    p, s = workloads.workload("c1")
    assert p["n_particles"] == 4096 and p["particles_per_feature"] == 200 and p["storage"] == 0
    sc = scene.generate(seed=p["seed"], **dict(s, n_frames=2))
    gpu = scphd_amd.Filter(p)
    orc = oracle.Oracle(dict(p, n_particles=K), threads=16)
    try:
        sg, wg = gpu.parents()
        so, wo = orc.parents()
        _close(sg[:K], so, 1e-14, 1e-15, "prior")
        for k in range(2):
            u, v, fixed = scene.frame_measurements(sc, k)
            gpu.predict(s["dt"])
            orc.predict(s["dt"])
            _close(gpu.parents()[0][:K], orc.parents()[0], 1e-13, 1e-13, f"frame {k} predicted parents")
            gpu.update(u, v, fixed)
            orc.update(u, v, fixed)
            dg, do = gpu.update_debug(), orc.update_debug()
            assert dg["eta"].shape == (4096, len(u))
            for key in ("nsurvivors", "nmerged", "nnondetect"):
                np.testing.assert_array_equal(dg[key][:K], do[key], err_msg=f"frame {k} {key}")
            _close(dg["eta"][:K], do["eta"], 1e-11, 1e-11, f"frame {k} eta")
            _close(dg["loglik"][:K], do["loglik"], 1e-11, 1e-9, f"frame {k} loglik")
            wg, wo = gpu.parents()[1], orc.parents()[1]
            assert np.isfinite(wg).all() and abs(wg.sum() - 1.0) < 1e-12
            _close(wg[:K] / wg[0], wo / wo[0], 1e-8, 0.0, f"frame {k} parent weight ratios")
            if k == 0:
                assert fixed
                rg, ro = gpu.resample(), orc.resample()
                assert rg == ro == 0, "frame 0 resampled: parents mixed, the slice comparison would not hold"
        c = gpu.counts()
        print(f"full-size step: {c['n_gaussians']} Gaussians over 4096 parents after frame 1, "
              f"{len(u)} measurements")
        assert c["n_gaussians"] > 4096 * 100
    finally:
        gpu.close()
