"""The plain filter step at full benchmark sizes, checked through a property
that holds at every size:
  * 4096 parents, S = 200, 200 landmarks, clutter lambda 50 (the bench
    workload), with FP32 and with FP64 sample storage;
  * 16384 parents, S = 100, 300 landmarks, clutter lambda 100 (the precision
    study), with FP32 and with FP64 sample storage.

Each parent's map update depends only on that parent's own state and on
noise counters that carry its parent index (and the Gaussian index within
the run's layout, which is the same for the first parents of any run). The
first K parents of the full-size GPU run therefore equal a K-parent oracle
run from the same seed, as long as no resampling mixes the parents. Frame 0
has a fixed camera and births only, so the weights stay uniform and nothing
is resampled. Frame 1 moves the camera and brings detection, non-detection,
merge and birth terms. Frame 2 follows without resampling, with an empty
measurement set (non-detection terms only). For parents 0 .. K-1 the test
compares after every frame:
  * bit-exact: survivor, merged and kept non-detection counts, and the
    resampling decision of frame 0;
  * within tolerance: predicted parent states (1e-13), normalisers eta
    (1e-11 relative), parent log-likelihoods (1e-11 relative, 1e-9
    absolute), and the parent weights as ratios to parent 0 (1e-8
    relative). Ratios, because the GPU normalises over all parents and the
    oracle over K.
Every launch, scan and compaction runs at its full-size shape (one
workgroup per parent in the per-parent kernels, 1-10 million Gaussians
after frame 1).

Resampling mixes all parents, so it is checked on its own: from the same
weights and counters at 4096, 16384 and 65536 parents, the decision and
every ancestor index equal the oracle's.
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


@pytest.mark.parametrize("name,storage,parents,samples", [("c1", 0, 4096, 200), ("c1", 1, 4096, 200),
                                                            ("c3", 0, 16384, 100), ("c3", 1, 16384, 100)],
                         ids=["c1-fp32", "c1-fp64", "c3-fp32", "c3-fp64"])
def test_first_parents_of_the_full_size_step_equal_a_small_oracle_run(name, storage, parents, samples):
    import scphd_amd

    p, s = workloads.workload(name, storage=storage)
    assert p["n_particles"] == parents and p["particles_per_feature"] == samples
    sc = scene.generate(seed=p["seed"], **dict(s, n_frames=2))
    gpu = scphd_amd.Filter(p)
    orc = oracle.Oracle(dict(p, n_particles=K), threads=16)
    try:
        sg, wg = gpu.parents()
        so, wo = orc.parents()
        _close(sg[:K], so, 1e-14, 1e-15, "prior")
        for k in range(3):
            u, v, fixed = scene.frame_measurements(sc, min(k, 1))
            if k == 2:  # an empty measurement set: non-detection terms only
                u, v = np.zeros(0), np.zeros(0)
            gpu.predict(s["dt"])
            orc.predict(s["dt"])
            _close(gpu.parents()[0][:K], orc.parents()[0], 1e-13, 1e-13, f"frame {k} predicted parents")
            gpu.update(u, v, fixed)
            orc.update(u, v, fixed)
            dg, do = gpu.update_debug(), orc.update_debug()
            assert dg["eta"].shape == (parents, len(u))
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
        print(f"full-size step: {c['n_gaussians']} Gaussians over {parents} parents after frame 2 "
              f"(empty measurement set)")
        assert c["n_gaussians"] > parents * 100
    finally:
        gpu.close()


@pytest.mark.parametrize("parents", [4096, 16384, 65536])
@pytest.mark.parametrize("weights", ["spread", "near-uniform", "zeros-and-one-heavy"])
def test_resampling_at_full_size_equals_the_oracle(parents, weights):
    """Systematic resampling over all parents at once (one FP64 prefix sum,
    one uniform from the resampling stream): the decision and every ancestor
    index equal the oracle's, from the same weights and counters.  The maps
    are empty, because resampling reads only the parent weights.  Weight
    sets: log-normal spread (resamples), near-uniform (n_eff above the
    threshold: no resampling), and a set with zero weights and one parent
    holding half the mass."""
    import scphd_amd

    p, _ = workloads.workload("c1", n_particles=parents)
    rng = np.random.default_rng(parents)
    if weights == "spread":
        w = rng.lognormal(0.0, 2.0, parents)
    elif weights == "near-uniform":
        w = 1.0 + 0.01 * rng.random(parents)
    else:
        w = rng.random(parents)
        w[rng.random(parents) < 0.3] = 0.0
        w[parents // 3] = w.sum()
    w /= w.sum()
    st = dict(states=rng.normal(0.0, 0.1, (parents, 12)), weights=w, gcount=np.zeros(parents, dtype=np.int32),
              wg=np.zeros(0), samples=np.zeros((7, 0)), n_predict=3, n_update=3, n_resample=2)
    gpu = scphd_amd.Filter(p)
    orc = oracle.Oracle(p, threads=16)
    try:
        gpu.set_state(st)
        orc.set_state(dict(st, wg=np.zeros(1), samples=np.zeros((7, 1))))
        rg, ro = gpu.resample(), orc.resample()
        assert rg == ro == (weights != "near-uniform")
        if ro:
            np.testing.assert_array_equal(gpu.ancestors(), orc.ancestors()[0])
        sg, wg = gpu.parents()
        so, wo = orc.parents()
        np.testing.assert_array_equal(sg, so)
        np.testing.assert_array_equal(wg, wo)
    finally:
        gpu.close()
