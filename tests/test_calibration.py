"""The filter calibrates: end-to-end accuracy of the FP64 oracle against the
scene truth on the 30-landmark scene (256 parents, S = 100, 50 frames at
dt = 0.5, cameras alternating as in src/main.cpp:83-91).

The reference's own accuracy measure is the weighted parent mean against
the true pose (load_disparity_results.m:27-28,66) and the map cardinality
(:34).  What the (u, v) measurements determine is the right camera's
rotation relative to the left one: with a free landmark map, scaling the
world and the baseline together leaves every projection unchanged, so the
translation's length is fixed only by the priors and is not asserted.

Asserted over scene seeds 1-4 (the first four; per-seed figures of seeds
1-8 ranged from 0.20 to 0.68 of the prior's error, 0.42 on average):
  * the rotation error of the weighted parent mean, averaged over the last
    10 frames, is below 0.85 of the prior mean's error for every seed, and
    below 0.6 of it on average over the seeds.  256 parents spread over
    the prior's +-3 sigma box leave the nearest parent about a sigma away;
    the error then falls by resampling and the camera model's random walk,
    not to the measurement precision (1 px ~ 0.1 deg at f = 500 px);
  * the best parent's map carries about one unit of weight per landmark in
    view: the sum of its Gaussian weights is within 15 % of the number of
    landmarks inside the left camera's field of view.
"""
import numpy as np

import oracle
import scene
import workloads


def _visible(sc, k, f=500.0):
    u, v, d = scene.project(sc["truth"][k], np.zeros(12), f, 320.0, 240.0)
    return int(((u >= 0) & (u <= 640) & (v >= 0) & (v <= 480) & (d >= 0)).sum())


def _run(seed):
    p, s = workloads.workload("c0")
    sc = scene.generate(seed=seed, **s)
    o = oracle.Oracle(p, threads=8)
    truth = sc["ext2"][[6, 7, 8]]
    errs = []
    for k in range(s["n_frames"]):
        u, v, fixed = scene.frame_measurements(sc, k)
        o.predict(s["dt"])
        o.update(u, v, fixed)
        st, w = o.parents()
        errs.append(np.linalg.norm((w[:, None] * st).sum(0)[[6, 7, 8]] - truth))
        if k == s["n_frames"] - 1:
            gc, wg, _ = o.maps()
            off = np.concatenate([[0], np.cumsum(gc)])
            b = int(np.argmax(w))
            card = wg[off[b]:off[b + 1]].sum()
        o.resample()
    mid = (np.array(p["prior_a"][6:9]) + np.array(p["prior_b"][6:9])) / 2
    prior_err = np.linalg.norm(mid - truth)
    return prior_err, float(np.mean(errs[-10:])), card, _visible(sc, s["n_frames"] - 1)


def test_oracle_calibrates_rotation_and_builds_the_map():
    ratios = []
    for seed in (1, 2, 3, 4):
        prior_err, late, card, vis = _run(seed)
        ratios.append(late / prior_err)
        print(f"seed {seed}: rotation error prior {np.degrees(prior_err):.3f} deg, last 10 frames "
              f"{np.degrees(late):.3f} deg; best map weight {card:.1f} with {vis} landmarks in view")
        assert late <= 0.85 * prior_err, f"seed {seed}"
        assert abs(card - vis) <= 0.15 * vis, f"seed {seed}"
    assert np.mean(ratios) <= 0.6
