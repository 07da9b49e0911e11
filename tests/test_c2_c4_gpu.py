"""The batched (C2) and served (C4) workloads on the GPU.

C2: the 64-rig ragged draw (parents in {256 .. 4096}, landmarks 20-300,
clutter lambda 5-80, baseline 0.1-0.5 m, f 400-800 px, S = 100), every rig
with its own seed from the master seed.  Two frames (births with the fixed
camera, then the moving camera with detection, non-detection, merge and
birth terms) run
  * in ensembles (one launch per kernel for a group of rigs), and
  * through the batch API (one handle per rig, frames of different rigs
    overlapping on the device).
A rig's parent p updates from its own state and noise counters only, and
min_neff = 0 keeps resampling from mixing parents, so each rig's first K
parents must equal a K-parent oracle run of that rig (its own f, baseline,
lambda, landmarks and seed): survivor, merged and non-detection counts
exactly, normalisers within 1e-11 relative, log-likelihoods within 1e-11
relative (1e-9 absolute), weights as ratios to parent 0 within 1e-8.

C4: rigs arriving as a Poisson process at 20 rigs/s for half a second,
parents log-uniform in [256, 8192], landmarks 20-400, lambda 5-100, S = 100,
frames at 30 Hz for 0.1-0.3 s each (3-9 frames), served with a 200 GB HBM
budget through the batch API (one handle per rig) and through phase cohorts
(shared launches, rigs joining and leaving ensembles between ticks; paced
in real time and saturated): every rig completes, and every per-frame
calibration estimate equals the one the rig gives run alone on a plain
filter, bit for bit.
"""
import numpy as np
import pytest

import oracle
import scene
import workloads

pytestmark = pytest.mark.gpu

K = 8


def _close(a, b, rtol, atol, what):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    assert a.shape == b.shape, f"{what}: shape {a.shape} vs {b.shape}"
    err = np.abs(a - b)
    bad = err > atol + rtol * np.maximum(np.abs(a), np.abs(b))
    assert not bad.any(), f"{what}: {bad.sum()} of {a.size} differ; worst {err.max():.3e}"


def _c2(frames=2):
    out = []
    for p, s in workloads.c2_rigs(master=1, n_rigs=64, samples=100, n_frames=frames):
        p = dict(p, min_neff=0.0)
        sc = scene.generate(seed=p["seed"], **s)
        out.append((p, s, [scene.frame_measurements(sc, k) for k in range(frames)]))
    return out


def _oracle_slices(rigs):
    """Per rig, per frame: the K-parent oracle's update debug and weights."""
    ref = []
    for p, s, sets in rigs:
        o = oracle.Oracle(dict(p, n_particles=K), threads=8)
        fr = []
        for u, v, fixed in sets:
            o.predict(s["dt"])
            o.update(u, v, fixed)
            fr.append((o.update_debug(), o.parents()[1].copy()))
            assert not o.resample()
        ref.append(fr)
    return ref


def _compare(what, dg, loglik, weights, M, do, wo):
    for key in ("nsurvivors", "nmerged", "nnondetect"):
        np.testing.assert_array_equal(dg[key], do[key], err_msg=f"{what} {key}")
    _close(dg["eta"][:, :M], do["eta"], 1e-11, 1e-11, f"{what} eta")
    _close(loglik, do["loglik"], 1e-11, 1e-9, f"{what} loglik")
    _close(weights / weights[0], wo / wo[0], 1e-8, 0.0, f"{what} weight ratios")


def test_c2_rigs_in_ensembles_equal_their_oracle_slices():
    import scphd_amd

    rigs = _c2()
    ref = _oracle_slices(rigs)
    total = sum(p["n_particles"] for p, _, _ in rigs)
    print(f"C2: 64 rigs, {total} parents")
    for g0 in range(0, 64, 16):  # four ensembles of 16 rigs
        grp = rigs[g0:g0 + 16]
        e = scphd_amd.Ensemble([p for p, _, _ in grp])
        try:
            poff = e.poff
            for k in range(2):
                e.predict(grp[0][1]["dt"])
                e.update([sets[k][0] for _, _, sets in grp], [sets[k][1] for _, _, sets in grp], grp[0][2][k][2])
                d = e.update_debug()
                w = e.parents()[1]
                for j, (p, s, sets) in enumerate(grp):
                    a = poff[j]
                    M = len(sets[k][0])
                    dj = {key: d[key][a:a + K] for key in ("nsurvivors", "nmerged", "nnondetect")}
                    dj["eta"] = d["eta"][a:a + K]
                    _compare(f"rig {g0 + j} frame {k}", dj, d["loglik"][a:a + K], w[a:poff[j + 1]][:K], M,
                             *ref[g0 + j][k])
                e.resample()
        finally:
            e.close()


def test_c2_rigs_in_the_batch_equal_their_oracle_slices():
    import scphd_amd

    rigs = _c2()
    ref = _oracle_slices(rigs)
    b = scphd_amd.Batch(n_workers=8)
    try:
        for g0 in range(0, 64, 16):  # admitted 16 at a time
            # reservations for two frames' maps (at most about two terms per
            # landmark and clutter point per parent)
            ids = [b.add_rig(p, scphd_amd.rig_footprint(p, 2 * (s["n_landmarks"] + int(s["lam"])) + 64))
                   for p, s, _ in rigs[g0:g0 + 16]]
            assert None not in ids
            for j, rid in enumerate(ids):
                p, s, sets = rigs[g0 + j]
                b.submit(rid, 0, s["dt"], *sets[0])
            assert b.step() == 16
            for j, rid in enumerate(ids):
                f = b.filter(rid)
                d, w = f.update_debug(), f.parents()[1]
                dj = {key: d[key][:K] for key in ("nsurvivors", "nmerged", "nnondetect", "eta")}
                _compare(f"rig {g0 + j} frame 0", dj, d["loglik"][:K], w[:K], len(rigs[g0 + j][2][0][0]),
                         *ref[g0 + j][0])
            for j, rid in enumerate(ids):
                p, s, sets = rigs[g0 + j]
                b.submit(rid, 1, s["dt"], *sets[1])
            assert b.step() == 16
            res = b.collect()
            assert all(x["status"] == 0 for x in res)
            for j, rid in enumerate(ids):
                f = b.filter(rid)
                d, w = f.update_debug(), f.parents()[1]
                dj = {key: d[key][:K] for key in ("nsurvivors", "nmerged", "nnondetect", "eta")}
                _compare(f"rig {g0 + j} frame 1", dj, d["loglik"][:K], w[:K], len(rigs[g0 + j][2][1][0]),
                         *ref[g0 + j][1])
                b.remove_rig(rid)
    finally:
        b.close()


_ALONE = {}


def _c4_burst():
    """A C4 arrival burst (20 rigs/s for half a second, C4's size ranges,
    3-9 frames each) and every rig's per-frame estimates run alone on a
    plain filter."""
    import scphd_amd
    import serve

    if "rigs" not in _ALONE:
        arrivals = workloads.c4_arrivals(master=3, rate=20.0, horizon=0.5, samples=100, dur=(0.1, 0.3))
        rigs = serve.prepare(arrivals)
        alone = []
        for r in rigs:
            f = scphd_amd.Filter(r["p"])
            est = []
            for u, v, fixed in r["frames"]:
                f.predict(r["s"]["dt"])
                f.update(u, v, fixed)
                assert f.check() == 0
                f.resample()
                est.append(f.estimate()[0])
            f.close()
            alone.append(np.array(est))
        _ALONE["rigs"], _ALONE["alone"] = rigs, alone
    return _ALONE["rigs"], _ALONE["alone"]


@pytest.mark.parametrize("paced", [False, True], ids=["saturated", "paced"])
def test_c4_cohorts_share_launches_and_equal_the_alone_runs(paced):
    """The served form with shared launches (serve.run_cohorts): the C4
    burst through four phase cohorts under a 200 GB budget, rigs joining
    and leaving their cohort's ensemble between ticks.  Every rig completes
    and every per-frame estimate equals its alone run's, bit for bit."""
    import serve

    rigs, alone = _c4_burst()
    out = serve.run_cohorts(rigs, hbm_budget_bytes=200e9, cohorts=4, paced=paced, keep_estimates=True,
                            max_wall=120.0)
    print(f"C4 cohorts ({'paced' if paced else 'saturated'}): {out['rigs']} rigs, {out['frames']} frames in "
          f"{out['ticks']} ticks, {out['updates_per_s']:.3e} updates/s, latency p50 {out['latency_ms']['p50']:.1f} "
          f"p99 {out['latency_ms']['p99']:.1f} ms, peak {out['peak_rigs_per_cohort']} rigs per cohort")
    assert out["frames"] == sum(len(r["frames"]) for r in rigs)
    assert out["ticks"] < out["frames"]  # launches are shared
    for i, est in enumerate(alone):
        np.testing.assert_array_equal(out["estimates"][i], est, err_msg=f"rig {i}")


def test_c4_served_rigs_complete_and_equal_their_alone_runs():
    import scphd_amd
    import serve

    rigs, alone = _c4_burst()
    assert len(rigs) >= 5
    # reservations at half the steady-state estimate: these rigs stream only
    # 3-9 frames, far from their steady-state maps
    out = serve.run_served(rigs, hbm_budget_bytes=200e9, n_workers=8, reserve_factor=0.5, paced=False,
                           max_wall=150.0, keep_estimates=True)
    print(f"C4 burst: {out['rigs']} rigs, {out['frames']} frames, peak {out['peak_active_rigs']} active, "
          f"{out['updates_per_s']:.3e} updates/s, budget {out['budget_gb']:.0f} GB, "
          f"peak reserved {out['peak_reserved_gb']:.1f} GB")
    assert out["rigs_done"] == len(rigs)
    assert out["frames"] == sum(len(r["frames"]) for r in rigs)
    for i, est in enumerate(alone):
        np.testing.assert_array_equal(out["estimates"][i], est, err_msg=f"rig {i}")
