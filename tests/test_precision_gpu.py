"""Precision study: GPU filters against the FP64 oracle.

Workload: the precision-study scene (300 landmarks, clutter lambda = 100,
S = 100) at a parent count the oracle finishes in seconds (128).  Three
filters start from one seed and run predict -> update -> resample freely for
FRAMES frames:
  * the oracle, FP64 samples (the reference arithmetic);
  * the GPU with FP64 sample storage ("all-FP64 GPU");
  * the GPU with FP32 sample storage (the plain path: FP32 bulk, FP64
    likelihood, log-sum-exp and resampling prefix sum).
Printed per frame and output (pytest -s): extrinsic estimate (weighted
parent mean) relative error, parent log-likelihood error, Gaussian counts,
ancestor agreement.

Asserted:
  * all-FP64 GPU, free running: identical map sizes and ancestors every
    frame, extrinsics within 1e-8 relative;
  * FP32 storage, free running: extrinsics within 1e-4 relative of the
    oracle in every frame before its first differing decision (a map size
    or an ancestor).  FP32 rounding of the clouds moves decisions that sit
    near a threshold (a survivor near min_weight, a merge pair near
    min_separation); once one differs, the parent likelihoods -- sums over
    hundreds of sharply peaked detection terms -- and then the resampled
    parents follow another trajectory: the filters agree in distribution,
    no longer sample by sample.  On MI355X, with this workload and seed, no
    decision differed in the 8 frames (log-likelihoods within 3.4e-2,
    extrinsics within 3.3e-10); with twice the camera roughening a map size
    differed at frame 4 and the ancestors at frame 7, after which the
    extrinsics differed by 2e-3;
  * FP32 storage, lock step: one step from the oracle's state (cast to FP32
    storage) gives the oracle's decisions and an extrinsic estimate within
    1e-4 relative, every frame.
"""
import numpy as np
import pytest

import oracle
import scene
import workloads

pytestmark = pytest.mark.gpu

FRAMES = 8
POSE = [0, 1, 2, 6, 7, 8]


def _summary(f, gpu, res):
    st, w = f.parents()
    wn = w / w.sum()
    anc = f.ancestors() if gpu else f.ancestors()[0]
    return dict(mean=wn @ st, ng=f.counts()["n_gaussians"], resampled=bool(res), anc=np.array(anc).copy())


def _run(p, s, sc, gpu):
    import scphd_amd

    f = scphd_amd.Filter(p) if gpu else oracle.Oracle(p, threads=16)
    rows = []
    for k in range(FRAMES):
        u, v, fixed = scene.frame_measurements(sc, k)
        f.predict(s["dt"])
        f.update(u, v, fixed)
        ll = f.update_debug()["loglik"].copy()
        res = f.resample()
        rows.append(dict(_summary(f, gpu, res), loglik=ll))
    if gpu:
        f.close()
    return rows


def _rel(a, b):
    return float(np.linalg.norm(a["mean"][POSE] - b["mean"][POSE]) / np.linalg.norm(b["mean"][POSE]))


def _same(a, b):
    return a["ng"] == b["ng"] and a["resampled"] == b["resampled"] and (
        not a["resampled"] or np.array_equal(a["anc"], b["anc"]))


def test_precision_fp32_and_fp64_storage_against_the_fp64_oracle():
    p, s = workloads.workload("c3", n_particles=128, n_frames=FRAMES)
    p["seed"] = 3
    sc = scene.generate(seed=p["seed"], **s)
    ref = _run(dict(p, storage=1), s, sc, gpu=False)
    out = {"fp64": _run(dict(p, storage=1), s, sc, gpu=True), "fp32": _run(dict(p, storage=0), s, sc, gpu=True)}
    print("\nframe  mode  extrinsic_rel_err  loglik_max_abs_err  gaussians(gpu/oracle)  decisions_equal")
    first_diff = {}
    for mode, rows in out.items():
        for k, (a, b) in enumerate(zip(rows, ref)):
            rel = _rel(a, b)
            lle = float(np.max(np.abs(a["loglik"] - b["loglik"])))
            same = _same(a, b)
            print(f"{k:5d}  {mode}  {rel:17.3e}  {lle:18.3e}  {a['ng']:>10d}/{b['ng']:<10d}  {same}")
            if not same and mode not in first_diff:
                first_diff[mode] = k
            if mode == "fp64":
                assert same, f"fp64 frame {k}: decisions differ"
                assert rel < 1e-8, f"fp64 frame {k}: extrinsics {rel:.3e}"
            elif mode not in first_diff:
                assert rel < 1e-4, f"fp32 frame {k}: extrinsics {rel:.3e} before any decision differed"
    print(f"first differing decision: fp32 frame {first_diff.get('fp32', 'none')}")


def test_precision_fp32_storage_lock_step():
    """One FP32-storage step from the oracle's state every frame."""
    import scphd_amd

    p, s = workloads.workload("c3", n_particles=128, n_frames=FRAMES)
    p["seed"] = 3
    sc = scene.generate(seed=p["seed"], **s)
    orc = oracle.Oracle(dict(p, storage=0), threads=16)
    gpu = scphd_amd.Filter(dict(p, storage=0))
    worst = 0.0
    try:
        for k in range(FRAMES):
            gpu.set_state(orc.state())
            u, v, fixed = scene.frame_measurements(sc, k)
            for f in (orc, gpu):
                f.predict(s["dt"])
                f.update(u, v, fixed)
            a = _summary(gpu, True, gpu.resample())
            b = _summary(orc, False, orc.resample())
            assert _same(a, b), f"frame {k}: decisions differ"
            worst = max(worst, _rel(a, b))
            assert _rel(a, b) < 1e-4, f"frame {k}"
    finally:
        gpu.close()
    print(f"fp32 lock step: worst extrinsic relative error {worst:.3e}")
