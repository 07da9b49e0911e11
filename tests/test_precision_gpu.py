"""Precision study: free-running GPU filters against the FP64 oracle.

Workload: the precision-study scene (300 landmarks, clutter lambda = 100,
S = 100) at a parent count the oracle finishes in seconds.  Three runs
start from the same seed and run predict -> update -> resample freely for
FRAMES frames:
  * the oracle, FP64 samples (the reference arithmetic);
  * the GPU with FP64 sample storage ("all-FP64 GPU");
  * the GPU with FP32 sample storage (the plain path: FP32 bulk, FP64
    likelihood, log-sum-exp and resampling prefix sum).
Per frame and output: extrinsic estimate (weighted parent mean) relative
error, parent log-likelihood error, Gaussian-count agreement, and
ancestor-index agreement.  Bounds asserted: all-FP64 GPU -- identical
map sizes and ancestors every frame and extrinsics within 1e-8 relative;
FP32 storage -- extrinsic estimate within 1e-4 relative of the FP64 oracle.
Free-running runs amplify last-bit differences wherever a decision sits at
a threshold (a merge pair at exactly min_separation, a survivor at exactly
min_weight); on MI355X the all-FP64 run agreed to ~1e-14 for 7 frames and
5.7e-10 after such a decision in frame 8, the FP32 run to 4.9e-7.  The
table is printed (pytest -s) for the record.
"""
import numpy as np
import pytest

import oracle
import scene
import workloads

pytestmark = pytest.mark.gpu

FRAMES = 8


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
        anc = f.ancestors() if gpu else f.ancestors()[0]
        st, w = f.parents()
        wn = w / w.sum()
        rows.append(dict(mean=wn @ st, loglik=ll, ng=f.counts()["n_gaussians"], resampled=bool(res),
                         anc=np.array(anc).copy()))
    if gpu:
        f.close()
    return rows


def test_precision_fp32_and_fp64_storage_against_the_fp64_oracle():
    p, s = workloads.workload("c3", n_particles=128, n_frames=FRAMES)
    p["seed"] = 3
    sc = scene.generate(seed=p["seed"], **s)
    ref = _run(dict(p, storage=1), s, sc, gpu=False)
    out = {"fp64": _run(dict(p, storage=1), s, sc, gpu=True), "fp32": _run(dict(p, storage=0), s, sc, gpu=True)}
    pose = [0, 1, 2, 6, 7, 8]
    print("\nframe  mode  extrinsic_rel_err  loglik_max_abs_err  gaussians(gpu/oracle)  ancestors_equal")
    worst = {}
    for mode, rows in out.items():
        for k, (a, b) in enumerate(zip(rows, ref)):
            rel = float(np.linalg.norm(a["mean"][pose] - b["mean"][pose]) / np.linalg.norm(b["mean"][pose]))
            lle = float(np.max(np.abs(a["loglik"] - b["loglik"])))
            same_anc = a["resampled"] == b["resampled"] and (not a["resampled"] or np.array_equal(a["anc"], b["anc"]))
            print(f"{k:5d}  {mode}  {rel:17.3e}  {lle:18.3e}  {a['ng']:>10d}/{b['ng']:<10d}  {same_anc}")
            worst[mode] = max(worst.get(mode, 0.0), rel)
            if mode == "fp64":
                assert a["ng"] == b["ng"] and same_anc, f"fp64 frame {k}: decisions differ"
    print(f"worst extrinsic relative error: fp64 {worst['fp64']:.3e}, fp32 {worst['fp32']:.3e}")
    assert worst["fp64"] < 1e-8
    assert worst["fp32"] < 1e-4
