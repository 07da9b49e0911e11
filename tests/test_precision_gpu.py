"""Precision study: GPU filters against the FP64 oracle.

Workload: the precision-study scene (300 landmarks, clutter lambda = 100,
S = 100) at a parent count the oracle finishes in seconds (128).  Four
filters start from one seed and run predict -> update -> resample freely for
FRAMES frames:
  * the oracle, FP64 samples (the reference arithmetic);
  * the GPU with FP64 sample storage ("all-FP64 GPU");
  * the GPU with FP32 sample storage (FP32 clouds, FP64 arithmetic);
  * the GPU with FP32 sample storage and FP32 per-sample arithmetic
    ("fp32a": prediction, reprojection, the moment sums and the posterior
    draws in single precision; the per-Gaussian algebra, likelihoods,
    log-sum-exp and resampling prefix sum in FP64).
Printed per frame and output (pytest -s): extrinsic estimate (weighted
parent mean) relative error, parent log-likelihood error, Gaussian counts,
ancestor agreement.

Asserted:
  * all-FP64 GPU, free running: identical map sizes and ancestors every
    frame, extrinsics within 1e-8 relative;
  * FP32 storage and FP32 arithmetic, free running: extrinsics within 1e-4
    relative of the oracle in every frame before its first differing decision (a map size
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
  * Per output, lock step, all three GPU variants: the error of every
    output of one step from the oracle's state -- the refit means (in units
    of the refit standard deviation) and variances (relative) of the
    Gaussians with at least half their samples in view, positions and
    velocities apart, the per-Gaussian Pd, the normalisers (relative), the parent log-likelihoods
    (absolute), the survivor / merged / non-detection counts (parents that
    differ), the parent weights (relative) and the extrinsic estimate after
    resampling (relative) -- printed per frame, with bounds asserted per
    variant.
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
    out = {"fp64": _run(dict(p, storage=1), s, sc, gpu=True), "fp32": _run(dict(p, storage=0), s, sc, gpu=True),
           "fp32a": _run(dict(p, storage=2), s, sc, gpu=True)}
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
    for mode in ("fp32", "fp32a"):
        print(f"first differing decision: {mode} frame {first_diff.get(mode, 'none')}")


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


VARIANTS = {"fp64": 1, "fp32": 0, "fp32a": 2}
# per variant: bounds on the refit mean error (refit standard deviations)
# and the refit variances (relative) of the positional coordinates (u, v, d),
# normalisers (relative), log-likelihoods (absolute), parents with a
# differing count (fraction) and the extrinsic estimate (relative)
BOUNDS = {"fp64": dict(mean=1e-8, var=1e-8, eta=1e-10, loglik=1e-8, counts=0.0, ext=1e-10),
          "fp32": dict(mean=1e-2, var=1e-2, eta=1e-5, loglik=1e-3, counts=0.05, ext=1e-4),
          "fp32a": dict(mean=1e-1, var=1e-1, eta=1e-4, loglik=1e-2, counts=0.05, ext=1e-4)}


def _fit_errors(fg, fo, pd):
    """Refit errors over the Gaussians with at least half their samples in
    view (pd_g >= Pd / 2; fewer samples give near-singular refits): the mean
    error in units of the oracle's refit standard deviation and the relative
    error of the refit variances, for the positional coordinates (u, v, d)
    and the velocities (vu, vv, vd) apart; and the largest pd_g error."""
    diag = np.array([0, 2, 5, 9, 14, 20])  # packed indices of the diagonal
    var = fo[6 + diag]
    ok = (fo[27] >= pd / 2) & (var > 0).all(axis=0)
    pde = float(np.max(np.abs(fg[27] - fo[27]), initial=0.0))
    out = {}
    for name, dims in (("pos", [0, 1, 2]), ("vel", [3, 4, 5])):
        if not ok.any():
            out[name] = (0.0, 0.0)
            continue
        v = var[dims][:, ok]
        m = float(np.max(np.abs(fg[dims][:, ok] - fo[dims][:, ok]) / np.sqrt(v)))
        r = float(np.max(np.abs(fg[6 + diag[dims]][:, ok] - v) / v))
        out[name] = (m, r)
    return out, pde


@pytest.mark.parametrize("mode", list(VARIANTS))
def test_precision_per_output_lock_step(mode):
    """Every output of one step from the oracle's state, per frame."""
    import scphd_amd

    p, s = workloads.workload("c3", n_particles=128, n_frames=FRAMES)
    p["seed"] = 3
    sc = scene.generate(seed=p["seed"], **s)
    orc = oracle.Oracle(dict(p, storage=1), threads=16)
    gpu = scphd_amd.Filter(dict(p, storage=VARIANTS[mode]))
    b = BOUNDS[mode]
    worst = {}
    print(f"\n{mode}: frame  mean_sd(pos vel)  var_rel(pos vel)  pd_abs  eta_rel  loglik_abs  counts_differ  "
          f"weight_rel  ext_rel")
    try:
        for k in range(FRAMES):
            gpu.set_state(orc.state())
            u, v, fixed = scene.frame_measurements(sc, k)
            for f in (orc, gpu):
                f.predict(s["dt"])
                f.update(u, v, fixed)
            n = gpu.update_stats()["n_gaussians_in"]
            fe, pde = _fit_errors(gpu.fit_debug(), orc.fit_debug(n), p["pd"])
            mean, vrel = fe["pos"]
            dg, do = gpu.update_debug(), orc.update_debug()
            eta = float(np.max(np.abs(dg["eta"] - do["eta"]) / np.maximum(np.abs(do["eta"]), 1e-300), initial=0.0))
            ll = float(np.max(np.abs(dg["loglik"] - do["loglik"])))
            nd = int(np.sum((dg["nsurvivors"] != do["nsurvivors"]) | (dg["nmerged"] != do["nmerged"])
                            | (dg["nnondetect"] != do["nnondetect"])))
            wg, wo = gpu.parents()[1], orc.parents()[1]
            wrel = float(np.max(np.abs(wg - wo) / np.maximum(wo, 1e-300)))
            a = _summary(gpu, True, gpu.resample())
            c = _summary(orc, False, orc.resample())
            ext = _rel(a, c)
            print(f"{mode}: {k:5d}  {mean:7.1e} {fe['vel'][0]:7.1e}  {vrel:7.1e} {fe['vel'][1]:7.1e}  {pde:6.1e}  "
                  f"{eta:7.1e}  {ll:10.2e}  {nd:6d} of {len(wo):<4d}  {wrel:10.2e}  {ext:7.1e}")
            for key, val in dict(mean=mean, var=vrel, eta=eta, loglik=ll, counts=nd / len(wo), ext=ext).items():
                worst[key] = max(worst.get(key, 0.0), val)
    finally:
        gpu.close()
    print(f"{mode}: worst {worst}")
    for key, lim in b.items():
        assert worst[key] <= lim, f"{mode}: {key} error {worst[key]:.3e} above {lim:.1e}"
