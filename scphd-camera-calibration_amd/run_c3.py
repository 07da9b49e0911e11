#!/usr/bin/env python3
"""Precision study at scale: one rig with 16384 parents, S = 100, 300
landmarks, clutter lambda = 100, run free on the GPU twice from one seed --
FP32 sample storage (the plain path) and FP64 sample storage (all FP64) --
and compared frame by frame.  At this size the FP64 CPU oracle is out of
reach; tests/test_precision_gpu.py compares both modes against it on the
same scene at 128 parents.

    python run_c3.py [--parents 16384] [--frames 200]

Per frame (JSON lines on stdout): extrinsic estimate of each mode and their
relative difference (x, y, z, theta, phi, psi), map sizes, and whether the
resampling ancestors agree; a summary line at the end.
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import numpy as np  # noqa: E402

import scene  # noqa: E402
import scphd_amd  # noqa: E402
import workloads  # noqa: E402

POSE = [0, 1, 2, 6, 7, 8]


def run(p, s, sc, frames, log):
    f = scphd_amd.Filter(p)
    rows = []
    t0 = time.perf_counter()
    for k in range(frames):
        u, v, fixed = scene.frame_measurements(sc, k)
        f.predict(s["dt"])
        f.update(u, v, fixed)
        res = f.resample()
        m, _, card = f.estimate()
        rows.append(dict(mean=m, ng=f.counts()["n_gaussians"], resampled=res, anc=f.ancestors() if res else None,
                         card=card))
        if log and k % 20 == 0:
            log(f"storage {p['storage']} frame {k} gaussians {rows[-1]['ng']} {time.perf_counter() - t0:.1f} s")
    f.close()
    return rows, time.perf_counter() - t0


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--parents", type=int, default=16384)
    ap.add_argument("--frames", type=int, default=200)
    args = ap.parse_args(argv)
    p, s = workloads.workload("c3", n_particles=args.parents, n_frames=args.frames)
    sc = scene.generate(seed=p["seed"], **s)
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    r32, t32 = run(dict(p, storage=0), s, sc, args.frames, log)
    r64, t64 = run(dict(p, storage=1), s, sc, args.frames, log)
    worst, same_anc, same_ng = 0.0, 0, 0
    for k, (a, b) in enumerate(zip(r32, r64)):
        rel = float(np.linalg.norm(a["mean"][POSE] - b["mean"][POSE]) / np.linalg.norm(b["mean"][POSE]))
        anc = a["resampled"] == b["resampled"] and (not a["resampled"] or np.array_equal(a["anc"], b["anc"]))
        worst = max(worst, rel)
        same_anc += anc
        same_ng += a["ng"] == b["ng"]
        print(json.dumps(dict(frame=k, rel_diff=rel, ng32=a["ng"], ng64=b["ng"], ancestors_equal=bool(anc),
                              card32=a["card"], card64=b["card"])))
    print(json.dumps(dict(summary=True, parents=args.parents, frames=args.frames, worst_extrinsic_rel_diff=worst,
                          frames_with_equal_ancestors=same_anc, frames_with_equal_map_size=same_ng,
                          seconds_fp32=t32, seconds_fp64=t64,
                          final_estimate_fp32=list(r32[-1]["mean"][POSE]), final_estimate_fp64=list(r64[-1]["mean"][POSE]))))
    return 0


if __name__ == "__main__":
    sys.exit(main())
