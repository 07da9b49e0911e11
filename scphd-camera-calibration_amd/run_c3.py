#!/usr/bin/env python3
"""Precision study at scale: one rig with 16384 parents, S = 100, 300
landmarks, clutter lambda = 100, run free on the GPU from one seed with FP64
sample storage (all FP64) and with each single-precision variant -- FP32
sample storage (FP64 arithmetic), and FP32 storage with FP32 per-sample
arithmetic -- each compared with the FP64 run frame by frame.  At this size
the FP64 CPU oracle is out of reach; tests/test_precision_gpu.py compares
all three against it on the same scene at 128 parents.

    python run_c3.py [--parents 16384] [--frames 30] [--variants fp32,fp32-arith] [--fp32-only]

Maps are uncapped (max_gaussians = 0, as in the reference).  They grow
through the run -- landmarks that leave the view keep their non-detection
clouds for ~180 frames (weight x ps per frame) -- from ~6 M Gaussians at
frame 0 to ~25 M at frame 80 at 16384 parents, so with FP64 storage
(56 B per sample, two generations, and a third while a generation grows)
the comparison fits at 8192 parents; --fp32-only runs the plain path alone
(e.g. all 200 frames at 16384 parents) and reports its throughput.

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


def run(p, s, sc, frames, log, every=5):
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
        if log and k % every == 0:
            log(f"storage {p['storage']} frame {k} gaussians {rows[-1]['ng']} {time.perf_counter() - t0:.1f} s")
    f.close()
    return rows, time.perf_counter() - t0


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--parents", type=int, default=16384)
    ap.add_argument("--frames", type=int, default=30)
    ap.add_argument("--fp32-only", action="store_true")
    ap.add_argument("--variants", default="fp32,fp32-arith",
                    help="single-precision variants compared with the FP64 run")
    args = ap.parse_args(argv)
    p, s = workloads.workload("c3", n_particles=args.parents, n_frames=args.frames, max_gaussians=0)
    sc = scene.generate(seed=p["seed"], **s)
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    storage = {"fp32": 0, "fp64": 1, "fp32-arith": 2}
    if args.fp32_only:
        r32, t32 = run(dict(p, storage=0), s, sc, args.frames, log, every=1)
        print(json.dumps(dict(summary=True, parents=args.parents, frames=args.frames, seconds_fp32=t32,
                              ms_per_frame=t32 / args.frames * 1e3, final_gaussians=r32[-1]["ng"],
                              final_estimate_fp32=list(r32[-1]["mean"][POSE]))))
        return 0
    r64, t64 = run(dict(p, storage=1), s, sc, args.frames, log)
    summary = dict(summary=True, parents=args.parents, frames=args.frames, seconds_fp64=t64,
                   final_estimate_fp64=list(r64[-1]["mean"][POSE]))
    for name in args.variants.split(","):
        rv, tv = run(dict(p, storage=storage[name]), s, sc, args.frames, log)
        worst, same_anc, same_ng, first = 0.0, 0, 0, None
        for k, (a, b) in enumerate(zip(rv, r64)):
            rel = float(np.linalg.norm(a["mean"][POSE] - b["mean"][POSE]) / np.linalg.norm(b["mean"][POSE]))
            anc = a["resampled"] == b["resampled"] and (not a["resampled"] or np.array_equal(a["anc"], b["anc"]))
            worst = max(worst, rel)
            same_anc += anc
            same_ng += a["ng"] == b["ng"]
            if first is None and not (anc and a["ng"] == b["ng"]):
                first = k
            print(json.dumps(dict(variant=name, frame=k, rel_diff=rel, ng=a["ng"], ng64=b["ng"],
                                  ancestors_equal=bool(anc), card=a["card"], card64=b["card"])))
        summary[name] = dict(worst_extrinsic_rel_diff=worst, frames_with_equal_ancestors=same_anc,
                             frames_with_equal_map_size=same_ng, first_differing_decision=first, seconds=tv,
                             final_estimate=list(rv[-1]["mean"][POSE]))
    print(json.dumps(summary))
    return 0


if __name__ == "__main__":
    sys.exit(main())
