#!/usr/bin/env python3
"""Served workload on one GPU: the served forms against the plain path.

    python run_c4.py [--workload c4|4096] [--rate 20] [--horizon 1.5] [--duration 1 3]
                     [--budget-gb 200] [--cohorts 4] [--workers 8]
                     [--modes plain,batch,cohorts] [--paced] [--saturated]

Workloads (arrivals as workloads.c4_arrivals, 30 Hz frames, S = 100):
* c4: C4's size mix -- parents log-uniform in [256, 8192], landmarks in
  [20, 400], clutter lambda in [5, 100] -- Poisson at --rate rigs/s over
  --horizon seconds, each rig streaming for a duration uniform in
  --duration seconds (C4 itself: 20 rigs/s, 2-10 s; one MI355X cannot keep
  up with that offered load, so the default bounds the run);
* 4096: every rig 4096 parents, 200 landmarks, lambda 50 (the bench rig at
  S = 100), same arrivals.

Modes, each over the same rigs and measurement sets:
* plain: one rig at a time through its own Filter, frames back to back;
* batch: the batch API (scphd_batch_*, one handle and stream per rig, frames
  of different rigs overlapping on the device, admission by HBM
  reservation, serve.run_served);
* cohorts: shared launches -- phase cohorts, each an ensemble that rigs
  join and leave between ticks, every tick one set of launches for all of
  its rigs (serve.run_cohorts);
each saturated (every frame offered at once: throughput) and, with --paced,
in real time (latency p50 / p99 from the time a frame is ready).

Prints one JSON line per mode and a summary with the ratios to plain.
"""
import argparse
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import serve  # noqa: E402
import workloads  # noqa: E402


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="c4", choices=["c4", "4096"])
    ap.add_argument("--master", type=int, default=3)
    ap.add_argument("--rate", type=float, default=20.0)
    ap.add_argument("--horizon", type=float, default=1.5)
    ap.add_argument("--duration", type=float, nargs=2, default=[1.0, 3.0])
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--cohorts", type=int, default=4)
    ap.add_argument("--workers", type=int, default=8)
    ap.add_argument("--modes", default="plain,batch,cohorts")
    ap.add_argument("--paced", action="store_true", help="also run the served modes in real time")
    ap.add_argument("--max-wall", type=float, default=300.0)
    args = ap.parse_args(argv)
    kw = dict(master=args.master, rate=args.rate, horizon=args.horizon, samples=100, dur=tuple(args.duration))
    if args.workload == "4096":
        kw.update(parents=(4096, 4096), landmarks=(200, 200), lam=(50.0, 50.0))
    rigs = serve.prepare(workloads.c4_arrivals(**kw))
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    frames = sum(len(r["frames"]) for r in rigs)
    log(f"{args.workload}: {len(rigs)} rigs, {frames} frames, parents "
        f"{min(r['p']['n_particles'] for r in rigs)}-{max(r['p']['n_particles'] for r in rigs)}")
    modes = args.modes.split(",")
    out = {}

    def emit(name, res):
        res = {k: v for k, v in res.items() if k != "estimates"}
        out[name] = res
        print(json.dumps(dict(mode=name, **res)), flush=True)

    budget = args.budget_gb * 1e9
    if "plain" in modes:
        emit("plain", serve.run_plain(rigs, progress=log))
    for paced in ([False, True] if args.paced else [False]):
        tag = "paced" if paced else "saturated"
        if "batch" in modes:
            emit(f"batch_{tag}", serve.run_served(rigs, hbm_budget_bytes=budget, n_workers=args.workers,
                                                  paced=paced, progress=log, max_wall=args.max_wall))
        if "cohorts" in modes:
            emit(f"cohorts_{tag}", serve.run_cohorts(rigs, hbm_budget_bytes=budget, cohorts=args.cohorts,
                                                     paced=paced, progress=log, max_wall=args.max_wall))
    summary = dict(summary=True, workload=args.workload, rigs=len(rigs), frames=frames, budget_gb=args.budget_gb,
                   cohorts=args.cohorts)
    if "plain" in out:
        base = out["plain"]["updates_per_s"]
        for name, res in out.items():
            if name != "plain" and base > 0:
                summary[f"ratio_{name}_over_plain"] = res["updates_per_s"] / base
    print(json.dumps(summary), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
