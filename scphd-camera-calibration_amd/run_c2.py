#!/usr/bin/env python3
"""Batched ragged workload: many independent rigs stepped together on one GPU.

    python run_c2.py [--rigs 64] [--frames 100] [--workers 8] [--plain-rigs N]
                     [--mode batch|ensemble|hybrid|both] [--ensemble-gb 120] [--workload c2|small]
                     [--alone-parents 2048] [--growth 2]

Rigs as workloads.c2_rigs: parents uniform in {256, ..., 4096}, landmarks
uniform in [20, 300], clutter lambda uniform in [5, 80], baseline uniform in
[0.1, 0.5] m, f uniform in [400, 800] px, S = 100, seeds from SplitMix64 of
a master seed (--workload small: workloads.small_ragged_rigs, parents in
{256, 512}, landmarks in [20, 100], lambda in [5, 40]).

* batch: every rig's frames are queued at once; rigs are admitted as the HBM
  budget allows (scphd_batch_*), and a rig leaves when its frames are done.
* ensemble: the rigs are packed, in order, into groups whose summed
  footprint (scphd_rig_footprint at --growth times the expected map size)
  fits --ensemble-gb and whose sample planes stay below 4 GiB (the kernels'
  32-bit offset form); each group is one ensemble (scphd_create_ensemble),
  its buffers reserved for that size before the first frame, whose
  frames step every rig of the group with one launch per kernel; the groups
  run one after the other.  The C2 maps keep growing over 100 frames (terms
  of landmarks that left the view decay only by the survival probability),
  to about twice the expected size: hence --growth 2, so that the maps grow
  into reserved memory rather than allocating large blocks near the
  device's capacity, which is slow.
* hybrid: rigs with at least --alone-parents parents run alone (one rig fills
  the GPU), the others in ensembles as above.

Prints one JSON line: component updates/s of each mode, of the plain path on
the first --plain-rigs rigs run one at a time, and the ratios.
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import scene  # noqa: E402
import scphd_amd  # noqa: E402
import serve  # noqa: E402
import workloads  # noqa: E402


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--rigs", type=int, default=64)
    ap.add_argument("--frames", type=int, default=100)
    ap.add_argument("--workers", type=int, default=8)
    ap.add_argument("--plain-rigs", type=int, default=8)
    ap.add_argument("--master", type=int, default=1)
    ap.add_argument("--mode", default="both", choices=["batch", "ensemble", "hybrid", "both"])
    ap.add_argument("--alone-parents", type=int, default=2048)
    ap.add_argument("--growth", type=float, default=2.0,
                    help="ensembles reserve and are packed for this multiple of the expected map size")
    ap.add_argument("--ensemble-gb", type=float, default=120.0)
    ap.add_argument("--workload", default="c2", choices=["c2", "small"])
    args = ap.parse_args(argv)

    rigs = []
    draw = workloads.c2_rigs if args.workload == "c2" else workloads.small_ragged_rigs
    for t, (p, s) in enumerate(draw(master=args.master, n_rigs=args.rigs, n_frames=args.frames)):
        sc = scene.generate(seed=p["seed"], **s)
        rigs.append(dict(t=0.0, p=p, s=s, frames=[scene.frame_measurements(sc, k) for k in range(s["n_frames"])]))
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    t0 = time.perf_counter()
    out = dict(workload=f"{args.workload}: {args.rigs} ragged rigs x {args.frames} frames, S = 100, FP32 storage",
               parents=[r["p"]["n_particles"] for r in rigs])
    if args.mode in ("batch", "both"):
        bat = serve.run_served(rigs, n_workers=args.workers, paced=False, progress=log)
        out["batched"] = dict(updates_per_s=bat["updates_per_s"], wall_s=bat["wall_s"], frames=bat["frames"],
                              peak_active_rigs=bat["peak_active_rigs"], peak_reserved_gb=bat["peak_reserved_gb"],
                              peak_used_gb=bat["peak_used_gb"], refused_admissions=bat["refused_admissions"])
        print(json.dumps(out), flush=True)
    if args.mode in ("ensemble", "hybrid", "both"):
        hybrid = args.mode == "hybrid"
        groups, alone = serve.pack_ensembles(rigs, args.ensemble_gb, args.growth,
                                             args.alone_parents if hybrid else None)
        pairs, wall, lat = 0, 0.0, []
        for i, g in enumerate(groups):
            e = serve.run_ensemble(g, progress=log, growth=args.growth)
            pairs += e["pairs"]
            wall += e["wall_s"]
            lat.append(e["latency_ms"])
            log(f"ensemble group {i + 1}/{len(groups)}: {len(g)} rigs, {e['updates_per_s']:.3g} updates/s")
        if alone:
            a = serve.run_plain(alone, progress=log)
            pairs += a["pairs"]
            wall += a["wall_s"]
            log(f"{len(alone)} rigs alone: {a['updates_per_s']:.3g} updates/s")
        out["hybrid" if hybrid else "ensemble"] = dict(updates_per_s=pairs / wall if wall > 0 else 0.0, wall_s=wall,
                                                       groups=[len(g) for g in groups], alone=len(alone),
                                                       group_latency_ms=lat)
        print(json.dumps(out), flush=True)
    if args.plain_rigs > 0:
        plain = serve.run_plain(rigs[:args.plain_rigs], progress=log)
        out["plain"] = dict(rigs=plain["rigs"], updates_per_s=plain["updates_per_s"], latency_ms=plain["latency_ms"])
        for k in ("batched", "ensemble", "hybrid"):
            if k in out and plain["updates_per_s"]:
                out[f"ratio_{k}"] = out[k]["updates_per_s"] / plain["updates_per_s"]
    out["total_s"] = time.perf_counter() - t0
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
