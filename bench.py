#!/usr/bin/env python3
"""Benchmark of the gfx950 SC-PHD calibration filter step on one MI355X.

Metric: GM-PHD component updates per second (one component update = one
(predicted Gaussian, measurement) Kalman term, the reference's
n_measure * n_features, cheesinglee/scphd-camera-calibration
src/scphdcameracalibration.cu:1617-1619), whole-step throughput
sum(N_g * M) / sum(frame time) over frames of predict -> update -> resample
(one iteration of the reference main loop, src/main.cpp:78-97), with the
measurement sets already resident in HBM.  Also reported: p50/p99 frame
latency, per-phase times, the achieved HBM bandwidth of the reproject/fit
kernel against the 8 TB/s peak, and the FP64 CPU oracle timed on the host's
cores on a bounded sample of the same workload.

Workload (default c1): one rig, 4096 parents, 200 samples per Gaussian,
200 landmarks, clutter lambda = 50, 640 x 480 image, f = 500 px, baseline
0.2 m, synthetic scene from a fixed seed.  Warm-up frames grow the maps to
their steady-state size before the timed frames.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c1]
"""
import argparse
import ctypes
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(REPO, "scphd-camera-calibration_amd")
sys.path.insert(0, PKG)

import numpy as np  # noqa: E402

import scene  # noqa: E402
import serve  # noqa: E402
import workloads  # noqa: E402

METRIC = "GM-PHD component updates/s at 4096 parents; p99 frame latency, served vs plain"
HBM_PEAK_GBS = 8000.0


class DeviceArray:
    """A device buffer through the HIP runtime (no framework needed)."""

    _hip = None

    @classmethod
    def hip(cls):
        if cls._hip is None:
            h = ctypes.CDLL("libamdhip64.so")
            h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
            h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
            h.hipFree.argtypes = [ctypes.c_void_p]
            cls._hip = h
        return cls._hip

    def __init__(self, host):
        host = np.ascontiguousarray(host, dtype=np.float64)
        self.ptr = ctypes.c_void_p()
        n = max(host.nbytes, 8)
        if self.hip().hipMalloc(ctypes.byref(self.ptr), n) != 0:
            raise RuntimeError("hipMalloc failed")
        if host.nbytes and self.hip().hipMemcpy(self.ptr, host.ctypes.data, host.nbytes, 1) != 0:
            raise RuntimeError("hipMemcpy failed")

    def at(self, i):
        return ctypes.c_void_p(int(self.ptr.value) + 8 * int(i))

    def free(self):
        if self.ptr:
            self.hip().hipFree(self.ptr)
            self.ptr = None


def fit_alg_bytes(st, S, esz):
    """Compulsory HBM bytes of one k_fit launch: every input sample read
    (7 planes), every kept non-detection sample written (7 planes), the
    28 FP64 fit planes written and the 24 B of table entry and flags per
    Gaussian."""
    g = st["n_gaussians_in"]
    return 7 * esz * S * (g + st["n_nondetect"]) + g * (28 * 8 + 24)


def cpu_baseline(p, s, parents, warm, timed, threads):
    """FP64 oracle (oracle/) on the same workload with `parents` parents:
    `warm` untimed frames grow the maps, `timed` frames are timed."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle  # noqa: E402

    q = dict(p)
    q["n_particles"] = parents
    sc = scene.generate(seed=q["seed"], **dict(s, n_frames=warm + timed))
    orc = oracle.Oracle(q, threads=threads)
    upd, sec = 0, 0.0
    for k in range(warm + timed):
        u, v, fixed = scene.frame_measurements(sc, k)
        t0 = time.perf_counter()
        ng = orc.counts()["n_gaussians"]
        orc.predict(s["dt"])
        orc.update(u, v, fixed)
        orc.resample()
        dt = time.perf_counter() - t0
        if k >= warm:
            upd += ng * len(u)
            sec += dt
    return upd / sec if sec > 0 else 0.0, sec, orc.counts()["n_gaussians"] / parents


TRAFFIC_FILE = os.path.join(REPO, "profiles", "k_fit_c1_traffic.txt")


def measured_traffic():
    """HBM bytes per k_fit launch from the committed rocprofv3 counter summary
    (FETCH_SIZE + WRITE_SIZE, each scaled by the copy-kernel factor), or None."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    vals = {}
    for line in open(TRAFFIC_FILE):
        if ":" in line and not line.startswith("#"):
            k, v = line.split(":", 1)
            vals[k.strip()] = v.strip()
    try:
        return float(vals["hbm_bytes_per_launch"]), float(vals["algorithmic_bytes_per_launch"])
    except (KeyError, ValueError):
        return None


def served_leg(args, seed):
    """Served calibration (scaled version of the 20 rigs/s workload): rigs
    arrive as a Poisson process, are admitted by HBM reservation, stream
    30 Hz frames -- through the batch API (one handle and stream per rig)
    and through phase cohorts (shared launches: ensembles that rigs join and
    leave between ticks), in real time and saturated; compared with the
    plain path running the same rigs one at a time.  200 GB budget."""
    import serve  # noqa: E402

    arr = workloads.c4_arrivals(master=seed, rate=args.served_rate, horizon=args.served_horizon,
                                parents=tuple(args.served_parents), landmarks=tuple(args.served_landmarks),
                                lam=tuple(args.served_lambda), dur=tuple(args.served_duration))
    rigs = serve.prepare(arr)
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    budget = 200e9
    rt = serve.run_served(rigs, hbm_budget_bytes=budget, n_workers=args.served_workers, progress=log)
    sat = serve.run_served(rigs, hbm_budget_bytes=budget, n_workers=args.served_workers, progress=log, paced=False)
    crt = serve.run_cohorts(rigs, hbm_budget_bytes=budget, cohorts=4, progress=log)
    csat = serve.run_cohorts(rigs, hbm_budget_bytes=budget, cohorts=4, progress=log, paced=False)
    pl = serve.run_plain(rigs, progress=log)
    return {
        "workload": f"Poisson {args.served_rate:g} rigs/s over {args.served_horizon:g} s, parents log-uniform "
                    f"{args.served_parents}, landmarks {args.served_landmarks}, lambda {args.served_lambda}, "
                    f"30 Hz for {args.served_duration} s, S = 100, FP32 storage",
        "hbm_budget_gb": budget / 1e9,
        "served_realtime": rt,
        "served_saturated": sat,
        "cohorts_realtime": crt,
        "cohorts_saturated": csat,
        "plain_one_rig_at_a_time": pl,
        "throughput_ratio_served_saturated_over_plain":
            sat["updates_per_s"] / pl["updates_per_s"] if pl["updates_per_s"] else None,
        "throughput_ratio_cohorts_saturated_over_plain":
            csat["updates_per_s"] / pl["updates_per_s"] if pl["updates_per_s"] else None,
    }


def ensemble_leg(args, seed):
    """Many small ragged rigs (workloads.small_ragged_rigs) stepped as one
    ensemble, one launch per kernel per frame, against the same rigs run one
    at a time through the plain path."""
    import serve  # noqa: E402

    arr = [(0.0, p, s) for p, s in workloads.small_ragged_rigs(master=seed, n_rigs=args.ensemble_rigs,
                                                               n_frames=args.ensemble_frames)]
    rigs = serve.prepare(arr)
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    ens = serve.run_ensemble(rigs, progress=log)
    pl = serve.run_plain(rigs, progress=log)
    return {
        "workload": f"{args.ensemble_rigs} ragged rigs x {args.ensemble_frames} frames: parents in {{256, 512}}, "
                    f"landmarks in [20, 100], lambda in [5, 40], f in [400, 800] px, baseline in [0.1, 0.5] m, "
                    f"S = 100, FP32 storage",
        "ensemble": ens,
        "plain_one_rig_at_a_time": pl,
        "throughput_ratio_ensemble_over_plain":
            ens["updates_per_s"] / pl["updates_per_s"] if pl["updates_per_s"] else None,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=12)
    ap.add_argument("--workload", default="c1")
    ap.add_argument("--storage", default="fp32", choices=["fp32", "fp64", "fp32-arith"],
                    help="sample clouds as float (FP64 arithmetic), as double, or as float with the per-sample "
                         "arithmetic in float")
    ap.add_argument("--cpu-parents", type=int, default=512)
    ap.add_argument("--cpu-warm", type=int, default=3, help="untimed oracle frames (maps reach their "
                                                             "steady size after 2)")
    ap.add_argument("--cpu-timed", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--served-rate", type=float, default=2.0, help="rig arrivals per second")
    ap.add_argument("--served-horizon", type=float, default=12.0, help="seconds of arrivals")
    ap.add_argument("--served-parents", type=int, nargs=2, default=[256, 2048])
    ap.add_argument("--served-landmarks", type=int, nargs=2, default=[20, 200])
    ap.add_argument("--served-lambda", type=float, nargs=2, default=[5.0, 60.0])
    ap.add_argument("--served-duration", type=float, nargs=2, default=[2.0, 6.0])
    ap.add_argument("--served-workers", type=int, default=8)
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--ensemble-rigs", type=int, default=64)
    ap.add_argument("--ensemble-frames", type=int, default=40)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box here has one MI355X and the filter step "
              "of one rig runs on one device; --gpus must be 1", file=sys.stderr)
        sys.exit(2)

    import scphd_amd  # noqa: E402

    p, s = workloads.workload(args.workload)
    p["storage"] = {"fp32": 0, "fp64": 1, "fp32-arith": 2}[args.storage]
    esz = 8 if args.storage == "fp64" else 4
    frames = args.warmup + args.steps
    s = dict(s, n_frames=max(s["n_frames"], frames))
    sc = scene.generate(seed=p["seed"], **s)

    # all measurement sets resident in HBM before the timed region
    sets = [scene.frame_measurements(sc, k) for k in range(frames)]
    offs = np.cumsum([0] + [len(u) for u, _, _ in sets])
    du = DeviceArray(np.concatenate([u for u, _, _ in sets]))
    dv = DeviceArray(np.concatenate([v for _, v, _ in sets]))

    f = scphd_amd.Filter(p)
    # buffers sized once, before the first frame, for 1.5 times the
    # steady-state map size (a generation holds a step's input map and its
    # posterior clouds): the steps then allocate nothing.  Growing the sample
    # generations frame by frame frees and allocates blocks of gigabytes, and
    # on these devices an allocation that follows large frees can wait
    # 0.5-2.5 s while the driver reclaims the freed memory.
    f.reserve(int(1.5 * p["n_particles"] * serve.expected_gaussians_per_parent(s)),
              max(len(u) for u, _, _ in sets))
    f.set_timing(True)
    S = p["particles_per_feature"]
    lat, updates, phase = [], [], []
    fit_bytes, fit_ms, smp_bytes, smp_ms = [], [], [], []
    counts = []
    step_bytes = []  # compulsory state traffic of each timed step (below)
    t_start = None
    for k in range(frames):
        u, v, fixed = sets[k]
        n = len(u)
        if k == args.warmup:
            f.sync()
            t_start = time.perf_counter()
        t0 = time.perf_counter()
        f.predict(s["dt"])
        f.update_device(du.at(offs[k]), dv.at(offs[k]), n, fixed)
        f.resample()
        f.sync()
        t1 = time.perf_counter()
        st = f.update_stats()
        fb = fit_alg_bytes(st, S, esz)
        print(f"frame {k} {'timed' if k >= args.warmup else 'warmup'} {(t1 - t0) * 1e3:.2f} ms "
              f"gaussians {f.counts()['n_gaussians']} k_fit_algorithmic_bytes {fb} phases_ms "
              f"{' '.join(f'{x:.2f}' for x in f.phase_ms())}", file=sys.stderr, flush=True)
        if k >= args.warmup:
            ph = f.phase_ms()
            lat.append((t1 - t0) * 1e3)
            updates.append(st["n_pairs"])
            phase.append(ph)
            fit_bytes.append(fb)
            fit_ms.append(ph[6])
            smp_bytes.append(7 * esz * st["n_merged"] * S)
            smp_ms.append(ph[7])
            counts.append(st)
            # read state k, write state k+1, read the measurements: 7 sample
            # values and a Gaussian weight per map entry in and out, 12 FP64
            # extrinsics + weight per parent each way, (u, v) per measurement
            n_out = f.counts()["n_gaussians"]
            step_bytes.append((7 * esz * S + esz) * (st["n_gaussians_in"] + n_out) + 2 * 104 * p["n_particles"]
                              + 16 * n)
    f.sync()
    total = time.perf_counter() - t_start
    du.free()
    dv.free()

    lat = np.array(lat)
    value = float(np.sum(updates) / total)
    phase = np.array(phase)
    fit_gbs = float(np.sum(fit_bytes) / (np.sum(fit_ms) * 1e-3) / 1e9) if np.sum(fit_ms) > 0 else 0.0
    smp_gbs = float(np.sum(smp_bytes) / (np.sum(smp_ms) * 1e-3) / 1e9) if np.sum(smp_ms) > 0 else 0.0
    last = counts[-1]
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "component updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": float(total / args.steps * 1e3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.storage == "fp32-arith" else "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{args.workload}: plain one-rig filter step, {p['n_particles']} parents, "
                        f"{S} samples per Gaussian, {s['n_landmarks']} landmarks, clutter lambda {s['lam']:g}, "
                        f"640x480 f={s['f']:g} baseline {s['baseline']:g} m, "
                        + ("FP32 sample storage, FP32 per-sample arithmetic (FP64 per-Gaussian algebra, "
                           "likelihoods, log-sum-exp and resampling)" if args.storage == "fp32-arith" else
                           f"{args.storage} sample storage, FP64 arithmetic"),
            "n_particles": p["n_particles"],
            "samples_per_gaussian": S,
            "landmarks": s["n_landmarks"],
            "clutter_lambda": s["lam"],
            "storage": args.storage,
            "parallelism": "one GPU, one rig",
        },
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                       "max": float(lat.max())},
        "phase_ms_mean": {name: float(phase[:, i].mean()) for i, name in enumerate(
            ["reproject_fit", "normalisers_survivors", "merge_parent_weights", "posterior_sampling",
             "map_assembly", "update_total", "k_fit", "k_sample"])},
        "maps": {"gaussians_in_last": last["n_gaussians_in"], "measurements_last": last["n_meas"],
                 "survivors_last": last["n_survivors"], "merged_last": last["n_merged"],
                 "max_survivors_per_parent_last": last["max_survivors_per_parent"],
                 "mean_gaussians_per_parent": float(np.mean([c["n_gaussians_in"] for c in counts]) /
                                                    p["n_particles"])},
        "roofline": {
            "kernel": "k_fit (reproject + Pd + refit + non-detection clouds)",
            "bound": "hbm",
            "achieved": fit_gbs,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": fit_gbs / HBM_PEAK_GBS,
            "traffic": (measured_traffic() or (None,))[0],
            "traffic_note": "HBM bytes per k_fit launch measured by rocprofv3 FETCH_SIZE + WRITE_SIZE (copy-kernel "
                            "scaled) at the same workload, profiles/k_fit_c1_traffic.txt; algorithmic bytes per "
                            f"launch there: {(measured_traffic() or (None, None))[1]}",
            "algorithmic_bytes": f"7*{esz} B per input sample read + 7*{esz} B per kept non-detection sample "
                                 f"written + 248 B per Gaussian (fit planes, table, flags); last launch "
                                 f"{fit_bytes[-1]} B",
            "k_sample_achieved_gbs": smp_gbs,
            "limiter_note": "hbm is the roofline the bytes are priced against; the kernel's limiter is VALU "
                            "issue (SQ counters in profiles/k_fit_c1_counters.txt), with HBM close behind",
        },
        "step_roofline": {
            "definition": "whole step: compulsory state bytes (read state k, write state k+1, read the "
                          "measurements: (7 sample values + weight) x esz per map entry in and out, 2 x 104 B "
                          "per parent, 16 B per measurement) over the whole timed frame time",
            "achieved": float(np.sum(step_bytes) / total / 1e9),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": float(np.sum(step_bytes) / total / 1e9) / HBM_PEAK_GBS,
            "bytes_last_step": int(step_bytes[-1]),
        },
    }
    if not args.no_served:
        out["served"] = served_leg(args, p["seed"])
    if not args.no_ensemble:
        out["ensemble"] = ensemble_leg(args, p["seed"])
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        cpu_v, cpu_s, cpu_gp = cpu_baseline(p, s, args.cpu_parents, args.cpu_warm, args.cpu_timed, threads)
        out["cpu_baseline"] = {
            "value": cpu_v,
            "unit": "component updates/s",
            "cores": threads,
            "kind": "port",
            "sample": f"FP64 oracle, same workload with {args.cpu_parents} parents: {args.cpu_warm} untimed "
                      f"frames then {args.cpu_timed} timed frames ({cpu_s:.1f} s, {cpu_gp:.0f} Gaussians/parent)",
        }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
