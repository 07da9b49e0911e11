#!/usr/bin/env python3
"""Monte-Carlo batch runner (the reference's batch_run.sh:1-20) on the batched
GPU mode.

    python batch_run.py CONFIG DATA_ROOT N_RUNS RESULTS_DIR [--workers 8] [--seed-stride 1]

Run n (1..N_RUNS) reads DATA_ROOT/<n>/data.mat (cells Z1, Z2), runs the
reference main loop on its own rig -- predict(0.5), update with Z1{k} on the
fixed camera for even k and Z2{k} on the moving camera for odd k (Z2 every
frame when slam is set), writeMat, checkStuff, resample, for
k < min(frames, max_steps) -- and writes RESULTS_DIR/<n>/prior.mat and
RESULTS_DIR/<n>/<k>.mat.  The configuration is copied to RESULTS_DIR.  Runs
share the GPU through scphd_batch_*: several rigs are resident at once (as
many as the HBM budget admits) and their frames overlap on the device; a run
gets seed = config seed + n * seed_stride, so that runs differ in their noise
and any run can be repeated alone.
"""
import argparse
import os
import shutil
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import scphd_amd  # noqa: E402


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("config")
    ap.add_argument("data_root")
    ap.add_argument("n_runs", type=int)
    ap.add_argument("results_dir")
    ap.add_argument("--workers", type=int, default=8)
    ap.add_argument("--seed-stride", type=int, default=1)
    ap.add_argument("--hbm-budget-gb", type=float, default=0.0)
    args = ap.parse_args(argv)

    base = scphd_amd.load_params(args.config)
    os.makedirs(args.results_dir, exist_ok=True)
    shutil.copy(args.config, args.results_dir)
    b = scphd_amd.Batch(hbm_budget_bytes=int(args.hbm_budget_gb * 1e9), n_workers=args.workers)
    pending = list(range(1, args.n_runs + 1))
    active = {}  # rig -> (run, frames submitted, frames finished)
    done, failed = 0, []
    t0 = time.time()
    try:
        while pending or active:
            while pending:
                n = pending[0]
                p = dict(base, seed=(base["seed"] + n * args.seed_stride) & 0xFFFFFFFFFFFFFFFF)
                rig = b.add_rig(p)
                if rig is None:
                    if not active:
                        raise RuntimeError(f"run {n} does not fit the HBM budget on its own")
                    break
                pending.pop(0)
                out = os.path.join(args.results_dir, str(n))
                os.makedirs(out, exist_ok=True)
                b.filter(rig).write_mat(os.path.join(out, "prior.mat"))
                b.set_output(rig, out)
                data = os.path.join(args.data_root, str(n), "data.mat")
                z1 = scphd_amd.read_measurements(data, "Z1")
                z2 = scphd_amd.read_measurements(data, "Z2")
                kmax = min(len(z1), len(z2), p["max_steps"])
                for k in range(kmax):
                    u, v = z2[k] if (p["slam"] or k % 2) else z1[k]
                    b.submit(rig, k, 0.5, u, v, not p["slam"] and k % 2 == 0)
                active[rig] = (n, kmax, 0)
                print(f"run {n}: {kmax} frames queued", flush=True)
            # wait for at least one run to finish, then free its slot
            finished = []
            while not finished:
                for r in b.collect():
                    n, kmax, got = active[r["rig"]]
                    active[r["rig"]] = (n, kmax, got + 1)
                    if r["status"] != 0:
                        failed.append((n, r["frame_id"], r["status"]))
                finished = [rig for rig, (n, kmax, got) in active.items() if got == kmax]
                if not finished:
                    time.sleep(0.005)
            for rig in finished:
                n = active.pop(rig)[0]
                b.remove_rig(rig)
                done += 1
                print(f"run {n} done ({done}/{args.n_runs}, {time.time() - t0:.1f} s)", flush=True)
    finally:
        b.close()
    if failed:
        print(f"failed frames: {failed[:10]}", file=sys.stderr)
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
