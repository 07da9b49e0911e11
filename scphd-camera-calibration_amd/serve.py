"""Served calibration: rigs arrive over time, are admitted by HBM reservation,
stream frames at a fixed rate, and share one GPU through the batch API
(scphd_batch_* in include/scphd.h).

run_served() replays an arrival list in real time:
  * an arriving rig waits in a FIFO until its reservation fits the HBM budget
    (scphd_batch_add_rig); the wait is its admission delay;
  * once admitted, its frame k is due at t_admit + k / fps and is submitted
    when due (frames that come due while earlier ones are still queued or on
    the GPU wait in the rig's queue: that wait is part of their latency);
  * a rig whose frames are all done leaves (scphd_batch_remove_rig) and frees
    its reservation for the rigs waiting to be admitted.
Per-frame latency = completion time - due time.  Throughput = component
updates (N_g * M per frame) / wall time.

run_plain() runs the same rigs through the plain path: one rig at a time, its
frames back to back on one Filter (buffers reserved at the expected map size
first), for the throughput and per-frame latency of one rig per launch.

run_ensemble() steps the same rigs together as one ensemble: one launch per
kernel per frame for all of them.

Neither loads anything but the library; scene generation happens before the
clock starts.
"""
import math
import time

import numpy as np

import scene
import scphd_amd


def expected_gaussians_per_parent(s):
    """Steady-state map size estimate: each landmark carries a detection
    term, a chain of 3-4 non-detection clouds (weights (1 - Pd)^k down to
    min_weight) and the terms of the landmarks that have left the view
    (weights ps^k, alive for many frames), and clutter births survive a frame
    or two: about 4.75 L + 1.5 lambda (measured: 150 Gaussians per parent for
    30 landmarks and lambda 10, 1000 for 200 landmarks and lambda 50)."""
    return int(4.75 * s["n_landmarks"] + 1.5 * s["lam"]) + 16


def expected_measurements(s):
    """A measurement count a frame rarely exceeds: the detected visible
    landmarks (0.95 * 0.9 * L) plus clutter lambda and four standard
    deviations of it."""
    return int(0.95 * 0.9 * s["n_landmarks"] + s["lam"] + 4.0 * math.sqrt(s["lam"])) + 8


def prepare(arrivals):
    """Draw every rig's measurement sets (before any timing)."""
    rigs = []
    for t, p, s in arrivals:
        sc = scene.generate(seed=p["seed"], **s)
        frames = [scene.frame_measurements(sc, k) for k in range(s["n_frames"])]
        rigs.append(dict(t=t, p=p, s=s, frames=frames))
    return rigs


def _pct(x, q):
    return float(np.percentile(x, q)) if len(x) else None


def run_served(rigs, hbm_budget_bytes=0, n_workers=8, reserve_factor=1.5, progress=None, max_wall=None,
               paced=True, keep_estimates=False):
    """Replay prepared rigs (see prepare) in real time; returns a summary
    dict.  paced=False submits all of a rig's frames when it is admitted and
    admits every rig at once (as the budget allows): the saturated
    throughput of the served form.  keep_estimates: the summary also holds
    "estimates", per rig (in the order of rigs) the per-frame calibration
    estimates (12 values each), in frame order."""
    est = [dict() for _ in rigs]
    b = scphd_amd.Batch(hbm_budget_bytes=hbm_budget_bytes, n_workers=n_workers)
    try:
        waiting = []          # indices of arrived, not yet admitted rigs (FIFO)
        active = {}           # rig id -> state
        nxt = 0               # next arrival
        blocked = False       # the head of the admission queue was refused and nothing has left since
        lat, svc, adm, pairs = [], [], [], 0
        done_rigs, peak_res, peak_used, peak_active, n_frames = 0, 0, 0, 0, 0
        t0 = b.now()
        last_print = t0
        while nxt < len(rigs) or waiting or active:
            now = b.now() - t0
            if max_wall is not None and now > max_wall:
                break
            while nxt < len(rigs) and (rigs[nxt]["t"] <= now or not paced):
                waiting.append(nxt)
                nxt += 1
            while waiting and not blocked:
                r = rigs[waiting[0]]
                res = int(reserve_factor * scphd_amd.rig_footprint(r["p"], expected_gaussians_per_parent(r["s"])))
                rid = b.add_rig(r["p"], res)
                if rid is None:
                    blocked = True  # offered again once a rig leaves
                    break
                idx = waiting.pop(0)
                adm.append(now - r["t"] if paced else 0.0)
                active[rid] = dict(r=r, idx=idx, t_admit=now, sent=0, done=0, due={})
            for rid, a in active.items():
                fps_dt = a["r"]["s"]["dt"]
                fr = a["r"]["frames"]
                while a["sent"] < len(fr) and (a["t_admit"] + a["sent"] * fps_dt <= now or not paced):
                    k = a["sent"]
                    u, v, fixed = fr[k]
                    a["due"][k] = t0 + (a["t_admit"] + k * fps_dt if paced else now)
                    b.submit(rid, k, fps_dt, u, v, fixed)
                    a["sent"] += 1
            for x in b.collect():
                a = active[x["rig"]]
                if x["status"] != 0:
                    raise RuntimeError(f"rig {x['rig']} frame {x['frame_id']} failed: status {x['status']}")
                lat.append(x["t_end"] - a["due"].pop(x["frame_id"]))
                svc.append(x["t_end"] - x["t_start"])
                if keep_estimates:
                    est[a["idx"]][x["frame_id"]] = np.array(x["estimate"])
                pairs += x["n_pairs"]
                n_frames += 1
                a["done"] += 1
            for rid in [rid for rid, a in active.items() if a["done"] == len(a["r"]["frames"])]:
                b.remove_rig(rid)
                del active[rid]
                done_rigs += 1
                blocked = False
            st = b.stats()
            peak_res = max(peak_res, st["reserved_bytes"])
            peak_used = max(peak_used, st["used_bytes"])
            peak_active = max(peak_active, len(active))
            if progress and now - (last_print - t0) >= 5.0:
                last_print = b.now()
                progress(f"served t={now:.1f}s rigs done {done_rigs}/{len(rigs)} active {len(active)} "
                         f"waiting {len(waiting)} frames {n_frames} queued {st['queued_frames']}")
            # next event: a due frame, an arrival, or a completion
            time.sleep(0.0005)
        wall = b.now() - t0
        st = b.stats()
    finally:
        b.close()
    lat = np.array(lat) * 1e3
    svc = np.array(svc) * 1e3
    adm = np.array(adm)
    out = dict(rigs=len(rigs), rigs_done=done_rigs, frames=n_frames, wall_s=wall,
                updates_per_s=pairs / wall if wall > 0 else 0.0, pairs=pairs,
                latency_ms=dict(p50=_pct(lat, 50), p99=_pct(lat, 99), max=float(lat.max()) if len(lat) else None),
                service_ms=dict(p50=_pct(svc, 50), p99=_pct(svc, 99)),
                admission_delay_s=dict(p50=_pct(adm, 50), p99=_pct(adm, 99), max=float(adm.max()) if len(adm) else None),
                peak_active_rigs=peak_active, peak_reserved_gb=peak_res / 1e9, peak_used_gb=peak_used / 1e9,
                refused_admissions=st["refused"], budget_gb=st["budget_bytes"] / 1e9, n_workers=n_workers,
                paced=paced)
    if keep_estimates:
        out["estimates"] = [np.array([e[k] for k in sorted(e)]) for e in est]
    return out


def run_plain(rigs, progress=None):
    """The same rigs one at a time through one Filter each, frames back to back."""
    lat, pairs, t_all = [], 0, 0.0
    for i, r in enumerate(rigs):
        f = scphd_amd.Filter(r["p"])
        try:  # steady-state sizes up front, as for an ensemble
            f.reserve(r["p"]["n_particles"] * expected_gaussians_per_parent(r["s"]), expected_measurements(r["s"]))
        except scphd_amd.ScphdError:
            pass
        dt = r["s"]["dt"]
        for u, v, fixed in r["frames"]:
            t0 = time.perf_counter()
            f.predict(dt)
            f.update(u, v, fixed)
            st = f.update_stats()
            if f.check() != 0:
                raise RuntimeError("nan weight")
            f.resample()
            f.estimate()
            t1 = time.perf_counter()
            lat.append(t1 - t0)
            pairs += st["n_pairs"]
            t_all += t1 - t0
        f.close()
        if progress and (i + 1) % 10 == 0:
            progress(f"plain rigs {i + 1}/{len(rigs)}")
    lat = np.array(lat) * 1e3
    return dict(rigs=len(rigs), frames=len(lat), wall_s=t_all, updates_per_s=pairs / t_all if t_all > 0 else 0.0,
                pairs=pairs, latency_ms=dict(p50=_pct(lat, 50), p99=_pct(lat, 99), max=float(lat.max())))


def run_ensemble(rigs, progress=None, growth=1.0):
    """The same rigs stepped together as one ensemble (scphd_create_ensemble):
    the rigs share the sample count, pruning, merging and motion settings
    (the library checks) and the frame sequence length and dt; an ensemble
    frame is one frame of every rig (predict, update, NaN check, resample,
    every rig's estimate).  Latency is per ensemble frame.  The buffers are
    reserved for `growth` times the expected map size before the first
    frame."""
    for r in rigs:
        if len(r["frames"]) != len(rigs[0]["frames"]) or r["s"]["dt"] != rigs[0]["s"]["dt"]:
            raise ValueError("an ensemble needs one frame sequence length and one dt")
    e = scphd_amd.Ensemble([r["p"] for r in rigs])
    try:  # the maps' steady-state size, allocated before the first frame
        e.reserve(int(growth * sum(r["p"]["n_particles"] * expected_gaussians_per_parent(r["s"]) for r in rigs)),
                  max(expected_measurements(r["s"]) for r in rigs))
    except scphd_amd.ScphdError:
        pass  # not enough memory to reserve up front: the buffers grow as the maps do
    dt = rigs[0]["s"]["dt"]
    lat, pairs, t_all = [], 0, 0.0
    try:
        for k in range(len(rigs[0]["frames"])):
            us = [r["frames"][k][0] for r in rigs]
            vs = [r["frames"][k][1] for r in rigs]
            fixed = rigs[0]["frames"][k][2]
            t0 = time.perf_counter()
            e.predict(dt)
            e.update(us, vs, fixed)
            st = e.update_stats()
            if e.check() != 0:
                raise RuntimeError("nan weight")
            e.resample()
            e.rig_estimates()
            t1 = time.perf_counter()
            lat.append(t1 - t0)
            pairs += st["n_pairs"]
            t_all += t1 - t0
            if progress and (k + 1) % 20 == 0:
                progress(f"ensemble frames {k + 1}/{len(rigs[0]['frames'])}")
    finally:
        e.close()
    lat = np.array(lat) * 1e3
    return dict(rigs=len(rigs), frames=len(lat), wall_s=t_all, updates_per_s=pairs / t_all if t_all > 0 else 0.0,
                pairs=pairs, latency_ms=dict(p50=_pct(lat, 50), p99=_pct(lat, 99), max=float(lat.max())))


def pack_ensembles(rigs, budget_gb, growth=2.0, alone_parents=None, plane_bytes=4 << 30):
    """Split prepared rigs into ensembles, in order: a group closes when the
    next rig would take its summed footprint (scphd_rig_footprint at `growth`
    times the expected map size) past budget_gb, or its sample planes (at
    that size, with the buffers' growth slack of 1.25) past plane_bytes, the
    limit of the kernels' 32-bit offsets.  Rigs with at least alone_parents
    parents (when given) are returned apart, to run alone.  Returns (groups,
    alone)."""
    alone = [r for r in rigs if alone_parents is not None and r["p"]["n_particles"] >= alone_parents]
    groups, cur, used, blocks = [], [], 0.0, 0.0
    for r in rigs:
        if alone_parents is not None and r["p"]["n_particles"] >= alone_parents:
            continue
        gp = int(growth * expected_gaussians_per_parent(r["s"]))
        need = scphd_amd.rig_footprint(r["p"], gp) / 1e9
        esz = 8 if r["p"].get("storage", 0) == 1 else 4
        max_blocks = plane_bytes / (r["p"]["particles_per_feature"] * esz) / 1.25
        nb = r["p"]["n_particles"] * gp
        if cur and (used + need > budget_gb or blocks + nb > max_blocks):
            groups.append(cur)
            cur, used, blocks = [], 0.0, 0.0
        cur.append(r)
        used += need
        blocks += nb
    if cur:
        groups.append(cur)
    return groups, alone


def offered_load(rigs, step_ms_per_gaussian_sample):
    """Rough device time the arrivals ask for per second of wall time, from a
    measured cost per (Gaussian x sample) of one step."""
    if not rigs:
        return 0.0
    horizon = max(r["t"] + len(r["frames"]) * r["s"]["dt"] for r in rigs)
    work = sum(len(r["frames"]) * r["p"]["n_particles"] * expected_gaussians_per_parent(r["s"]) *
               r["p"]["particles_per_feature"] for r in rigs)
    return work * step_ms_per_gaussian_sample * 1e-3 / horizon if horizon > 0 else math.inf
