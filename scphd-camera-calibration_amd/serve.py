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

run_cohorts() serves the arrivals with shared launches: phase cohorts of
rigs, each an ensemble that rigs join and leave between its ticks, every
tick one set of launches over all of its rigs.

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


def cohort_schedule(rigs, cohorts, paced=True):
    """Phase cohorts of the served form with shared launches.  With frames
    every T seconds, cohort c steps at the ticks (n K + c) T / K; a rig
    arriving at t belongs to the cohort of the first tick at or after t
    (paced) -- its frame k, ready at t + k T, is then stepped at the k-th
    tick of that cohort after its first, at most T / K after it is ready.
    Not paced: every frame is ready at the start and rigs go round-robin.
    Returns, per rig, (cohort, index of its first tick in units of T / K)."""
    if not rigs:
        return []
    T = rigs[0]["s"]["dt"]
    if any(r["s"]["dt"] != T for r in rigs):
        raise ValueError("the cohorts need one frame period")
    out = []
    for i, r in enumerate(rigs):
        if paced:
            n = int(math.ceil(r["t"] / (T / cohorts) - 1e-9))
            out.append((n % cohorts, n))
        else:
            out.append((i % cohorts, i % cohorts))
    return out


def _per_gaussian_bytes(p):
    """Device bytes per map entry of a rig's footprint (scphd_rig_footprint
    between two map sizes)."""
    q = dict(p, n_particles=1024)
    return (scphd_amd.rig_footprint(q, 1024) - scphd_amd.rig_footprint(q, 0)) / (1024 * 1024)


def run_cohorts(rigs, hbm_budget_bytes=200e9, cohorts=4, paced=True, reserve_factor=1.5, keep_estimates=False,
                max_wall=None, progress=None):
    """The served form with shared launches: rigs stream through `cohorts`
    ensembles (cohort_schedule), one host thread each.  Every tick of a
    cohort is one predict / update / resample of all of its rigs, each with
    its own next frame and its own fixed-camera flag
    (scphd_update_ensemble_rigs); a rig joins its cohort's ensemble between
    ticks (scphd_ensemble_add_rig) once its reservation fits the cohort's
    share of the HBM budget, and leaves after its last frame
    (scphd_ensemble_remove_rig).  A rig that has to wait for admission
    starts late and catches up one frame per tick.  Each cohort's ensemble
    is created, sized for its share of the budget and emptied before the
    clock starts, so neither ticks nor joins allocate.  Per-frame latency = completion time - the time the frame
    was ready (arrival + k T); throughput = component updates / wall time.
    Rigs draw their noise from counters that start at zero when they join,
    so every rig's estimates equal those of its plain run."""
    import threading

    scphd_amd.lib()  # loaded once, before the threads
    sched = cohort_schedule(rigs, cohorts, paced)
    T = rigs[0]["s"]["dt"] if rigs else 1.0
    share = hbm_budget_bytes / cohorts
    est = [dict() for _ in rigs]
    lat = [[] for _ in range(cohorts)]
    pairs = [0] * cohorts
    ticks = [0] * cohorts
    peak = [0] * cohorts
    refused = [0] * cohorts
    adm_delay = [[] for _ in range(cohorts)]
    errors = []
    queues = []
    ens = []
    max_meas = max((expected_measurements(x["s"]) for x in rigs), default=1)
    for c in range(cohorts):
        q = sorted((i for i, (cc, _) in enumerate(sched) if cc == c), key=lambda i: sched[i][1])
        queues.append(q)
        e = None
        if q:  # created with its first rig, sized for the share, emptied
            e = scphd_amd.Ensemble([rigs[q[0]]["p"]])
            try:
                e.reserve(int(share / _per_gaussian_bytes(rigs[q[0]]["p"])), max_meas)
            except scphd_amd.ScphdError:
                pass  # the buffers grow as the maps do
            e.remove_rig(0)
        ens.append(e)
    t0 = time.perf_counter()

    def worker(c):
        queue = queues[c]
        e, members, used = ens[c], [], 0.0
        n = c
        try:
            while queue or members:
                tick_t = n * T / cohorts
                if paced:
                    wait = t0 + tick_t - time.perf_counter()
                    if wait > 0:
                        time.sleep(wait)
                now = time.perf_counter() - t0
                if max_wall is not None and now > max_wall:
                    break
                # admissions: rigs whose first tick has come, in order, while the share allows
                while queue and (not paced or sched[queue[0]][1] <= n):
                    i = queue[0]
                    r = rigs[i]
                    res = reserve_factor * scphd_amd.rig_footprint(r["p"], expected_gaussians_per_parent(r["s"]))
                    if members and used + res > share:
                        refused[c] += 1
                        break
                    e.add_rig(r["p"])
                    queue.pop(0)
                    used += res
                    members.append(dict(i=i, k=0, res=res, t_join=now))
                    adm_delay[c].append(now - (r["t"] if paced else 0.0))
                if not members:
                    n += cohorts
                    continue
                us, vs, fx = [], [], []
                for m in members:
                    u, v, fixed = rigs[m["i"]]["frames"][m["k"]]
                    us.append(u)
                    vs.append(v)
                    fx.append(fixed)
                e.predict(T)
                e.update(us, vs, fx)
                st = e.update_stats()
                if e.check() != 0:
                    raise RuntimeError(f"cohort {c}: nan weight")
                e.resample()
                mean, _, _ = e.rig_estimates()
                done_t = time.perf_counter() - t0
                pairs[c] += st["n_pairs"]
                ticks[c] += 1
                peak[c] = max(peak[c], len(members))
                for j, m in enumerate(members):
                    r = rigs[m["i"]]
                    ready = (r["t"] + m["k"] * T) if paced else 0.0
                    lat[c].append(done_t - ready)
                    if keep_estimates:
                        est[m["i"]][m["k"]] = mean[j].copy()
                    m["k"] += 1
                for j in reversed(range(len(members))):
                    m = members[j]
                    if m["k"] == len(rigs[m["i"]]["frames"]):
                        e.remove_rig(j)
                        used -= m["res"]
                        del members[j]
                n += cohorts
        except Exception as ex:  # reported by the caller
            errors.append(ex)
        finally:
            if e is not None:
                e.close()

    threads = [threading.Thread(target=worker, args=(c,)) for c in range(cohorts)]
    for th in threads:
        th.start()
    last = time.perf_counter()
    while any(th.is_alive() for th in threads):
        time.sleep(0.05)
        if progress and time.perf_counter() - last >= 5.0:
            last = time.perf_counter()
            progress(f"cohorts t={last - t0:.1f}s frames {sum(len(x) for x in lat)} ticks {sum(ticks)}")
    for th in threads:
        th.join()
    if errors:
        raise errors[0]
    wall = time.perf_counter() - t0
    allat = np.array([x for xs in lat for x in xs]) * 1e3
    adm = np.array([x for xs in adm_delay for x in xs])
    out = dict(rigs=len(rigs), frames=int(allat.size), wall_s=wall, cohorts=cohorts, paced=paced,
               updates_per_s=sum(pairs) / wall if wall > 0 else 0.0, pairs=int(sum(pairs)), ticks=int(sum(ticks)),
               latency_ms=dict(p50=_pct(allat, 50), p99=_pct(allat, 99),
                               max=float(allat.max()) if allat.size else None),
               admission_delay_s=dict(p50=_pct(adm, 50), p99=_pct(adm, 99),
                                      max=float(adm.max()) if adm.size else None),
               peak_rigs_per_cohort=int(max(peak)) if peak else 0, refused_admissions=int(sum(refused)),
               budget_gb=hbm_budget_bytes / 1e9)
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
