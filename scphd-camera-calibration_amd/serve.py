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


def cohort_admit(pending, arrive, need, tick_t, left, empty):
    """The rigs a cohort takes at a tick at time tick_t: in the order of
    `pending` (arrival order), every rig that has arrived (arrive[i] <=
    tick_t) and whose reservation need[i] fits `left`, what is left of the
    budget; with nothing admitted (`empty`) the first arrived rig is taken
    whatever its size, so that no rig waits forever.  Rigs that have not
    arrived end the scan.  Returns (taken, passed over)."""
    take, skip = [], []
    for i in pending:
        if arrive[i] > tick_t:
            break
        if need[i] <= left or (empty and not take):
            take.append(i)
            left -= need[i]
        else:
            skip.append(i)
    return take, skip


def _per_gaussian_bytes(p):
    """Device bytes per map entry of a rig's footprint (scphd_rig_footprint
    between two map sizes)."""
    q = dict(p, n_particles=1024)
    return (scphd_amd.rig_footprint(q, 1024) - scphd_amd.rig_footprint(q, 0)) / (1024 * 1024)


def run_cohorts(rigs, hbm_budget_bytes=200e9, cohorts=4, paced=True, reserve_factor=1.5, keep_estimates=False,
                max_wall=None, progress=None):
    """The served form with shared launches: rigs stream through K
    ensembles (cohorts), one host thread each, cohort c stepping at the
    ticks (n K + c) T / K (frames every T).  K is the budget over the
    largest rig's reservation, at most `cohorts`, so that every rig fits a
    cohort's share.  Every tick is one
    predict / update / resample of all of the cohort's rigs, each with its
    own next frame and its own fixed-camera flag (scphd_update_ensemble_rigs).

    Memory and admission: each cohort's ensemble is created, its buffers
    sized for an equal share of the HBM budget, and emptied before the
    clock starts, so that ticks and joins do not allocate.  At each of its
    ticks a cohort takes, in arrival order, the arrived rigs whose
    reservations (reserve_factor times scphd_rig_footprint at the expected
    map size) fit what is left of its share (cohort_admit; smaller rigs may
    pass one that does not fit).  Should a cohort's maps outgrow the
    estimates, its buffers grow; once the rigs that grew them have left,
    they are compacted (scphd_compact) and sized for the share again.

    A rig is stepped from the first tick that takes it -- its phase
    cohort's, at most T / K after it arrives, when that one has room.  It
    joins between ticks (scphd_ensemble_add_rig), its frame k -- ready k T
    after its arrival -- is stepped k periods after its first tick (a rig
    that waited keeps its delay), and it leaves after its last frame
    (scphd_ensemble_remove_rig).

    Per-frame latency = completion time - the time the frame was ready
    (arrival + k T; not paced: 0); throughput = component updates / wall
    time.  Rigs draw their noise from counters that start at zero when they
    join, so every rig's estimates equal those of its plain run."""
    import threading

    scphd_amd.lib()  # loaded once, before the threads
    T = rigs[0]["s"]["dt"] if rigs else 1.0
    if any(r["s"]["dt"] != T for r in rigs):
        raise ValueError("the cohorts need one frame period")
    booked = [0.0]  # reservations of the admitted rigs, over all cohorts
    arrive = [r["t"] if paced else 0.0 for r in rigs]
    need = [reserve_factor * scphd_amd.rig_footprint(r["p"], expected_gaussians_per_parent(r["s"])) for r in rigs]
    cohorts = max(1, min(cohorts, int(hbm_budget_bytes // max(need, default=1.0))))
    share = hbm_budget_bytes / cohorts
    pending = sorted(range(len(rigs)), key=lambda i: arrive[i])
    est = [dict() for _ in rigs]
    lat = [[] for _ in range(cohorts)]
    pairs = [0] * cohorts
    ticks = [0] * cohorts
    peak = [0] * cohorts
    waited = set()  # rigs that found no room at some tick after arriving
    adm_delay = [[] for _ in range(cohorts)]
    errors = []
    lock = threading.Lock()
    held = [0] * cohorts  # device bytes each cohort's ensemble holds
    peak_held = [0]
    ens = []
    max_meas = max((expected_measurements(x["s"]) for x in rigs), default=1)

    def size_for_share(e, p):
        try:
            e.reserve(int(share / _per_gaussian_bytes(p)), max_meas)
        except scphd_amd.ScphdError:
            pass  # the buffers grow as the maps do

    for c in range(min(cohorts, len(rigs))):  # created with a rig, sized, emptied
        e = scphd_amd.Ensemble([rigs[c]["p"]])
        size_for_share(e, rigs[c]["p"])
        e.remove_rig(0)
        ens.append(e)
    t0 = time.perf_counter()

    def worker(c):
        e, members = ens[c], []
        n = c
        try:
            while True:
                tick_t = n * T / cohorts
                if paced:
                    wait = t0 + tick_t - time.perf_counter()
                    if wait > 0:
                        time.sleep(wait)
                now = time.perf_counter() - t0
                if max_wall is not None and now > max_wall:
                    break
                with lock:
                    if not pending and not members:
                        break
                    left = min(share - sum(need[m["i"]] for m in members), hbm_budget_bytes - booked[0])
                    # (a rig larger than the budget runs alone in a cohort)
                    take, skip = cohort_admit(pending, arrive, need, tick_t, left, not members and booked[0] == 0.0)
                    waited.update(skip)
                    for i in take:
                        pending.remove(i)
                        booked[0] += need[i]
                for i in take:
                    e.add_rig(rigs[i]["p"])
                    members.append(dict(i=i, k=0, first=n))
                    adm_delay[c].append(tick_t - arrive[i] if paced else now)
                if not members:
                    if not paced:
                        time.sleep(0.001)  # idle: leave the host to the working cohorts
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
                with lock:
                    held[c] = e.memory()
                    peak_held[0] = max(peak_held[0], sum(held))
                for j, m in enumerate(members):
                    lat[c].append(done_t - (arrive[m["i"]] + m["k"] * T if paced else 0.0))
                    if keep_estimates:
                        est[m["i"]][m["k"]] = mean[j].copy()
                    m["k"] += 1
                gone = None
                for j in reversed(range(len(members))):
                    m = members[j]
                    if m["k"] == len(rigs[m["i"]]["frames"]):
                        e.remove_rig(j)
                        with lock:
                            booked[0] -= need[m["i"]]
                        del members[j]
                        gone = rigs[m["i"]]["p"]
                # buffers keep their largest size: once the rig that grew
                # them past the share has left, they are sized for it again
                if gone is not None and e.memory() > 1.25 * share and sum(need[m["i"]] for m in members) <= share:
                    e.compact()
                    size_for_share(e, gone)
                    with lock:
                        held[c] = e.memory()
                n += cohorts
        except Exception as ex:  # reported by the caller, with the cohorts' sizes
            with lock:
                sizes = [(h.n_rigs, h.Np, h.counts()["n_gaussians"], round(h.memory() / 1e9, 1)) for h in ens]
            errors.append(RuntimeError(f"cohort {c}: {ex}; per cohort (rigs, parents, Gaussians, GB held): {sizes}; "
                                       f"reserved {booked[0] / 1e9:.1f} GB of {hbm_budget_bytes / 1e9:.0f}"))

    threads = [threading.Thread(target=worker, args=(c,)) for c in range(len(ens))]
    try:
        for th in threads:
            th.start()
        last = time.perf_counter()
        while any(th.is_alive() for th in threads):
            time.sleep(0.05)
            if progress and time.perf_counter() - last >= 5.0:
                last = time.perf_counter()
                progress(f"cohorts t={last - t0:.1f}s frames {sum(len(x) for x in lat)} ticks {sum(ticks)} "
                         f"waiting {len(pending)} reserved {booked[0] / 1e9:.1f} GB held {sum(held) / 1e9:.1f} GB")
        for th in threads:
            th.join()
    finally:
        for e in ens:
            e.close()
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
               peak_rigs_per_cohort=int(max(peak)) if peak else 0, rigs_that_waited=len(waited),
               budget_gb=hbm_budget_bytes / 1e9, peak_held_gb=peak_held[0] / 1e9)
    if keep_estimates:
        out["estimates"] = [np.array([e[k] for k in sorted(e)]) for e in est]
    return out


def run_plain(rigs, progress=None):
    """The same rigs one at a time through one Filter each, frames back to back."""
    lat, pairs, t_all, gpp = [], 0, 0.0, []
    for i, r in enumerate(rigs):
        f = scphd_amd.Filter(r["p"])
        try:  # steady-state sizes up front (a generation holds the input map and its posterior clouds)
            f.reserve(int(1.5 * r["p"]["n_particles"] * expected_gaussians_per_parent(r["s"])),
                      expected_measurements(r["s"]))
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
        gpp.append(f.counts()["n_gaussians"] / r["p"]["n_particles"] / expected_gaussians_per_parent(r["s"]))
        f.close()
        if progress and (i + 1) % 10 == 0:
            progress(f"plain rigs {i + 1}/{len(rigs)}")
    lat = np.array(lat) * 1e3
    return dict(rigs=len(rigs), frames=len(lat), wall_s=t_all, updates_per_s=pairs / t_all if t_all > 0 else 0.0,
                pairs=pairs, latency_ms=dict(p50=_pct(lat, 50), p99=_pct(lat, 99), max=float(lat.max())),
                final_map_over_expected=dict(p50=_pct(gpp, 50), max=float(max(gpp)) if gpp else None))


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
