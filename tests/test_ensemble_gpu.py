"""Ensembles (scphd_create_ensemble): many rigs stepped by the same kernel
launches.

Each rig of an ensemble must evolve bit for bit as a plain filter created
with that rig's parameters and fed the same measurement sets: parents,
weights, maps, samples, per-rig estimates and per-rig resampling decisions.
The rigs are ragged: their parent counts, focal lengths, baselines (prior),
clutter rates and birth models differ, they see different scenes, so their
measurement counts differ every frame, and one rig gets an empty set on one
frame (M = 0 beside M > 0).
"""
import numpy as np
import pytest

import scene
import workloads

pytestmark = pytest.mark.gpu

SAMPLES = 64
RIGS = [dict(seed=101, n_particles=64, n_landmarks=25, lam=8.0, f=500.0, baseline=0.2),
        dict(seed=202, n_particles=96, n_landmarks=40, lam=15.0, f=650.0, baseline=0.3),
        dict(seed=303, n_particles=32, n_landmarks=15, lam=4.0, f=420.0, baseline=0.15),
        dict(seed=404, n_particles=48, n_landmarks=30, lam=10.0, f=560.0, baseline=0.25)]
FRAMES = 6
EMPTY = (2, 3)  # rig 2 sees no measurement on frame 3


def frames():
    """Per rig: params and the measurement sets (u, v) of every frame; the
    fixed-camera flag of rig 0's scene for every rig."""
    rigs = []
    for r, cfg in enumerate(RIGS):
        p, s = workloads.rig(n_frames=FRAMES, scene_rig=r, samples=SAMPLES, **cfg)
        sc = scene.generate(seed=cfg["seed"], **s)
        sets = []
        for k in range(FRAMES):
            u, v, fixed = scene.frame_measurements(sc, k)
            if (r, k) == EMPTY:
                u, v = np.zeros(0), np.zeros(0)
            sets.append((np.asarray(u, dtype=np.float64), np.asarray(v, dtype=np.float64), fixed))
        rigs.append((p, s, sets))
    fixed = [rigs[0][2][k][2] for k in range(FRAMES)]
    return rigs, fixed


def run_alone(p, dt, sets, fixed):
    import scphd_amd

    f = scphd_amd.Filter(p)
    est, rs = [], []
    try:
        for k, (u, v, _) in enumerate(sets):
            f.predict(dt)
            f.update(u, v, fixed[k])
            dbg = f.update_debug()
            assert f.check() == 0
            rs.append(f.resample())
            m, c, card = f.estimate()
            est.append(np.concatenate([m, c.ravel(), [card]]))
        return f.state(), np.array(est), rs, dbg
    finally:
        f.close()


def test_ensemble_rigs_match_plain_filters_bit_for_bit():
    import scphd_amd

    rigs, fixed = frames()
    dt = rigs[0][1]["dt"]
    ref = [run_alone(p, dt, sets, fixed) for p, _, sets in rigs]

    e = scphd_amd.Ensemble([r[0] for r in rigs])
    try:
        est = [[] for _ in RIGS]
        rs = [[] for _ in RIGS]
        for k in range(FRAMES):
            e.predict(dt)
            e.update([r[2][k][0] for r in rigs], [r[2][k][1] for r in rigs], fixed[k])
            dbg = e.update_debug()
            assert e.check() == 0
            e.resample()
            flags, neff = e.rig_resample()
            assert np.all(np.isfinite(neff))
            m, c, card = e.rig_estimates()
            for r in range(len(RIGS)):
                rs[r].append(bool(flags[r]))
                est[r].append(np.concatenate([m[r], c[r].ravel(), [card[r]]]))
        st = e.state()
    finally:
        e.close()

    S = SAMPLES
    g0, p0 = 0, 0
    for r, (rst, rest, rrs, rdbg) in enumerate(ref):
        n = RIGS[r]["n_particles"]
        sl = slice(p0, p0 + n)
        p0 += n
        m = rdbg["eta"].shape[1]  # the last frame's normalisers and parent log-likelihoods
        np.testing.assert_array_equal(dbg["eta"][sl, :m], rdbg["eta"], err_msg=f"rig {r} eta")
        np.testing.assert_array_equal(dbg["loglik"][sl], rdbg["loglik"], err_msg=f"rig {r} loglik")
        np.testing.assert_array_equal(st["states"][sl], rst["states"], err_msg=f"rig {r} states")
        np.testing.assert_array_equal(st["weights"][sl], rst["weights"], err_msg=f"rig {r} weights")
        np.testing.assert_array_equal(st["gcount"][sl], rst["gcount"], err_msg=f"rig {r} gcount")
        ng = int(rst["gcount"].sum())
        np.testing.assert_array_equal(st["wg"][g0:g0 + ng], rst["wg"], err_msg=f"rig {r} map weights")
        np.testing.assert_array_equal(st["samples"][:, g0 * S:(g0 + ng) * S], rst["samples"],
                                      err_msg=f"rig {r} samples")
        np.testing.assert_array_equal(np.array(est[r]), rest, err_msg=f"rig {r} estimates")
        assert rs[r] == rrs, f"rig {r} resampling decisions"
        g0 += ng
    assert g0 == int(st["gcount"].sum())


# the step at which each rig joins the ensemble, and the frames it streams
# before it leaves: rig 2 joins beside rigs 0 and 1, rig 0 leaves from the
# front while the others stay, all leave by step 6, rig 3 joins the empty
# ensemble at step 7
JOIN = [0, 0, 2, 7]
RUN = [3, 5, 4, 3]


def _rig_slice(st, poff, i, S):
    """Rig i's parents, weights and map out of an ensemble state."""
    sl = slice(poff[i], poff[i + 1])
    gc = st["gcount"]
    g0, g1 = int(gc[:poff[i]].sum()), int(gc[:poff[i + 1]].sum())
    return dict(states=st["states"][sl], weights=st["weights"][sl], gcount=gc[sl], wg=st["wg"][g0:g1],
                samples=st["samples"][:, g0 * S:g1 * S])


def test_rigs_joining_and_leaving_between_frames_match_plain_filters():
    """Served form: rigs join an ensemble between frames (scphd_ensemble_add_rig)
    and leave when their stream ends (scphd_ensemble_remove_rig), each step
    covering every member with its own next frame and its own fixed-camera
    flag (scphd_update_ensemble_rigs).  Rig 0 leaves from the front while
    others stay; rig 3 joins after every other rig has left.  Every rig's per-frame estimates and
    resampling decisions, and its parents, weights and map when it leaves,
    are bit-identical to a plain filter of that rig fed the same frames."""
    import scphd_amd

    rigs, _ = frames()
    dt = rigs[0][1]["dt"]
    ref = []
    for r, (p, _, sets) in enumerate(rigs):
        ref.append(run_alone(p, dt, sets[:RUN[r]], [sets[k][2] for k in range(RUN[r])]))
    e = None
    members, done = [], 0
    est = [[] for _ in rigs]
    rs = [[] for _ in rigs]
    final = {}
    step = 0
    try:
        while done < len(rigs):
            # leave: members whose stream has ended (the ensemble's last rig
            # closes it instead)
            for r in [m for m in members if len(est[m]) == RUN[m]]:
                i = members.index(r)
                final[r] = _rig_slice(e.state(), e.poff, i, SAMPLES)
                e.remove_rig(i)
                members.remove(r)
                done += 1
            for r in range(len(rigs)):
                if JOIN[r] == step:
                    if e is None:
                        e = scphd_amd.Ensemble([rigs[r][0]])
                    else:  # joins, also after every rig has left
                        assert e.add_rig(rigs[r][0]) == len(members)
                    members.append(r)
            if not members:
                step += 1
                continue
            ks = [len(est[r]) for r in members]
            e.predict(dt)
            e.update([rigs[r][2][k][0] for r, k in zip(members, ks)], [rigs[r][2][k][1] for r, k in zip(members, ks)],
                     [rigs[r][2][k][2] for r, k in zip(members, ks)])
            assert e.check() == 0
            e.resample()
            flags, _ = e.rig_resample()
            m, c, card = e.rig_estimates()
            for i, r in enumerate(members):
                rs[r].append(bool(flags[i]))
                est[r].append(np.concatenate([m[i], c[i].ravel(), [card[i]]]))
            step += 1
    finally:
        if e is not None:
            e.close()
    for r, (rst, rest, rrs, _) in enumerate(ref):
        for key in ("states", "weights", "gcount", "wg", "samples"):
            np.testing.assert_array_equal(final[r][key], rst[key], err_msg=f"rig {r} {key}")
        np.testing.assert_array_equal(np.array(est[r]), rest, err_msg=f"rig {r} estimates")
        assert rs[r] == rrs, f"rig {r} resampling decisions"


def test_ensemble_membership_refusals():
    """Membership changes only between frames and only with the parameters
    the rigs share; an ensemble whose rigs have all left takes new rigs and
    refuses steps."""
    import scphd_amd

    p, _ = workloads.rig(samples=SAMPLES, **RIGS[0])
    q, _ = workloads.rig(samples=SAMPLES, **RIGS[1])
    e = scphd_amd.Ensemble([p])
    try:
        e.remove_rig(0)
        assert e.Np == 0 and e.n_rigs == 0
        with pytest.raises(scphd_amd.ScphdError, match="STATE"):
            e.predict(0.1)
        assert e.add_rig(p) == 0
        with pytest.raises(scphd_amd.ScphdError, match="ARG"):
            e.add_rig(dict(q, min_weight=1e-3))
        e.predict(0.1)
        with pytest.raises(scphd_amd.ScphdError, match="STATE"):
            e.add_rig(q)
        e.update([np.zeros(0)], [np.zeros(0)], True)
        assert e.add_rig(q) == 1
        with pytest.raises(scphd_amd.ScphdError, match="ARG"):
            e.remove_rig(2)
        e.remove_rig(0)
        assert e.n_rigs == 1 and e.Np == q["n_particles"]
    finally:
        e.close()


def test_ensemble_refuses_plain_updates_and_bad_offsets():
    import scphd_amd

    p, _ = workloads.rig(samples=SAMPLES, **RIGS[0])
    q, _ = workloads.rig(samples=SAMPLES, **RIGS[1])
    bad_rig = dict(q, min_weight=1e-3)  # a shared parameter that differs
    with pytest.raises(scphd_amd.ScphdError):
        scphd_amd.Ensemble([p, bad_rig])
    e = scphd_amd.Ensemble([p, q])
    try:
        with pytest.raises(scphd_amd.ScphdError):
            scphd_amd.Filter.update(e, np.zeros(3), np.zeros(3), True)
        bad = np.array([0, 5, 3], dtype=np.int32)
        u = np.zeros(5)
        rc = scphd_amd.lib().scphd_update_ensemble(e.h, scphd_amd._dp(u), scphd_amd._dp(u), scphd_amd._ip(bad), 1)
        assert scphd_amd.ERRORS.get(rc) == "SCPHD_ERR_ARG"
        with pytest.raises(scphd_amd.ScphdError):
            e.rig_resample()  # no resampling step yet
    finally:
        e.close()


def test_reserve_changes_no_result():
    """scphd_reserve pre-sizes the buffers; the steps that follow give the
    same bits as steps that grow the buffers as they go."""
    import scphd_amd

    rigs, fixed = frames()
    p, s, sets = rigs[1]
    out = []
    for reserve in (False, True):
        f = scphd_amd.Filter(p)
        try:
            if reserve:
                before = f.memory()
                f.reserve(p["n_particles"] * 200, 64)
                assert f.memory() > before
            for k, (u, v, _) in enumerate(sets):
                f.predict(s["dt"])
                f.update(u, v, fixed[k])
                f.resample()
            out.append(f.state())
        finally:
            f.close()
    for key in ("states", "weights", "gcount", "wg", "samples"):
        np.testing.assert_array_equal(out[0][key], out[1][key], err_msg=key)


def test_compact_changes_no_result_and_hands_memory_back():
    """scphd_compact between frames (the map's clouds into a generation of
    their own size, the spare generation and work arrays released): the
    frames that follow give the same bits as without it, and the handle
    holds less memory right after it."""
    import scphd_amd

    rigs, fixed = frames()
    p, s, sets = rigs[1]
    out = []
    for compact in (False, True):
        f = scphd_amd.Filter(p)
        try:
            f.reserve(p["n_particles"] * 400, 64)  # far more than the map needs
            for k, (u, v, _) in enumerate(sets):
                f.predict(s["dt"])
                f.update(u, v, fixed[k])
                f.resample()
                if compact and k in (1, 3):
                    before = f.memory()
                    f.compact()
                    assert f.memory() < before
                    with pytest.raises(scphd_amd.ScphdError, match="STATE"):
                        f.update_debug()  # released with the work arrays
            out.append(f.state())
        finally:
            f.close()
    for key in ("states", "weights", "gcount", "wg", "samples"):
        np.testing.assert_array_equal(out[0][key], out[1][key], err_msg=key)


BIG = [dict(seed=505, n_particles=4096, n_landmarks=200, lam=50.0, f=500.0, baseline=0.2),
       dict(seed=606, n_particles=1024, n_landmarks=120, lam=30.0, f=600.0, baseline=0.25)]


def test_ensemble_of_full_size_rigs_matches_plain_filters_frame_by_frame():
    """An ensemble of a 4096-parent rig with S = 200 (the bench workload's
    size) and a 1024-parent rig, stepped beside one plain filter per rig:
    every frame, parents, weights, normalisers, log-likelihoods, per-parent
    survivor / merged / non-detection counts, per-rig estimates and
    resampling decisions are bit-identical.  The maps themselves (tens of GB
    at this size) are covered through those counts and the normalisers,
    which depend on every sample."""
    import scphd_amd

    frames_n = 4
    rigs = []
    for r, cfg in enumerate(BIG):
        p, s = workloads.rig(n_frames=frames_n, scene_rig=r, samples=200, **cfg)
        sc = scene.generate(seed=cfg["seed"], **s)
        rigs.append((p, s, [scene.frame_measurements(sc, k) for k in range(frames_n)]))
    dt = rigs[0][1]["dt"]
    plain = [scphd_amd.Filter(p) for p, _, _ in rigs]
    e = scphd_amd.Ensemble([p for p, _, _ in rigs])
    try:
        for k in range(frames_n):
            fixed = rigs[0][2][k][2]
            us = [np.asarray(r[2][k][0], dtype=np.float64) for r in rigs]
            vs = [np.asarray(r[2][k][1], dtype=np.float64) for r in rigs]
            e.predict(dt)
            e.update(us, vs, fixed)
            de = e.update_debug()
            rse = e.resample()
            flags, _ = e.rig_resample()
            me, ce, carde = e.rig_estimates()
            se, we = e.parents()
            p0 = 0
            for r, f in enumerate(plain):
                f.predict(dt)
                f.update(us[r], vs[r], fixed)
                dp = f.update_debug()
                rs = f.resample()
                m, c, card = f.estimate()
                s_, w_ = f.parents()
                n = BIG[r]["n_particles"]
                sl = slice(p0, p0 + n)
                p0 += n
                M = len(us[r])
                what = f"frame {k} rig {r}"
                np.testing.assert_array_equal(de["eta"][sl, :M], dp["eta"], err_msg=f"{what} eta")
                for key in ("loglik", "nsurvivors", "nmerged", "nnondetect"):
                    np.testing.assert_array_equal(de[key][sl], dp[key], err_msg=f"{what} {key}")
                np.testing.assert_array_equal(se[sl], s_, err_msg=f"{what} states")
                np.testing.assert_array_equal(we[sl], w_, err_msg=f"{what} weights")
                assert bool(flags[r]) == rs, f"{what} resampling decision"
                np.testing.assert_array_equal(me[r], m, err_msg=f"{what} mean")
                np.testing.assert_array_equal(ce[r], c, err_msg=f"{what} covariance")
                assert carde[r] == card, f"{what} cardinality"
            assert rse == any(bool(x) for x in flags)
        assert e.counts()["n_gaussians"] > 4096 * 100
    finally:
        e.close()
        for f in plain:
            f.close()
