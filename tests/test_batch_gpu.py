"""Batched / served operation (scphd_batch_*) on the GPU.

* A rig's results are the same alone and in a mix: each rig is run alone
  through a plain Filter, then together with other rigs through one Batch
  with its frames interleaved with theirs; parents, weights, maps and the
  per-frame estimates must be bit-identical; also at full size (a
  4096-parent rig with S = 200 beside a 2048-parent rig).
* Admission by HBM reservation, removal, and re-admission.
* The device-side estimate (scphd_get_estimate) against numpy on the
  parents read back.
"""
import numpy as np
import pytest

import scene
import workloads

pytestmark = pytest.mark.gpu

RIGS = [dict(seed=11, n_particles=64, n_landmarks=25, lam=8.0, f=500.0, baseline=0.2, samples=64),
        dict(seed=12, n_particles=96, n_landmarks=40, lam=15.0, f=650.0, baseline=0.3, samples=50),
        dict(seed=13, n_particles=48, n_landmarks=15, lam=4.0, f=420.0, baseline=0.15, samples=80)]
FRAMES = 6


def make(i, frames=FRAMES):
    p, s = workloads.rig(n_frames=frames, scene_rig=i, **RIGS[i])
    sc = scene.generate(seed=p["seed"], **s)
    return p, s, [scene.frame_measurements(sc, k) for k in range(frames)]


def alone(i):
    import scphd_amd

    p, s, sets = make(i)
    f = scphd_amd.Filter(p)
    est = []
    for u, v, fixed in sets:
        f.predict(s["dt"])
        f.update(u, v, fixed)
        assert f.check() == 0
        f.resample()
        est.append(f.estimate()[0])
    out = f.state(), np.array(est)
    f.close()
    return out


def test_rig_results_are_identical_alone_and_in_a_mix():
    import scphd_amd

    ref = [alone(i) for i in range(len(RIGS))]
    b = scphd_amd.Batch(n_workers=3)
    try:
        rigs, sets = [], []
        for i in range(len(RIGS)):
            p, s, fs = make(i)
            rigs.append(b.add_rig(p))
            sets.append((s, fs))
        assert None not in rigs
        # interleave: frame k of every rig before frame k+1 of any
        for k in range(FRAMES):
            for r, (s, fs) in zip(rigs, sets):
                u, v, fixed = fs[k]
                b.submit(r, k, s["dt"], u, v, fixed)
        assert b.step() == FRAMES * len(RIGS)
        res = b.collect()
        assert len(res) == FRAMES * len(RIGS)
        assert all(x["status"] == 0 for x in res)
        for i, r in enumerate(rigs):
            mine = sorted((x for x in res if x["rig"] == r), key=lambda x: x["frame_id"])
            assert [x["frame_id"] for x in mine] == list(range(FRAMES))
            # a rig's frames run in order
            assert all(mine[j]["t_start"] >= mine[j - 1]["t_end"] for j in range(1, FRAMES))
            np.testing.assert_array_equal(np.array([x["estimate"] for x in mine]), ref[i][1])
            st = b.filter(r).state()
            for key in ("states", "weights", "gcount", "wg", "samples"):
                np.testing.assert_array_equal(st[key], ref[i][0][key], err_msg=f"rig {i} {key}")
        s = b.stats()
        assert s["frames_done"] == FRAMES * len(RIGS) and s["queued_frames"] == 0 and s["rigs"] == len(RIGS)
        assert s["pairs_done"] == sum(x["n_pairs"] for x in res) > 0
        assert 0 < s["used_bytes"] <= s["reserved_bytes"]
    finally:
        b.close()


def test_admission_removal_and_readmission():
    import scphd_amd

    p0, _, _ = make(0)
    need = scphd_amd.rig_footprint(p0, 200)
    b = scphd_amd.Batch(hbm_budget_bytes=2 * need + need // 2, n_workers=2)
    try:
        r0 = b.add_rig(p0, need)
        r1 = b.add_rig(make(1)[0], need)
        assert r0 is not None and r1 is not None
        assert b.add_rig(make(2)[0], need) is None
        assert b.stats()["refused"] == 1
        b.remove_rig(r0)
        with pytest.raises(scphd_amd.ScphdError):  # a removed rig takes no frames
            b.submit(r0, 0, 0.5, [1.0], [1.0], True)
        r2 = b.add_rig(make(2)[0], need)
        assert r2 == r0  # the freed slot is reused
        s = b.stats()
        assert s["rigs"] == 2 and s["reserved_bytes"] == 2 * need
        # the re-admitted rig runs
        p2, s2, fs = make(2)
        for k, (u, v, fixed) in enumerate(fs[:2]):
            b.submit(r2, k, s2["dt"], u, v, fixed)
        assert b.step() == 2
        assert [x["status"] for x in b.collect()] == [0, 0]
        with pytest.raises(scphd_amd.ScphdError):
            b.submit(999, 0, 0.5, [1.0], [1.0], True)
    finally:
        b.close()


def test_estimate_matches_parents_read_back():
    import scphd_amd

    p, s, sets = make(1)
    f = scphd_amd.Filter(p)
    for u, v, fixed in sets[:4]:
        f.predict(s["dt"])
        f.update(u, v, fixed)
        f.resample()
    m, c, card = f.estimate()
    st, w = f.parents()
    wn = w / w.sum()
    mu = wn @ st
    d = st - mu
    np.testing.assert_allclose(m, mu, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(c, (d * wn[:, None]).T @ d, rtol=1e-10, atol=1e-18)
    gc, wg, _ = f.maps()
    best = int(np.argmax(w))
    off = np.concatenate([[0], np.cumsum(gc)])
    assert abs(card - wg[off[best]:off[best + 1]].sum()) <= 1e-12 * max(1.0, card)
    f.close()


def test_set_noise_rekeys_the_streams():
    import scphd_amd

    p, s, sets = make(2, frames=2)
    runs = []
    for reseed in (None, 9, 9):
        f = scphd_amd.Filter(p)
        if reseed is not None:
            f.set_noise(reseed)
        for u, v, fixed in sets:
            f.predict(s["dt"])
            f.update(u, v, fixed)
        runs.append(f.state())
        if reseed is not None:
            with pytest.raises(scphd_amd.ScphdError, match="ARG"):
                f.set_noise(1, mode=1)
        f.close()
    np.testing.assert_array_equal(runs[1]["samples"], runs[2]["samples"])
    assert not np.array_equal(runs[0]["states"], runs[1]["states"])


def test_evicted_rig_resumes_exactly_where_it_stopped():
    """Served-mode eviction: snapshot a rig after 3 frames, remove it, admit it
    again (in another slot, next to another rig) and run the remaining frames;
    the result equals the uninterrupted run bit for bit."""
    import scphd_amd

    ref = alone(1)
    p, s, sets = make(1)
    b = scphd_amd.Batch(n_workers=2)
    try:
        r = b.add_rig(p)
        for k in range(3):
            b.submit(r, k, s["dt"], *sets[k])
        b.step()
        snap = b.snapshot(r)
        other = b.add_rig(make(0)[0])  # takes the freed slot
        r2 = b.resume(snap)
        assert r2 is not None and r2 != other
        for k in range(3, FRAMES):
            b.submit(r2, k, s["dt"], *sets[k])
        b.step()
        st = b.filter(r2).state()
        for key in ("states", "weights", "gcount", "wg", "samples", "n_update", "n_resample"):
            np.testing.assert_array_equal(st[key], ref[0][key], err_msg=key)
    finally:
        b.close()


def test_check_reports_a_nan_parent_weight():
    """checkStuff (src/scphdcameracalibration.cu:812-821): a NaN parent
    weight is an error code, counted on the device."""
    import scphd_amd

    p, s, sets = make(0)
    f = scphd_amd.Filter(p)
    u, v, fixed = sets[0]
    f.predict(s["dt"])
    f.update(u, v, fixed)
    assert f.check() == 0
    st = f.state()
    w = st["weights"].copy()
    w[len(w) // 2] = np.nan
    st["weights"] = w
    f.set_state(st)
    assert scphd_amd.ERRORS.get(f.check()) == "SCPHD_ERR_NAN"
    w[len(w) // 2] = 1.0 / len(w)
    st["weights"] = w
    f.set_state(st)
    assert f.check() == 0
    f.close()


def test_full_size_rigs_are_identical_alone_and_in_a_batch():
    """A 4096-parent rig with S = 200 (the bench workload's size) and a
    2048-parent rig with S = 100, served together by one Batch with their
    frames interleaved: per-frame estimates, and the parents and weights at
    the end, are bit-identical to each rig run alone.  The maps (tens of GB
    here) are left out; the estimates depend on every map through the
    parent weights."""
    import scphd_amd

    big = [dict(seed=21, n_particles=4096, n_landmarks=200, lam=50.0, samples=200),
           dict(seed=22, n_particles=2048, n_landmarks=150, lam=40.0, f=600.0, baseline=0.25, samples=100)]
    frames = 4
    made = []
    for i, cfg in enumerate(big):
        p, s = workloads.rig(n_frames=frames, scene_rig=i, **cfg)
        sc = scene.generate(seed=p["seed"], **s)
        made.append((p, s, [scene.frame_measurements(sc, k) for k in range(frames)]))
    ref = []
    for p, s, sets in made:
        f = scphd_amd.Filter(p)
        est = []
        for u, v, fixed in sets:
            f.predict(s["dt"])
            f.update(u, v, fixed)
            assert f.check() == 0
            f.resample()
            est.append(f.estimate()[0])
        ref.append((f.parents(), np.array(est)))
        f.close()
    b = scphd_amd.Batch(n_workers=2)
    try:
        rigs = [b.add_rig(p) for p, _, _ in made]
        assert None not in rigs
        for k in range(frames):
            for r, (_, s, sets) in zip(rigs, made):
                u, v, fixed = sets[k]
                b.submit(r, k, s["dt"], u, v, fixed)
        assert b.step() == frames * len(big)
        res = b.collect()
        assert len(res) == frames * len(big) and all(x["status"] == 0 for x in res)
        for i, r in enumerate(rigs):
            mine = sorted((x for x in res if x["rig"] == r), key=lambda x: x["frame_id"])
            np.testing.assert_array_equal(np.array([x["estimate"] for x in mine]), ref[i][1],
                                          err_msg=f"rig {i} estimates")
            states, weights = b.filter(r).parents()
            np.testing.assert_array_equal(states, ref[i][0][0], err_msg=f"rig {i} states")
            np.testing.assert_array_equal(weights, ref[i][0][1], err_msg=f"rig {i} weights")
    finally:
        b.close()
