"""Lock-step parity of the gfx950 filter step against the FP64 oracle.

Every frame starts both sides from the oracle's state (the GPU gets it
through scphd_set_state), runs predict -> update -> resample with the same
Philox noise streams, and compares:
  * bit-exact: survivor counts, merged counts, kept non-detection counts,
    map sizes per parent, the resampling decision and ancestor indices;
  * within tolerance: parent states and weights, normalisers eta,
    parent log-likelihoods, Gaussian weights, sample clouds.
Tolerances: FP64 storage: 1e-9 relative on weights / normalisers, 1e-7 on
samples (sample positions go through 1/d); FP32 storage: samples within 4
float ulps (2.5e-7 relative) plus 1e-6 absolute.  The refit of every cloud
(disparity-space mean, covariance, pd_g) is compared too: means within
1e-12 relative, covariances within 1e-9 of sqrt(P_aa P_bb) (the GPU sums
the moments in one pass, the oracle in two), pd_g within 1e-13.
"""
import numpy as np
import pytest

import oracle
import scene
import workloads

pytestmark = pytest.mark.gpu


def small(storage, n_particles=48, samples=100, seed=7, **kw):
    p, s = workloads.workload("c0", n_particles=n_particles, particles_per_feature=samples, **kw)
    p["storage"] = storage
    p["seed"] = seed
    return p, s


def _close(a, b, rtol, atol, what):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    assert a.shape == b.shape, f"{what}: shape {a.shape} vs {b.shape}"
    if a.size == 0:
        return 0.0
    err = np.abs(a - b)
    lim = atol + rtol * np.maximum(np.abs(a), np.abs(b))
    bad = err > lim
    assert not bad.any(), (f"{what}: {bad.sum()} of {a.size} differ; worst {err.max():.3e} at "
                           f"{int(np.argmax(err))} ({a.flat[np.argmax(err)]!r} vs {b.flat[np.argmax(err)]!r})")
    return float((err / np.maximum(np.maximum(np.abs(a), np.abs(b)), 1e-300)).max())


def compare_fit(fg, fo, what):
    """Per-cloud refit planes (mean 6, packed covariance 21, pd_g)."""
    assert fg.shape == fo.shape, f"{what}: fit shape {fg.shape} vs {fo.shape}"
    if fo.shape[1] == 0:
        return
    _close(fg[:6], fo[:6], 1e-12, 1e-12, f"{what} fit means")
    diag = {}
    e = 6
    for a in range(6):
        for b in range(a + 1):
            if a == b:
                diag[a] = np.abs(fo[e])
            e += 1
    e = 6
    for a in range(6):
        for b in range(a + 1):
            scale = np.sqrt(diag[a] * diag[b])
            err = np.abs(fg[e] - fo[e])
            bad = err > 1e-9 * scale + 1e-12
            assert not bad.any(), (f"{what} fit covariance ({a},{b}): {bad.sum()} of {err.size} differ, worst "
                                   f"{err.max():.3e}")
            e += 1
    _close(fg[27], fo[27], 1e-13, 1e-15, f"{what} fit pd_g")


def run_lockstep(p, s, frames, warm=0, limits=None):
    import scphd_amd

    sc = scene.generate(seed=p["seed"], **s)
    orc = oracle.Oracle(p, threads=8)
    gpu = scphd_amd.Filter(p)
    if limits:
        gpu.set_debug_limits(**limits)
    # same prior draw
    so, wo = orc.parents()
    sg, wg = gpu.parents()
    _close(sg, so, 1e-14, 1e-15, "prior")
    for k in range(warm):
        u, v, fixed = scene.frame_measurements(sc, k)
        orc.predict(s["dt"])
        orc.update(u, v, fixed)
        orc.resample()
    f32 = p["storage"] == 0
    rs = 1e-6 if f32 else 1e-7
    report = []
    for k in range(warm, warm + frames):
        st = orc.state()
        gpu.set_state(st)
        u, v, fixed = scene.frame_measurements(sc, k)
        orc.predict(s["dt"])
        gpu.predict(s["dt"])
        _close(gpu.parents()[0], orc.parents()[0], 1e-13, 1e-13, f"frame {k} predicted parents")
        orc.update(u, v, fixed)
        gpu.update(u, v, fixed)
        do, dg = orc.update_debug(), gpu.update_debug()
        for key in ("nsurvivors", "nmerged", "nnondetect"):
            np.testing.assert_array_equal(dg[key], do[key], err_msg=f"frame {k} {key}")
        ng = gpu.update_stats()["n_gaussians_in"]
        compare_fit(gpu.fit_debug(), orc.fit_debug(ng), f"frame {k}")
        _close(dg["eta"], do["eta"], 1e-11, 1e-11, f"frame {k} eta")
        _close(dg["loglik"], do["loglik"], 1e-11, 1e-9, f"frame {k} loglik")
        gco, wgo, smo = orc.maps()
        gcg, wgg, smg = gpu.maps()
        np.testing.assert_array_equal(gcg, gco, err_msg=f"frame {k} map sizes")
        _close(wgg, wgo, 1e-9, 1e-15, f"frame {k} gaussian weights")
        e = _close(smg, smo, rs, 1e-6 if f32 else 1e-9, f"frame {k} samples")
        _close(gpu.parents()[1], orc.parents()[1], 1e-9, 1e-15, f"frame {k} parent weights")
        ro, rg = orc.resample(), gpu.resample()
        assert ro == rg, f"frame {k} resample decision"
        if ro:
            np.testing.assert_array_equal(gpu.ancestors(), orc.ancestors()[0], err_msg=f"frame {k} ancestors")
        report.append((k, int(gco.sum()), e, int(do["nsurvivors"].max())))
    gpu.close()
    return report


@pytest.mark.parametrize("storage", [1, 0], ids=["fp64", "fp32"])
def test_lockstep_small(storage):
    """c0 geometry with 48 parents: births-only first frame, then frames
    alternating fixed / moving camera with detection, merge, non-detection
    and resampling terms."""
    p, s = small(storage)
    rep = run_lockstep(p, s, frames=6)
    assert rep[-1][1] > 0


def test_lockstep_c0_full_size():
    """The whole C0 workload, 256 parents, S = 100, 30 landmarks, lambda 10,
    50 frames with the cameras alternating: every frame's update from the
    oracle's state, compared after every frame (maps of ~150 Gaussians per
    parent, ~40,000 in all)."""
    p, s = workloads.workload("c0")
    assert p["n_particles"] == 256 and p["particles_per_feature"] == 100 and s["n_frames"] == 50
    rep = run_lockstep(p, s, frames=50)
    assert min(r[1] for r in rep[5:]) > 20000


def test_lockstep_after_warmup_fp32():
    """Lock-step frames from a state grown by 8 oracle frames (maps of
    ~100 Gaussians per parent)."""
    p, s = small(0, n_particles=32)
    run_lockstep(p, s, frames=3, warm=8)


def test_lockstep_resampling_path():
    """min_neff = 1.01 forces systematic resampling every frame, so the
    ancestor indices and the re-indexed maps are compared every frame."""
    p, s = small(1, n_particles=40, min_neff=1.01)
    run_lockstep(p, s, frames=4)


def test_lockstep_samples_per_gaussian_200():
    """S = 200 (four samples per lane in the reproject/fit kernel)."""
    p, s = small(0, n_particles=16, samples=200)
    run_lockstep(p, s, frames=4)


def test_lockstep_birth_term_dominates():
    """Birth weight 0.05: log(M w0 + kappa) exceeds every detection term's
    log weight (a 2-D density is at most 1 / (2 pi) with S >= R = I), so the
    normaliser's maximum is the birth/clutter term for every measurement and
    the max pass visits no Gaussian."""
    p, s = small(1, n_particles=24, birth_w0=0.05)
    rep = run_lockstep(p, s, frames=4)
    assert rep[-1][1] > 0


def test_lockstep_pd_one_refits_all_zero_clouds():
    """Pd = 1: a kept non-detection cloud weighs 0 wherever it is in view, so
    in the next update every w_i pd_i is 0 while pd_g > 0 -- the cloud is
    refitted with unit weights (k_fit_fix), as the reference normalises an
    all-zero cloud to weight 1 per sample."""
    p, s = small(1, n_particles=24, pd=1.0)
    rep = run_lockstep(p, s, frames=5)
    assert rep[-1][1] > 0


def test_empty_measurement_set():
    """M = 0: no births, no detection terms; only non-detection terms and
    the cardinality term of the parent likelihood remain."""
    import scphd_amd

    p, s = small(1, n_particles=16)
    sc = scene.generate(seed=p["seed"], **s)
    orc = oracle.Oracle(p)
    gpu = scphd_amd.Filter(p)
    for k in range(3):
        u, v, fixed = scene.frame_measurements(sc, k)
        orc.predict(0.5)
        orc.update(u, v, fixed)
        orc.resample()
    gpu.set_state(orc.state())
    orc.predict(0.5)
    gpu.predict(0.5)
    orc.update([], [], False)
    gpu.update([], [], False)
    gco, wgo, _ = orc.maps()
    gcg, wgg, _ = gpu.maps()
    np.testing.assert_array_equal(gcg, gco)
    _close(wgg, wgo, 1e-9, 1e-15, "weights")
    _close(gpu.parents()[1], orc.parents()[1], 1e-9, 1e-15, "parent weights")
    gpu.close()


def test_device_noise_equals_the_oracle_bit_for_bit():
    """The Philox4x32-10 + Box-Muller normals drawn on the GPU equal the
    oracle's to the last bit, for every stream (single-precision polynomial
    log / sin / cos, correctly rounded sqrt, explicit fma)."""
    import scphd_amd

    for n, seed, stream, step, parent, gauss in [(4096, 1, 2, 0, 0, 0), (4096, 2**64 - 59, 3, 17, 4095, 1999),
                                                 (4096, 7, 1, 123456, 9, 0), (4096, 5, 2, 99, 17, 311),
                                                 (4096, 9, 3, 3, 0, 7),
                                                 # 393,216 Box-Muller pairs: the square root's correction step
                                                 # (hardware root moved by one ulp) against the CPU's sqrtf
                                                 (1 << 17, 3, 1, 5, 2, 3)]:
        dev = scphd_amd.device_normals(seed, stream, step, parent, gauss, n)
        if stream in (2, 3):  # the samples' normals, in groups sharing Philox blocks
            ref = np.stack([np.pad(oracle.stream_normals(seed, stream, step, parent, gauss, s), (0, 3 if stream == 2 else 0))
                            for s in range(n)])
        else:
            ref = np.stack([oracle.normals(seed, stream, step, parent, gauss, e, 6) for e in range(n)])
        np.testing.assert_array_equal(dev, ref)


# ---- general code paths, forced at small sizes through the verification
# ---- limits (scphd_debug_set_limits); each must give the oracle's results

def _grown(n_particles=24, storage=0):
    """c0 geometry, 24 parents, from a state grown by 6 oracle frames: maps
    of ~150 Gaussians and ~100-300 survivors per parent."""
    return small(storage, n_particles=n_particles)


def test_lockstep_merge_in_global_scratch():
    """Survivor lists longer than 16 merge in global scratch
    (k_merge_groups<true>), the path of parents beyond the LDS workspace."""
    p, s = _grown()
    rep = run_lockstep(p, s, frames=3, warm=6, limits=dict(merge_lds_cap=16))
    assert max(r[3] for r in rep) > 16


def test_lockstep_merge_round_loop_lds():
    """Every parent merges in the round loop (k_merge_rounds<false>), the
    path of candidate or neighbour-list overflow."""
    p, s = _grown(storage=1)
    run_lockstep(p, s, frames=3, warm=6, limits=dict(merge_force_rounds=True))


def test_lockstep_merge_round_loop_global_scratch():
    """Round loop with its workspace in global scratch (k_merge_rounds<true>)."""
    p, s = _grown()
    run_lockstep(p, s, frames=3, warm=6, limits=dict(merge_lds_cap=16, merge_force_rounds=True))


def test_lockstep_normaliser_whole_map_max_pass():
    """No list of the Gaussians above the birth/clutter term: the normaliser
    max pass sweeps the whole map (the path of more than 2048 such
    Gaussians)."""
    p, s = _grown()
    run_lockstep(p, s, frames=3, warm=6, limits=dict(eta_heavy_cap=0))


def test_lockstep_whole_map_candidate_sweeps():
    """No candidate lists: the normaliser's kept-term count and the survivor
    pass sweep the whole map for every measurement (the path of more than
    eight candidate terms per measurement)."""
    p, s = _grown(storage=1)
    run_lockstep(p, s, frames=3, warm=6, limits=dict(eta_ncand_cap=0))


def test_lockstep_merge_candidates_from_all_pairs():
    """Every survivor treated as ill-conditioned: merge candidates come from
    all pairs (in LDS and in global scratch) instead of windows in u, the
    path of survivors whose (u, v) covariance is near singular."""
    p, s = _grown(storage=1)
    run_lockstep(p, s, frames=2, warm=6, limits=dict(merge_all_pairs=True))
    run_lockstep(p, s, frames=2, warm=6, limits=dict(merge_all_pairs=True, merge_lds_cap=16))
