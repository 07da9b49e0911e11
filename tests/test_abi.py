"""The C ABI (include/scphd.h): exports, struct layouts, configuration
parsing and argument checking -- everything that runs without a GPU."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

import oracle
import scphd_amd
import workloads

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "scphd.h")
GOLDEN = os.path.join(REPO, "tests", "golden")


def declared():
    src = open(HEADER).read()
    return sorted(set(re.findall(r"^[a-z0-9_ ]*\**\s*(scphd_\w+)\s*\(", src, flags=re.M)))


def test_every_declared_symbol_is_exported():
    names = declared()
    assert len(names) >= 20
    L = ctypes.CDLL(scphd_amd.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(scphd_amd.EXPORTED) == names


def test_params_layout_matches_oracle_mirror():
    assert ctypes.sizeof(scphd_amd.Params) == oracle.lib().orc_params_size() == ctypes.sizeof(oracle.OracleParams)
    for (n1, t1), (n2, t2) in zip(scphd_amd.Params._fields_, oracle.OracleParams._fields_):
        assert n1 == n2
        assert getattr(scphd_amd.Params, n1).offset == getattr(oracle.OracleParams, n2).offset


def test_defaults_are_reference_config():
    p = scphd_amd.Params()
    assert scphd_amd.lib().scphd_params_init(ctypes.byref(p)) == 0
    assert p.n_particles == 100 and p.particles_per_feature == 100
    assert p.min_separation == 6.0 and p.min_weight == 1e-4 and p.ps == 0.9
    assert p.fx == -895.6561 and p.v0 == 300.0 and p.lambda_ == 10.0001
    assert p.birth_w0 == 1e-6 and p.birth_var_vd0 == 0.05


def test_reference_config_file_parses_with_mean_std_mapping():
    d = scphd_amd.load_params(os.path.join(GOLDEN, "calibration.cfg"))
    assert d["n_particles"] == 100 and d["max_steps"] == 100 and d["verbosity"] == 0
    assert d["slam"] == 0  # absent in the file: false
    s3 = math.sqrt(3.0)
    # prior.cartesian.z: mean -3, std 3 -> uniform [-3 - 3 sqrt 3, -3 + 3 sqrt 3]
    assert d["prior_a"][2] == pytest.approx(-3 - 3 * s3) and d["prior_b"][2] == pytest.approx(-3 + 3 * s3)
    assert d["prior_a"][1] == -50.0 and d["prior_b"][1] == -50.0
    assert d["prior_a"][7] == pytest.approx(0.1 - 0.1 * s3)
    assert d["cam_accel_std"] == [0.0, 0.0, 0.15, 0.0, 0.001, 0.0]
    assert d["feat_accel_std"] == [0.05, 0.05, 0.1]
    assert d["std_u"] == 5.0 and d["pd"] == 0.95 and d["lambda"] == 10.0001
    assert d["birth_w0"] == 1e-6 and d["birth_d0"] == 8.0 and d["birth_var_vd0"] == 0.05
    assert d["min_neff"] == 0.75 and d["min_range"] == 0.5 and d["max_velocity"] == 5.0


def test_generated_config_round_trips(tmp_path):
    p, _ = workloads.workload("c1")
    path = tmp_path / "c1.cfg"
    path.write_text(workloads.to_cfg(p))
    d = scphd_amd.load_params(str(path))
    for k, v in p.items():
        if k == "storage":
            continue
        if isinstance(v, list):
            np.testing.assert_allclose(d[k], v, rtol=1e-15, err_msg=k)
        else:
            assert d[k] == pytest.approx(v, rel=1e-15), k


def test_config_errors_name_the_problem(tmp_path):
    good = open(os.path.join(GOLDEN, "calibration.cfg")).read()
    bad = tmp_path / "missing.cfg"
    bad.write_text(good.replace("min_separation = 6.0", ""))
    with pytest.raises(scphd_amd.ScphdError, match="min_separation"):
        scphd_amd.load_params(str(bad))
    broken = tmp_path / "broken.cfg"
    broken.write_text("n_particles = = 3\n")
    with pytest.raises(scphd_amd.ScphdError, match="CONFIG"):
        scphd_amd.load_params(str(broken))
    with pytest.raises(scphd_amd.ScphdError, match="cannot open"):
        scphd_amd.load_params(str(tmp_path / "nope.cfg"))


def test_config_grammar_variants(tmp_path):
    """Comments of three kinds, ':' and '=', ',' and ';' terminators,
    hex and L-suffixed integers, booleans in any case, @include."""
    inc = tmp_path / "birth.cfg"
    inc.write_text('birth : { w0 = 1e-6; d0 = 8.0, vu0 = 0; vv0 = 0\n vd0 = 0 var_d0 = 2.0 var_vu0 = 6 '
                   'var_vv0 = 6 var_vd0 = 0.05 }\n')
    body = open(os.path.join(GOLDEN, "calibration.cfg")).read()
    body = re.sub(r"birth : \{.*?\}\s*$", '@include "birth.cfg"\n', body, flags=re.S)
    body = body.replace("n_particles = 100", "n_particles = 0x40L /* hex */ // comment")
    body = body.replace("save_all_maps = false", "save_all_maps = TRUE")
    main = tmp_path / "main.cfg"
    main.write_text(body)
    d = scphd_amd.load_params(str(main))
    assert d["n_particles"] == 64 and d["save_all_maps"] == 1 and d["birth_var_vu0"] == 6.0


def test_storage_key_names_the_three_variants(tmp_path):
    """storage: "fp32" / 0 (FP32 clouds, FP64 arithmetic), "fp64" / 1,
    "fp32-arith" / 2 (FP32 clouds and per-sample arithmetic)."""
    body = open(os.path.join(GOLDEN, "calibration.cfg")).read()
    for val, want in [('"fp32"', 0), ('"fp64"', 1), ('"fp32-arith"', 2), ("0", 0), ("1", 1), ("2", 2)]:
        path = tmp_path / "s.cfg"
        path.write_text(body + f"\nstorage = {val};\n")
        assert scphd_amd.load_params(str(path))["storage"] == want, val


def test_create_rejects_bad_arguments():
    p, _ = workloads.workload("c0")
    for k, v in [("n_particles", 0), ("particles_per_feature", 0), ("particles_per_feature", 1000), ("storage", 7)]:
        q = dict(p)
        q[k] = v
        with pytest.raises(scphd_amd.ScphdError, match="ARG"):
            scphd_amd.Filter(q)


# ------------------------------------------------------------ batch / served
def test_batch_options_and_argument_checks():
    o = scphd_amd.BatchOpts()
    assert scphd_amd.lib().scphd_batch_opts_init(ctypes.byref(o)) == 0
    assert (o.hbm_budget_bytes, o.max_rigs, o.n_workers) == (0, 1024, 8)
    bad = scphd_amd.BatchOpts(-1, 0, 0)
    h = ctypes.c_void_p()
    assert scphd_amd.lib().scphd_batch_create(ctypes.byref(bad), ctypes.byref(h)) == -1
    bad = scphd_amd.BatchOpts(1 << 30, 0, 65)
    assert scphd_amd.lib().scphd_batch_create(ctypes.byref(bad), ctypes.byref(h)) == -1


def test_rig_footprint_scales_with_the_state():
    p, _ = workloads.workload("c0")
    base = scphd_amd.rig_footprint(p, 100)
    assert scphd_amd.rig_footprint(dict(p, n_particles=2 * p["n_particles"]), 100) > 1.9 * base - (128 << 20)
    assert scphd_amd.rig_footprint(p, 200) > base
    assert scphd_amd.rig_footprint(dict(p, particles_per_feature=2 * p["particles_per_feature"]), 100) > base
    assert scphd_amd.rig_footprint(dict(p, storage=1), 100) > base
    # both sample generations at FP32: 2 * N_p * G * S * 28 B is a lower bound
    assert base >= 2 * p["n_particles"] * 100 * p["particles_per_feature"] * 28
    assert scphd_amd.rig_footprint(dict(p, n_particles=0), 100) == 0


def test_batch_admission_is_refused_beyond_the_budget():
    """Runs without a GPU: a reservation above the budget is refused before any
    device work, and the refusal is counted."""
    p, _ = workloads.workload("c0")
    need = scphd_amd.rig_footprint(p, 100)
    b = scphd_amd.Batch(hbm_budget_bytes=need - 1, max_rigs=4, n_workers=2)
    try:
        assert b.add_rig(p, reserve_bytes=need) is None
        s = b.stats()
        assert s["refused"] == 1 and s["rigs"] == 0 and s["reserved_bytes"] == 0 and s["budget_bytes"] == need - 1
        with pytest.raises(scphd_amd.ScphdError, match="ARG"):
            b.submit(0, 0, 0.5, [1.0], [2.0], True)
        assert b.step() == 0
        assert b.collect() == []
    finally:
        b.close()


# ------------------------------------------------------------------ MAT I/O
def _cells(frames):
    c = np.empty((1, len(frames)), dtype=object)
    for k, (u, v) in enumerate(frames):
        c[0, k] = np.vstack([u, v]) if len(u) else np.zeros((2, 0))
    return c


@pytest.mark.parametrize("compress", [False, True])
def test_mat_measurement_reader_matches_scipy(tmp_path, compress):
    import scipy.io

    rng = np.random.default_rng(5)
    z1 = [(rng.uniform(0, 640, n), rng.uniform(0, 480, n)) for n in (3, 0, 7, 1)]
    z2 = [(rng.uniform(0, 640, n), rng.uniform(0, 480, n)) for n in (2, 5, 0, 4)]
    path = str(tmp_path / "seq.mat")
    scipy.io.savemat(path, {"Z1": _cells(z1), "Z2": _cells(z2), "other": np.eye(3)}, do_compression=compress)
    for name, ref in (("Z1", z1), ("Z2", z2)):
        got = scphd_amd.read_measurements(path, name)
        assert len(got) == len(ref)
        for (gu, gv), (ru, rv) in zip(got, ref):
            np.testing.assert_array_equal(gu, ru)
            np.testing.assert_array_equal(gv, rv)
    with pytest.raises(scphd_amd.ScphdError, match="CONFIG"):
        scphd_amd.read_measurements(path, "Z3")
    with pytest.raises(scphd_amd.ScphdError, match="CONFIG"):
        scphd_amd.read_measurements(path, "other")  # not a cell array


REF_DATA = "/root/reference/data"


@pytest.mark.skipif(not os.path.isdir(REF_DATA), reason="the reference's data files are not present")
def test_mat_reader_reads_the_reference_sequences():
    """The reference's own input sequences (data/*.mat) read as scipy reads them."""
    import scipy.io

    n = 0
    for fn in sorted(os.listdir(REF_DATA)):
        path = os.path.join(REF_DATA, fn)
        try:
            ref = scipy.io.loadmat(path)
        except (NotImplementedError, ValueError):
            continue  # not MAT level 5
        for name in ("Z1", "Z2"):
            if name not in ref:
                continue
            got = scphd_amd.read_measurements(path, name)
            cells = ref[name].ravel(order="F")
            assert len(got) == len(cells)
            for (gu, gv), c in zip(got, cells):
                c = np.asarray(c, dtype=np.float64).reshape(2, -1) if c.size else np.zeros((2, 0))
                np.testing.assert_array_equal(gu, c[0])
                np.testing.assert_array_equal(gv, c[1])
            n += 1
    assert n > 0


# --------------------------------------------------------------- evaluation
def test_ospa_matches_its_definition_on_known_cases():
    import evaluate

    X = np.array([[0.0, 0, 0], [1, 0, 0]])
    assert evaluate.ospa(X, X, c=2, p=1) == 0
    assert evaluate.ospa([], [], c=2) == 0
    assert evaluate.ospa(X, np.zeros((0, 3)), c=2, p=2) == 2  # cardinality error only
    # one missing point: ((0 + c^p) / 2)^(1/p)
    assert abs(evaluate.ospa(X[:1], X, c=3, p=2) - math.sqrt(9 / 2)) < 1e-12
    # assignment is optimal, not greedy: swapped order costs nothing
    assert evaluate.ospa(X, X[::-1], c=5) == 0
    # cut-off: a far point contributes c
    Y = np.array([[0.0, 0, 0], [100, 0, 0]])
    assert abs(evaluate.ospa(X, Y, c=2, p=1) - (0 + 2) / 2) < 1e-12
    # symmetric and brute-force equal on random sets
    rng = np.random.default_rng(3)
    from itertools import permutations
    for _ in range(20):
        A, B = rng.normal(size=(3, 2)), rng.normal(size=(4, 2))
        c, p = 1.5, 2.0
        best = min(sum(min(c, np.linalg.norm(A[i] - B[j])) ** p for i, j in enumerate(perm))
                   for perm in permutations(range(4), 3))
        ref = ((best + c ** p * 1) / 4) ** (1 / p)
        assert abs(evaluate.ospa(A, B, c, p) - ref) < 1e-12
        assert abs(evaluate.ospa(B, A, c, p) - ref) < 1e-12


def test_calibration_error_wraps_angles():
    import evaluate

    est = np.zeros(12)
    est[[0, 1, 2]] = [0.21, 0.0, -0.01]
    est[[6, 7, 8]] = [math.pi - 0.01, 0.0, -0.02]
    err = evaluate.calibration_error(est, [0.2, 0.0, 0.0, -math.pi + 0.01, 0.0, 0.0])
    np.testing.assert_allclose(err, [0.01, 0.0, -0.01, -0.02, 0.0, -0.02], atol=1e-12)


def test_map_points_keep_heavy_gaussians():
    import evaluate

    S = 4
    gc = np.array([1, 2])
    wg = np.array([0.9, 0.4, 0.7])
    smp = np.zeros((7, 3 * S))
    for g in range(3):
        smp[:3, g * S:(g + 1) * S] = g + 1
    np.testing.assert_array_equal(evaluate.map_points(gc, wg, smp, 1, S), [[3.0, 3.0, 3.0]])
    np.testing.assert_array_equal(evaluate.map_points(gc, wg, smp, 0, S), [[1.0, 1.0, 1.0]])


def test_seed_is_parsed_exactly_to_64_bits(tmp_path):
    p, _ = workloads.workload("c0")
    for sd in (0, 2 ** 53 + 1, 2 ** 64 - 59):
        cfg = tmp_path / f"s{sd}.cfg"
        cfg.write_text(workloads.to_cfg(dict(p, seed=sd)))
        assert scphd_amd.load_params(str(cfg))["seed"] == sd
    cfg = tmp_path / "hex.cfg"
    cfg.write_text(workloads.to_cfg(p).replace(f"seed = {p['seed']}", "seed = 0xFFFFFFFFFFFFFFC5L"))
    assert scphd_amd.load_params(str(cfg))["seed"] == 2 ** 64 - 59


def test_ensemble_packing_respects_budget_and_plane_limit():
    """serve.pack_ensembles (host logic only): every rig in exactly one group,
    in order; each group's footprint within the budget and its sample planes
    within the 32-bit offset limit (a single rig may exceed either alone);
    rigs from the alone threshold up are set apart."""
    import serve
    import scphd_amd

    rigs = [dict(t=0.0, p=p, s=s, frames=[]) for p, s in workloads.c2_rigs(master=1, n_rigs=64, n_frames=1)]
    budget, growth = 120.0, 2.0
    for alone_parents in (None, 2048):
        groups, alone = serve.pack_ensembles(rigs, budget, growth, alone_parents)
        flat = [r for g in groups for r in g]
        assert [id(r) for r in flat] == [id(r) for r in rigs if r not in alone]
        if alone_parents is None:
            assert alone == []
        else:
            assert alone and all(r["p"]["n_particles"] >= alone_parents for r in alone)
            assert all(r["p"]["n_particles"] < alone_parents for r in flat)
        for g in groups:
            gps = [int(growth * serve.expected_gaussians_per_parent(r["s"])) for r in g]
            foot = sum(scphd_amd.rig_footprint(r["p"], gp) for r, gp in zip(g, gps)) / 1e9
            blocks = sum(r["p"]["n_particles"] * gp for r, gp in zip(g, gps))
            if len(g) > 1:
                assert foot <= budget
                assert blocks * 100 * 4 * 1.25 <= 4 << 30


def test_cohort_admission_is_first_fit_in_arrival_order():
    """serve.cohort_admit (host logic only): arrived rigs in arrival order,
    each taken when it fits what is left of the budget, smaller ones passing
    a rig that does not fit; the scan ends at the first rig not yet
    arrived; with nothing admitted the first arrived rig is taken whatever its size."""
    import serve

    arrive = [0.0, 0.1, 0.2, 0.3, 5.0]
    need = [60.0, 50.0, 20.0, 10.0, 1.0]
    pending = [0, 1, 2, 3, 4]
    take, skip = serve.cohort_admit(pending, arrive, need, 1.0, 100.0, False)
    assert take == [0, 2, 3] and skip == [1]
    take, skip = serve.cohort_admit(pending, arrive, need, 0.15, 100.0, False)
    assert take == [0] and skip == [1]
    take, skip = serve.cohort_admit([1, 2], arrive, need, 1.0, 10.0, True)
    assert take == [1] and skip == [2]  # empty: the head, although it exceeds the share
    take, skip = serve.cohort_admit([4], arrive, need, 1.0, 10.0, True)
    assert take == [] and skip == []  # not arrived yet
