"""Pins the FP64 oracle (oracle/scphd_oracle.cpp) without the reference,
which cannot be built here (CUDA + thrust + libconfig++ + matio): published
known answers, closed forms, invariants, and cross-checks against
independent numpy/scipy evaluations, chosen so that a dropped term, a wrong
sign or index, or a transposed operand fails one of them."""
import math

import numpy as np
import pytest
import scipy.special
import scipy.stats

import oracle
import workloads


def params(**kw):
    p, _ = workloads.workload("c0", **kw)
    return p


def rot_np(th, ph, ps):
    """Rx(th) Ry(ph) Rz(ps) as an explicit product of elementary rotations."""
    c, s = math.cos, math.sin
    Rx = np.array([[1, 0, 0], [0, c(th), -s(th)], [0, s(th), c(th)]])
    Ry = np.array([[c(ph), 0, s(ph)], [0, 1, 0], [-s(ph), 0, c(ph)]])
    Rz = np.array([[c(ps), -s(ps), 0], [s(ps), c(ps), 0], [0, 0, 1]])
    return Rx @ Ry @ Rz


def spd(rng, n=6, scale=1.0, cond=100.0):
    Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    ev = scale * np.exp(rng.uniform(0, math.log(cond), n))
    return (Q * ev) @ Q.T


# ------------------------------------------------------------------ noise
def test_philox_known_answers():
    """Random123 kat_vectors for philox4x32-10."""
    assert list(oracle.philox([0, 0, 0, 0], [0, 0])) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert list(oracle.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2)) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6,
                                                                         0x6D5451FD]
    assert list(oracle.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0])) == [
        0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_normals_are_standard_normal_and_keyed():
    z = np.concatenate([oracle.normals(11, 2, 3, p, k, s, 4) for p in range(8) for k in range(25) for s in range(25)])
    assert abs(z.mean()) < 0.03 and abs(z.std() - 1) < 0.03
    assert scipy.stats.kstest(z, "norm").pvalue > 1e-3
    # distinct counters (stream, step, parent, gaussian, element) give distinct draws
    a = oracle.normals(11, 2, 3, 1, 1, 1, 3)
    for args in [(12, 2, 3, 1, 1, 1), (11, 3, 3, 1, 1, 1), (11, 2, 4, 1, 1, 1), (11, 2, 3, 2, 1, 1),
                 (11, 2, 3, 1, 2, 1), (11, 2, 3, 1, 1, 2)]:
        assert not np.allclose(oracle.normals(*args, 3), a)
    # the first normals of a block do not depend on how many are asked for
    np.testing.assert_array_equal(oracle.normals(5, 1, 0, 0, 0, 0, 6)[:3], oracle.normals(5, 1, 0, 0, 0, 0, 3))


def _box_muller_exact(x0, x1):
    """Box-Muller with libm on the same inputs: a = (x0 + 1) 2^-32 to 24
    significant bits (truncated), c = (x1 >> 8) 2^-24."""
    n = x0 + 1
    if n >= 1 << 32:
        a = 1.0
    else:
        sh = max(n.bit_length() - 24, 0)
        a = ((n >> sh) << sh) / 2.0**32
    c = (x1 >> 8) / 2.0**24
    r = math.sqrt(-2.0 * math.log(a))
    return r * math.cos(2 * math.pi * c), r * math.sin(2 * math.pi * c)


def test_box_muller_polynomials_match_libm():
    rng = np.random.default_rng(7)
    words = [(0, 0), (0xFFFFFFFF, 0xFFFFFFFF), (0xFFFFFFFE, 0x00FFFFFF), (1, 0x20000000), (0x7FFFFFFF, 0x40000000),
             (0xB504F333, 0x60000000), (0xB504F334, 0x80000000), (12345, 0xA0000000), (99, 0xC0000000),
             (0x80000000, 0xE0000000), (0x0000FFFF, 0x1FFFFFFF), (0x00FFFFFF, 0x3FFFFF00)]
    words += [(int(a), int(b)) for a, b in rng.integers(0, 1 << 32, size=(4000, 2), dtype=np.uint64)]
    for x0, x1 in words:
        z = oracle.box_muller(x0, x1)
        e = _box_muller_exact(x0, x1)
        assert abs(z[0] - e[0]) < 4e-7 * (1 + abs(e[0])) and abs(z[1] - e[1]) < 4e-7 * (1 + abs(e[1])), (x0, x1, z, e)
    # a = 1 gives r = 0; the smallest a gives the largest radius
    assert np.all(oracle.box_muller(0xFFFFFFFF, 0x12345678) == 0)
    assert abs(np.hypot(*oracle.box_muller(0, 0)) - math.sqrt(64 * math.log(2))) < 1e-6


def _sqrt_corrected(x, s):
    """The device's root correction (sqrt_rn_normal): from a hardware root s
    within one ulp, the neighbour below when x - dn s <= 0, the one above
    when x - up s > 0.  Each fma residual is formed exactly in float64 (the
    product of two floats is exact there) and only its sign is used."""
    x = np.float32(x)
    dn = np.frombuffer(np.uint32(np.float32(s).view(np.uint32) - 1).tobytes(), np.float32)[0]
    up = np.frombuffer(np.uint32(np.float32(s).view(np.uint32) + 1).tobytes(), np.float32)[0]
    r = dn if float(x) - float(dn) * float(s) <= 0.0 else np.float32(s)
    return up if float(x) - float(up) * float(s) > 0.0 else r


def test_square_root_correction_recovers_the_correctly_rounded_root():
    """Whichever of the three roots within one ulp the hardware returns,
    the correction gives numpy's correctly rounded float32 sqrt, over the
    range of -2 log a (2^-23 .. 45), with exact squares and the values next
    to them; +-0 pass through unchanged."""
    rng = np.random.default_rng(3)
    xs = np.exp(rng.uniform(np.log(2.0**-23), np.log(45.0), 20000)).astype(np.float32)
    sq = (np.arange(1, 6000, dtype=np.float32) / np.float32(128.0)) ** 2  # exact squares below 45
    xs = np.concatenate([xs, sq, np.nextafter(sq, np.float32(0)), np.nextafter(sq, np.float32(64)),
                         np.float32([2.0**-23, 45.0, 1.0, 2.0, 4.0])])
    for x in xs:
        t = np.sqrt(x)
        for s in (np.nextafter(t, np.float32(0)), t, np.nextafter(t, np.float32(64))):
            assert _sqrt_corrected(x, s) == t, (x, s, t)
    for z in (np.float32(0.0), np.float32(-0.0)):
        r = _sqrt_corrected(z, z)
        assert r == 0 and np.signbit(r) == np.signbit(z)


def test_normals_are_box_muller_of_philox_words():
    w = oracle.philox([(7 << 2) | 1, 3, 2, (5 << 4) | 2], [11, 0])
    z = oracle.normals(11, 2, 5, 2, 3, 7, 8)
    np.testing.assert_array_equal(z[4:6], oracle.box_muller(int(w[0]), int(w[1])))
    np.testing.assert_array_equal(z[6:8], oracle.box_muller(int(w[2]), int(w[3])))


# ----------------------------------------------------------------- scalars
def test_safe_log_and_wrap_angle():
    assert oracle.safe_log(0.0) == -np.finfo(np.float32).max
    assert oracle.safe_log(-1.0) == -np.finfo(np.float32).max
    assert oracle.safe_log(math.e) == pytest.approx(1.0, abs=1e-15)
    assert oracle.wrap_angle(0.0) == 0.0
    assert oracle.wrap_angle(math.pi) == math.pi  # fmod keeps pi, which is not > pi
    assert oracle.wrap_angle(-math.pi) == -math.pi
    assert oracle.wrap_angle(1.5 * math.pi) == pytest.approx(-0.5 * math.pi, abs=1e-15)
    assert oracle.wrap_angle(-1.5 * math.pi) == pytest.approx(0.5 * math.pi, abs=1e-15)
    for k in (-3, -1, 2, 7):
        assert oracle.wrap_angle(0.3 + 2 * math.pi * k) == pytest.approx(0.3, abs=1e-13)


def test_lse2_is_logaddexp():
    for a, b in [(0.0, 0.0), (-3.0, 2.0), (10.0, -700.0), (-1e3, -1e3 - 1e-9)]:
        assert oracle.lse2(a, b) == pytest.approx(np.logaddexp(a, b), rel=1e-15, abs=1e-15)


# --------------------------------------------------------------- geometry
def test_rotation_is_xyz_product_and_orthonormal():
    rng = np.random.default_rng(1)
    for _ in range(20):
        e = np.zeros(12)
        e[6:9] = rng.uniform(-math.pi, math.pi, 3)
        R = oracle.rotation(e)
        np.testing.assert_allclose(R, rot_np(*e[6:9]), atol=1e-15)
        np.testing.assert_allclose(R.T @ R, np.eye(3), atol=1e-14)
        assert np.linalg.det(R) == pytest.approx(1.0, abs=1e-14)


def test_projection_closed_form_and_inverse():
    P = params()
    f = -P["fx"]
    e = np.zeros(12)
    q, pd = oracle.world_to_disparity(P, e, [0, 0, 4.0, 0, 0, 0])
    np.testing.assert_allclose(q, [P["u0"], P["v0"], f / 4.0, 0, 0, 0], atol=1e-12)
    assert pd == P["pd"]
    # behind the camera: d < 0 -> outside the field of view
    q, pd = oracle.world_to_disparity(P, e, [0, 0, -4.0, 0, 0, 0])
    assert q[2] < 0 and pd == 0.0
    rng = np.random.default_rng(2)
    for _ in range(50):
        e = np.concatenate([rng.normal(0, 0.3, 6), rng.normal(0, 0.2, 3), rng.normal(0, 0.1, 3)])
        pw = np.concatenate([rng.uniform(-2, 2, 2), rng.uniform(2, 10, 1), rng.normal(0, 0.1, 3)])
        q, _ = oracle.world_to_disparity(P, e, pw)
        np.testing.assert_allclose(oracle.disparity_to_world(P, e, q), pw, rtol=1e-11, atol=1e-11)
        # camera-frame point from an independent R^T (p - t)
        R = rot_np(*e[6:9])
        pc = R.T @ (pw[:3] - e[:3])
        assert q[0] == pytest.approx(P["u0"] + f * pc[0] / pc[2], rel=1e-12)
        assert q[1] == pytest.approx(P["v0"] - P["fy"] * pc[1] / pc[2], rel=1e-12)
        assert q[2] == pytest.approx(f / pc[2], rel=1e-12)


def test_disparity_rates_are_time_derivatives():
    """(u', v', d') equal the finite-difference derivative of (u, v, d) along
    the camera-frame velocity."""
    P = params()
    rng = np.random.default_rng(3)
    e = np.concatenate([rng.normal(0, 0.1, 3), np.zeros(3), rng.normal(0, 0.1, 3), np.zeros(3)])
    for _ in range(20):
        pw = np.concatenate([rng.uniform(-1, 1, 2), rng.uniform(3, 8, 1), rng.normal(0, 0.5, 3)])
        q0, _ = oracle.world_to_disparity(P, e, pw)
        h = 1e-6
        p1 = pw.copy()
        p1[:3] += h * pw[3:]
        pm = pw.copy()
        pm[:3] -= h * pw[3:]
        d = (oracle.world_to_disparity(P, e, p1)[0][:3] - oracle.world_to_disparity(P, e, pm)[0][:3]) / (2 * h)
        np.testing.assert_allclose(q0[3:], d, rtol=1e-6, atol=1e-6)


def test_field_of_view_boundaries():
    """0 <= u <= 2 u0, 0 <= v <= 2 v0, d >= 0 (equality inside)."""
    P = params()
    f = -P["fx"]
    e = np.zeros(12)
    z = 5.0
    for u, v, inside in [(0.0, 240.0, True), (640.0, 240.0, True), (320.0, 0.0, True), (320.0, 480.0, True),
                         (-1e-6, 240.0, False), (640.0 + 1e-6, 240.0, False), (320.0, 480.001, False)]:
        x = (u - P["u0"]) * z / f
        y = (v - P["v0"]) * z / (-P["fy"])
        q, pd = oracle.world_to_disparity(P, e, [x, y, z, 0, 0, 0])
        assert abs(q[0] - u) < 1e-9 and abs(q[1] - v) < 1e-9
        if abs(q[0] - u) < 1e-12 and abs(q[1] - v) < 1e-12 and q[0] in (0.0, 640.0) or not inside:
            assert (pd > 0) == inside, (u, v)


# ------------------------------------------------------------------ 6x6 algebra
def test_cholesky_matches_numpy():
    rng = np.random.default_rng(4)
    for _ in range(20):
        A = spd(rng, cond=1e4)
        np.testing.assert_allclose(oracle.cholesky(A), np.linalg.cholesky(A), rtol=1e-10, atol=1e-12)


def test_positive_definiteness_rule():
    rng = np.random.default_rng(5)
    A = spd(rng)
    assert oracle.is_pos_def(A)
    B = A.copy()
    B[2, 2] = -1.0
    assert not oracle.is_pos_def(B)
    # rank-deficient (numerically singular) matrices are not PD under the
    # relative-pivot rule, whatever the sign of the rounding residue
    X = rng.normal(size=(6, 4))
    assert not oracle.is_pos_def(X @ X.T)
    # the rule is scale-free: tiny variances are fine, near-collinearity is
    # not; relative pivot of a correlation rho is 1 - rho^2
    assert oracle.is_pos_def(np.diag([1.0, 1.0, 1.0, 1.0, 1.0, 1e-12]))
    C = np.eye(6)
    C[4, 5] = C[5, 4] = 1 - 1e-7  # 1 - rho^2 ~ 2e-7
    assert oracle.is_pos_def(C)
    C[4, 5] = C[5, 4] = 1 - 1e-12  # ~ 2e-12, below 1e-9
    assert not oracle.is_pos_def(C)
    assert not oracle.is_pos_def(np.diag([1.0, 1.0, 1.0, 1.0, 1.0, np.nan]))


def test_mahalanobis_matches_solve():
    rng = np.random.default_rng(6)
    for _ in range(20):
        Pa, Pb = spd(rng), spd(rng)
        ma, mb = rng.normal(size=6), rng.normal(size=6)
        d = ma - mb
        ref = d @ np.linalg.solve(Pa + Pb, d)
        assert oracle.mahal(ma, Pa, mb, Pb) == pytest.approx(ref, rel=1e-10)
        assert oracle.mahal(ma, Pa, mb, Pb) == oracle.mahal(mb, Pb, ma, Pa)


def test_kalman_detection_term_is_textbook():
    """log w = log pd + log w_g + log N(z; H mu, H P H^T + R); mean and
    Joseph covariance from an independent numpy Kalman update; the Joseph
    covariance equals (I - K H) P for the optimal gain."""
    P0 = params()
    R = np.diag([P0["std_u"] ** 2, P0["std_v"] ** 2])
    H = np.hstack([np.eye(2), np.zeros((2, 4))])
    rng = np.random.default_rng(7)
    for _ in range(30):
        P = spd(rng, scale=3.0, cond=1e3)
        mu = rng.normal(0, 5, 6) + np.array([300, 200, 100, 0, 0, 0])
        z = mu[:2] + rng.normal(0, 3, 2)
        w, pd = rng.uniform(0.01, 1), rng.uniform(0.1, 1)
        lw, m2, P2, ok = oracle.kalman(P0, w, pd, mu, P, z[0], z[1])
        S = H @ P @ H.T + R
        K = P @ H.T @ np.linalg.inv(S)
        ref_lw = math.log(pd) + math.log(w) + scipy.stats.multivariate_normal(H @ mu, S).logpdf(z)
        assert ok
        assert lw == pytest.approx(ref_lw, rel=1e-12, abs=1e-12)
        np.testing.assert_allclose(m2, mu + K @ (z - H @ mu), rtol=1e-12)
        I = np.eye(6)
        np.testing.assert_allclose(P2, (I - K @ H) @ P @ (I - K @ H).T + K @ R @ K.T, rtol=1e-9, atol=1e-10)
        np.testing.assert_allclose(P2, (I - K @ H) @ P, rtol=1e-7, atol=1e-8)
        np.testing.assert_array_equal(P2, P2.T)


def test_fit_is_weighted_moments_with_reference_normaliser():
    """mean = sum w x; cov = sum w (x - m)(x - m)^T (W^2 + sum w^2) / W with
    w normalised to sum 1 (src/kernels.cu:196-205 normaliser)."""
    rng = np.random.default_rng(8)
    S = 50
    q = rng.normal(0, [3, 2, 10, 1, 1, 0.1], size=(S, 6)) + [300, 200, 100, 0, 0, 0]
    w = rng.uniform(0, 1, S)
    w[::7] = 0.0
    mean, cov = oracle.fit(q, w)
    wn = w / w.sum()
    m = wn @ q
    np.testing.assert_allclose(mean, m, rtol=1e-13)
    c = ((q - m).T * wn) @ (q - m) * (1 + np.sum(wn ** 2))
    np.testing.assert_allclose(cov, c, rtol=1e-10, atol=1e-12)
    # all-zero weights normalise to 1 each (src/kernels.cu:1012-1016)
    mean0, _ = oracle.fit(q, np.zeros(S))
    np.testing.assert_allclose(mean0, q.sum(0), rtol=1e-12)


# ------------------------------------------------------------------- merge
def _mix(rng, n, spread):
    ws = rng.uniform(0.01, 1, n)
    ms = rng.normal(0, spread, (n, 6)) + [300, 200, 100, 0, 0, 0]
    Ps = [spd(rng, scale=1.0, cond=10) for _ in range(n)]
    return ws, ms, Ps


def test_merge_invariants():
    P0 = params()
    rng = np.random.default_rng(9)
    for spread in (0.5, 3.0, 50.0):
        ws, ms, Ps = _mix(rng, 12, spread)
        ow, om, oP = oracle.merge(P0, ws, ms, Ps)
        assert 1 <= len(ow) <= len(ws)
        assert ow.sum() == pytest.approx(ws.sum(), rel=1e-12)
        for C in oP:
            assert np.all(np.linalg.eigvalsh(C) > 0)
        # mass-weighted mean is preserved
        np.testing.assert_allclose((ow[:, None] * om).sum(0), (ws[:, None] * ms).sum(0), rtol=1e-12)


def test_merge_far_apart_is_sorted_identity_and_close_pair_merges():
    P0 = params()
    rng = np.random.default_rng(10)
    ws, ms, Ps = _mix(rng, 6, 0.0)
    ms = ms + np.arange(6)[:, None] * 1e3  # far apart
    ow, om, oP = oracle.merge(P0, ws, ms, Ps)
    order = np.argsort(-ws, kind="stable")
    np.testing.assert_allclose(ow, ws[order], rtol=1e-15)
    np.testing.assert_allclose(om, ms[order], rtol=1e-15)
    # two components at the same mean merge into one: moment matching
    P1, P2 = spd(rng), spd(rng)
    m = np.array([300.0, 200, 100, 0, 0, 0])
    ow, om, oP = oracle.merge(P0, [0.3, 0.7], [m, m + 1e-3], [P1, P2])
    assert len(ow) == 1 and ow[0] == pytest.approx(1.0)
    mbar = (0.3 * m + 0.7 * (m + 1e-3))
    d1, d2 = mbar - m, mbar - (m + 1e-3)
    Pbar = 0.3 * (P1 + np.outer(d1, d1)) + 0.7 * (P2 + np.outer(d2, d2))
    np.testing.assert_allclose(om[0], mbar, rtol=1e-14)
    np.testing.assert_allclose(oP[0], Pbar, rtol=1e-12)


def test_merge_greedy_order_and_ties():
    """The heaviest term leads; a chain a-b-c with a, c out of range of each
    other merges {a, b} first (leader a heaviest) and leaves c; equal weights
    lead in index order."""
    P0 = params(min_separation=4.0)
    I = np.eye(6)
    m0 = np.array([300.0, 200, 100, 0, 0, 0])
    step = np.zeros(6)
    step[0] = 2.5  # d^2 between neighbours = 2.5^2 / 2 = 3.125 < 4; two steps: 12.5 > 4
    ms = [m0, m0 + step, m0 + 2 * step]
    ow, om, _ = oracle.merge(P0, [0.5, 0.3, 0.2], ms, [I, I, I])
    assert len(ow) == 2
    assert ow[0] == pytest.approx(0.8) and ow[1] == pytest.approx(0.2)
    ow, om, _ = oracle.merge(P0, [0.2, 0.3, 0.5], ms, [I, I, I])
    assert len(ow) == 2 and ow[0] == pytest.approx(0.8) and ow[1] == pytest.approx(0.2)
    # ties: both at weight 0.4, far apart -> emitted in index order
    ow, om, _ = oracle.merge(P0, [0.4, 0.4], [m0 + 100 * step, m0], [I, I])
    np.testing.assert_allclose(om[0], m0 + 100 * step)


def _walk_groups(w, m, P, thr):
    """Greedy merge groups formed by one pass over the (weight desc, index
    asc) order with precomputed pairwise membership -- the phased form the
    GPU merge uses: a leader takes its unmerged members; a leader that is not
    a member of its own group (non-PD covariance) stays selectable, is taken
    again with nothing left and ends the loop."""
    n = len(w)
    mem = [[oracle.mahal(m[i], P[i], m[j], P[j]) < thr for j in range(n)] for i in range(n)]
    order = sorted(range(n), key=lambda i: (-w[i], i))
    flag = [False] * n
    groups, r = [], 0
    while r < n:
        L = order[r]
        if flag[L]:
            r += 1
            continue
        g = [j for j in range(n) if not flag[j] and mem[L][j]]
        if not g:
            break
        for j in g:
            flag[j] = True
        groups.append((L, g))
        if mem[L][L]:
            r += 1
    return groups


def test_merge_groups_from_a_single_walk_equal_the_round_loop():
    """The oracle's merge is the reference's round loop (re-evaluating every
    unmerged term against each new leader); the walk over pairwise
    membership must give the same groups: same count, same weights, same
    moment-matched means.  Mixtures with clusters, chains, weight ties and a
    non-PD covariance."""
    P0 = params(min_separation=4.0)
    rng = np.random.default_rng(12)
    for trial in range(12):
        n = int(rng.integers(2, 14))
        centres = rng.normal(0, 3.0, size=(3, 6))
        ms = centres[rng.integers(0, 3, n)] + rng.normal(0, 1.0, size=(n, 6))
        Ps = [spd(rng) for _ in range(n)]
        ws = rng.choice([0.1, 0.2, 0.3], size=n)  # ties are common
        if trial % 3 == 0:
            Ps[int(rng.integers(0, n))] = -np.eye(6)  # not PD: never a member of anything
        ow, om, _ = oracle.merge(P0, ws, ms, Ps)
        groups = _walk_groups(ws, ms, Ps, 4.0)
        assert len(ow) == len(groups), trial
        for k, (L, g) in enumerate(groups):
            W = sum(ws[j] for j in g)
            mu = sum(ws[j] * ms[j] for j in g) / W
            C = sum(ws[j] * (Ps[j] + np.outer(mu - ms[j], mu - ms[j])) for j in g) / W
            if np.all(np.linalg.eigvalsh((C + C.T) / 2) > 0):
                assert ow[k] == pytest.approx(W, rel=1e-12), (trial, k)
                np.testing.assert_allclose(om[k], mu, rtol=1e-9, atol=1e-12)
            else:  # non-PD merge: the leader is emitted
                assert ow[k] == ws[L]
                np.testing.assert_array_equal(om[k], ms[L])


# ------------------------------------------------------------- resampling
def test_systematic_resampling_definition_and_counts():
    rng = np.random.default_rng(11)
    for N in (1, 5, 64, 1000):
        w = rng.gamma(0.3, size=N)
        w /= w.sum()
        u = rng.uniform(0, 1.0 / N)
        a = oracle.systematic(w, u)
        assert np.all(np.diff(a) >= 0)
        c = np.cumsum(w)
        r = u + np.arange(N) / N
        brute = np.array([np.searchsorted(c, x, side="left") for x in r])
        brute = np.minimum(brute, N - 1)
        np.testing.assert_array_equal(a, brute)
        counts = np.bincount(a, minlength=N)
        assert np.all((counts >= np.floor(N * w) - 1) & (counts <= np.ceil(N * w) + 1))
        assert counts.sum() == N


def test_systematic_resampling_overrun_fallback():
    """Weights summing below the last point: the remaining draws take the
    first maximum weight (src/scphdcameracalibration.cu:2675-2697)."""
    w = np.array([0.1, 0.3, 0.3, 0.1])  # sums to 0.8
    a = oracle.systematic(w, 0.2)
    # cumulative 0.1 0.4 0.7 0.8; points 0.2, 0.45, 0.7, 0.95 -> 1, 2, 2,
    # overrun -> first maximum weight = index 1
    np.testing.assert_array_equal(a, [1, 2, 2, 1])


# ----------------------------------------------------------- whole step
def test_prior_uniform_in_bounds_and_predict_closed_form():
    p = params(n_particles=64)
    p["cam_accel_std"] = [0.0] * 6
    p["feat_accel_std"] = [0.0] * 3
    o = oracle.Oracle(p)
    st0, w0 = o.parents()
    a, b = np.array(p["prior_a"]), np.array(p["prior_b"])
    assert np.all(st0 >= a - 1e-15) and np.all(st0 <= b + 1e-15)
    np.testing.assert_allclose(w0, 1.0 / 64)
    st0[:, 3:6] = 0.01
    st0[:, 9:12] = 0.002
    o.set_state(dict(states=st0, weights=w0, gcount=np.zeros(64, np.int32), wg=np.zeros(0),
                     samples=np.zeros((7, 0)), n_predict=0, n_update=0, n_resample=0))
    o.predict(0.5)
    st1, _ = o.parents()
    exp = st0.copy()
    exp[:, 0:3] += 0.5 * st0[:, 3:6]
    exp[:, 6:9] += 0.5 * st0[:, 9:12]
    np.testing.assert_allclose(st1, exp, atol=1e-15)


def test_first_update_is_births_only_and_fixed_camera_keeps_weights():
    p = params(n_particles=8)
    o = oracle.Oracle(p)
    u = np.array([100.0, 300.0, 500.0])
    v = np.array([100.0, 200.0, 400.0])
    o.predict(0.5)
    o.update(u, v, True)
    d = o.update_debug()
    M = 3
    bpc = math.log(M * p["birth_w0"] + p["lambda"] / (2 * p["v0"]) ** 2)
    np.testing.assert_allclose(d["eta"], bpc, rtol=1e-15)
    gc, wg, smp = o.maps()
    np.testing.assert_array_equal(gc, 3)
    np.testing.assert_allclose(wg, p["birth_w0"] / math.exp(bpc), rtol=1e-14)
    np.testing.assert_allclose(o.parents()[1], 1.0 / 8)
    np.testing.assert_allclose(smp[6], 1.0 / p["particles_per_feature"])


def test_normaliser_bounds_and_likelihood_sums():
    """loglik_p = sum_m eta_{p,m} + sum_g w_g over the predicted map
    (reference :1813-1849), and every normaliser is at least the
    birth/clutter term log(M w0 + kappa).  eta itself is checked against
    scipy's logsumexp over the real detection terms in
    tests/test_oracle_update.py."""
    import scene

    p, s = workloads.workload("c0", n_particles=4)
    sc = scene.generate(seed=3, **s)
    o = oracle.Oracle(p)
    for k in range(3):
        u, v, fixed = scene.frame_measurements(sc, k)
        o.predict(0.5)
        o.update(u, v, fixed)
    gc, wg, _ = o.maps()
    u, v, fixed = scene.frame_measurements(sc, 3)
    o.predict(0.5)
    wpred = o.maps()[1]
    o.update(u, v, fixed)
    d = o.update_debug()
    off = np.concatenate([[0], np.cumsum(gc)])
    for pp in range(4):
        assert d["loglik"][pp] == pytest.approx(d["eta"][pp].sum() + wpred[off[pp]:off[pp + 1]].sum(), rel=1e-12)
    M = len(u)
    bpc = math.log(M * p["birth_w0"] + p["lambda"] / (2 * p["v0"]) ** 2)
    assert np.all(d["eta"] >= bpc - 1e-12)


def test_log_normaliser_equals_scipy_logsumexp():
    # exercised through a births-only parent plus detection terms of known
    # log-weights: eta = logsumexp(l_g, bpc)
    rng = np.random.default_rng(12)
    for _ in range(10):
        l = rng.normal(-5, 3, 7)
        bpc = rng.normal(-3, 1)
        ref = scipy.special.logsumexp(np.append(l, bpc))
        acc = l[0]
        for x in l[1:]:
            acc = oracle.lse2(acc, x)
        assert oracle.lse2(acc, bpc) == pytest.approx(ref, rel=1e-14)


def test_grouped_sample_normals_partition_the_group_blocks():
    """Landmark prediction: samples s, s + 16, s + 32, s + 48 of a 64-sample
    stretch take consecutive triples of their group's twelve normals (three
    Philox blocks of element (s >> 6) << 4 | s & 15); posterior sampling:
    samples s, s + 16 of a 32-sample stretch take the two halves.  No two
    samples draw the same normals, and the pooled values are standard
    normal."""
    seed, step, p, k = 77, 5, 3, 11
    for s0 in (0, 7, 15, 64 + 9, 128 + 15):
        g = oracle.normals(seed, 2, step, p, k, ((s0 >> 6) << 4) | (s0 & 15), 12)
        for m in range(4):
            np.testing.assert_array_equal(oracle.stream_normals(seed, 2, step, p, k, s0 + 16 * m), g[3 * m:3 * m + 3])
    for s0 in (0, 5, 15, 32 + 3, 96 + 15):
        g = oracle.normals(seed, 3, step, p, k, ((s0 >> 5) << 4) | (s0 & 15), 12)
        for m in range(2):
            np.testing.assert_array_equal(oracle.stream_normals(seed, 3, step, p, k, s0 + 16 * m), g[6 * m:6 * m + 6])
    for stream, n in ((2, 3), (3, 6)):
        z = np.stack([oracle.stream_normals(seed, stream, step, p, k, s) for s in range(4096)])
        assert z.shape == (4096, n)
        assert len(np.unique(z, axis=0)) == len(z)  # no two samples draw the same normals
        assert abs(z.mean()) < 0.03 and abs(z.std() - 1.0) < 0.03
