"""Independent pin of the oracle's composition of one whole update.

A numpy brute force of one SC-PHD update, written from the reference's
description and textbook forms rather than from the oracle, runs on tiny
hand-made states (2 parents, 3 Gaussians each, 24 samples per Gaussian,
5 measurements) and is compared with oracle.update() on every output:

* the per-sample detection probabilities, the reweight and the refit
  (fitGaussiansKernel, src/kernels.cu:102-438, including the covariance
  normaliser (W^2 + sum w^2) / W of :189-205);
* the normalisers eta_{p,m} = scipy logsumexp over the real detection terms
  and the birth/clutter term log(M w0 + kappa)
  (src/scphdcameracalibration.cu:1729-1768, src/disparitymeasurementmodel.cu:161);
* the survivor set and its pack order: detection terms measurement-major,
  then the births (updateKernel, src/kernels.cu:608-659), with the Kalman
  update in its textbook Joseph form;
* the greedy merge (src/kernels.cu:662-874) by a plain argmax loop;
* the non-detection keep rule w_g (1 - pd_g) >= min_weight and its sample
  weights 1 - pd_i (computeNonDetections, :1290-1420);
* the single-cluster parent reweight w_p <- w_p exp(sum_m eta + sum_g w_g),
  normalised (:1813-1876), only for the moving camera;
* posterior samples x = L z + mu (L = numpy Cholesky, z the injected
  normals of the posterior stream) unprojected to the world
  (:640-690, :1972-2010, src/kernels.cu:49-80);
* the [posterior | non-detection] layout of each parent's new map
  (:1422-1574, :2038-2059).

The rotation is composed from the three elementary rotations
R = Rx(theta) Ry(phi) Rz(psi) (the convention of transformCameraToWorld,
src/disparitymeasurementmodel.cu:82-144), not from the oracle's expanded
coefficients.  Positive-definiteness is decided by numpy's Cholesky; the
states are built so that no decision is near a threshold.
"""
import math

import numpy as np
import pytest
import scipy.special

import oracle
import workloads

ST_POSTERIOR = 3


def rot(th, ph, ps):
    cx, sx, cy, sy, cz, sz = math.cos(th), math.sin(th), math.cos(ph), math.sin(ph), math.cos(ps), math.sin(ps)
    rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    return rx @ ry @ rz


def to_disparity(P, R, t, pw):
    """World samples (n x 6) -> (u, v, d, du, dv, dd) of the camera (R, t)."""
    pc = (pw[:, :3] - t) @ R  # rows R^T (p - t)
    vc = pw[:, 3:] @ R
    x, y, z = pc.T
    vx, vy, vz = vc.T
    fx, fy = P["fx"], P["fy"]
    u = P["u0"] - fx * x / z
    v = P["v0"] - fy * y / z
    d = -fx / z
    # time derivatives of u, v, d by the quotient rule
    du = -fx * (vx * z - x * vz) / z ** 2
    dv = -fy * (vy * z - y * vz) / z ** 2
    dd = fx * vz / z ** 2
    return np.column_stack([u, v, d, du, dv, dd])


def to_world(P, R, t, q):
    """Inverse of to_disparity: camera point from (u, v, d) and its rates,
    then world = R p + t."""
    u, v, d, du, dv, dd = q.T
    fx, fy = P["fx"], P["fy"]
    z = -fx / d
    x = (u - P["u0"]) * z / (-fx)
    y = (v - P["v0"]) * z / (-fy)
    # the rates by inverting the quotient rule: d = -fx / z gives z' = fx d' / d^2
    vz = fx * dd / d ** 2
    vx = (du * z / (-fx)) + x * vz / z
    vy = (dv * z / (-fy)) + y * vz / z
    pc = np.column_stack([x, y, z])
    vc = np.column_stack([vx, vy, vz])
    return np.column_stack([pc @ R.T + t, vc @ R.T])


MARGINS = []  # smallest relative Cholesky pivot of every PD decision taken


def chol_pd(A):
    """Positive-definite by numpy's Cholesky; records the decision's margin
    (smallest pivot L_ii^2 / A_ii) so the test can check that none is near
    the oracle's relative-pivot rule (1e-9)."""
    try:
        L = np.linalg.cholesky(A)
    except np.linalg.LinAlgError:
        MARGINS.append(0.0)
        return False
    MARGINS.append(float(np.min(np.diag(L) ** 2 / np.diag(A))))
    return True


def brute_update(P, st, zu, zv, fixed):
    """One update of the state dict st (after predict) -> expected outputs."""
    Np, S = P["n_particles"], P["particles_per_feature"]
    M = len(zu)
    Rm = np.diag([P["std_u"] ** 2, P["std_v"] ** 2])
    H = np.zeros((2, 6))
    H[0, 0] = H[1, 1] = 1.0
    kappa = P["lambda"] / (2 * P["v0"]) ** 2
    bpc = math.log(M * P["birth_w0"] + kappa)
    off = np.concatenate([[0], np.cumsum(st["gcount"])])
    smp = st["samples"]
    out = dict(eta=np.zeros((Np, M)), loglik=np.zeros(Np), nsurv=np.zeros(Np, int), nmerged=np.zeros(Np, int),
               nnd=np.zeros(Np, int), gcount=[], wg=[], samples=[], fit=[], surv=[])
    for p in range(Np):
        ext = np.zeros(12) if fixed else st["states"][p]
        R = rot(*ext[6:9])
        t = ext[0:3]
        fits = []
        for g in range(off[p], off[p + 1]):
            cl = smp[:, g * S:(g + 1) * S].T  # S x 7
            q = to_disparity(P, R, t, cl[:, :6])
            vis = (q[:, 0] >= 0) & (q[:, 0] <= 2 * P["u0"]) & (q[:, 1] >= 0) & (q[:, 1] <= 2 * P["v0"]) & (q[:, 2] >= 0)
            pdi = np.where(vis, P["pd"], 0.0)
            pdg = pdi.mean()
            w = cl[:, 6] * pdi
            if pdg == 0:
                mu, C = q[0].copy(), np.zeros((6, 6))
            else:
                w = w / w.sum() if w.sum() > 0 else np.ones(S)
                mu = (w[:, None] * q).sum(0)
                W, W2 = w.sum(), (w * w).sum()
                dq = q - mu
                C = (dq.T * w) @ dq * (W * W + W2) / W
            fits.append(dict(g=g, w=st["wg"][g], q=q, pdi=pdi, pdg=pdg, mu=mu, P=C))
        out["fit"].append(fits)
        # detection terms
        terms = []  # [m][k] -> (logw, mean, cov, pd)
        for m in range(M):
            row = []
            z = np.array([zu[m], zv[m]])
            for f in fits:
                Sm = H @ f["P"] @ H.T + Rm
                Sm[0, 1] = Sm[1, 0] = 0.5 * (Sm[0, 1] + Sm[1, 0])
                nu = z - H @ f["mu"]
                Si = np.linalg.inv(Sm)
                K = f["P"] @ H.T @ Si
                A = np.eye(6) - K @ H
                Pn = A @ f["P"] @ A.T + K @ Rm @ K.T
                with np.errstate(divide="ignore"):
                    lw = (np.log(f["pdg"]) + np.log(f["w"]) - 0.5 * nu @ Si @ nu - math.log(2 * math.pi)
                          - 0.5 * math.log(np.linalg.det(Sm)))
                pd_ok = chol_pd(Pn)
                row.append((lw if pd_ok else -np.finfo(np.float32).max, f["mu"] + K @ nu, Pn, pd_ok))
            terms.append(row)
        eta = np.array([scipy.special.logsumexp([r[0] for r in terms[m]] + [bpc]) for m in range(M)])
        out["eta"][p] = eta
        out["loglik"][p] = eta.sum() + sum(f["w"] for f in fits)
        # pack: detection terms measurement-major, then births
        surv = []
        for m in range(M):
            for k, (lw, mu, C, ok) in enumerate(terms[m]):
                wt = math.exp(lw - eta[m])
                if ok and wt >= P["min_weight"]:
                    surv.append((wt, mu, C, ("det", m, k)))
        bcov = np.diag([P["std_u"] ** 2, P["std_v"] ** 2, P["birth_var_d0"], P["birth_var_vu0"],
                        P["birth_var_vv0"], P["birth_var_vd0"]])
        for m in range(M):
            wt = P["birth_w0"] / math.exp(eta[m])
            if wt >= P["min_weight"]:
                bm = np.array([zu[m], zv[m], P["birth_d0"], P["birth_vu0"], P["birth_vv0"], P["birth_vd0"]])
                surv.append((wt, bm, bcov, ("birth", m)))
        out["surv"].append([s_[3] for s_ in surv])
        out["nsurv"][p] = len(surv)
        # greedy merge: heaviest unmerged leader (lowest index on ties)
        used = [False] * len(surv)
        merged = []
        while True:
            cand = [i for i in range(len(surv)) if not used[i]]
            if not cand:
                break
            j = max(cand, key=lambda i: (surv[i][0], -i))
            mem = []
            for i in cand:
                dl = surv[j][1] - surv[i][1]
                Sg = surv[j][2] + surv[i][2]
                if dl @ np.linalg.solve(Sg, dl) < P["min_separation"]:
                    mem.append(i)
            W = sum(surv[i][0] for i in mem)
            if W == 0:
                break
            mu = sum(surv[i][0] * surv[i][1] for i in mem) / W
            C = sum(surv[i][0] * (surv[i][2] + np.outer(mu - surv[i][1], mu - surv[i][1])) for i in mem) / W
            C = 0.5 * (C + C.T)
            merged.append((W, mu, C) if chol_pd(C) else surv[j][:3])
            for i in mem:
                used[i] = True
        out["nmerged"][p] = len(merged)
        # new map: posterior clouds then kept non-detection clouds
        wg_new, smp_new = [], []
        for k, (W, mu, C) in enumerate(merged):
            L = np.linalg.cholesky(C)
            z = np.stack([oracle.stream_normals(P["seed"], ST_POSTERIOR, st["n_update"], p, k, s_) for s_ in range(S)])
            q = z @ L.T + mu
            pw = to_world(P, R, t, q)
            wg_new.append(W)
            smp_new.append(np.column_stack([pw, np.full(S, 1.0 / S)]))
        nnd = 0
        for f in fits:
            wnd = f["w"] * (1 - f["pdg"])
            if wnd >= P["min_weight"]:
                nnd += 1
                wg_new.append(wnd)
                g = f["g"]
                smp_new.append(np.column_stack([smp[:6, g * S:(g + 1) * S].T, 1.0 - f["pdi"]]))
        out["nnd"][p] = nnd
        out["gcount"].append(len(wg_new))
        out["wg"] += wg_new
        out["samples"] += smp_new
    out["gcount"] = np.array(out["gcount"])
    out["wg"] = np.array(out["wg"])
    out["samples"] = np.concatenate(out["samples"]).T if out["samples"] else np.zeros((7, 0))
    lw = np.log(st["weights"]) + out["loglik"]
    out["pw"] = st["weights"].copy() if fixed else np.exp(lw - scipy.special.logsumexp(lw))
    return out


def tiny_state(P, rng, fixed):
    """Two parents with three Gaussians each, 24 samples per cloud: a landmark
    seen in the middle of the image (weight 0.9), one straddling the right
    image border of the camera used (weight 0.4: part of its cloud is out of
    view) and a faint one (weight 2e-3: its non-detection term falls below
    min_weight)."""
    S = P["particles_per_feature"]
    xb = 2.56 + (0.0 if fixed else 0.2)  # u = 640 at 4 m for the left camera, 0.2 m further for the right
    centres = [np.array([0.3, -0.2, 4.0]), np.array([xb, 0.1, 4.0]), np.array([-0.8, 0.5, 6.0])]
    weights = [0.9, 0.4, 2e-3]
    spread = [0.01, 0.05, 0.02]
    gcount, wg, cols = [], [], []
    for p in range(P["n_particles"]):
        gcount.append(3)
        for c, w, sd in zip(centres, weights, spread):
            pos = c + rng.normal(0, sd, (S, 3))
            vel = rng.normal(0, 0.01, (S, 3))
            cols.append(np.column_stack([pos, vel, np.full(S, 1.0 / S)]))
            wg.append(w)
    states = np.zeros((P["n_particles"], 12))
    states[:, 0] = 0.2
    states[1, 0] += 0.01
    states[1, 7] += 0.002
    return dict(states=states, weights=np.full(P["n_particles"], 1.0 / P["n_particles"]),
                gcount=np.array(gcount, dtype=np.int32), wg=np.array(wg), samples=np.concatenate(cols).T.copy(),
                n_predict=3, n_update=3, n_resample=3)


def measurements(P, st, fixed):
    """A measurement at each of the first two landmarks as parent 0's camera
    sees them (0.3 px off), a second one near the first landmark (so two
    detection terms of one Gaussian survive and merge), and two clutter
    points over 150 px from every landmark: their detection terms underflow
    to exp() = 0, so both normalisers equal the birth/clutter term exactly
    and the two birth terms tie in weight -- the merge then emits them in
    pack order (lowest index first on ties)."""
    ext = np.zeros(12) if fixed else st["states"][0]
    R, t = rot(*ext[6:9]), ext[0:3]
    S = P["particles_per_feature"]
    zu, zv = [], []
    for g in (0, 1):
        q = to_disparity(P, R, t, st["samples"][:6, g * S:(g + 1) * S].T).mean(0)
        zu.append(q[0] + 0.3)
        zv.append(q[1] - 0.3)
    zu.append(zu[0] + 1.2)
    zv.append(zv[0] + 0.4)
    zu += [100.0, 600.0]
    zv += [400.0, 40.0]
    return np.array(zu), np.array(zv)


@pytest.mark.parametrize("fixed", [False, True], ids=["moving", "fixed"])
def test_whole_update_equals_numpy_brute_force(fixed):
    p, _ = workloads.workload("c0", n_particles=2, particles_per_feature=24)
    p["storage"] = 1
    rng = np.random.default_rng(11)
    st0 = tiny_state(p, rng, fixed)
    o = oracle.Oracle(p, threads=1)
    o.set_state(st0)
    o.predict(0.5)
    st = o.state()
    # the predicted state is the update's input on both sides
    zu, zv = measurements(p, st, fixed)
    MARGINS.clear()
    exp = brute_update(p, st, zu, zv, fixed)
    assert all(m == 0.0 or m > 1e-6 for m in MARGINS), "a PD decision of the fixture is near the threshold"
    o.update(zu, zv, fixed)
    d = o.update_debug()
    gc, wg, smp = o.maps()
    # the refit: means, covariances, pd_g
    fit = o.fit_debug(int(st["gcount"].sum()))
    k = 0
    for fits in exp["fit"]:
        for f in fits:
            np.testing.assert_allclose(fit[:6, k], f["mu"], rtol=1e-12, atol=1e-12)
            e = 6
            for a in range(6):
                for b in range(a + 1):
                    assert fit[e, k] == pytest.approx(f["P"][a, b], rel=1e-9, abs=1e-12)
                    e += 1
            assert fit[27, k] == pytest.approx(f["pdg"], rel=1e-15)
            k += 1
    # the decisions the fixture is built to exercise
    assert (exp["nnd"] == 2).all(), "two of three non-detection terms kept per parent"
    pdg_border = exp["fit"][0][1]["pdg"]
    assert 0 < pdg_border < p["pd"], "the border cloud is partly out of view"
    assert (exp["nmerged"] < exp["nsurv"]).all(), "some survivors merge"
    n0 = exp["nmerged"][0]
    assert np.sum(exp["wg"][:n0] == exp["wg"][n0 - 1]) >= 2, "two birth terms tie in weight"
    # counts and normalisers
    np.testing.assert_array_equal(d["nsurvivors"], exp["nsurv"])
    np.testing.assert_array_equal(d["nmerged"], exp["nmerged"])
    np.testing.assert_array_equal(d["nnondetect"], exp["nnd"])
    np.testing.assert_allclose(d["eta"], exp["eta"], rtol=1e-12)
    np.testing.assert_allclose(d["loglik"], exp["loglik"], rtol=1e-12)
    # the new maps: sizes, [posterior | non-detection] weights and samples
    np.testing.assert_array_equal(gc, exp["gcount"])
    np.testing.assert_allclose(wg, exp["wg"], rtol=1e-10)
    np.testing.assert_allclose(smp, exp["samples"], rtol=1e-9, atol=1e-12)
    # parent weights: reweighted only for the moving camera
    np.testing.assert_allclose(o.parents()[1], exp["pw"], rtol=1e-12)
    if fixed:
        np.testing.assert_array_equal(o.parents()[1], st["weights"])
    else:
        assert abs(o.parents()[1][0] - 0.5) > 1e-3, "the moving camera's likelihood tells the parents apart"
