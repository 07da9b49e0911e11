#!/usr/bin/env python3
"""Evaluation of calibration runs: the quantities of the reference's MATLAB
scripts (load_disparity_results.m, compute_disparity_crlb.m,
compute_batch_results.m), computed from the MAT files the filter writes.

    python evaluate.py RESULTS_DIR TRUTH.mat [--runs N] [--config CFG] [--steps K] [--out summary.json]

Pieces:

* calibration_error(estimate12, truth6): error of the estimated right-camera
  pose (x, y, z, theta, phi, psi = estimate components 0-2, 6-8) against the
  true pose, as cam_error in load_disparity_results.m:66; angle differences
  wrapped into [-pi, pi).
* ospa_dist(X, Y, c, p, metric, covs): the optimal sub-pattern assignment
  distance of Schuhmacher, Vo and Vo, "A consistent metric for performance
  evaluation of multi-object filters", IEEE Trans. Signal Processing 56(8),
  2008, with its localisation and cardinality components: with
  |X| = m <= |Y| = n and d_c = min(c, d),
      dist = ((min_pi sum_i d_c(x_i, y_pi(i))^p + c^p (n - m)) / n)^(1/p),
      loc  = (min_pi sum_i d_c(x_i, y_pi(i))^p / n)^(1/p),
      card = (c^p (n - m) / n)^(1/p),
  all 0 for two empty sets.  The base distance is Euclidean, or the
  Hellinger distance between Gaussians (the reference calls the un-vendored
  ospa_dist with 'DistanceMetric', 'Hellinger' and the maps' covariances,
  load_disparity_results.m:90-98).  Assignment by
  scipy.optimize.linear_sum_assignment.  ospa(X, Y, c, p) is the total alone.
* hellinger(m1, P1, m2, P2): sqrt(1 - BC) with the Bhattacharyya
  coefficient of two Gaussians, BC = det(P1)^(1/4) det(P2)^(1/4)
  det(Pm)^(-1/2) exp(-dm^T Pm^-1 dm / 8), Pm = (P1 + P2) / 2.
* DisparityModel: the stereo disparity measurement model of the filter
  (include/scphd_models.h, the reference's DisparityMeasurementModel) over
  numpy arrays of 6-row states, for poses [x, y, z, theta, phi, psi].
* disparity_map(...): the estimated map of one result file seen by one
  camera, as load_disparity_results.m:27-52: n = the number of Gaussians
  with weight >= 0.33, the first n sample clouds (storage order) projected
  into disparity space, and each cloud's mean and sample covariance.
* crlb_map_truth(...): the reference's CRLB-style map truth
  (compute_disparity_crlb.m:30-140): per target and camera a disparity-space
  Gaussian, predicted by sampling it, mapping the samples to the world with
  the previous camera, moving them with the constant-velocity model and
  projecting them into both cameras (sample mean and covariance), and
  updated at every step the target is in view: the first time with the
  birth covariance (and, for camera 1, the prior disparity d0), then by the
  Kalman covariance update (I - K H) P with H = [I2 0], R = diag(nu^2, nv^2),
  its mean set to the true measurement.
* load_results(...) / batch_results(...): the per-run and Monte-Carlo
  summaries of load_disparity_results.m and compute_batch_results.m:
  weighted mean and covariance of the parents, cardinality (sum of map
  weights), camera error, and the Hellinger OSPA (total, localisation,
  cardinality) of both cameras' maps against the CRLB truth (c = 1, p = 1).
  The estimated camera-2 pose of step k is step k's weighted mean
  (load_disparity_results.m:41 indexes the first column for every step).
* map_points(...): landmark estimates of one parent's map, the sample mean
  of every Gaussian whose weight reaches `threshold`.
"""
import argparse
import json
import math
import os
import sys

import numpy as np


def wrap(a):
    return (np.asarray(a) + math.pi) % (2 * math.pi) - math.pi


def calibration_error(estimate12, truth6):
    e = np.asarray(estimate12, dtype=np.float64)
    t = np.asarray(truth6, dtype=np.float64)
    pose = np.concatenate([e[0:3], e[6:9]])
    err = pose - t
    err[3:] = wrap(err[3:])
    return err


def pose_of_state(x12):
    """[x, y, z, theta, phi, psi] of a 12-element camera state
    (x, y, z, vx, vy, vz, theta, phi, psi, rates)."""
    x = np.asarray(x12, dtype=np.float64).ravel()
    return np.concatenate([x[0:3], x[6:9]])


# ------------------------------------------------------------ distances
def hellinger(m1, P1, m2, P2):
    m1, m2 = np.asarray(m1, dtype=np.float64), np.asarray(m2, dtype=np.float64)
    P1, P2 = np.atleast_2d(P1).astype(np.float64), np.atleast_2d(P2).astype(np.float64)
    Pm = 0.5 * (P1 + P2)
    dm = m1 - m2
    s1, l1 = np.linalg.slogdet(P1)
    s2, l2 = np.linalg.slogdet(P2)
    sm, lm = np.linalg.slogdet(Pm)
    if s1 <= 0 or s2 <= 0 or sm <= 0:
        return 1.0
    lbc = 0.25 * l1 + 0.25 * l2 - 0.5 * lm - 0.125 * float(dm @ np.linalg.solve(Pm, dm))
    return float(math.sqrt(max(0.0, 1.0 - math.exp(lbc))))


def _as_rows(X):
    X = np.asarray(X, dtype=np.float64)
    if X.size == 0:
        return X.reshape(0, X.shape[-1] if X.ndim > 1 else 1)
    return X.reshape(-1, X.shape[-1])


def ospa_dist(X, Y, c=1.0, p=1.0, metric="euclidean", covs=None):
    """(dist, loc, card) between point sets X (m x k) and Y (n x k); with
    metric 'hellinger', covs = (CX, CY), arrays of m and n k x k covariances."""
    from scipy.optimize import linear_sum_assignment

    X, Y = _as_rows(X), _as_rows(Y)
    CX, CY = (None, None) if covs is None else (np.asarray(covs[0]), np.asarray(covs[1]))
    m, n = len(X), len(Y)
    if m == 0 and n == 0:
        return 0.0, 0.0, 0.0
    if m > n:
        X, Y, m, n, CX, CY = Y, X, n, m, CY, CX
    if m == 0:
        return float(c), 0.0, float(c)
    if metric == "euclidean":
        D = np.linalg.norm(X[:, None, :] - Y[None, :, :], axis=2)
    elif metric == "hellinger":
        D = np.array([[hellinger(X[i], CX[i], Y[j], CY[j]) for j in range(n)] for i in range(m)])
    else:
        raise ValueError(f"unknown metric {metric!r}")
    D = np.minimum(c, D) ** p
    r, k = linear_sum_assignment(D)
    cost = float(D[r, k].sum())
    return (((cost + c ** p * (n - m)) / n) ** (1.0 / p), (cost / n) ** (1.0 / p),
            ((c ** p * (n - m)) / n) ** (1.0 / p))


def ospa(X, Y, c=1.0, p=1.0):
    return ospa_dist(X, Y, c, p)[0]


# ------------------------------------------------------ disparity model
def rotation(th, ph, ps):
    """Camera->world rotation R = Rx(th) Ry(ph) Rz(ps)."""
    ct, st, cp, sp, cs, ss = (math.cos(th), math.sin(th), math.cos(ph), math.sin(ph),
                              math.cos(ps), math.sin(ps))
    return np.array([
        [cp * cs, -cp * ss, sp],
        [ct * ss + st * sp * cs, ct * cs - st * sp * ss, -st * cp],
        [st * ss - ct * sp * cs, st * cs + ct * sp * ss, ct * cp],
    ])


class DisparityModel:
    """u = u0 - fx x/z, v = v0 - fy y/z, d = -fx/z and their rates, for
    camera-frame points p_c = R^T (p_w - t), v_c = R^T v_w."""

    def __init__(self, fx, fy, u0, v0):
        self.fx, self.fy, self.u0, self.v0 = float(fx), float(fy), float(u0), float(v0)

    @classmethod
    def from_params(cls, p):
        return cls(p["fx"], p["fy"], p["u0"], p["v0"])

    def compute_measurement(self, X, pose):
        X = np.asarray(X, dtype=np.float64).reshape(6, -1)
        pose = np.asarray(pose, dtype=np.float64).ravel()
        R = rotation(pose[3], pose[4], pose[5])
        pc = R.T @ (X[0:3] - pose[0:3, None])
        vc = R.T @ X[3:6]
        x, y, z = pc
        vx, vy, vz = vc
        return np.vstack([self.u0 - self.fx * x / z, self.v0 - self.fy * y / z, -self.fx / z,
                          -self.fx * (vx * z - vz * x) / z ** 2, -self.fy * (vy * z - vz * y) / z ** 2,
                          self.fx * vz / z ** 2])

    def invert_measurement(self, Q, pose):
        Q = np.asarray(Q, dtype=np.float64).reshape(6, -1)
        pose = np.asarray(pose, dtype=np.float64).ravel()
        u, v, d, vu, vv, vd = Q
        r = self.fx / self.fy
        pc = np.vstack([(u - self.u0) / d, r * (v - self.v0) / d, -self.fx / d])
        vc = np.vstack([(vu * d - vd * (u - self.u0)) / d ** 2, r * (vv * d - vd * (v - self.v0)) / d ** 2,
                        self.fx * vd / d ** 2])
        R = rotation(pose[3], pose[4], pose[5])
        return np.vstack([R @ pc + pose[0:3, None], R @ vc])

    def in_range(self, X, pose):
        """The filter's field-of-view test: 0 <= u <= 2 u0, 0 <= v <= 2 v0, d >= 0."""
        q = self.compute_measurement(X, pose)
        return (q[0] >= 0) & (q[0] <= 2 * self.u0) & (q[1] >= 0) & (q[1] <= 2 * self.v0) & (q[2] >= 0)


def cv_motion(X, dt, std3, rng):
    """Constant-velocity motion with N(0, std^2) accelerations per axis
    (LinearCVMotionModel3D::computeNoisyMotion)."""
    X = np.array(X, dtype=np.float64).reshape(6, -1)
    a = rng.standard_normal(X[0:3].shape) * np.asarray(std3, dtype=np.float64)[:, None]
    X[0:3] += X[3:6] * dt + 0.5 * a * dt * dt
    X[3:6] += a * dt
    return X


# ------------------------------------------------------------ map truth
def kalman_cov_update(P, nu, nv):
    """(I - K H) P with H = [I2 0], R = diag(nu^2, nv^2), K = P H^T S^-1,
    S = H P H^T + R (compute_disparity_crlb.m:75-78)."""
    P = np.asarray(P, dtype=np.float64)
    H = np.hstack([np.eye(2), np.zeros((2, 4))])
    S = H @ P @ H.T + np.diag([nu ** 2, nv ** 2])
    K = P @ H.T @ np.linalg.inv(S)
    return (np.eye(6) - K @ H) @ P


def crlb_map_truth(true_traj, cam_traj, campose_1, model, dt, std3, n_particles=256, nu=3.0, nv=3.0, d0=8.0,
                   seed=0):
    """Per step k: {'gaussians': [(means1, covs1), (means2, covs2)] of the
    targets seen so far and in view (6 x n, 6 x 6 x n), 'particles': the
    predicted world samples of the targets kept for camera 1, 'targets'}.
    true_traj is 6 x K x T, cam_traj 12 x K (camera-2 states), campose_1 the
    fixed camera's pose."""
    rng = np.random.default_rng(seed)
    true_traj = np.asarray(true_traj, dtype=np.float64)
    K, T = true_traj.shape[1], true_traj.shape[2]
    pose1 = np.asarray(campose_1, dtype=np.float64).ravel()[:6]
    birth_cov = np.diag([nu, nv, 3.0, 1e-4, 1e-4, 1e-4]) ** 2
    mean = [np.full((6, T), np.nan), np.full((6, T), np.nan)]
    cov = [np.full((6, 6, T), np.nan), np.full((6, 6, T), np.nan)]
    truth = []
    for k in range(K):
        pose2 = pose_of_state(np.asarray(cam_traj, dtype=np.float64)[:, k])
        poses = [pose1, pose2]
        # odd steps (1-based) predict from camera 2's Gaussians, seen last with camera 2
        src, pose_old = (1, pose2) if (k + 1) % 2 == 1 else (0, pose1)
        world = None
        if k > 0:
            world = np.full((6, n_particles, T), np.nan)
            for i in range(T):
                m, P = mean[src][:, i], cov[src][:, :, i]
                if np.any(np.isnan(m)) or np.any(np.isnan(P)):
                    continue
                L = np.linalg.cholesky(0.5 * (P + P.T) + 1e-12 * np.eye(6))
                q = m[:, None] + L @ rng.standard_normal((6, n_particles))
                Xw = cv_motion(model.invert_measurement(q, pose_old), dt, std3, rng)
                for c in range(2):
                    qd = model.compute_measurement(Xw, poses[c])
                    mean[c][:, i] = qd.mean(axis=1)
                    cov[c][:, :, i] = np.cov(qd)
                world[:, :, i] = Xw
        targets = true_traj[:, k, :]
        seen = []
        for c in range(2):
            inr = model.in_range(targets, poses[c])
            zt = model.compute_measurement(targets, poses[c])
            for i in range(T):
                if not inr[i]:
                    continue
                P = cov[c][:, :, i]
                mean[c][:, i] = zt[:, i]
                if np.any(np.isnan(P)):  # seen for the first time
                    cov[c][:, :, i] = birth_cov
                    if c == 0:
                        mean[c][2, i] = d0
                else:
                    cov[c][:, :, i] = kalman_cov_update(P, nu, nv)
            keep = ~(np.isnan(mean[c][0]) | ~inr)
            seen.append((mean[c][:, keep].copy(), cov[c][:, :, keep].copy(), keep))
        particles = (world[:, :, seen[0][2]].reshape(6, -1) if world is not None
                     else np.full((3, 1), np.nan))
        truth.append({"gaussians": [(s[0], s[1]) for s in seen], "particles": particles, "targets": targets})
    return truth


# -------------------------------------------------------- run summaries
def weighted_mean_cov(X, w):
    """Weighted mean and covariance of the columns of X (k x N) with weights w."""
    X = np.asarray(X, dtype=np.float64)
    w = np.asarray(w, dtype=np.float64).ravel()
    w = w / w.sum()
    mu = X @ w
    D = X - mu[:, None]
    return mu, (D * w) @ D.T


def disparity_map(map_weights, map_particles, S, pose, model, threshold=0.33):
    w = np.asarray(map_weights, dtype=np.float64).ravel()
    n = int(np.sum(w >= threshold))
    if n == 0:
        return np.zeros((6, 0)), np.zeros((6, 6, 0))
    q = model.compute_measurement(np.asarray(map_particles, dtype=np.float64)[:, :n * S], pose)
    means = np.stack([q[:, i * S:(i + 1) * S].mean(axis=1) for i in range(n)], axis=1)
    covs = np.stack([np.cov(q[:, i * S:(i + 1) * S]) for i in range(n)], axis=2)
    return means, covs


def load_results(results_dir, truth, campose_1, campose_2, model, n_steps=None):
    """Summary of one run's MAT files results_dir/<k>.mat against `truth`
    (crlb_map_truth's list)."""
    from scipy.io import loadmat

    K = len(truth) if n_steps is None else n_steps
    exp_traj = np.zeros((12, K))
    cov_traj = np.zeros((12, 12, K))
    cardinality = np.zeros(K)
    ospa1 = np.zeros((3, K))
    ospa2 = np.zeros((3, K))
    pose1 = np.asarray(campose_1, dtype=np.float64).ravel()[:6]
    for k in range(K):
        d = loadmat(os.path.join(results_dir, f"{k}.mat"))
        particles, w = d["particles"], d["weights"].ravel()
        exp_traj[:, k], cov_traj[:, :, k] = weighted_mean_cov(particles, w)
        mw = d["map_weights"].ravel()
        mp = d["map_particles"]
        cardinality[k] = mw.sum()
        S = mp.shape[1] // max(len(mw), 1) if len(mw) else 1
        pose2 = pose_of_state(exp_traj[:, k])
        for c, (pose, out) in enumerate(((pose1, ospa1), (pose2, ospa2))):
            mm, cc = disparity_map(mw, mp, S, pose, model)
            gm, gc = truth[k]["gaussians"][c]
            out[:, k] = ospa_dist(mm[0:3].T, gm[0:3].T, 1.0, 1.0, "hellinger",
                                  (np.moveaxis(cc[0:3, 0:3], 2, 0), np.moveaxis(gc[0:3, 0:3], 2, 0)))
    cam_error = np.stack([calibration_error(exp_traj[:, k], campose_2) for k in range(K)], axis=1)
    return {"exp_traj": exp_traj, "cov_traj": cov_traj, "cardinality": cardinality, "cam_error": cam_error,
            "ospa1": ospa1, "ospa2": ospa2}


def batch_results(runs):
    """Means over the runs that loaded (compute_batch_results.m:21-40)."""
    ok = [r for r in runs if r is not None]
    if not ok:
        return None
    return {k: np.mean([r[k] for r in ok], axis=0) for k in ("cam_error", "ospa1", "ospa2", "cardinality")} | \
        {"runs": len(ok)}


def map_points(gcount, wg, samples, parent, S, threshold=0.5):
    off = np.concatenate([[0], np.cumsum(gcount)])
    g0, g1 = int(off[parent]), int(off[parent + 1])
    pts = []
    for g in range(g0, g1):
        if wg[g] >= threshold:
            pts.append(samples[:3, g * S:(g + 1) * S].mean(axis=1))
    return np.array(pts).reshape(-1, 3)


# ------------------------------------------------------------------ CLI
def main(argv=None):
    from scipy.io import loadmat

    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("results_dir", help="a run's directory of <k>.mat files, or the root of runs 1..N")
    ap.add_argument("truth", help="the scene MAT file (true_traj, cam_traj, campose_1, campose_2, dt, std_a*)")
    ap.add_argument("--runs", type=int, default=0, help="evaluate results_dir/1 .. results_dir/N")
    ap.add_argument("--config", default=None,
                    help="filter configuration (camera intrinsics); default: the .cfg in results_dir")
    ap.add_argument("--steps", type=int, default=0)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)
    here = os.path.dirname(os.path.abspath(__file__))
    sys.path.insert(0, here)
    import scphd_amd

    cfg_path = args.config
    if cfg_path is None:  # batch_run.py copies the configuration into the results root
        cands = sorted(f for f in os.listdir(args.results_dir) if f.endswith(".cfg"))
        if not cands:
            ap.error("no --config given and no .cfg in the results directory")
        cfg_path = os.path.join(args.results_dir, cands[0])
    model = DisparityModel.from_params(scphd_amd.load_params(cfg_path))
    t = loadmat(args.truth)
    K = args.steps or t["true_traj"].shape[1]
    std3 = [float(np.ravel(t.get(k, [[0.0]]))[0]) for k in ("std_ax", "std_ay", "std_az")]
    dt = float(np.ravel(t.get("dt", [[0.5]]))[0])
    campose_1 = np.ravel(t["campose_1"]).astype(np.float64)
    campose_2 = np.ravel(t["campose_2"]).astype(np.float64)
    truth = crlb_map_truth(t["true_traj"][:, :K], t["cam_traj"][:, :K], campose_1, model, dt, std3, seed=args.seed)
    dirs = [os.path.join(args.results_dir, str(n)) for n in range(1, args.runs + 1)] if args.runs else \
        [args.results_dir]
    runs = []
    for d in dirs:
        try:
            runs.append(load_results(d, truth, campose_1, campose_2, model, K))
        except (OSError, KeyError, ValueError) as e:
            print(f"{d}: not evaluated ({e})", file=sys.stderr)
            runs.append(None)
    summary = batch_results(runs)
    out = json.dumps({k: (v.tolist() if isinstance(v, np.ndarray) else v) for k, v in (summary or {}).items()})
    if args.out:
        with open(args.out, "w") as f:
            f.write(out + "\n")
    else:
        print(out)
    return 0 if summary else 1


if __name__ == "__main__":
    sys.exit(main())
