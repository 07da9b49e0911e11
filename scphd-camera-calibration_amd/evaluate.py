"""Evaluation of calibration runs: the quantities of the reference's MATLAB
scripts (load_disparity_results.m:27-34,66, compute_disparity_ospa.m),
computed from the filter's state or from the MAT files it writes.

* calibration_error(estimate12, truth6): error of the estimated right-camera
  pose (x, y, z, theta, phi, psi = estimate components 0-2, 6-8) against the
  true pose, as cam_error in load_disparity_results.m:66; angle differences
  wrapped into [-pi, pi).
* ospa(X, Y, c, p): the optimal sub-pattern assignment distance of
  Schuhmacher, Vo and Vo, "A consistent metric for performance evaluation of
  multi-object filters", IEEE Trans. Signal Processing 56(8), 2008: with
  |X| = m <= |Y| = n, d = ( (min over assignments of sum_i
  min(c, |x_i - y_pi(i)|)^p + c^p (n - m)) / n )^(1/p), and 0 for two
  empty sets.  The reference calls an un-vendored ospa_dist; this is an
  implementation of the published definition (assignment by
  scipy.optimize.linear_sum_assignment).
* map_points(gcount, wg, samples, parent, S, threshold): landmark estimates
  of one parent's map -- the sample mean of every Gaussian whose weight
  reaches `threshold` (load_disparity_results.m:36-37 keeps features with
  weight >= 0.5).
"""
import math

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


def ospa(X, Y, c=1.0, p=1.0):
    from scipy.optimize import linear_sum_assignment

    X = np.asarray(X, dtype=np.float64).reshape(-1, np.shape(X)[-1] if np.size(X) else 1)
    Y = np.asarray(Y, dtype=np.float64).reshape(-1, np.shape(Y)[-1] if np.size(Y) else 1)
    m, n = len(X), len(Y)
    if m == 0 and n == 0:
        return 0.0
    if m > n:
        X, Y, m, n = Y, X, n, m
    if m == 0:
        return float(c)
    D = np.minimum(c, np.linalg.norm(X[:, None, :] - Y[None, :, :], axis=2)) ** p
    r, k = linear_sum_assignment(D)
    return float(((D[r, k].sum() + c ** p * (n - m)) / n) ** (1.0 / p))


def map_points(gcount, wg, samples, parent, S, threshold=0.5):
    off = np.concatenate([[0], np.cumsum(gcount)])
    g0, g1 = int(off[parent]), int(off[parent + 1])
    pts = []
    for g in range(g0, g1):
        if wg[g] >= threshold:
            pts.append(samples[:3, g * S:(g + 1) * S].mean(axis=1))
    return np.array(pts).reshape(-1, 3)
