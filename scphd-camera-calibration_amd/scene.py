"""Seeded synthetic stereo scenes for the calibration filter.

Plays the role of the reference's MATLAB scene generator
(cheesinglee/scphd-camera-calibration gen_disparity_scene.m) with its own,
documented choices:

* world frame = left (fixed) camera frame;
* static landmarks uniform in x, y in [-2, 2] m, z in [2, 10] m, seen from a
  rig translating at constant velocity v_rig, so in the filter frame the
  landmarks move at -v_rig;
* right camera true extrinsics [baseline, 0, 0, 0, 0, 0] plus
  N(0, diag(0.02 m, 0.02 m, 0.02 m, 1 deg, 1 deg, 1 deg)^2), rates zero;
* every landmark inside the field of view (the range test of the reference's
  worldToDisparityKernel, src/kernels.cu:37-41) is detected with
  probability pd and N(0, sigma^2) pixel noise on (u, v);
* Poisson(lambda) clutter uniform over the image.  A disparity in [1, 64] is
  drawn for each clutter point and recorded but not consumed, because the
  reference update uses (u, v) only (H = [I2 0], src/kernels.cu:501-511).

All draws come from numpy's PCG64 seeded by SplitMix64(master, rig).
"""
import math

import numpy as np

MASK64 = (1 << 64) - 1


def splitmix64(master, index):
    z = (master + (index + 1) * 0x9E3779B97F4A7C15) & MASK64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
    return z ^ (z >> 31)


def rotation(th, ph, ps):
    """Camera->world rotation R = Rx(th) Ry(ph) Rz(ps) (the coefficients of
    transformCameraToWorld, src/disparitymeasurementmodel.cu:82-144)."""
    ct, st, cp, sp, cs, ss = (math.cos(th), math.sin(th), math.cos(ph), math.sin(ph),
                              math.cos(ps), math.sin(ps))
    return np.array([
        [cp * cs, -cp * ss, sp],
        [ct * ss + st * sp * cs, ct * cs - st * sp * ss, -st * cp],
        [st * ss - ct * sp * cs, st * cs + ct * sp * ss, ct * cp],
    ])


def project(points, ext, f, u0, v0):
    """(u, v, d) of world points (n x 3) seen by a camera with extrinsics ext."""
    R = rotation(ext[6], ext[7], ext[8])
    pc = (points - np.asarray(ext[:3])) @ R  # rows: R^T (p - t)
    z = pc[:, 2]
    fx = fy = -f
    u = u0 - fx / z * pc[:, 0]
    v = v0 - fy / z * pc[:, 1]
    d = -fx / z
    return u, v, d


def generate(seed=1, rig=0, n_landmarks=30, lam=10.0, f=500.0, baseline=0.2, n_frames=50, dt=0.5,
             pd=0.95, sigma=1.0, width=640.0, height=480.0, v_rig=(0.04, 0.0, -0.04)):
    rng = np.random.Generator(np.random.PCG64(splitmix64(seed, rig)))
    L = n_landmarks
    p0 = np.column_stack([rng.uniform(-2, 2, L), rng.uniform(-2, 2, L), rng.uniform(2, 10, L)])
    vel = -np.asarray(v_rig, dtype=float)
    ext2 = np.array([baseline, 0.0, 0.0, 0, 0, 0, 0, 0, 0, 0, 0, 0], dtype=float)
    ext2[0:3] += rng.normal(0.0, 0.02, 3)
    ext2[6:9] += rng.normal(0.0, math.radians(1.0), 3)
    ext1 = np.zeros(12)
    u0, v0 = width / 2, height / 2
    Z1, Z2, truth = [], [], []
    for k in range(n_frames):
        pts = p0 + vel * (k * dt)
        truth.append(pts)
        sets = []
        for ext in (ext1, ext2):
            u, v, d = project(pts, ext, f, u0, v0)
            vis = (u >= 0) & (u <= width) & (v >= 0) & (v <= height) & (d >= 0)
            det = vis & (rng.random(L) < pd)
            zu = u[det] + rng.normal(0.0, sigma, det.sum())
            zv = v[det] + rng.normal(0.0, sigma, det.sum())
            nc = rng.poisson(lam)
            cu = rng.uniform(0, width, nc)
            cv = rng.uniform(0, height, nc)
            rng.uniform(1, 64, nc)  # clutter disparity, recorded by the draw order only
            zz = np.vstack([np.concatenate([zu, cu]), np.concatenate([zv, cv])])
            perm = rng.permutation(zz.shape[1])
            sets.append(np.ascontiguousarray(zz[:, perm]))
        Z1.append(sets[0])
        Z2.append(sets[1])
    return dict(Z1=Z1, Z2=Z2, truth=truth, ext2=ext2, dt=dt, seed=seed, rig=rig)


def frame_measurements(scene, k, slam=False):
    """Measurement set and camera flag of frame k, as the reference main loop
    alternates cameras (src/main.cpp:83-91): even k -> camera 1 (fixed),
    odd k -> camera 2 (moving)."""
    if slam or k % 2 == 1:
        z = scene["Z2"][k]
        return z[0], z[1], False
    z = scene["Z1"][k]
    return z[0], z[1], True
