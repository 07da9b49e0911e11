#!/usr/bin/env python3
"""Writes tests/golden/reference_sequences.npz from the reference's input
sequences (cheesinglee/scphd-camera-calibration data/*.mat): for each file,
cells Z1 and Z2 (2 x n_k matrices of image measurements u, v per frame, read
by readMatfile, src/main.cpp:19-49), the first FRAMES frames (the reference
configuration's max_steps, config/calibration.cfg).

Stored per file and variable: <file>_<var>_counts (int32, one per frame) and
<file>_<var>_uv (float64, 2 x total, frames concatenated).  Read with
scipy.io.loadmat, which parses the file as data.

    python tests/golden/make_reference_sequences.py [/root/reference/data]
"""
import os
import sys

import numpy as np
import scipy.io

FRAMES = 100
FILES = ["disparity_sim", "disparity_sim_noclutter", "fov_test", "spiral"]


def main(src):
    out = {}
    for name in FILES:
        m = scipy.io.loadmat(os.path.join(src, name + ".mat"))
        for var in ("Z1", "Z2"):
            cells = m[var].ravel()[:FRAMES]
            mats = [np.asarray(c, dtype=np.float64).reshape(2, -1) if np.size(c) else np.zeros((2, 0))
                    for c in cells]
            out[f"{name}_{var}_counts"] = np.array([a.shape[1] for a in mats], dtype=np.int32)
            out[f"{name}_{var}_uv"] = np.concatenate(mats, axis=1)
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_sequences.npz")
    np.savez_compressed(dst, **out)
    print(f"wrote {dst}: {len(FILES)} files x 2 cells x {FRAMES} frames")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/data")
