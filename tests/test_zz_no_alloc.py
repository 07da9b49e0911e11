"""scphd_resample_noalloc, the resampling entry point that allocates no
device memory (its scratch is a static device array shared by every call
on the GPU), against scphd_resample through the C ABI: two handles stepped
identically, one resampled by each entry point, must agree bit for bit --
the decision, the ancestors, the parents, the weights and the maps -- on
calls in a row, the second run with fewer parents than the first (the
scratch then holds the larger call's contents)."""
import ctypes

import numpy as np
import pytest

import scene
import workloads

pytestmark = pytest.mark.gpu


def _run_pair(n_particles, seed, frames, min_neff):
    import scphd_amd

    p, s = workloads.workload("c0", n_particles=n_particles, particles_per_feature=64, min_neff=min_neff)
    p["seed"] = seed
    sc = scene.generate(seed=seed, **s)
    a, b = scphd_amd.Filter(p), scphd_amd.Filter(p)
    for f in (a, b):
        f.reserve(n_particles * 800, 256)
    decisions = []
    for k in range(frames):
        u, v, fixed = scene.frame_measurements(sc, k)
        for f in (a, b):
            f.predict(s["dt"])
            f.update(u, v, fixed)
        ra = a.resample()
        rb = b.resample_noalloc()
        assert ra == rb, f"frame {k}: resampling decision"
        decisions.append(ra)
        np.testing.assert_array_equal(b.ancestors(), a.ancestors(), err_msg=f"frame {k} ancestors")
        sa, wa = a.parents()
        sb, wb = b.parents()
        np.testing.assert_array_equal(sb, sa, err_msg=f"frame {k} parents")
        np.testing.assert_array_equal(wb, wa, err_msg=f"frame {k} weights")
        ga, wga, xa = a.maps()
        gb, wgb, xb = b.maps()
        np.testing.assert_array_equal(gb, ga, err_msg=f"frame {k} map sizes")
        np.testing.assert_array_equal(wgb, wga, err_msg=f"frame {k} map weights")
        np.testing.assert_array_equal(xb, xa, err_msg=f"frame {k} samples")
    a.close()
    b.close()
    return decisions


def test_noalloc_resampling_equals_resample_on_calls_in_a_row():
    # min_neff = 1.01: every frame resamples
    assert all(_run_pair(48, 11, 3, 1.01))
    # the second run, with fewer parents, reuses the scratch the first left
    assert all(_run_pair(16, 5, 2, 1.01))
    # min_neff = 0: no frame resamples (identity ancestors, weights kept)
    assert not any(_run_pair(16, 6, 2, 0.0))


def test_noalloc_resampling_refusals():
    import scphd_amd

    L = scphd_amd.lib()
    r = ctypes.c_int32(0)
    # more parents than the static scratch holds
    p, _ = workloads.workload("c0", n_particles=80000, particles_per_feature=8)
    f = scphd_amd.Filter(p)
    assert L.scphd_resample_noalloc(f.h, ctypes.byref(r)) == -4  # SCPHD_ERR_CAPACITY
    assert b"static scratch" in L.scphd_last_error(f.h)
    f.close()
    # ensembles are refused
    rigs = [workloads.rig(seed=1, n_particles=8, n_landmarks=10, lam=5.0)[0],
            workloads.rig(seed=2, n_particles=8, n_landmarks=10, lam=5.0)[0]]
    e = scphd_amd.Ensemble(rigs)
    assert L.scphd_resample_noalloc(e.h, ctypes.byref(r)) == -1  # SCPHD_ERR_ARG
    e.close()
