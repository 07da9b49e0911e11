"""Python front end of the gfx950 SC-PHD calibration filter.

Thin ctypes binding of the C ABI in include/scphd.h (libscphd_hip.so, built
in this directory by the Makefile).  The class Filter mirrors the reference
driver SCPHDCameraCalibration (cheesinglee/scphd-camera-calibration
src/scphdcameracalibration.cuh:29-122): predict(dt), update(u, v, fixed),
resample(), check() plus state read-back.  There is no CPU fallback: a
missing library raises ImportError.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# SCPHD_LIBRARY selects another build of the library (e.g. a variant under test)
LIB_PATH = os.environ.get("SCPHD_LIBRARY") or os.path.join(HERE, "libscphd_hip.so")

SCPHD_OK = 0
ERRORS = {-1: "SCPHD_ERR_ARG", -2: "SCPHD_ERR_HIP", -3: "SCPHD_ERR_NAN", -4: "SCPHD_ERR_CAPACITY",
          -5: "SCPHD_ERR_CONFIG", -6: "SCPHD_ERR_STATE", -7: "SCPHD_ERR_NODEVICE"}

_D = ctypes.c_double
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_U64 = ctypes.c_uint64
_P = ctypes.c_void_p

EXPORTED = [
    "scphd_params_init", "scphd_params_load", "scphd_create", "scphd_destroy", "scphd_last_error",
    "scphd_predict", "scphd_update", "scphd_update_device", "scphd_resample", "scphd_check", "scphd_sync",
    "scphd_get_counts", "scphd_get_parents", "scphd_get_maps", "scphd_set_state", "scphd_get_update_stats",
    "scphd_get_update_debug", "scphd_get_ancestors", "scphd_set_timing", "scphd_get_phase_ms",
    "scphd_get_fit_debug", "scphd_get_estimate", "scphd_get_memory", "scphd_set_noise",
    "scphd_debug_normals",
    "scphd_batch_opts_init", "scphd_batch_create", "scphd_batch_destroy", "scphd_batch_last_error",
    "scphd_batch_now", "scphd_rig_footprint", "scphd_batch_add_rig", "scphd_batch_remove_rig",
    "scphd_batch_submit", "scphd_batch_step", "scphd_batch_collect", "scphd_batch_get_stats",
    "scphd_batch_filter", "scphd_mat_read_measurements", "scphd_write_mat", "scphd_batch_set_output",
    "scphd_create_ensemble", "scphd_update_ensemble", "scphd_get_rig_estimates", "scphd_get_rig_resample",
    "scphd_reserve", "scphd_debug_set_limits", "scphd_resample_noalloc", "scphd_update_ensemble_rigs",
    "scphd_ensemble_add_rig", "scphd_ensemble_remove_rig", "scphd_compact",
]


class Params(ctypes.Structure):
    """scphd_params (include/scphd.h)."""

    _fields_ = [
        ("n_particles", _I32), ("particles_per_feature", _I32), ("max_gaussians", _I32),
        ("storage", _I32), ("verbosity", _I32), ("slam", _I32), ("max_steps", _I32),
        ("save_all_maps", _I32),
        ("min_weight", _D), ("min_separation", _D), ("min_neff", _D), ("ps", _D),
        ("prior_a", _D * 12), ("prior_b", _D * 12),
        ("cam_accel_std", _D * 6), ("feat_accel_std", _D * 3),
        ("fx", _D), ("fy", _D), ("u0", _D), ("v0", _D),
        ("std_u", _D), ("std_v", _D), ("pd", _D), ("lambda_", _D),
        ("birth_w0", _D), ("birth_d0", _D), ("birth_vu0", _D), ("birth_vv0", _D), ("birth_vd0", _D),
        ("birth_var_d0", _D), ("birth_var_vu0", _D), ("birth_var_vv0", _D), ("birth_var_vd0", _D),
        ("max_velocity", _D), ("min_range", _D),
        ("seed", _U64),
    ]

    @classmethod
    def from_dict(cls, d):
        p = cls()
        lib().scphd_params_init(ctypes.byref(p))
        for name, _ in cls._fields_:
            key = "lambda" if name == "lambda_" else name
            if key not in d:
                continue
            fld = getattr(p, name)
            if isinstance(fld, ctypes.Array):
                for i, x in enumerate(d[key]):
                    fld[i] = x
            else:
                setattr(p, name, d[key])
        return p

    def to_dict(self):
        out = {}
        for name, _ in self._fields_:
            key = "lambda" if name == "lambda_" else name
            v = getattr(self, name)
            out[key] = list(v) if isinstance(v, ctypes.Array) else v
        return out


class Counts(ctypes.Structure):
    _fields_ = [(n, _I64) for n in ("n_particles", "particles_per_feature", "n_gaussians", "n_samples",
                                    "n_predict", "n_update", "n_resample", "pending_predict")]


class UpdateStats(ctypes.Structure):
    _fields_ = [(n, _I64) for n in ("n_meas", "n_gaussians_in", "n_pairs", "n_survivors", "n_merged",
                                    "n_nondetect", "n_gaussians_out", "max_survivors_per_parent",
                                    "cap_exceeded")]


class BatchOpts(ctypes.Structure):
    """scphd_batch_opts (include/scphd.h)."""

    _fields_ = [("hbm_budget_bytes", _I64), ("max_rigs", _I32), ("n_workers", _I32)]


class FrameResult(ctypes.Structure):
    """scphd_frame_result (include/scphd.h)."""

    _fields_ = [("rig", _I32), ("status", _I32), ("frame_id", _I64), ("t_submit", _D), ("t_start", _D),
                ("t_end", _D), ("n_meas", _I64), ("n_pairs", _I64), ("n_gaussians", _I64), ("resampled", _I32),
                ("reserved0", _I32), ("estimate", _D * 12), ("cardinality", _D)]


class BatchStats(ctypes.Structure):
    _fields_ = [(n, _I64) for n in ("rigs", "reserved_bytes", "budget_bytes", "used_bytes", "queued_frames",
                                    "frames_done", "pairs_done", "refused")]


_lib = None


def lib():
    """Load libscphd_hip.so; raises ImportError when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `make -C {HERE}` (no CPU fallback exists)")
    L = ctypes.CDLL(LIB_PATH)
    dp = ctypes.POINTER(_D)
    ip = ctypes.POINTER(_I32)
    pp = ctypes.POINTER(Params)
    sig = {
        "scphd_params_init": (_I32, [pp]),
        "scphd_params_load": (_I32, [ctypes.c_char_p, pp, ctypes.c_char_p, ctypes.c_size_t]),
        "scphd_create": (_I32, [pp, ctypes.POINTER(_P)]),
        "scphd_destroy": (None, [_P]),
        "scphd_last_error": (ctypes.c_char_p, [_P]),
        "scphd_predict": (_I32, [_P, _D]),
        "scphd_update": (_I32, [_P, dp, dp, _I32, _I32]),
        "scphd_update_device": (_I32, [_P, _P, _P, _I32, _I32]),
        "scphd_resample": (_I32, [_P, ip]),
        "scphd_resample_noalloc": (_I32, [_P, ip]),
        "scphd_check": (_I32, [_P]),
        "scphd_sync": (_I32, [_P]),
        "scphd_get_counts": (_I32, [_P, ctypes.POINTER(Counts)]),
        "scphd_get_parents": (_I32, [_P, dp, dp]),
        "scphd_get_maps": (_I32, [_P, ip, dp, dp]),
        "scphd_set_state": (_I32, [_P, dp, dp, ip, _I64, dp, dp, _I64, _I64, _I64]),
        "scphd_get_update_stats": (_I32, [_P, ctypes.POINTER(UpdateStats)]),
        "scphd_get_update_debug": (_I32, [_P, dp, dp, ip, ip, ip]),
        "scphd_get_ancestors": (_I32, [_P, ip]),
        "scphd_set_timing": (_I32, [_P, _I32]),
        "scphd_get_phase_ms": (_I32, [_P, dp]),
        "scphd_get_fit_debug": (_I32, [_P, dp]),
        "scphd_debug_set_limits": (_I32, [_P, _I32, _I32, _I32, _I32, _I32]),
        "scphd_get_estimate": (_I32, [_P, dp, dp, dp]),
        "scphd_set_noise": (_I32, [_P, _I32, _U64, dp, ctypes.c_size_t]),
        "scphd_debug_normals": (_I32, [_U64, ctypes.c_uint32, _U64, ctypes.c_uint32, ctypes.c_uint32, _I32, dp]),
        "scphd_get_memory": (_I32, [_P, ctypes.POINTER(_I64)]),
        "scphd_batch_opts_init": (_I32, [ctypes.POINTER(BatchOpts)]),
        "scphd_batch_create": (_I32, [ctypes.POINTER(BatchOpts), ctypes.POINTER(_P)]),
        "scphd_batch_destroy": (None, [_P]),
        "scphd_batch_last_error": (ctypes.c_char_p, [_P]),
        "scphd_batch_now": (_D, [_P]),
        "scphd_rig_footprint": (_I64, [pp, _I32]),
        "scphd_batch_add_rig": (_I32, [_P, pp, _I64, ip]),
        "scphd_batch_remove_rig": (_I32, [_P, _I32]),
        "scphd_batch_submit": (_I32, [_P, _I32, _I64, _D, dp, dp, _I32, _I32]),
        "scphd_batch_step": (_I32, [_P, ip]),
        "scphd_batch_collect": (_I32, [_P, ctypes.POINTER(FrameResult), _I32, ip]),
        "scphd_batch_get_stats": (_I32, [_P, ctypes.POINTER(BatchStats)]),
        "scphd_batch_filter": (_P, [_P, _I32]),
        "scphd_mat_read_measurements": (_I32, [ctypes.c_char_p, ctypes.c_char_p, ip, ctypes.POINTER(_I64), ip, dp,
                                               dp]),
        "scphd_write_mat": (_I32, [_P, ctypes.c_char_p]),
        "scphd_batch_set_output": (_I32, [_P, _I32, ctypes.c_char_p]),
        "scphd_create_ensemble": (_I32, [pp, _I32, ctypes.POINTER(_P)]),
        "scphd_update_ensemble": (_I32, [_P, dp, dp, ip, _I32]),
        "scphd_update_ensemble_rigs": (_I32, [_P, dp, dp, ip, ip]),
        "scphd_ensemble_add_rig": (_I32, [_P, pp, ip]),
        "scphd_ensemble_remove_rig": (_I32, [_P, _I32]),
        "scphd_compact": (_I32, [_P]),
        "scphd_get_rig_estimates": (_I32, [_P, dp, dp, dp]),
        "scphd_get_rig_resample": (_I32, [_P, ip, dp]),
        "scphd_reserve": (_I32, [_P, _I64, _I32]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(_D))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(_I32))


class ScphdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


def read_measurements(path, varname):
    """Measurement sequence of a MAT-file cell array (scphd_mat_read_measurements):
    a list of (u, v) arrays, one per frame."""
    L = lib()
    n = _I32(0)
    tot = _I64(0)
    rc = L.scphd_mat_read_measurements(path.encode(), varname.encode(), ctypes.byref(n), ctypes.byref(tot), None,
                                       None, None)
    if rc != SCPHD_OK:
        raise ScphdError(rc, f"cannot read {varname} from {path}")
    cnt = np.zeros(max(n.value, 1), dtype=np.int32)
    u = np.zeros(max(tot.value, 1))
    v = np.zeros(max(tot.value, 1))
    rc = L.scphd_mat_read_measurements(path.encode(), varname.encode(), ctypes.byref(n), ctypes.byref(tot), _ip(cnt),
                                       _dp(u), _dp(v))
    if rc != SCPHD_OK:
        raise ScphdError(rc, f"cannot read {varname} from {path}")
    off = np.concatenate([[0], np.cumsum(cnt[:n.value])])
    return [(u[off[k]:off[k + 1]].copy(), v[off[k]:off[k + 1]].copy()) for k in range(n.value)]


def device_normals(seed, stream, step, parent, gauss, n_elem):
    """Six normals per element 0..n_elem-1 drawn on the GPU (scphd_debug_normals)."""
    out = np.zeros(6 * max(n_elem, 1))
    rc = lib().scphd_debug_normals(seed, stream, step, parent, gauss, n_elem, _dp(out))
    if rc != SCPHD_OK:
        raise ScphdError(rc, "scphd_debug_normals failed")
    return out[:6 * n_elem].reshape(n_elem, 6)


def load_params(path):
    """Parse a libconfig file into a parameter dict (scphd_params_load)."""
    p = Params()
    err = ctypes.create_string_buffer(512)
    rc = lib().scphd_params_load(path.encode(), ctypes.byref(p), err, 512)
    if rc != SCPHD_OK:
        raise ScphdError(rc, err.value.decode())
    return p.to_dict()


class Filter:
    """One calibration filter (one stereo rig) resident on the GPU."""

    def __init__(self, params):
        L = lib()
        self.params = Params.from_dict(params) if isinstance(params, dict) else params
        h = _P()
        rc = L.scphd_create(ctypes.byref(self.params), ctypes.byref(h))
        if rc != SCPHD_OK:
            raise ScphdError(rc, "scphd_create failed")
        self.h = h
        self.Np = self.params.n_particles
        self.S = self.params.particles_per_feature

    def close(self):
        h = getattr(self, "h", None)
        if h:
            lib().scphd_destroy(h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, rc, what):
        if rc != SCPHD_OK:
            msg = lib().scphd_last_error(self.h)
            raise ScphdError(rc, f"{what}: {msg.decode() if msg else ''}")

    def predict(self, dt):
        self._check(lib().scphd_predict(self.h, dt), "predict")

    def update(self, u, v, fixed):
        u = np.ascontiguousarray(u, dtype=np.float64)
        v = np.ascontiguousarray(v, dtype=np.float64)
        if u.shape != v.shape:
            raise ValueError("u and v differ in length")
        self._check(lib().scphd_update(self.h, _dp(u), _dp(v), len(u), int(bool(fixed))), "update")

    def update_device(self, u_ptr, v_ptr, n, fixed):
        self._check(lib().scphd_update_device(self.h, u_ptr, v_ptr, n, int(bool(fixed))), "update_device")

    def resample(self):
        r = _I32(0)
        self._check(lib().scphd_resample(self.h, ctypes.byref(r)), "resample")
        return bool(r.value)

    def resample_noalloc(self):
        """scphd_resample with no device allocation (static scratch)."""
        r = _I32(0)
        self._check(lib().scphd_resample_noalloc(self.h, ctypes.byref(r)), "resample_noalloc")
        return bool(r.value)

    def check(self):
        return lib().scphd_check(self.h)

    def sync(self):
        self._check(lib().scphd_sync(self.h), "sync")

    def counts(self):
        c = Counts()
        self._check(lib().scphd_get_counts(self.h, ctypes.byref(c)), "get_counts")
        return {n: getattr(c, n) for n, _ in Counts._fields_}

    def parents(self):
        st = np.zeros((self.Np, 12))
        w = np.zeros(self.Np)
        self._check(lib().scphd_get_parents(self.h, _dp(st), _dp(w)), "get_parents")
        return st, w

    def maps(self):
        self._check(lib().scphd_get_maps(self.h, None, None, None), "get_maps")  # materialise first
        c = self.counts()
        gc = np.zeros(self.Np, dtype=np.int32)
        wg = np.zeros(max(c["n_gaussians"], 1))
        smp = np.zeros((7, max(c["n_samples"], 1)))
        self._check(lib().scphd_get_maps(self.h, _ip(gc), _dp(wg), _dp(smp)), "get_maps")
        return gc, wg[:c["n_gaussians"]], smp[:, :c["n_samples"]]

    def state(self):
        st, w = self.parents()
        gc, wg, smp = self.maps()
        c = self.counts()
        return dict(states=st, weights=w, gcount=gc, wg=wg, samples=np.ascontiguousarray(smp),
                    n_predict=c["n_predict"], n_update=c["n_update"], n_resample=c["n_resample"])

    def set_state(self, s):
        st = np.ascontiguousarray(s["states"], dtype=np.float64)
        w = np.ascontiguousarray(s["weights"], dtype=np.float64)
        gc = np.ascontiguousarray(s["gcount"], dtype=np.int32)
        wg = np.ascontiguousarray(s["wg"], dtype=np.float64)
        smp = np.ascontiguousarray(s["samples"], dtype=np.float64)
        n = int(gc.sum())
        if wg.size == 0:
            wg = np.zeros(1)
        if smp.size == 0:
            smp = np.zeros((7, 1))
        self._check(lib().scphd_set_state(self.h, _dp(st), _dp(w), _ip(gc), n, _dp(wg), _dp(smp),
                                          int(s["n_predict"]), int(s["n_update"]), int(s["n_resample"])),
                    "set_state")

    def update_stats(self):
        s = UpdateStats()
        self._check(lib().scphd_get_update_stats(self.h, ctypes.byref(s)), "get_update_stats")
        return {n: getattr(s, n) for n, _ in UpdateStats._fields_}

    def update_debug(self):
        M = self.update_stats()["n_meas"]
        eta = np.zeros(max(self.Np * M, 1))
        ll = np.zeros(self.Np)
        nm = np.zeros(self.Np, dtype=np.int32)
        nn = np.zeros(self.Np, dtype=np.int32)
        ns = np.zeros(self.Np, dtype=np.int32)
        self._check(lib().scphd_get_update_debug(self.h, _dp(eta), _dp(ll), _ip(nm), _ip(nn), _ip(ns)),
                    "get_update_debug")
        return dict(eta=eta[:self.Np * M].reshape(self.Np, M), loglik=ll, nmerged=nm, nnondetect=nn,
                    nsurvivors=ns)

    def fit_debug(self):
        n = self.update_stats()["n_gaussians_in"]
        out = np.zeros(28 * max(n, 1))
        self._check(lib().scphd_get_fit_debug(self.h, _dp(out)), "get_fit_debug")
        return out[:28 * n].reshape(28, n)

    def set_debug_limits(self, merge_lds_cap=-1, eta_heavy_cap=-1, eta_ncand_cap=-1, merge_force_rounds=False,
                         merge_all_pairs=False):
        """Lower the capacities at which the update takes its general paths
        (scphd_debug_set_limits); -1 keeps the normal limit."""
        self._check(lib().scphd_debug_set_limits(self.h, int(merge_lds_cap), int(eta_heavy_cap), int(eta_ncand_cap),
                                                 int(bool(merge_force_rounds)), int(bool(merge_all_pairs))),
                    "debug_set_limits")

    def ancestors(self):
        a = np.zeros(self.Np, dtype=np.int32)
        self._check(lib().scphd_get_ancestors(self.h, _ip(a)), "get_ancestors")
        return a

    def set_timing(self, on):
        self._check(lib().scphd_set_timing(self.h, int(bool(on))), "set_timing")

    def phase_ms(self):
        out = np.zeros(8)
        self._check(lib().scphd_get_phase_ms(self.h, _dp(out)), "get_phase_ms")
        return out

    def estimate(self):
        """Weighted parent mean (12), covariance (12 x 12) and the highest-weight
        parent's cardinality (scphd_get_estimate)."""
        m = np.zeros(12)
        c = np.zeros((12, 12))
        card = np.zeros(1)
        self._check(lib().scphd_get_estimate(self.h, _dp(m), _dp(c), _dp(card)), "get_estimate")
        return m, c, float(card[0])

    def set_noise(self, seed, mode=0):
        """Re-key the Philox noise streams (scphd_set_noise)."""
        self._check(lib().scphd_set_noise(self.h, int(mode), int(seed) & 0xFFFFFFFFFFFFFFFF, None, 0), "set_noise")

    def write_mat(self, path):
        """writeMat: parents and the highest-weight parent's map (scphd_write_mat)."""
        self._check(lib().scphd_write_mat(self.h, path.encode()), "write_mat")

    def reserve(self, n_gaussians, max_meas):
        """Pre-size the device buffers (scphd_reserve)."""
        self._check(lib().scphd_reserve(self.h, int(n_gaussians), int(max_meas)), "reserve")

    def compact(self):
        """Hand back device memory beyond the map (scphd_compact), between frames."""
        self._check(lib().scphd_compact(self.h), "compact")

    def memory(self):
        b = _I64(0)
        self._check(lib().scphd_get_memory(self.h, ctypes.byref(b)), "get_memory")
        return b.value


class Ensemble(Filter):
    """Rigs stepped by the same kernel launches (scphd_create_ensemble).
    Rig r evolves as Filter(rigs[r]) fed the same measurements; its parents
    follow those of rigs 0 .. r-1 in parents() and maps()."""

    def __init__(self, rigs):
        L = lib()
        self.rig_params = [Params.from_dict(p) if isinstance(p, dict) else p for p in rigs]
        arr = (Params * len(rigs))(*self.rig_params)
        h = _P()
        rc = L.scphd_create_ensemble(arr, len(rigs), ctypes.byref(h))
        if rc != SCPHD_OK:
            raise ScphdError(rc, "scphd_create_ensemble failed")
        self.h = h
        self.params = self.rig_params[0]
        self.n_rigs = len(rigs)
        self.rig_np = [p.n_particles for p in self.rig_params]
        self.poff = np.concatenate([[0], np.cumsum(self.rig_np)]).astype(int)
        self.Np = int(self.poff[-1])
        self.S = self.params.particles_per_feature

    def _set_rigs(self):
        self.n_rigs = len(self.rig_params)
        self.rig_np = [p.n_particles for p in self.rig_params]
        self.poff = np.concatenate([[0], np.cumsum(self.rig_np)]).astype(int)
        self.Np = int(self.poff[-1])

    def add_rig(self, params):
        """Append a rig between frames (scphd_ensemble_add_rig); returns its index."""
        p = Params.from_dict(params) if isinstance(params, dict) else params
        out = np.zeros(1, dtype=np.int32)
        self._check(lib().scphd_ensemble_add_rig(self.h, ctypes.byref(p), _ip(out)), "ensemble_add_rig")
        self.rig_params.append(p)
        self._set_rigs()
        return int(out[0])

    def remove_rig(self, rig):
        """Drop rig `rig` between frames (scphd_ensemble_remove_rig); later rigs move down one index."""
        self._check(lib().scphd_ensemble_remove_rig(self.h, int(rig)), "ensemble_remove_rig")
        del self.rig_params[rig]
        self._set_rigs()

    def update(self, us, vs, fixed):
        """us[r], vs[r]: rig r's measurement set; fixed: one flag for every
        rig, or a sequence of one per rig (scphd_update_ensemble_rigs)."""
        if len(us) != self.n_rigs or len(vs) != self.n_rigs:
            raise ValueError("one measurement set per rig")
        moff = np.zeros(self.n_rigs + 1, dtype=np.int32)
        moff[1:] = np.cumsum([len(x) for x in us])
        u = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float64) for x in us] + [np.zeros(1)]))
        v = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float64) for x in vs] + [np.zeros(1)]))
        if any(len(a) != len(b) for a, b in zip(us, vs)):
            raise ValueError("u and v differ in length")
        if np.ndim(fixed) == 0:
            self._check(lib().scphd_update_ensemble(self.h, _dp(u), _dp(v), _ip(moff), int(bool(fixed))),
                        "update_ensemble")
        else:
            fx = np.ascontiguousarray([int(bool(x)) for x in fixed], dtype=np.int32)
            if len(fx) != self.n_rigs:
                raise ValueError("one fixed-camera flag per rig")
            self._check(lib().scphd_update_ensemble_rigs(self.h, _dp(u), _dp(v), _ip(moff), _ip(fx)),
                        "update_ensemble_rigs")
        self._mstride = max(len(x) for x in us) if us else 0

    def update_debug(self):
        """As Filter.update_debug; eta rows have the largest rig's measurement
        count (a rig with fewer uses the first moff[r+1]-moff[r] columns)."""
        M = getattr(self, "_mstride", 0)
        eta = np.zeros(max(self.Np * M, 1))
        ll = np.zeros(self.Np)
        nm = np.zeros(self.Np, dtype=np.int32)
        nn = np.zeros(self.Np, dtype=np.int32)
        ns = np.zeros(self.Np, dtype=np.int32)
        self._check(lib().scphd_get_update_debug(self.h, _dp(eta), _dp(ll), _ip(nm), _ip(nn), _ip(ns)),
                    "get_update_debug")
        return dict(eta=eta[:self.Np * M].reshape(self.Np, M), loglik=ll, nmerged=nm, nnondetect=nn,
                    nsurvivors=ns)

    def rig_estimates(self):
        """Per rig: mean (n_rigs, 12), covariance (n_rigs, 12, 12), cardinality (n_rigs)."""
        m = np.zeros((self.n_rigs, 12))
        c = np.zeros((self.n_rigs, 12, 12))
        card = np.zeros(self.n_rigs)
        self._check(lib().scphd_get_rig_estimates(self.h, _dp(m), _dp(c), _dp(card)), "get_rig_estimates")
        return m, c, card

    def rig_resample(self):
        """Per rig: resampled at the last resample() call, and n_eff."""
        r = np.zeros(self.n_rigs, dtype=np.int32)
        ne = np.zeros(self.n_rigs)
        self._check(lib().scphd_get_rig_resample(self.h, _ip(r), _dp(ne)), "get_rig_resample")
        return r.astype(bool), ne


def rig_footprint(params, gaussians_per_parent):
    """Device bytes a rig needs at about gaussians_per_parent Gaussians per parent."""
    p = Params.from_dict(params) if isinstance(params, dict) else params
    return lib().scphd_rig_footprint(ctypes.byref(p), int(gaussians_per_parent))


class _Borrowed(Filter):
    """A rig's filter handle owned by a Batch (read-back only; never destroyed here)."""

    def __init__(self, h, params):
        self.params = Params.from_dict(params) if isinstance(params, dict) else params
        self.h = h
        self.Np = self.params.n_particles
        self.S = self.params.particles_per_feature

    def close(self):
        self.h = None


class Batch:
    """Many rigs on one GPU (scphd_batch_*): admission by HBM reservation,
    per-rig frame queues, frames of different rigs overlapping on the device."""

    def __init__(self, hbm_budget_bytes=0, max_rigs=0, n_workers=0):
        L = lib()
        o = BatchOpts()
        L.scphd_batch_opts_init(ctypes.byref(o))
        o.hbm_budget_bytes = int(hbm_budget_bytes)
        if max_rigs:
            o.max_rigs = int(max_rigs)
        if n_workers:
            o.n_workers = int(n_workers)
        h = _P()
        rc = L.scphd_batch_create(ctypes.byref(o), ctypes.byref(h))
        if rc != SCPHD_OK:
            raise ScphdError(rc, "scphd_batch_create failed")
        self.h = h
        self.params = {}

    def close(self):
        h = getattr(self, "h", None)
        if h:
            lib().scphd_batch_destroy(h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, rc, what):
        if rc != SCPHD_OK:
            msg = lib().scphd_batch_last_error(self.h)
            raise ScphdError(rc, f"{what}: {msg.decode() if msg else ''}")

    def now(self):
        return lib().scphd_batch_now(self.h)

    def add_rig(self, params, reserve_bytes=0):
        """Admit a rig; returns its id, or None when the HBM budget or the slots
        are exhausted (offer it again later)."""
        p = Params.from_dict(params) if isinstance(params, dict) else params
        r = _I32(-1)
        rc = lib().scphd_batch_add_rig(self.h, ctypes.byref(p), int(reserve_bytes), ctypes.byref(r))
        if rc == -4:
            return None
        self._check(rc, "add_rig")
        self.params[r.value] = p
        return r.value

    def remove_rig(self, rig):
        self._check(lib().scphd_batch_remove_rig(self.h, int(rig)), "remove_rig")
        self.params.pop(rig, None)

    def set_output(self, rig, directory):
        """Write <directory>/<frame_id>.mat after each of the rig's frames (None stops)."""
        d = directory.encode() if directory else None
        self._check(lib().scphd_batch_set_output(self.h, int(rig), d), "set_output")

    def submit(self, rig, frame_id, dt, u, v, fixed):
        u = np.ascontiguousarray(u, dtype=np.float64)
        v = np.ascontiguousarray(v, dtype=np.float64)
        if u.shape != v.shape:
            raise ValueError("u and v differ in length")
        self._check(lib().scphd_batch_submit(self.h, int(rig), int(frame_id), float(dt), _dp(u), _dp(v), len(u),
                                             int(bool(fixed))), "submit")

    def step(self):
        """Wait for every submitted frame; returns the number finished."""
        n = _I32(0)
        self._check(lib().scphd_batch_step(self.h, ctypes.byref(n)), "step")
        return n.value

    def collect(self, cap=4096):
        """Finished-frame results, oldest first, as a list of dicts."""
        out = []
        buf = (FrameResult * cap)()
        while True:
            n = _I32(0)
            self._check(lib().scphd_batch_collect(self.h, buf, cap, ctypes.byref(n)), "collect")
            for i in range(n.value):
                r = buf[i]
                d = {name: getattr(r, name) for name, _ in FrameResult._fields_ if name != "reserved0"}
                d["estimate"] = np.array(list(r.estimate))
                out.append(d)
            if n.value < cap:
                return out

    def stats(self):
        s = BatchStats()
        self._check(lib().scphd_batch_get_stats(self.h, ctypes.byref(s)), "get_stats")
        return {n: getattr(s, n) for n, _ in BatchStats._fields_}

    def snapshot(self, rig):
        """Evict a rig: its complete state (parents, maps, noise counters) as
        host arrays, then the rig is removed and its reservation freed.  Call
        while none of its frames is queued or running."""
        state = self.filter(rig).state()
        params = self.params[rig]
        self.remove_rig(rig)
        return dict(params=params, state=state)

    def resume(self, snap, reserve_bytes=0):
        """Re-admit an evicted rig with its state; returns the new rig id, or
        None when the budget does not admit it yet.  The rig continues
        exactly where it stopped."""
        rig = self.add_rig(snap["params"], reserve_bytes)
        if rig is not None:
            self.filter(rig).set_state(snap["state"])
        return rig

    def filter(self, rig):
        """The rig's filter for read-back while none of its frames is queued or running."""
        h = lib().scphd_batch_filter(self.h, int(rig))
        if not h:
            raise ScphdError(-1, f"unknown rig {rig}")
        return _Borrowed(_P(h), self.params[rig])
