"""ctypes front end of the FP64 CPU oracle (oracle/scphd_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product path.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libscphd_oracle.so")

_D = ctypes.c_double
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_U32 = ctypes.c_uint32
_U64 = ctypes.c_uint64
_P = ctypes.c_void_p


class OracleParams(ctypes.Structure):
    """Mirror of scphd_params (include/scphd.h)."""

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


def params_from_dict(d):
    p = OracleParams()
    for name, _ in OracleParams._fields_:
        key = "lambda" if name == "lambda_" else name
        if key not in d:
            continue
        val = d[key]
        fld = getattr(p, name)
        if isinstance(fld, ctypes.Array):
            for i, x in enumerate(val):
                fld[i] = x
        else:
            setattr(p, name, val)
    return p


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = ctypes.CDLL(LIB_PATH)
    dp = ctypes.POINTER(_D)
    ip = ctypes.POINTER(_I32)
    lp = ctypes.POINTER(_I64)
    up = ctypes.POINTER(_U32)
    pp = ctypes.POINTER(OracleParams)
    sig = {
        "orc_params_size": (_I32, []),
        "orc_create": (_P, [pp]),
        "orc_destroy": (None, [_P]),
        "orc_set_threads": (_I32, [_I32]),
        "orc_predict": (_I32, [_P, _D]),
        "orc_update": (_I32, [_P, dp, dp, _I32, _I32]),
        "orc_resample": (_I32, [_P, ip]),
        "orc_check": (_I32, [_P]),
        "orc_get_counts": (None, [_P, lp]),
        "orc_get_parents": (None, [_P, dp, dp]),
        "orc_get_maps": (None, [_P, ip, dp, dp]),
        "orc_set_state": (_I32, [_P, dp, dp, ip, _I64, dp, dp, _I64, _I64, _I64]),
        "orc_get_update_debug": (None, [_P, dp, dp, ip, ip, ip]),
        "orc_get_ancestors": (None, [_P, ip, dp]),
        "orc_get_fit_debug": (None, [_P, dp]),
        "orc_philox": (None, [up, up, up]),
        "orc_normals": (None, [_U64, _U32, _U64, _U32, _U32, _U32, _I32, dp]),
        "orc_stream_normals": (None, [_U64, _U32, _U64, _U32, _U32, _U32, dp]),
        "orc_box_muller": (None, [_U32, _U32, dp]),
        "orc_wrap_angle": (_D, [_D]),
        "orc_safe_log": (_D, [_D]),
        "orc_lse2": (_D, [_D, _D]),
        "orc_rotation": (None, [dp, dp]),
        "orc_world_to_disparity": (_D, [pp, dp, dp, dp]),
        "orc_disparity_to_world": (None, [pp, dp, dp, dp]),
        "orc_cholesky": (None, [dp, dp]),
        "orc_is_pos_def": (_I32, [dp]),
        "orc_mahal": (_D, [dp, dp, dp, dp]),
        "orc_kalman": (_I32, [pp, _D, _D, dp, dp, _D, _D, dp]),
        "orc_merge": (_I32, [pp, _I32, dp, dp, dp, dp, dp, dp]),
        "orc_systematic": (None, [dp, _I32, _D, ip]),
        "orc_fit": (None, [_I32, dp, dp, dp, dp]),
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


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Oracle:
    """FP64 oracle filter with the same operations as the GPU handle."""

    def __init__(self, params: dict, threads: int = 0):
        L = lib()
        self.params = params_from_dict(params)
        if threads:
            L.orc_set_threads(threads)
        self.h = L.orc_create(ctypes.byref(self.params))
        if not self.h:
            raise ValueError("oracle: invalid parameters")
        self.Np = params["n_particles"]
        self.S = params["particles_per_feature"]

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            lib().orc_destroy(h)
            self.h = None

    def predict(self, dt):
        lib().orc_predict(self.h, dt)

    def update(self, u, v, fixed):
        u = f64(u)
        v = f64(v)
        rc = lib().orc_update(self.h, _dp(u), _dp(v), len(u), int(bool(fixed)))
        if rc != 0:
            raise RuntimeError(f"oracle update failed rc={rc}")

    def resample(self):
        r = _I32(0)
        lib().orc_resample(self.h, ctypes.byref(r))
        return bool(r.value)

    def check(self):
        return lib().orc_check(self.h)

    def counts(self):
        out = np.zeros(8, dtype=np.int64)
        lib().orc_get_counts(self.h, out.ctypes.data_as(ctypes.POINTER(_I64)))
        keys = ["n_particles", "particles_per_feature", "n_gaussians", "n_samples", "n_predict",
                "n_update", "n_resample", "last_M"]
        return dict(zip(keys, (int(x) for x in out)))

    def parents(self):
        st = np.zeros((self.Np, 12))
        w = np.zeros(self.Np)
        lib().orc_get_parents(self.h, _dp(st), _dp(w))
        return st, w

    def maps(self):
        c = self.counts()
        gc = np.zeros(self.Np, dtype=np.int32)
        wg = np.zeros(c["n_gaussians"])
        smp = np.zeros((7, c["n_samples"]))
        lib().orc_get_maps(self.h, _ip(gc), _dp(wg), _dp(smp))
        return gc, wg, smp

    def state(self):
        st, w = self.parents()
        gc, wg, smp = self.maps()
        c = self.counts()
        return dict(states=st, weights=w, gcount=gc, wg=wg, samples=smp,
                    n_predict=c["n_predict"], n_update=c["n_update"], n_resample=c["n_resample"])

    def set_state(self, s):
        st = f64(s["states"])
        w = f64(s["weights"])
        gc = np.ascontiguousarray(s["gcount"], dtype=np.int32)
        wg = f64(s["wg"])
        smp = f64(s["samples"])
        rc = lib().orc_set_state(self.h, _dp(st), _dp(w), _ip(gc), int(gc.sum()), _dp(wg), _dp(smp),
                                 int(s["n_predict"]), int(s["n_update"]), int(s["n_resample"]))
        if rc != 0:
            raise ValueError("oracle set_state: bad sizes")

    def update_debug(self):
        M = self.counts()["last_M"]
        eta = np.zeros(self.Np * M)
        ll = np.zeros(self.Np)
        nm = np.zeros(self.Np, dtype=np.int32)
        nn = np.zeros(self.Np, dtype=np.int32)
        ns = np.zeros(self.Np, dtype=np.int32)
        lib().orc_get_update_debug(self.h, _dp(eta), _dp(ll), _ip(nm), _ip(nn), _ip(ns))
        return dict(eta=eta.reshape(self.Np, M), loglik=ll, nmerged=nm, nnondetect=nn, nsurvivors=ns)

    def fit_debug(self, n):
        out = np.zeros(28 * max(n, 1))
        lib().orc_get_fit_debug(self.h, _dp(out))
        return out[:28 * n].reshape(28, n)

    def ancestors(self):
        a = np.zeros(self.Np, dtype=np.int32)
        neff = _D(0)
        lib().orc_get_ancestors(self.h, _ip(a), ctypes.byref(neff))
        return a, neff.value


# ---- unit-level helpers ---------------------------------------------------

def philox(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    P = ctypes.POINTER(_U32)
    lib().orc_philox(c.ctypes.data_as(P), k.ctypes.data_as(P), o.ctypes.data_as(P))
    return o


def box_muller(x0, x1):
    z = np.zeros(2)
    lib().orc_box_muller(x0, x1, _dp(z))
    return z


def normals(seed, stream, step, parent, gauss, elem, n):
    z = np.zeros(n)
    lib().orc_normals(seed, stream, step, parent, gauss, elem, n, _dp(z))
    return z


def stream_normals(seed, stream, step, parent, gauss, s):
    """The normals of sample s of stream 2 (three: landmark prediction) or 3
    (six: posterior sampling), as the filter draws them."""
    z = np.zeros(6)
    lib().orc_stream_normals(seed, stream, step, parent, gauss, s, _dp(z))
    return z[:3] if stream == 2 else z


def rotation(e):
    e = f64(e)
    R = np.zeros(9)
    lib().orc_rotation(_dp(e), _dp(R))
    return R.reshape(3, 3)


def world_to_disparity(params, e, pw):
    p = params_from_dict(params)
    e = f64(e)
    pw = f64(pw)
    q = np.zeros(6)
    pd = lib().orc_world_to_disparity(ctypes.byref(p), _dp(e), _dp(pw), _dp(q))
    return q, pd


def disparity_to_world(params, e, q):
    p = params_from_dict(params)
    e = f64(e)
    q = f64(q)
    pw = np.zeros(6)
    lib().orc_disparity_to_world(ctypes.byref(p), _dp(e), _dp(q), _dp(pw))
    return pw


def cholesky(A):
    A = f64(np.asarray(A).T.reshape(36))  # column-major
    L = np.zeros(36)
    lib().orc_cholesky(_dp(A), _dp(L))
    return L.reshape(6, 6).T


def is_pos_def(A):
    A = f64(np.asarray(A).T.reshape(36))
    return bool(lib().orc_is_pos_def(_dp(A)))


def mahal(ma, Pa, mb, Pb):
    return lib().orc_mahal(_dp(f64(ma)), _dp(f64(np.asarray(Pa).T.reshape(36))),
                           _dp(f64(mb)), _dp(f64(np.asarray(Pb).T.reshape(36))))


def kalman(params, w, pd, m, P, zu, zv):
    p = params_from_dict(params)
    out = np.zeros(43)
    ok = lib().orc_kalman(ctypes.byref(p), w, pd, _dp(f64(m)), _dp(f64(np.asarray(P).T.reshape(36))),
                          zu, zv, _dp(out))
    return out[0], out[1:7], out[7:].reshape(6, 6).T, bool(ok)


def merge(params, w, m, P):
    p = params_from_dict(params)
    n = len(w)
    Pc = f64(np.stack([np.asarray(x).T.reshape(36) for x in P])) if n else np.zeros(0)
    ow = np.zeros(max(n, 1))
    om = np.zeros(max(n, 1) * 6)
    oP = np.zeros(max(n, 1) * 36)
    k = lib().orc_merge(ctypes.byref(p), n, _dp(f64(w)), _dp(f64(np.asarray(m).reshape(-1))), _dp(Pc),
                        _dp(ow), _dp(om), _dp(oP))
    return ow[:k], om[:6 * k].reshape(k, 6), np.stack([oP[36 * i:36 * i + 36].reshape(6, 6).T for i in range(k)]) if k else np.zeros((0, 6, 6))


def systematic(w, u):
    w = f64(w)
    a = np.zeros(len(w), dtype=np.int32)
    lib().orc_systematic(_dp(w), len(w), u, _ip(a))
    return a


def fit(q, w):
    q = f64(q)
    w = f64(w)
    mean = np.zeros(6)
    cov = np.zeros(36)
    lib().orc_fit(len(w), _dp(q), _dp(w), _dp(mean), _dp(cov))
    return mean, cov.reshape(6, 6).T


def wrap_angle(a):
    return lib().orc_wrap_angle(a)


def safe_log(x):
    return lib().orc_safe_log(x)


def lse2(a, b):
    return lib().orc_lse2(a, b)
