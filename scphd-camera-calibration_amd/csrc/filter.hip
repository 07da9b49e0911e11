/*
 * filter.hip -- C ABI of the SC-PHD calibration filter (include/scphd.h):
 * device state ownership and the host-side sequencing of the kernels of
 * one filter step on the handle's stream.
 *
 * The step follows SCPHDCameraCalibration (cheesinglee/scphd-camera-calibration
 * src/scphdcameracalibration.cu): predict :1034-1151, update :1576-2094,
 * resample :2651-2817.  Host<->device traffic per update is the measurement
 * set and two small count read-backs (per-parent survivor offsets and the
 * output totals) used to size buffers.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "devmem.h"
#include "kernels.h"
#include "scphd.h"

using namespace scphd;

namespace {

/* Device buffers come from the process-wide block cache (devmem.cpp),
 * which enforces the allocation floor.  A block a handle replaces during a
 * call is retired to the handle and reaches the cache only once the
 * handle's streams have drained, so another handle's stream never receives
 * a block that queued work still uses. */

/* a released block: to the owner's retired list while its streams may still
 * use it, else straight to the cache */
void retire(std::vector<Block>* retired, void* p, size_t bytes) {
    if (!p) return;
    if (retired)
        retired->push_back(Block{p, bytes});
    else
        cache_put(Block{p, bytes});
}

template <typename T>
struct DBuf {
    T* p = nullptr;
    int64_t n = 0;
    size_t bytes = 0;
    std::vector<Block>* retired = nullptr; /* the owning handle's list */
    hipError_t ensure(int64_t need) {
        if (need <= n) return hipSuccess;
        release();
        const int64_t cap = std::max<int64_t>(need + need / 4, 64);
        size_t b = (size_t)cap * sizeof(T);
        void* q = nullptr;
        hipError_t e = cache_alloc(&q, &b);
        if (e != hipSuccess) return e;
        p = (T*)q;
        bytes = b;
        n = (int64_t)(b / sizeof(T));
        return hipSuccess;
    }
    void release() {
        retire(retired, p, bytes);
        p = nullptr;
        n = 0;
        bytes = 0;
    }
};

/* Grow b to at least need elements keeping its first keep (at most its
 * current size); the old block is retired to the owner. */
template <typename T>
hipError_t grow_keep(DBuf<T>& b, int64_t need, int64_t keep, hipStream_t s) {
    if (need <= b.n) return hipSuccess;
    DBuf<T> nb;
    nb.retired = b.retired;
    hipError_t e = nb.ensure(need);
    if (e != hipSuccess) return e;
    keep = std::min<int64_t>(keep, b.n);
    if (keep > 0) {
        e = hipMemcpyAsync(nb.p, b.p, (size_t)keep * sizeof(T), hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) {
            nb.release();
            return e;
        }
    }
    b.release();
    b.p = nb.p;
    b.n = nb.n;
    b.bytes = nb.bytes;
    return hipSuccess;
}

/* One generation of sample blocks: 7 planes of float or double. */
struct Generation {
    void* pl[7] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    size_t pbytes[7] = {};
    int64_t cap_blocks = 0;
    std::vector<Block>* retired = nullptr; /* the owning handle's list */
    void release() {
        for (int c = 0; c < 7; ++c) {
            retire(retired, pl[c], pbytes[c]);
            pl[c] = nullptr;
            pbytes[c] = 0;
        }
        cap_blocks = 0;
    }
    /* grow to at least need blocks keeping the first keep blocks.  Planes
     * are replaced one at a time (allocate, copy the kept blocks, wait, hand
     * the old plane back), so the transient is one plane, not a whole
     * generation; if an allocation fails, every plane still holds the kept
     * blocks and at least the old capacity, which stays the recorded one.
     * Keeping blocks happens only while no other stream works on the
     * generation, so after the wait an old plane goes straight to the cache.
     * drained: no queued work touches the generation, so with nothing to
     * keep its planes go straight to the cache too, where the allocations
     * that follow can reuse or free them. */
    hipError_t ensure(int64_t need, int64_t keep, int S, size_t esz, hipStream_t s, bool drained = false) {
        if (need <= cap_blocks) return hipSuccess;
        const int64_t cap = std::max<int64_t>(need + need / 4, 16);
        if (keep <= 0 && drained) {
            for (int c = 0; c < 7; ++c) {
                if (pl[c]) cache_put(Block{pl[c], pbytes[c]});
                pl[c] = nullptr;
                pbytes[c] = 0;
            }
            cap_blocks = 0;
        } else if (keep <= 0) {
            release(); /* nothing to keep: release first, then allocate */
        }
        int64_t got = INT64_MAX;
        for (int c = 0; c < 7; ++c) {
            void* np = nullptr;
            size_t b = (size_t)cap * S * esz;
            hipError_t e = cache_alloc(&np, &b);
            if (e != hipSuccess) {
                if (keep <= 0) release();
                return e;
            }
            if (keep > 0 && pl[c]) {
                e = hipMemcpyAsync(np, pl[c], (size_t)keep * S * esz, hipMemcpyDeviceToDevice, s);
                if (e == hipSuccess) e = hipStreamSynchronize(s);
                if (e != hipSuccess) {
                    retire(retired, np, b);
                    return e;
                }
                cache_put(Block{pl[c], pbytes[c]});
            } else {
                retire(retired, pl[c], pbytes[c]);
            }
            pl[c] = np;
            pbytes[c] = b;
            got = std::min<int64_t>(got, (int64_t)(b / ((size_t)S * esz)));
        }
        cap_blocks = got;
        return hipSuccess;
    }
    template <typename T>
    Planes<T> planes() const {
        Planes<T> r;
        for (int c = 0; c < 7; ++c) r.p[c] = (T*)pl[c];
        return r;
    }
};

}  // namespace

/* mapped read-back layout: int32 slots, an FP64 pair count at byte
 * MAP_PAIRS, EST_DOUBLES per rig from byte MAP_EST */
static constexpr int MAP_PAIRS = 128, MAP_EST = 256;
static size_t map_bytes(int n_rigs) {
    return (MAP_EST + (size_t)n_rigs * EST_DOUBLES * sizeof(double) + 4095) / 4096 * 4096;
}

struct scphd_filter {
    scphd_params prm;
    int dev = 0;
    hipStream_t stream = nullptr;
    /* second stream: non-detection clouds (k_nd_fill) overlap the
     * normalisers; joined before the next generation can grow */
    hipStream_t side = nullptr;
    hipEvent_t ev_fit = nullptr, ev_nd = nullptr;
    std::string err;
    FilterModels models;
    BirthModel birth;
    NoiseKey key;
    int Np = 0, S = 0;
    /* ensemble: n_rigs rigs stepped by the same launches; rig r owns
     * parents [poff[r], poff[r+1]) (prig: the rig of each parent), draws its
     * noise from dkeys[r], projects with dmeas[r] and reads its measurements
     * at [moff[r], moff[r+1]) */
    bool ensemble = false;
    int n_rigs = 1;
    std::vector<scphd_params> rig_prm;
    std::vector<DisparityMeasurementModel> rig_meas;
    std::vector<int32_t> hpoff;
    DBuf<NoiseKey> dkeys;
    DBuf<DisparityMeasurementModel> dmeas;
    DBuf<int32_t> dpoff, dprig, dmoff, drflag;
    DBuf<double> dbpcs;
    DBuf<BirthModel> dbirths; /* per rig; one for a plain filter */
    std::vector<int32_t> hmoff;
    std::vector<double> hbpcs;
    /* per rig, SC_COUNT step counts of the handle when the rig joined (a rig
     * added between steps draws its noise with its own counts) */
    std::vector<uint32_t> hsbase;
    DBuf<uint32_t> dsbase;
    bool sbase_used = false; /* some rig joined after steps were taken */
    /* per rig, the fixed camera at this update (scphd_update_ensemble_rigs) */
    std::vector<int32_t> hfixed;
    DBuf<int32_t> dfixed;
    bool rig_fixed = false;
    int64_t ens_meas = 0;
    size_t esz = 4;
    bool arith32 = false; /* per-sample arithmetic in single precision (SCPHD_STORAGE_FP32_ARITH) */

    /* parents */
    DBuf<double> ext, pw, pose, loglik, cum, ext_tmp;
    DBuf<int32_t> anc, newcnt;
    /* map table and its successor */
    int64_t Ng = 0;
    DBuf<int32_t> goff, gparent, gblock;
    DBuf<double> gw;
    DBuf<int32_t> goff2, gparent2, gblock2;
    DBuf<double> gw2;
    /* per entry: pd_g of its source term at the last update, 1 for posterior
     * terms (valid once an update built the table; set_state clears it) */
    DBuf<float> gpd, gpd2;
    bool hint_ok = false;
    /* sample generations */
    Generation gen[2];
    int cur = 0;
    /* every block the table references is below this index of gen[cur] */
    int64_t blocks_used = 0;
    /* per-Gaussian intermediates */
    DBuf<double> fit, pre, lik;
    DBuf<int32_t> keep, keepscan, ndlist, holes, redo, ndfill;
    /* measurements, normalisers, survivors, merge */
    DBuf<double> du, dv, eta;
    int capM = 0;
    DBuf<int32_t> nsurv, survoff, nmerged, mergeoff, outcnt, scantmp, sg, sm, dflag, mcnt, mncand, mcand;
    DBuf<int32_t> ngroups, glead, gsta, gend, gmem;
    DBuf<double> sw, smu, mw, mmu, mP, dneff;
    DBuf<char> mscratch;
    /* mapped host memory the device writes read-back values into: int32
     * [0, 16) for k_readout and k_nan_count, an FP64 pair count at byte
     * MAP_PAIRS, the FP64 estimates from byte MAP_EST (EST_DOUBLES a rig) */
    int32_t* hmap = nullptr;
    int32_t* dmap = nullptr; /* its device address */
    std::vector<int32_t> hoff;

    /* counters and deferred prediction */
    int64_t n_predict = 0, n_update = 0, n_resample = 0;
    bool pending = false;
    double pdt = 0.0;
    uint64_t pstep = 0;

    /* blocks replaced during a call, handed to the cache once the streams
     * have drained */
    std::vector<Block> retired;

    scphd_update_stats stats{};
    int64_t last_M = 0;
    int64_t kt_hint = 0; /* blocks the last update's posterior clouds took beyond the dropped terms' */
    bool has_outputs = false; /* the last update's per-parent outputs are held (scphd_compact releases them) */
    bool timing = false;
    hipEvent_t ev[10] = {};  /* 0..5 phase boundaries, 6..7 fit kernel, 8..9 sample kernel */
    double phase_ms[8] = {};
    int lds_cap = 0;
    /* the open profiler range of the update's current phase (0: none) */
    roctx_range_id_t phase_range = 0;
    /* verification limits (scphd_debug_set_limits); -1 = the normal limit */
    int dbg_lds_cap = -1, dbg_heavy = -1, dbg_ncand = -1, dbg_rounds = 0, dbg_all_pairs = 0;
};

namespace {

int fail(scphd_filter* f, int code, const std::string& msg) {
    if (f) f->err = msg;
    return code;
}

int hipfail(scphd_filter* f, hipError_t e, const char* what) {
    return fail(f, SCPHD_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

/* the noise keys the kernels draw with: one, or one per rig (with the
 * rigs' step counts at joining, for counter class `which`) */
RigKeys rig_keys(const scphd_filter* f, int which) {
    RigKeys r;
    r.key = f->key;
    if (f->ensemble) {
        r.keys = f->dkeys.p;
        r.prig = f->dprig.p;
        r.poff = f->dpoff.p;
        r.sbase = f->sbase_used ? f->dsbase.p : nullptr;
        r.which = which;
    }
    return r;
}

/* every device buffer of a handle */
template <typename Fn>
void for_each_buf(scphd_filter* f, Fn&& fn) {
    fn(f->dkeys); fn(f->dmeas); fn(f->dpoff); fn(f->dprig); fn(f->dmoff); fn(f->drflag); fn(f->dbpcs);
    fn(f->dbirths); fn(f->dsbase); fn(f->dfixed); fn(f->ext); fn(f->pw); fn(f->pose); fn(f->loglik); fn(f->cum); fn(f->ext_tmp); fn(f->anc);
    fn(f->newcnt); fn(f->goff); fn(f->gparent); fn(f->gblock); fn(f->gw); fn(f->goff2); fn(f->gparent2);
    fn(f->gblock2); fn(f->gw2); fn(f->gpd); fn(f->gpd2); fn(f->fit); fn(f->pre); fn(f->lik); fn(f->keep);
    fn(f->keepscan); fn(f->ndlist); fn(f->holes); fn(f->redo); fn(f->ndfill); fn(f->du); fn(f->dv); fn(f->eta); fn(f->nsurv);
    fn(f->survoff); fn(f->nmerged); fn(f->mergeoff); fn(f->outcnt); fn(f->scantmp); fn(f->sg); fn(f->sm);
    fn(f->dflag); fn(f->mcnt); fn(f->mncand); fn(f->mcand); fn(f->ngroups); fn(f->glead); fn(f->gsta);
    fn(f->gend); fn(f->gmem); fn(f->sw); fn(f->smu); fn(f->mw); fn(f->mmu); fn(f->mP); fn(f->dneff);
    fn(f->mscratch);
}

/* hand the retired blocks to the cache: call only when no queued work on
 * either stream can still use them (after a wait on the main stream that
 * follows the join of the second stream) */
void flush_retired(scphd_filter* f) {
    for (const Block& b : f->retired) cache_put(b);
    f->retired.clear();
}

DisparityMeasurementModel meas_of(const scphd_params& p) {
    return DisparityMeasurementModel(p.fx, p.fy, p.u0, p.v0, p.std_u, p.std_v, p.pd, p.lambda);
}

BirthModel birth_of(const scphd_params& p) {
    BirthModel b;
    b.mean[0] = b.mean[1] = 0.0;
    b.mean[2] = p.birth_d0;
    b.mean[3] = p.birth_vu0;
    b.mean[4] = p.birth_vv0;
    b.mean[5] = p.birth_vd0;
    for (double& x : b.P) x = 0.0;
    b.P[pidx(0, 0)] = p.std_u * p.std_u;
    b.P[pidx(1, 1)] = p.std_v * p.std_v;
    b.P[pidx(2, 2)] = p.birth_var_d0;
    b.P[pidx(3, 3)] = p.birth_var_vu0;
    b.P[pidx(4, 4)] = p.birth_var_vv0;
    b.P[pidx(5, 5)] = p.birth_var_vd0;
    b.logw0 = safe_log(p.birth_w0);
    b.pd_ok = is_pd6(b.P) ? 1 : 0;
    return b;
}

#define HIPTRY(expr)                                              \
    do {                                                          \
        hipError_t e__ = (expr);                                  \
        if (e__ != hipSuccess) return hipfail(f, e__, #expr);     \
    } while (0)

void set_models(scphd_filter* f) {
    const scphd_params& p = f->prm;
    f->models.meas = meas_of(p);
    f->models.cam = OrientedLinearCVMotionModel3D(p.cam_accel_std[0], p.cam_accel_std[1], p.cam_accel_std[2],
                                                  p.cam_accel_std[3], p.cam_accel_std[4], p.cam_accel_std[5]);
    f->models.feat = LinearCVMotionModel3D(p.feat_accel_std[0], p.feat_accel_std[1], p.feat_accel_std[2]);
    f->birth = birth_of(p);
    f->key = NoiseKey::from_seed(p.seed);
    /* largest survivor count whose merge workspace fits 60 KB of LDS */
    int cap = 64;
    while (merge_lds_bytes(cap * 2) <= 61440) cap *= 2;
    while (merge_lds_bytes(cap + 64) <= 61440) cap += 64;
    f->lds_cap = cap;
}

/* Apply a deferred sample prediction into fresh, unshared blocks. */
template <typename T>
int materialize_t(scphd_filter* f) {
    if (!f->pending) return SCPHD_OK;
    f->pending = false;
    if (f->Ng == 0) return SCPHD_OK;
    const int nxt = 1 - f->cur;
    HIPTRY(f->gen[nxt].ensure(f->Ng, 0, f->S, f->esz, f->stream));
    HIPTRY(launch_materialize<T>(f->stream, f->gen[f->cur].planes<T>(), f->gen[nxt].planes<T>(), f->gblock.p,
                                 f->gparent.p, f->goff.p, f->Ng, f->S, f->models.feat, f->pdt, rig_keys(f, SC_PREDICT), f->pstep,
                                 f->arith32));
    HIPTRY(launch_iota(f->stream, f->gblock.p, f->Ng));
    f->cur = nxt;
    f->blocks_used = f->Ng;
    return SCPHD_OK;
}

int materialize(scphd_filter* f) {
    return f->esz == 8 ? materialize_t<double>(f) : materialize_t<float>(f);
}

/* A profiler range over a call (predict, resample): pushed at construction,
 * popped however the call returns. */
struct ScopedRange {
    explicit ScopedRange(const char* m) { roctxRangePushA(m); }
    ~ScopedRange() { roctxRangePop(); }
};

/* Phase boundary i of an update: a HIP event when timing is on, and a
 * profiler (roctx) range per phase, as the reference marks its phases with
 * NVTX ranges (src/scphdcameracalibration.cu:1062-2636).  The ranges span the
 * host-side enqueueing of each phase; a failed update's open range is closed
 * at the next one. */
void record(scphd_filter* f, int i) {
    if (f->timing) (void)hipEventRecord(f->ev[i], f->stream);
    static const char* const names[5] = {"scphd update: reproject + fit", "scphd update: normalisers + survivors",
                                         "scphd update: merge + parent weights", "scphd update: posterior sampling",
                                         "scphd update: map assembly"};
    if (i <= 5) {
        if (f->phase_range) roctxRangeStop(f->phase_range);
        f->phase_range = i < 5 ? roctxRangeStartA(names[i]) : 0;
    }
}

template <typename T>
int update_t(scphd_filter* f, int M, int fixed) {
    hipStream_t s = f->stream;
    const int Np = f->Np, S = f->S;
    const int64_t Ng = f->Ng;
    const scphd_params& P = f->prm;
    const int nxt = 1 - f->cur;
    record(f, 0);

    HIPTRY(launch_pose(s, f->ext.p, f->pose.p, Np, fixed, f->rig_fixed ? f->dprig.p : nullptr,
                       f->rig_fixed ? f->dfixed.p : nullptr));
    HIPTRY(f->keepscan.ensure(Ng + 1));
    HIPTRY(f->scantmp.ensure(scan_tmp_size(std::max<int64_t>(Ng, Np) + 1) + 16));
    if (Ng > 0) {
        /* room for this step's posterior clouds too (the last step's count,
         * scaled by this step's measurement count, plus a quarter), so that
         * sizing them after the merge rarely has to grow the generation and
         * copy the non-detection clouds already in it.  The estimate is only
         * a head start: a generation that holds Ng plus the estimate is not
         * grown for the quarter, and where the estimate does not fit, the
         * generation is sized for the non-detection clouds alone and grows
         * after the merge. */
        const int64_t hint = f->last_M > 0 ? f->kt_hint * (int64_t)M / f->last_M : 0;
        if (f->gen[nxt].cap_blocks < Ng + hint) {
            /* the other generation holds nothing the next map needs, and
             * between updates only the parent prediction is queued (the
             * second stream is idle): once that has run, its planes can be
             * handed back before the larger ones are allocated */
            HIPTRY(hipStreamSynchronize(s));
            flush_retired(f);
            if (f->gen[nxt].ensure(Ng + hint + hint / 4, 0, S, f->esz, s, true) != hipSuccess) {
                (void)hipGetLastError();
                HIPTRY(f->gen[nxt].ensure(Ng, 0, S, f->esz, s, true));
            }
        }
        HIPTRY(f->fit.ensure((int64_t)FP_COUNT * Ng));
        HIPTRY(f->pre.ensure((int64_t)PP_COUNT * Ng));
        HIPTRY(f->lik.ensure((int64_t)LK_COUNT * (Ng + 3))); /* 3 records of padding: read four at a time */
        HIPTRY(f->keep.ensure(Ng));
        HIPTRY(f->redo.ensure(Ng + 1)); /* [Ng]: the list length */
        HIPTRY(f->ndfill.ensure(Ng + 1)); /* [Ng]: the list length */
        HIPTRY(f->ndlist.ensure(Ng));
        HIPTRY(f->holes.ensure(Ng));
        FitArgs fa;
        fa.gblock = f->gblock.p;
        fa.gparent = f->gparent.p;
        fa.goff = f->goff.p;
        fa.gw = f->gw.p;
        fa.pose = f->pose.p;
        fa.Ng = Ng;
        fa.S = S;
        fa.meas = f->models.meas;
        fa.rig_meas = f->ensemble ? f->dmeas.p : nullptr;
        fa.feat = f->models.feat;
        fa.pending = f->pending ? 1 : 0;
        fa.dt = f->pdt;
        fa.key = rig_keys(f, SC_PREDICT); /* the landmark prediction deferred from predict */
        fa.pstep = f->pstep;
        fa.min_weight = P.min_weight;
        fa.fit = f->fit.p;
        fa.keep = f->keep.p;
        fa.redo = f->redo.p;
        fa.redo_n = f->redo.p + Ng;
        HIPTRY(hipMemsetAsync(fa.redo_n, 0, sizeof(int32_t), s));
        fa.ndfill = f->ndfill.p;
        fa.ndfill_n = f->ndfill.p + Ng;
        HIPTRY(hipMemsetAsync(fa.ndfill_n, 0, sizeof(int32_t), s));
        fa.nd_base = 0;
        fa.gpd = f->hint_ok ? f->gpd.p : nullptr;
        /* 32-bit byte offsets when every block read (below blocks_used of
         * the current generation) and written (non-detection clouds below
         * Ng of the next) lies below 4 GiB */
        fa.off32 = std::max(f->blocks_used, Ng) * (int64_t)S * (int64_t)f->esz <
                   (int64_t)0xFFFFFFFFll - 64 * (int64_t)f->esz;
        fa.arith32 = f->arith32 ? 1 : 0;
        record(f, 6);
        HIPTRY(launch_fit<T>(s, fa, f->gen[f->cur].planes<T>(), f->gen[nxt].planes<T>()));
        record(f, 7);
        HIPTRY(hipEventRecord(f->ev_fit, s));
        HIPTRY(hipStreamWaitEvent(f->side, f->ev_fit, 0));
        HIPTRY(launch_nd_fill<T>(f->side, fa, f->gen[f->cur].planes<T>(), f->gen[nxt].planes<T>()));
        HIPTRY(hipEventRecord(f->ev_nd, f->side));
        HIPTRY(launch_fit_fix<T>(s, fa, f->gen[f->cur].planes<T>(), f->gen[nxt].planes<T>()));
        HIPTRY(launch_precompute(s, f->fit.p, f->gw.p, Ng, P.std_u * P.std_u, P.std_v * P.std_v, f->pre.p,
                                 f->lik.p));
    }
    f->pending = false;
    record(f, 1);

    /* normalisers, parent log-likelihood, survivor counts */
    HIPTRY(f->eta.ensure((int64_t)Np * std::max(M, 1)));
    HIPTRY(f->nsurv.ensure(Np + 1));
    HIPTRY(f->survoff.ensure(Np + 1));
    HIPTRY(f->nmerged.ensure(Np + 1));
    HIPTRY(f->mergeoff.ensure(Np + 1));
    HIPTRY(f->outcnt.ensure(Np + 1));
    HIPTRY(f->goff2.ensure(Np + 1));
    HIPTRY(f->dflag.ensure(4));
    EtaArgs ea;
    ea.goff = f->goff.p;
    ea.gw = f->gw.p;
    ea.pre = f->pre.p;
    ea.lik = f->lik.p;
    ea.Ng = Ng;
    ea.zu = f->du.p;
    ea.zv = f->dv.p;
    ea.M = M;
    ea.Np = Np;
    ea.bpc = safe_log(M * P.birth_w0 + f->models.meas.kappa());
    if (f->ensemble) { /* M is the largest rig's count, the stride of eta, mcnt, mncand and mcand */
        ea.moff = f->dmoff.p;
        ea.bpcs = f->dbpcs.p;
        ea.prig = f->dprig.p;
        HIPTRY(launch_pair_count(s, f->goff.p, f->dmoff.p, f->dpoff.p, f->n_rigs,
                                 (double*)((char*)f->dmap + MAP_PAIRS)));
    }
    ea.min_weight = P.min_weight;
    ea.births = f->dbirths.p;
    ea.eta = f->eta.p;
    ea.loglik = f->loglik.p;
    ea.nsurv = f->nsurv.p;
    HIPTRY(f->mcnt.ensure((int64_t)Np * std::max(M, 1)));
    HIPTRY(f->mncand.ensure((int64_t)Np * std::max(M, 1)));
    HIPTRY(f->mcand.ensure((int64_t)Np * std::max(M, 1) * ETA_NCAND));
    ea.mcnt = f->mcnt.p;
    ea.mncand = f->mncand.p;
    ea.mcand = f->mcand.p;
    ea.lw_gate = P.min_weight > 0 ? std::log(P.min_weight) - 1e-3 : -INFINITY;
    ea.heavy_cap = f->dbg_heavy >= 0 ? std::min(f->dbg_heavy, ETA_HEAVY) : ETA_HEAVY;
    ea.ncand_cap = f->dbg_ncand >= 0 ? std::min(f->dbg_ncand, ETA_NCAND) : ETA_NCAND;
    HIPTRY(launch_eta(s, ea));
    HIPTRY(launch_scan(s, f->nsurv.p, f->survoff.p, Np, f->scantmp.p));
    HIPTRY(launch_readout(s, f->survoff.p, Np, 0, nullptr, nullptr, f->dmap));
    HIPTRY(hipStreamSynchronize(s));
    const int64_t Stot = f->hmap[0];
    const int maxU = f->hmap[1];

    HIPTRY(f->sw.ensure(Stot + 1));
    HIPTRY(f->sg.ensure(Stot + 1));
    HIPTRY(f->sm.ensure(Stot + 1));
    const int64_t scap = f->sw.n;
    HIPTRY(f->smu.ensure(6 * scap));
    HIPTRY(f->mw.ensure(scap));
    HIPTRY(f->mmu.ensure(6 * scap));
    HIPTRY(f->mP.ensure(21 * scap));
    HIPTRY(hipMemsetAsync(f->dflag.p, 0, 4 * sizeof(int32_t), s));
    if (Stot > 0) {
        SurvArgs sa;
        sa.e = ea;
        sa.fit = f->fit.p;
        sa.survoff = f->survoff.p;
        sa.sw = f->sw.p;
        sa.sg = f->sg.p;
        sa.sm = f->sm.p;
        sa.smu = f->smu.p;
        sa.cap = scap;
        sa.err = f->dflag.p;
        HIPTRY(launch_survivors(s, sa));
    }
    record(f, 2);

    /* merge */
    MergeArgs ma;
    ma.survoff = f->survoff.p;
    ma.sw = f->sw.p;
    ma.sg = f->sg.p;
    ma.sm = f->sm.p;
    ma.smu = f->smu.p;
    ma.scap = scap;
    ma.pre = f->pre.p;
    ma.Ng = Ng;
    ma.births = f->dbirths.p;
    ma.min_separation = P.min_separation;
    ma.Np = Np;
    const int lds_cap = f->dbg_lds_cap >= 0 ? std::min(f->dbg_lds_cap, f->lds_cap) : f->lds_cap;
    ma.lds_cap = lds_cap;
    ma.force_rounds = f->dbg_rounds;
    ma.all_pairs = f->dbg_all_pairs;
    ma.scratch = nullptr;
    ma.mw = f->mw.p;
    ma.mmu = f->mmu.p;
    ma.mP = f->mP.p;
    ma.nmerged = f->nmerged.p;
    HIPTRY(f->ngroups.ensure(Np + 1));
    HIPTRY(f->glead.ensure(Stot + 1));
    HIPTRY(f->gsta.ensure(Stot + 1));
    HIPTRY(f->gend.ensure(Stot + 1));
    HIPTRY(f->gmem.ensure(Stot + 1));
    ma.ngroups = f->ngroups.p;
    ma.glead = f->glead.p;
    ma.gsta = f->gsta.p;
    ma.gend = f->gend.p;
    ma.gmem = f->gmem.p;
    ma.Stot = Stot;
    ma.big_lds_up = 0;
    ma.big_lds_bytes = 0;
    if (maxU > lds_cap) {
        HIPTRY(f->mscratch.ensure(merge_scratch_bytes(Stot, Np)));
        ma.scratch = f->mscratch.p;
        /* the big parents' sort keys in LDS: 12 B per slot, at most 48 KB */
        int up = 1;
        while (up < maxU) up <<= 1;
        ma.big_lds_up = up <= 4096 ? up : 0;
        /* the sort keys (12 B per slot of next_pow2(U)), then, in the same
         * LDS, 28 B per survivor for the walk's counts, offsets, labels,
         * buckets and group records */
        ma.big_lds_bytes = std::max(12 * ma.big_lds_up, ma.big_lds_up ? std::min(28 * maxU + 16, 65536) : 0);
    }
    HIPTRY(launch_merge(s, ma, merge_lds_bytes(std::max(1, std::min(maxU, lds_cap)))));
    if (!fixed)
        HIPTRY(launch_parent_weights(s, f->pw.p, f->loglik.p, Np, f->ensemble ? f->dpoff.p : nullptr, f->n_rigs,
                                     f->rig_fixed ? f->dfixed.p : nullptr));
    record(f, 3);

    /* output counts */
    HIPTRY(launch_scan(s, f->nmerged.p, f->mergeoff.p, Np, f->scantmp.p));
    if (Ng > 0)
        HIPTRY(launch_scan(s, f->keep.p, f->keepscan.p, Ng, f->scantmp.p));
    else
        HIPTRY(hipMemsetAsync(f->keepscan.p, 0, sizeof(int32_t), s));
    HIPTRY(launch_outcount(s, f->nmerged.p, f->goff.p, f->keepscan.p, Np, f->outcnt.p));
    HIPTRY(launch_scan(s, f->outcnt.p, f->goff2.p, Np, f->scantmp.p));
    /* the non-detection clouds in [0, Ng) of the next generation are
     * complete before any return or growth of the generation (which copies
     * them) */
    HIPTRY(hipStreamWaitEvent(s, f->ev_nd, 0));
    HIPTRY(launch_readout(s, f->goff2.p, Np, P.max_gaussians, f->mergeoff.p + Np, f->dflag.p, f->dmap));
    HIPTRY(hipStreamSynchronize(s));
    flush_retired(f); /* both streams have drained */
    const int64_t Nout = f->hmap[0];
    const int64_t Kt = f->hmap[3];
    if (f->hmap[4] != 0) return fail(f, SCPHD_ERR_STATE, "survivor count mismatch between passes");
    const int64_t capx = f->hmap[2];
    f->stats.cap_exceeded = capx;
    if (capx > 0) return fail(f, SCPHD_ERR_CAPACITY, "a parent map exceeds max_gaussians");

    /* posterior clouds: first into the blocks of the dropped non-detection
     * terms (k_fit wrote nothing there that is kept), the rest from block
     * Ng on; the next generation holds Ng + max(0, Kt - holes) blocks */
    const int64_t nholes = Ng - (Nout - Kt);
    const int64_t over = std::max<int64_t>(0, Kt - nholes);
    HIPTRY(launch_ndlist(s, f->keep.p, f->keepscan.p, Ng, f->ndlist.p, f->holes.p));
    HIPTRY(f->gen[nxt].ensure(Ng + over, Ng, S, f->esz, s));
    SampleArgs pa;
    pa.mergeoff = f->mergeoff.p;
    pa.survoff = f->survoff.p;
    pa.mmu = f->mmu.p;
    pa.mP = f->mP.p;
    pa.scap = scap;
    pa.pose = f->pose.p;
    pa.Np = Np;
    pa.Kt = Kt;
    pa.S = S;
    pa.meas = f->models.meas;
    pa.rig_meas = f->ensemble ? f->dmeas.p : nullptr;
    pa.key = rig_keys(f, SC_UPDATE);
    pa.ustep = (uint64_t)f->n_update;
    pa.post_base = Ng;
    pa.holes = f->holes.p;
    pa.nholes = nholes;
    /* 32-bit byte offsets while every destination block lies below 4 GiB */
    pa.off32 = (Ng + over) * (int64_t)S * (int64_t)f->esz < (int64_t)0xFFFFFFFFll - 64 * (int64_t)f->esz;
    pa.arith32 = f->arith32 ? 1 : 0;
    record(f, 8);
    HIPTRY(launch_sample<T>(s, pa, f->gen[nxt].planes<T>()));
    record(f, 9);
    record(f, 4);

    /* next map table */
    HIPTRY(f->gparent2.ensure(Nout + 1));
    HIPTRY(f->gblock2.ensure(Nout + 1));
    HIPTRY(f->gw2.ensure(Nout + 1));
    HIPTRY(f->gpd2.ensure(Nout + 1));
    TableArgs ta;
    ta.outoff = f->goff2.p;
    ta.mergeoff = f->mergeoff.p;
    ta.survoff = f->survoff.p;
    ta.goff = f->goff.p;
    ta.keepscan = f->keepscan.p;
    ta.ndlist = f->ndlist.p;
    ta.mw = f->mw.p;
    ta.gw = f->gw.p;
    ta.fit = f->fit.p;
    ta.Ng_in = Ng;
    ta.Np = Np;
    ta.Nout = Nout;
    ta.post_base = Ng;
    ta.holes = f->holes.p;
    ta.nholes = nholes;
    ta.nd_base = 0;
    ta.gparent = f->gparent2.p;
    ta.gblock = f->gblock2.p;
    ta.gwo = f->gw2.p;
    ta.gpd = f->gpd2.p;
    HIPTRY(launch_table(s, ta));
    record(f, 5);
    std::swap(f->goff, f->goff2);
    std::swap(f->gparent, f->gparent2);
    std::swap(f->gblock, f->gblock2);
    std::swap(f->gw, f->gw2);
    std::swap(f->gpd, f->gpd2);
    f->hint_ok = true;
    f->cur = nxt;
    f->Ng = Nout;
    f->blocks_used = Ng + over; /* non-detection clouds and posteriors in holes below Ng, the rest from Ng */
    f->n_update++;

    f->stats.n_meas = f->ensemble ? f->ens_meas : M;
    f->stats.n_gaussians_in = Ng;
    f->stats.n_pairs = f->ensemble ? (int64_t)*(const double*)((const char*)f->hmap + MAP_PAIRS) : Ng * M;
    f->stats.n_survivors = Stot;
    f->stats.n_merged = Kt;
    f->stats.n_nondetect = Nout - Kt;
    f->stats.n_gaussians_out = Nout;
    f->stats.max_survivors_per_parent = maxU;
    f->last_M = M;
    f->kt_hint = over;
    f->has_outputs = true;
    if (f->timing) {
        HIPTRY(hipEventSynchronize(f->ev[5]));
        for (int i = 0; i < 5; ++i) {
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, f->ev[i], f->ev[i + 1]);
            f->phase_ms[i] = ms;
        }
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, f->ev[0], f->ev[5]);
        f->phase_ms[5] = ms;
        ms = 0.f;
        if (Ng > 0) (void)hipEventElapsedTime(&ms, f->ev[6], f->ev[7]);
        f->phase_ms[6] = ms;
        ms = 0.f;
        if (Kt > 0) (void)hipEventElapsedTime(&ms, f->ev[8], f->ev[9]);
        f->phase_ms[7] = ms;
    }
    return SCPHD_OK;
}

int update_common(scphd_filter* f, int M, int fixed) {
    int rc = f->esz == 8 ? update_t<double>(f, M, fixed) : update_t<float>(f, M, fixed);
    return rc;
}

int check_handle(scphd_filter* f) { return f ? SCPHD_OK : SCPHD_ERR_ARG; }

/* a handle with parents: an ensemble whose rigs have all left takes new
 * rigs and nothing else */
int check_live(scphd_filter* f) {
    if (!f) return SCPHD_ERR_ARG;
    if (f->Np == 0) return fail(f, SCPHD_ERR_STATE, "an ensemble with no rigs: add a rig first");
    return SCPHD_OK;
}

}  // namespace

/* ================================================================== ABI */
extern "C" {

}  /* extern "C" */

namespace {

/* Parameters every rig of an ensemble shares: the sample count, storage,
 * capacity, pruning, merging, resampling, survival, motion models and
 * measurement noise.  Rigs may differ in parent count, prior, camera
 * intrinsics, detection probability, clutter rate, birth model and seed. */
bool same_shared(const scphd_params& a, const scphd_params& b) {
    bool eq = a.particles_per_feature == b.particles_per_feature && a.max_gaussians == b.max_gaussians &&
              a.storage == b.storage && a.min_weight == b.min_weight && a.min_separation == b.min_separation &&
              a.min_neff == b.min_neff && a.ps == b.ps && a.std_u == b.std_u && a.std_v == b.std_v;
    for (int i = 0; i < 6; ++i) eq = eq && a.cam_accel_std[i] == b.cam_accel_std[i];
    for (int i = 0; i < 3; ++i) eq = eq && a.feat_accel_std[i] == b.feat_accel_std[i];
    return eq;
}

/* One filter of n_rigs rigs; ensemble false: the plain filter of p[0]. */
int create_filter(const scphd_params* p, int32_t n_rigs, bool ensemble, scphd_t** out) {
    if (!p || !out || n_rigs <= 0 || (!ensemble && n_rigs != 1)) return SCPHD_ERR_ARG;
    *out = nullptr;
    int64_t Np_total = 0;
    for (int r = 0; r < n_rigs; ++r) {
        const scphd_params& q = p[r];
        if (q.n_particles <= 0 || q.particles_per_feature <= 0 || q.particles_per_feature > 512)
            return SCPHD_ERR_ARG;
        if (q.storage != SCPHD_STORAGE_FP32 && q.storage != SCPHD_STORAGE_FP64 &&
            q.storage != SCPHD_STORAGE_FP32_ARITH)
            return SCPHD_ERR_ARG;
        if (!same_shared(q, p[0])) return SCPHD_ERR_ARG;
        Np_total += q.n_particles;
    }
    if (Np_total > (int64_t)1 << 30) return SCPHD_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return SCPHD_ERR_NODEVICE;
    /* the device with the most compute units (src/scphdcameracalibration.cu:1013-1024) */
    int best = 0, best_cu = -1;
    for (int d = 0; d < ndev; ++d) {
        int cu = 0;
        if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, d) == hipSuccess && cu > best_cu) {
            best_cu = cu;
            best = d;
        }
    }
    scphd_filter* f = new scphd_filter();
    for_each_buf(f, [f](auto& b) { b.retired = &f->retired; });
    f->gen[0].retired = f->gen[1].retired = &f->retired;
    f->prm = *p;
    f->dev = best;
    f->ensemble = ensemble;
    f->n_rigs = n_rigs;
    f->Np = (int)Np_total;
    f->S = p->particles_per_feature;
    f->esz = p->storage == SCPHD_STORAGE_FP64 ? 8 : 4;
    f->arith32 = p->storage == SCPHD_STORAGE_FP32_ARITH;
    set_models(f);
    auto bail = [&](int code) {
        scphd_destroy(f);
        return code;
    };
    if (hipSetDevice(f->dev) != hipSuccess) return bail(SCPHD_ERR_NODEVICE);
    if (hipStreamCreateWithFlags(&f->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&f->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&f->ev_fit, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&f->ev_nd, hipEventDisableTiming) != hipSuccess)
        return bail(SCPHD_ERR_HIP);
    for (auto& e : f->ev)
        if (hipEventCreate(&e) != hipSuccess) return bail(SCPHD_ERR_HIP);
    const int Np = f->Np;
    if (f->ext.ensure(12 * (int64_t)Np) || f->ext_tmp.ensure(12 * (int64_t)Np) || f->pw.ensure(Np) ||
        f->pose.ensure(12 * (int64_t)Np) || f->loglik.ensure(Np) || f->cum.ensure(Np) || f->anc.ensure(Np) ||
        f->newcnt.ensure(Np + 1) || f->goff.ensure(Np + 1) || f->goff2.ensure(Np + 1) || f->dneff.ensure(n_rigs) ||
        f->drflag.ensure(n_rigs) || f->dflag.ensure(4) || f->gparent.ensure(1) || f->gblock.ensure(1) ||
        f->gw.ensure(1))
        return bail(SCPHD_ERR_HIP);
    std::vector<BirthModel> births(n_rigs);
    f->hpoff.assign(n_rigs + 1, 0);
    for (int r = 0; r < n_rigs; ++r) {
        births[r] = birth_of(p[r]);
        f->hpoff[r + 1] = f->hpoff[r] + p[r].n_particles;
    }
    if (f->dbirths.ensure(n_rigs) ||
        hipMemcpy(f->dbirths.p, births.data(), n_rigs * sizeof(BirthModel), hipMemcpyHostToDevice) != hipSuccess)
        return bail(SCPHD_ERR_HIP);
    if (f->ensemble) {
        f->rig_prm.assign(p, p + n_rigs);
        f->hsbase.assign((size_t)SC_COUNT * n_rigs, 0u); /* every rig joins at the first step */
        std::vector<NoiseKey> keys(n_rigs);
        std::vector<int32_t> prig(Np);
        for (int r = 0; r < n_rigs; ++r) {
            keys[r] = NoiseKey::from_seed(p[r].seed);
            f->rig_meas.push_back(meas_of(p[r]));
            for (int q = f->hpoff[r]; q < f->hpoff[r + 1]; ++q) prig[q] = r;
        }
        if (f->dkeys.ensure(n_rigs) || f->dmeas.ensure(n_rigs) || f->dpoff.ensure(n_rigs + 1) ||
            f->dprig.ensure(Np) || f->dmoff.ensure(n_rigs + 1) || f->dbpcs.ensure(n_rigs) ||
            hipMemcpy(f->dkeys.p, keys.data(), n_rigs * sizeof(NoiseKey), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(f->dmeas.p, f->rig_meas.data(), n_rigs * sizeof(DisparityMeasurementModel),
                      hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(f->dpoff.p, f->hpoff.data(), (n_rigs + 1) * sizeof(int32_t), hipMemcpyHostToDevice) !=
                hipSuccess ||
            hipMemcpy(f->dprig.p, prig.data(), Np * sizeof(int32_t), hipMemcpyHostToDevice) != hipSuccess)
            return bail(SCPHD_ERR_HIP);
    }
    if (hipHostMalloc((void**)&f->hmap, map_bytes(n_rigs), hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess ||
        hipHostGetDevicePointer((void**)&f->dmap, f->hmap, 0) != hipSuccess)
        return bail(SCPHD_ERR_HIP);
    for (int r = 0; r < n_rigs; ++r) { /* each rig's prior, as a filter of its own draws it */
        const int64_t P0 = f->hpoff[r];
        if (launch_prior(f->stream, f->ext.p + 12 * P0, f->pw.p + P0, p[r].n_particles, p[r].prior_a, p[r].prior_b,
                         NoiseKey::from_seed(p[r].seed)) != hipSuccess)
            return bail(SCPHD_ERR_HIP);
    }
    if (hipMemsetAsync(f->goff.p, 0, (size_t)(Np + 1) * sizeof(int32_t), f->stream) != hipSuccess)
        return bail(SCPHD_ERR_HIP);
    if (launch_iota(f->stream, f->anc.p, Np) != hipSuccess) return bail(SCPHD_ERR_HIP);
    if (hipStreamSynchronize(f->stream) != hipSuccess) return bail(SCPHD_ERR_HIP);
    *out = f;
    return SCPHD_OK;
}

}  // namespace

extern "C" {

int scphd_create(const scphd_params* p, scphd_t** out) { return create_filter(p, 1, false, out); }

int scphd_create_ensemble(const scphd_params* rigs, int32_t n_rigs, scphd_t** out) {
    return create_filter(rigs, n_rigs, true, out);
}

}  /* extern "C" */

namespace {

/* the per-rig device arrays of an ensemble from its host copies: keys,
 * measurement models, birth models, parent offsets, parent -> rig and the
 * step counts at joining */
int upload_rigs(scphd_filter* f) {
    const int R = f->n_rigs, Np = f->Np;
    std::vector<NoiseKey> keys(R);
    std::vector<BirthModel> births(R);
    std::vector<int32_t> prig(std::max(Np, 1));
    for (int r = 0; r < R; ++r) {
        keys[r] = NoiseKey::from_seed(f->rig_prm[r].seed);
        births[r] = birth_of(f->rig_prm[r]);
        for (int q = f->hpoff[r]; q < f->hpoff[r + 1]; ++q) prig[q] = r;
    }
    hipStream_t s = f->stream;
    f->sbase_used = false;
    if (R == 0) return SCPHD_OK;
    HIPTRY(f->dkeys.ensure(R));
    HIPTRY(f->dmeas.ensure(R));
    HIPTRY(f->dbirths.ensure(R));
    HIPTRY(f->dpoff.ensure(R + 1));
    HIPTRY(f->dprig.ensure(std::max(Np, 1)));
    HIPTRY(f->dmoff.ensure(R + 1));
    HIPTRY(f->dbpcs.ensure(R));
    HIPTRY(f->dsbase.ensure((int64_t)SC_COUNT * R));
    HIPTRY(hipMemcpyAsync(f->dkeys.p, keys.data(), R * sizeof(NoiseKey), hipMemcpyHostToDevice, s));
    HIPTRY(hipMemcpyAsync(f->dmeas.p, f->rig_meas.data(), R * sizeof(DisparityMeasurementModel),
                          hipMemcpyHostToDevice, s));
    HIPTRY(hipMemcpyAsync(f->dbirths.p, births.data(), R * sizeof(BirthModel), hipMemcpyHostToDevice, s));
    HIPTRY(hipMemcpyAsync(f->dpoff.p, f->hpoff.data(), (R + 1) * sizeof(int32_t), hipMemcpyHostToDevice, s));
    HIPTRY(hipMemcpyAsync(f->dprig.p, prig.data(), Np * sizeof(int32_t), hipMemcpyHostToDevice, s));
    HIPTRY(hipMemcpyAsync(f->dsbase.p, f->hsbase.data(), (size_t)SC_COUNT * R * sizeof(uint32_t),
                          hipMemcpyHostToDevice, s));
    /* the host vectors are read by the copies: wait for them */
    HIPTRY(hipStreamSynchronize(s));
    for (uint32_t b : f->hsbase) f->sbase_used = f->sbase_used || b != 0;
    return SCPHD_OK;
}

/* rigs join and leave between frames: every predict has had its update */
int between_frames(scphd_filter* f, const char* what) {
    if (f->n_predict != f->n_update || f->pending)
        return fail(f, SCPHD_ERR_STATE, std::string(what) + ": between frames only (after update or resample, "
                                                             "before predict)");
    return SCPHD_OK;
}

}  // namespace

extern "C" {

int scphd_ensemble_add_rig(scphd_t* f, const scphd_params* p, int32_t* rig_out) {
    if (check_handle(f) || !p) return SCPHD_ERR_ARG;
    if (!f->ensemble) return fail(f, SCPHD_ERR_ARG, "scphd_ensemble_add_rig: not an ensemble");
    if (p->n_particles <= 0 || !same_shared(*p, f->rig_prm.empty() ? f->prm : f->rig_prm[0]))
        return fail(f, SCPHD_ERR_ARG, "scphd_ensemble_add_rig: parameters the ensemble's rigs must share differ");
    if (int rc = between_frames(f, "scphd_ensemble_add_rig")) return rc;
    const int64_t Np_new = (int64_t)f->Np + p->n_particles;
    if (Np_new > (int64_t)1 << 30) return SCPHD_ERR_ARG;
    HIPTRY(hipSetDevice(f->dev));
    hipStream_t s = f->stream;
    const int P0 = f->Np, N = p->n_particles, R = f->n_rigs + 1;
    /* per-parent arrays, keeping what the current parents hold */
    HIPTRY(grow_keep(f->ext, 12 * Np_new, 12 * (int64_t)P0, s));
    HIPTRY(grow_keep(f->pw, Np_new, P0, s));
    HIPTRY(grow_keep(f->loglik, Np_new, P0, s));
    HIPTRY(grow_keep(f->anc, Np_new, P0, s));
    HIPTRY(grow_keep(f->goff, Np_new + 1, (int64_t)P0 + 1, s));
    HIPTRY(grow_keep(f->dneff, R, R - 1, s));
    HIPTRY(grow_keep(f->drflag, R, R - 1, s));
    HIPTRY(f->ext_tmp.ensure(12 * Np_new));
    HIPTRY(f->pose.ensure(12 * Np_new));
    HIPTRY(f->cum.ensure(Np_new));
    HIPTRY(f->newcnt.ensure(Np_new + 1));
    HIPTRY(f->goff2.ensure(Np_new + 1));
    if (map_bytes(R) > map_bytes(f->n_rigs)) { /* room for the new rig's estimates */
        HIPTRY(hipStreamSynchronize(s));
        int32_t* h = nullptr;
        int32_t* d = nullptr;
        HIPTRY(hipHostMalloc((void**)&h, map_bytes(R), hipHostMallocMapped | hipHostMallocCoherent));
        if (hipHostGetDevicePointer((void**)&d, h, 0) != hipSuccess) {
            (void)hipHostFree(h);
            return fail(f, SCPHD_ERR_HIP, "scphd_ensemble_add_rig: mapped read-back memory");
        }
        (void)hipHostFree(f->hmap);
        f->hmap = h;
        f->dmap = d;
    }
    /* the rig, its parents (no Gaussians) and its prior, drawn as a filter
     * of its own draws it; its noise counters start at zero */
    f->rig_prm.push_back(*p);
    f->rig_meas.push_back(meas_of(*p));
    f->hpoff.push_back((int32_t)Np_new);
    const uint32_t base[SC_COUNT] = {(uint32_t)f->n_predict, (uint32_t)f->n_update, (uint32_t)f->n_resample};
    f->hsbase.insert(f->hsbase.end(), base, base + SC_COUNT);
    f->n_rigs = R;
    f->Np = (int)Np_new;
    HIPTRY(launch_prior(s, f->ext.p + 12 * (int64_t)P0, f->pw.p + P0, N, p->prior_a, p->prior_b,
                        NoiseKey::from_seed(p->seed)));
    HIPTRY(launch_fill_i32(s, f->goff.p + P0 + 1, (int32_t)f->Ng, N));
    HIPTRY(launch_iota(s, f->anc.p + P0, N, P0));
    HIPTRY(hipMemsetAsync(f->loglik.p + P0, 0, (size_t)N * sizeof(double), s));
    if (int rc = upload_rigs(f)) return rc;
    flush_retired(f); /* upload_rigs waited for the stream */
    if (rig_out) *rig_out = R - 1;
    return SCPHD_OK;
}

int scphd_ensemble_remove_rig(scphd_t* f, int32_t rig) {
    if (check_handle(f)) return SCPHD_ERR_ARG;
    if (!f->ensemble) return fail(f, SCPHD_ERR_ARG, "scphd_ensemble_remove_rig: not an ensemble");
    if (rig < 0 || rig >= f->n_rigs) return fail(f, SCPHD_ERR_ARG, "scphd_ensemble_remove_rig: no such rig");
    if (int rc = between_frames(f, "scphd_ensemble_remove_rig")) return rc;
    HIPTRY(hipSetDevice(f->dev));
    hipStream_t s = f->stream;
    const int P0 = f->hpoff[rig], P1 = f->hpoff[rig + 1], Nr = P1 - P0, Np = f->Np, Np_new = Np - Nr;
    int32_t g01[2] = {0, 0};
    HIPTRY(hipMemcpyAsync(&g01[0], f->goff.p + P0, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIPTRY(hipMemcpyAsync(&g01[1], f->goff.p + P1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIPTRY(hipStreamSynchronize(s));
    const int64_t G0 = g01[0], Gr = g01[1] - g01[0], Ng_new = f->Ng - Gr;
    /* the map table without the rig's Gaussians (their sample blocks stay
     * in the generation, unreferenced, until the next update writes the
     * next one) */
    HIPTRY(f->gparent2.ensure(Ng_new + 1));
    HIPTRY(f->gblock2.ensure(Ng_new + 1));
    HIPTRY(f->gw2.ensure(Ng_new + 1));
    if (f->hint_ok) HIPTRY(f->gpd2.ensure(Ng_new + 1));
    HIPTRY(launch_table_remove(s, f->gparent.p, f->gblock.p, f->gw.p, f->hint_ok ? f->gpd.p : nullptr, G0, Gr,
                               Ng_new, P1, Nr, f->gparent2.p, f->gblock2.p, f->gw2.p,
                               f->hint_ok ? f->gpd2.p : nullptr));
    HIPTRY(launch_goff_remove(s, f->goff.p, P0, Nr, (int32_t)Gr, Np_new, f->goff2.p));
    std::swap(f->goff, f->goff2);
    std::swap(f->gparent, f->gparent2);
    std::swap(f->gblock, f->gblock2);
    std::swap(f->gw, f->gw2);
    if (f->hint_ok) std::swap(f->gpd, f->gpd2);
    /* per-parent arrays: the parents after the rig move down (through the
     * scratch arrays, the ranges overlap) */
    const int tail = Np - P1;
    if (tail > 0) {
        HIPTRY(hipMemcpyAsync(f->ext_tmp.p, f->ext.p + 12 * (int64_t)P1, (size_t)tail * 12 * sizeof(double),
                              hipMemcpyDeviceToDevice, s));
        HIPTRY(hipMemcpyAsync(f->ext.p + 12 * (int64_t)P0, f->ext_tmp.p, (size_t)tail * 12 * sizeof(double),
                              hipMemcpyDeviceToDevice, s));
        for (DBuf<double>* b : {&f->pw, &f->loglik}) {
            HIPTRY(hipMemcpyAsync(f->cum.p, b->p + P1, (size_t)tail * sizeof(double), hipMemcpyDeviceToDevice, s));
            HIPTRY(hipMemcpyAsync(b->p + P0, f->cum.p, (size_t)tail * sizeof(double), hipMemcpyDeviceToDevice, s));
        }
        HIPTRY(hipMemcpyAsync(f->newcnt.p, f->anc.p + P1, (size_t)tail * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
        HIPTRY(hipMemcpyAsync(f->anc.p + P0, f->newcnt.p, (size_t)tail * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
        HIPTRY(launch_index_shift(s, f->anc.p + P0, tail, P1, Nr));
    }
    /* per-rig values of the last resampling */
    const int R = f->n_rigs;
    std::vector<double> neff(R);
    std::vector<int32_t> rflag(R);
    HIPTRY(hipMemcpyAsync(neff.data(), f->dneff.p, R * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPTRY(hipMemcpyAsync(rflag.data(), f->drflag.p, R * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIPTRY(hipStreamSynchronize(s));
    neff.erase(neff.begin() + rig);
    rflag.erase(rflag.begin() + rig);
    if (R > 1) {
        HIPTRY(hipMemcpyAsync(f->dneff.p, neff.data(), (R - 1) * sizeof(double), hipMemcpyHostToDevice, s));
        HIPTRY(hipMemcpyAsync(f->drflag.p, rflag.data(), (R - 1) * sizeof(int32_t), hipMemcpyHostToDevice, s));
    }
    f->rig_prm.erase(f->rig_prm.begin() + rig);
    f->rig_meas.erase(f->rig_meas.begin() + rig);
    f->hsbase.erase(f->hsbase.begin() + (size_t)SC_COUNT * rig, f->hsbase.begin() + (size_t)SC_COUNT * (rig + 1));
    f->hpoff.erase(f->hpoff.begin() + rig + 1);
    for (int r = rig + 1; r < (int)f->hpoff.size(); ++r) f->hpoff[r] -= Nr;
    f->n_rigs = R - 1;
    f->Np = Np_new;
    f->Ng = Ng_new;
    if (!f->rig_prm.empty()) f->prm = f->rig_prm[0]; /* else the shared values of the last rig stay */
    if (int rc = upload_rigs(f)) return rc; /* waits for the stream */
    flush_retired(f);
    return SCPHD_OK;
}

void scphd_destroy(scphd_t* f) {
    if (!f) return;
    (void)hipSetDevice(f->dev);
    if (f->stream) (void)hipStreamSynchronize(f->stream);
    if (f->side) (void)hipStreamSynchronize(f->side);
    for_each_buf(f, [](auto& b) { b.release(); });
    f->gen[0].release();
    f->gen[1].release();
    flush_retired(f); /* the streams have drained above */
    if (f->hmap) (void)hipHostFree(f->hmap);
    for (auto& e : f->ev)
        if (e) (void)hipEventDestroy(e);
    if (f->ev_fit) (void)hipEventDestroy(f->ev_fit);
    if (f->ev_nd) (void)hipEventDestroy(f->ev_nd);
    if (f->side) (void)hipStreamDestroy(f->side);
    if (f->stream) (void)hipStreamDestroy(f->stream);
    delete f;
}

const char* scphd_last_error(const scphd_t* f) { return f ? f->err.c_str() : "null handle"; }

int scphd_predict(scphd_t* f, double dt) {
    if (int rc_live = check_live(f)) return rc_live;
    const ScopedRange range("scphd predict");
    HIPTRY(hipSetDevice(f->dev));
    int rc = materialize(f);
    if (rc) return rc;
    HIPTRY(launch_parent_predict(f->stream, f->ext.p, f->Np, f->models.cam, dt, rig_keys(f, SC_PREDICT),
                                 (uint64_t)f->n_predict));
    HIPTRY(launch_scale(f->stream, f->gw.p, f->Ng, f->prm.ps));
    f->pending = f->Ng > 0;
    f->pdt = dt;
    f->pstep = (uint64_t)f->n_predict;
    f->n_predict++;
    return SCPHD_OK;
}

int scphd_update_device(scphd_t* f, const double* u_device, const double* v_device, int32_t n_meas,
                        int32_t fixed_camera) {
    if (int rc_live = check_live(f)) return rc_live;
    if (n_meas < 0 || (n_meas > 0 && (!u_device || !v_device))) return SCPHD_ERR_ARG;
    if (f->ensemble) return fail(f, SCPHD_ERR_ARG, "an ensemble: use scphd_update_ensemble");
    HIPTRY(hipSetDevice(f->dev));
    HIPTRY(f->du.ensure(std::max(n_meas, 1)));
    HIPTRY(f->dv.ensure(std::max(n_meas, 1)));
    if (n_meas > 0 && u_device != f->du.p) {
        HIPTRY(hipMemcpyAsync(f->du.p, u_device, n_meas * sizeof(double), hipMemcpyDeviceToDevice, f->stream));
        HIPTRY(hipMemcpyAsync(f->dv.p, v_device, n_meas * sizeof(double), hipMemcpyDeviceToDevice, f->stream));
    }
    return update_common(f, n_meas, fixed_camera ? 1 : 0);
}

int scphd_update(scphd_t* f, const double* u, const double* v, int32_t n_meas, int32_t fixed_camera) {
    if (int rc_live = check_live(f)) return rc_live;
    if (n_meas < 0 || (n_meas > 0 && (!u || !v))) return SCPHD_ERR_ARG;
    if (f->ensemble) return fail(f, SCPHD_ERR_ARG, "an ensemble: use scphd_update_ensemble");
    HIPTRY(hipSetDevice(f->dev));
    HIPTRY(f->du.ensure(std::max(n_meas, 1)));
    HIPTRY(f->dv.ensure(std::max(n_meas, 1)));
    if (n_meas > 0) {
        HIPTRY(hipMemcpyAsync(f->du.p, u, n_meas * sizeof(double), hipMemcpyHostToDevice, f->stream));
        HIPTRY(hipMemcpyAsync(f->dv.p, v, n_meas * sizeof(double), hipMemcpyHostToDevice, f->stream));
    }
    return update_common(f, n_meas, fixed_camera ? 1 : 0);
}

}  /* extern "C" */

namespace {

/* an ensemble update; rig_fixed (may be null): the fixed-camera flag per
 * rig, else fixed_camera for every rig */
int update_ensemble(scphd_t* f, const double* u, const double* v, const int32_t* moff, int32_t fixed_camera,
                    const int32_t* rig_fixed) {
    if (check_handle(f) || !moff) return SCPHD_ERR_ARG;
    if (!f->ensemble) return fail(f, SCPHD_ERR_ARG, "not an ensemble: use scphd_update");
    if (int rc = check_live(f)) return rc;
    const int R = f->n_rigs;
    /* per-rig flags that differ go to the kernels; equal ones act as one */
    bool mixed = false;
    if (rig_fixed) {
        for (int r = 1; r < R; ++r) mixed = mixed || ((rig_fixed[r] != 0) != (rig_fixed[0] != 0));
        fixed_camera = rig_fixed[0] != 0;
    }
    if (moff[0] != 0) return fail(f, SCPHD_ERR_ARG, "moff[0] must be 0");
    int Mmax = 0;
    for (int r = 0; r < R; ++r) {
        if (moff[r + 1] < moff[r]) return fail(f, SCPHD_ERR_ARG, "moff must not decrease");
        Mmax = std::max(Mmax, moff[r + 1] - moff[r]);
    }
    const int32_t total = moff[R];
    if (total > 0 && (!u || !v)) return SCPHD_ERR_ARG;
    HIPTRY(hipSetDevice(f->dev));
    f->hmoff.assign(moff, moff + R + 1);
    f->hbpcs.resize(R);
    for (int r = 0; r < R; ++r)
        f->hbpcs[r] = safe_log((moff[r + 1] - moff[r]) * f->rig_prm[r].birth_w0 + f->rig_meas[r].kappa());
    f->ens_meas = total;
    HIPTRY(f->du.ensure(std::max(total, 1)));
    HIPTRY(f->dv.ensure(std::max(total, 1)));
    if (total > 0) {
        HIPTRY(hipMemcpyAsync(f->du.p, u, total * sizeof(double), hipMemcpyHostToDevice, f->stream));
        HIPTRY(hipMemcpyAsync(f->dv.p, v, total * sizeof(double), hipMemcpyHostToDevice, f->stream));
    }
    HIPTRY(hipMemcpyAsync(f->dmoff.p, f->hmoff.data(), (R + 1) * sizeof(int32_t), hipMemcpyHostToDevice,
                          f->stream));
    HIPTRY(hipMemcpyAsync(f->dbpcs.p, f->hbpcs.data(), R * sizeof(double), hipMemcpyHostToDevice, f->stream));
    if (mixed) {
        f->hfixed.resize(R);
        for (int r = 0; r < R; ++r) f->hfixed[r] = rig_fixed[r] != 0;
        HIPTRY(f->dfixed.ensure(R));
        HIPTRY(hipMemcpyAsync(f->dfixed.p, f->hfixed.data(), R * sizeof(int32_t), hipMemcpyHostToDevice, f->stream));
    }
    f->rig_fixed = mixed;
    /* mixed flags: the update runs the moving-camera path, and the kernels
     * give each rig its own camera and weight rule */
    const int rc = update_common(f, Mmax, mixed ? 0 : (fixed_camera ? 1 : 0));
    f->rig_fixed = false;
    return rc;
}

}  // namespace

extern "C" {

int scphd_update_ensemble(scphd_t* f, const double* u, const double* v, const int32_t* moff,
                          int32_t fixed_camera) {
    return update_ensemble(f, u, v, moff, fixed_camera, nullptr);
}

int scphd_update_ensemble_rigs(scphd_t* f, const double* u, const double* v, const int32_t* moff,
                               const int32_t* fixed_camera) {
    if (!fixed_camera) return SCPHD_ERR_ARG;
    return update_ensemble(f, u, v, moff, 0, fixed_camera);
}

int scphd_resample(scphd_t* f, int32_t* resampled) {
    if (int rc_live = check_live(f)) return rc_live;
    const ScopedRange range("scphd resample");
    HIPTRY(hipSetDevice(f->dev));
    int rc = materialize(f);
    if (rc) return rc;
    hipStream_t s = f->stream;
    const int Np = f->Np;
    ResampleArgs ra;
    ra.pw = f->pw.p;
    ra.ext = f->ext.p;
    ra.ext_tmp = f->ext_tmp.p;
    ra.cum = f->cum.p;
    ra.anc = f->anc.p;
    ra.newcnt = f->newcnt.p;
    ra.goff = f->goff.p;
    ra.Np = Np;
    ra.poff = f->ensemble ? f->dpoff.p : nullptr;
    ra.nrigs = f->n_rigs;
    ra.min_neff = f->prm.min_neff;
    ra.key = f->key;
    ra.keys = f->ensemble ? f->dkeys.p : nullptr;
    ra.sbase = f->ensemble && f->sbase_used ? f->dsbase.p : nullptr;
    ra.rstep = (uint64_t)f->n_resample;
    ra.flag = f->dflag.p;
    ra.rflag = f->drflag.p;
    ra.neff = f->dneff.p;
    HIPTRY(hipMemsetAsync(f->dflag.p, 0, sizeof(int32_t), s));
    HIPTRY(launch_resample(s, ra));
    HIPTRY(f->scantmp.ensure(scan_tmp_size(Np + 1) + 16));
    HIPTRY(launch_scan(s, f->newcnt.p, f->goff2.p, Np, f->scantmp.p));
    HIPTRY(launch_readout(s, f->goff2.p, Np, 0, f->dflag.p, nullptr, f->dmap));
    HIPTRY(hipStreamSynchronize(s));
    flush_retired(f);
    f->n_resample++;
    const int go = f->hmap[3];
    if (resampled) *resampled = go;
    if (!go) return SCPHD_OK;
    const int64_t Nnew = f->hmap[0];
    HIPTRY(f->gparent2.ensure(Nnew + 1));
    HIPTRY(f->gblock2.ensure(Nnew + 1));
    HIPTRY(f->gw2.ensure(Nnew + 1));
    if (f->hint_ok) HIPTRY(f->gpd2.ensure(Nnew + 1));
    HIPTRY(launch_table_resample(s, f->anc.p, f->goff.p, f->goff2.p, f->gblock.p, f->gw.p, Np, Nnew,
                                 f->gparent2.p, f->gblock2.p, f->gw2.p, f->hint_ok ? f->gpd.p : nullptr,
                                 f->hint_ok ? f->gpd2.p : nullptr));
    std::swap(f->goff, f->goff2);
    std::swap(f->gparent, f->gparent2);
    std::swap(f->gblock, f->gblock2);
    std::swap(f->gw, f->gw2);
    if (f->hint_ok) std::swap(f->gpd, f->gpd2);
    f->Ng = Nnew;
    return SCPHD_OK;
}

/* ---- resampling without device allocation ----
 * The scratch of scphd_resample_noalloc: a static device array, one per GPU
 * the library runs on, held until the process ends.  Host code never passes
 * the array's name to a kernel (on the host it names a stand-in, not a
 * device address): the device address is looked up once per device with
 * hipGetSymbolAddress and kept here. */
static constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)8 << 20;
static __device__ __align__(256) char scratch[NOALLOC_SCRATCH_BYTES];

struct ScratchAddrs {
    std::mutex mu;
    std::map<int, char*> by_dev;
};

static ScratchAddrs& scratch_addrs() {
    static ScratchAddrs* a = new ScratchAddrs();
    return *a;
}

/* device address of `scratch` on dev, which must be the current device */
static hipError_t scratch_address(int dev, char** out) {
    ScratchAddrs& sa = scratch_addrs();
    std::lock_guard<std::mutex> lk(sa.mu);
    auto it = sa.by_dev.find(dev);
    if (it == sa.by_dev.end()) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        it = sa.by_dev.emplace(dev, (char*)p).first;
    }
    *out = it->second;
    return hipSuccess;
}

/* byte offsets into the scratch for n parents: cumulative weights,
 * gathered extrinsics, ancestors, new map sizes, scan partials, flags */
struct NoAllocLayout {
    size_t cum, ext_tmp, anc, newcnt, scantmp, flag, bytes;
    explicit NoAllocLayout(int n) {
        auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
        cum = 0;
        ext_tmp = up(cum + (size_t)n * sizeof(double));
        anc = up(ext_tmp + (size_t)n * 12 * sizeof(double));
        newcnt = up(anc + (size_t)n * sizeof(int32_t));
        scantmp = up(newcnt + ((size_t)n + 1) * sizeof(int32_t));
        flag = up(scantmp + ((size_t)scan_tmp_size((int64_t)n + 1) + 16) * sizeof(int32_t));
        bytes = up(flag + 4 * sizeof(int32_t));
    }
};

int scphd_resample_noalloc(scphd_t* f, int32_t* resampled) {
    if (check_handle(f)) return SCPHD_ERR_ARG;
    if (f->ensemble) return fail(f, SCPHD_ERR_ARG, "scphd_resample_noalloc takes a plain filter, not an ensemble");
    const int Np = f->Np;
    const NoAllocLayout ly(Np);
    if (ly.bytes > NOALLOC_SCRATCH_BYTES)
        return fail(f, SCPHD_ERR_CAPACITY, "scphd_resample_noalloc: " + std::to_string(Np) +
                                               " parents need more than the static scratch of 8 MiB (" +
                                               std::to_string(ly.bytes) + " bytes)");
    HIPTRY(hipSetDevice(f->dev));
    char* base = nullptr;
    HIPTRY(scratch_address(f->dev, &base));
    /* a deferred predict is applied into the other generation, which must
     * already hold the map */
    if (f->pending && f->Ng > 0 && f->gen[1 - f->cur].cap_blocks < f->Ng)
        return fail(f, SCPHD_ERR_CAPACITY,
                    "scphd_resample_noalloc: the deferred predict needs a larger sample generation (scphd_reserve)");
    int rc = materialize(f);
    if (rc) return rc;
    hipStream_t s = f->stream;
    ResampleArgs ra;
    ra.pw = f->pw.p;
    ra.ext = f->ext.p;
    ra.ext_tmp = (double*)(base + ly.ext_tmp);
    ra.cum = (double*)(base + ly.cum);
    ra.anc = (int32_t*)(base + ly.anc);
    ra.newcnt = (int32_t*)(base + ly.newcnt);
    ra.goff = f->goff.p;
    ra.Np = Np;
    ra.min_neff = f->prm.min_neff;
    ra.key = f->key;
    ra.rstep = (uint64_t)f->n_resample;
    ra.flag = (int32_t*)(base + ly.flag);
    ra.rflag = f->drflag.p;
    ra.neff = f->dneff.p;
    ra.defer = 1;
    /* the scratch keeps the last call's contents: clear the flag */
    HIPTRY(launch_fill_i32(s, ra.flag, 0, 1));
    HIPTRY(launch_resample(s, ra));
    HIPTRY(launch_scan(s, ra.newcnt, f->goff2.p, Np, (int32_t*)(base + ly.scantmp)));
    HIPTRY(launch_readout(s, f->goff2.p, Np, 0, ra.flag, nullptr, f->dmap));
    HIPTRY(hipStreamSynchronize(s));
    const int go = f->hmap[3];
    const int64_t Nnew = f->hmap[0];
    if (go && (Nnew + 1 > f->gparent2.n || Nnew + 1 > f->gblock2.n || Nnew + 1 > f->gw2.n ||
               (f->hint_ok && Nnew + 1 > f->gpd2.n)))
        return fail(f, SCPHD_ERR_CAPACITY, "scphd_resample_noalloc: the resampled map table needs " +
                                               std::to_string(Nnew + 1) + " entries, more than the handle holds " +
                                               "(scphd_reserve); the filter is unchanged");
    HIPTRY(launch_resample_commit(s, f->ext.p, ra.ext_tmp, f->pw.p, f->anc.p, ra.anc, Np, go));
    f->n_resample++;
    if (resampled) *resampled = go;
    if (go) {
        HIPTRY(launch_table_resample(s, ra.anc, f->goff.p, f->goff2.p, f->gblock.p, f->gw.p, Np, Nnew,
                                     f->gparent2.p, f->gblock2.p, f->gw2.p, f->hint_ok ? f->gpd.p : nullptr,
                                     f->hint_ok ? f->gpd2.p : nullptr));
        std::swap(f->goff, f->goff2);
        std::swap(f->gparent, f->gparent2);
        std::swap(f->gblock, f->gblock2);
        std::swap(f->gw, f->gw2);
        if (f->hint_ok) std::swap(f->gpd, f->gpd2);
        f->Ng = Nnew;
    }
    /* the caller may reuse the scratch (another handle on this GPU) once
     * this returns */
    HIPTRY(hipStreamSynchronize(s));
    return SCPHD_OK;
}

int scphd_check(scphd_t* f) {
    if (int rc_live = check_live(f)) return rc_live;
    HIPTRY(hipSetDevice(f->dev));
    HIPTRY(launch_nan_count(f->stream, f->pw.p, f->Np, f->dmap + 8));
    HIPTRY(hipStreamSynchronize(f->stream));
    if (f->hmap[8] != 0) return fail(f, SCPHD_ERR_NAN, "nan weight detected");
    return SCPHD_OK;
}

int scphd_sync(scphd_t* f) {
    if (check_handle(f)) return SCPHD_ERR_ARG;
    HIPTRY(hipSetDevice(f->dev));
    HIPTRY(hipStreamSynchronize(f->stream));
    return SCPHD_OK;
}

int scphd_get_counts(scphd_t* f, scphd_counts* c) {
    if (check_handle(f) || !c) return SCPHD_ERR_ARG;
    c->n_particles = f->Np;
    c->particles_per_feature = f->S;
    c->n_gaussians = f->Ng;
    c->n_samples = f->Ng * f->S;
    c->n_predict = f->n_predict;
    c->n_update = f->n_update;
    c->n_resample = f->n_resample;
    c->pending_predict = f->pending ? 1 : 0;
    return SCPHD_OK;
}

int scphd_get_parents(scphd_t* f, double* states, double* weights) {
    if (int rc_live = check_live(f)) return rc_live;
    HIPTRY(hipSetDevice(f->dev));
    if (states)
        HIPTRY(hipMemcpyAsync(states, f->ext.p, (size_t)f->Np * 12 * sizeof(double), hipMemcpyDeviceToHost,
                              f->stream));
    if (weights)
        HIPTRY(hipMemcpyAsync(weights, f->pw.p, (size_t)f->Np * sizeof(double), hipMemcpyDeviceToHost, f->stream));
    HIPTRY(hipStreamSynchronize(f->stream));
    return SCPHD_OK;
}

int scphd_get_maps(scphd_t* f, int32_t* gcount, double* wg, double* samples) {
    if (int rc_live = check_live(f)) return rc_live;
    HIPTRY(hipSetDevice(f->dev));
    int rc = materialize(f);
    if (rc) return rc;
    hipStream_t s = f->stream;
    if (gcount) {
        std::vector<int32_t> off(f->Np + 1);
        HIPTRY(hipMemcpyAsync(off.data(), f->goff.p, off.size() * sizeof(int32_t), hipMemcpyDeviceToHost, s));
        HIPTRY(hipStreamSynchronize(s));
        for (int p = 0; p < f->Np; ++p) gcount[p] = off[p + 1] - off[p];
    }
    if (wg && f->Ng > 0)
        HIPTRY(hipMemcpyAsync(wg, f->gw.p, (size_t)f->Ng * sizeof(double), hipMemcpyDeviceToHost, s));
    if (samples && f->Ng > 0) {
        const int64_t n = f->Ng * f->S;
        void* dv = nullptr;
        size_t db = (size_t)n * 7 * sizeof(double);
        HIPTRY(cache_alloc(&dv, &db));
        double* d = (double*)dv;
        hipError_t e = f->esz == 8 ? launch_extract<double>(s, f->gen[f->cur].planes<double>(), f->gblock.p, f->Ng,
                                                            f->S, d)
                                   : launch_extract<float>(s, f->gen[f->cur].planes<float>(), f->gblock.p, f->Ng,
                                                           f->S, d);
        if (e == hipSuccess)
            e = hipMemcpyAsync(samples, d, (size_t)n * 7 * sizeof(double), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        retire(&f->retired, dv, db);
        if (e != hipSuccess) return hipfail(f, e, "get_maps");
    }
    HIPTRY(hipStreamSynchronize(s));
    flush_retired(f);
    return SCPHD_OK;
}

int scphd_set_state(scphd_t* f, const double* states, const double* weights, const int32_t* gcount,
                    int64_t n_gaussians, const double* wg, const double* samples, int64_t n_predict,
                    int64_t n_update, int64_t n_resample) {
    if (int rc_live = check_live(f)) return rc_live;
    if (!states || !weights || !gcount || n_gaussians < 0) return SCPHD_ERR_ARG;
    if (n_gaussians > 0 && (!wg || !samples)) return SCPHD_ERR_ARG;
    HIPTRY(hipSetDevice(f->dev));
    const int Np = f->Np, S = f->S;
    std::vector<int32_t> off(Np + 1, 0);
    std::vector<int32_t> par((size_t)std::max<int64_t>(n_gaussians, 1));
    for (int p = 0; p < Np; ++p) {
        if (gcount[p] < 0) return fail(f, SCPHD_ERR_ARG, "negative gaussian count");
        off[p + 1] = off[p] + gcount[p];
        for (int k = 0; k < gcount[p] && off[p] + k < n_gaussians; ++k) par[off[p] + k] = p;
    }
    if (off[Np] != n_gaussians) return fail(f, SCPHD_ERR_ARG, "gcount does not sum to n_gaussians");
    hipStream_t s = f->stream;
    const int64_t Ng = n_gaussians;
    HIPTRY(f->goff.ensure(Np + 1));
    HIPTRY(f->gparent.ensure(Ng + 1));
    HIPTRY(f->gblock.ensure(Ng + 1));
    HIPTRY(f->gw.ensure(Ng + 1));
    f->hint_ok = false;
    HIPTRY(hipMemcpyAsync(f->ext.p, states, (size_t)Np * 12 * sizeof(double), hipMemcpyHostToDevice, s));
    HIPTRY(hipMemcpyAsync(f->pw.p, weights, (size_t)Np * sizeof(double), hipMemcpyHostToDevice, s));
    HIPTRY(hipMemcpyAsync(f->goff.p, off.data(), (size_t)(Np + 1) * sizeof(int32_t), hipMemcpyHostToDevice, s));
    if (Ng > 0) {
        HIPTRY(hipMemcpyAsync(f->gparent.p, par.data(), (size_t)Ng * sizeof(int32_t), hipMemcpyHostToDevice, s));
        HIPTRY(hipMemcpyAsync(f->gw.p, wg, (size_t)Ng * sizeof(double), hipMemcpyHostToDevice, s));
        HIPTRY(launch_iota(s, f->gblock.p, Ng));
        HIPTRY(f->gen[f->cur].ensure(Ng, 0, S, f->esz, s));
        const int64_t n = Ng * S;
        void* dv = nullptr;
        size_t db = (size_t)n * 7 * sizeof(double);
        HIPTRY(cache_alloc(&dv, &db));
        double* d = (double*)dv;
        hipError_t e = hipMemcpyAsync(d, samples, (size_t)n * 7 * sizeof(double), hipMemcpyHostToDevice, s);
        if (e == hipSuccess)
            e = f->esz == 8 ? launch_inject<double>(s, d, Ng, S, f->gen[f->cur].planes<double>())
                            : launch_inject<float>(s, d, Ng, S, f->gen[f->cur].planes<float>());
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        retire(&f->retired, dv, db);
        if (e != hipSuccess) return hipfail(f, e, "set_state");
    }
    HIPTRY(hipStreamSynchronize(s));
    flush_retired(f);
    f->Ng = Ng;
    f->blocks_used = Ng;
    f->n_predict = n_predict;
    f->n_update = n_update;
    f->n_resample = n_resample;
    f->pending = false;
    return SCPHD_OK;
}

int scphd_get_update_stats(scphd_t* f, scphd_update_stats* st) {
    if (check_handle(f) || !st) return SCPHD_ERR_ARG;
    *st = f->stats;
    return SCPHD_OK;
}

int scphd_get_update_debug(scphd_t* f, double* eta, double* loglik, int32_t* nmerged, int32_t* nnondetect,
                           int32_t* nsurvivors) {
    if (int rc_live = check_live(f)) return rc_live;
    HIPTRY(hipSetDevice(f->dev));
    hipStream_t s = f->stream;
    const int Np = f->Np;
    /* an update of the current parents has run, and scphd_compact has not
     * released its outputs since */
    if (!f->has_outputs || f->nmerged.n < Np || f->outcnt.n < Np || f->nsurv.n < Np ||
        (f->last_M > 0 && f->eta.n < (int64_t)Np * f->last_M))
        return fail(f, SCPHD_ERR_STATE, "no update outputs for the current parents");
    if (eta && f->last_M > 0)
        HIPTRY(hipMemcpyAsync(eta, f->eta.p, (size_t)Np * f->last_M * sizeof(double), hipMemcpyDeviceToHost, s));
    if (loglik) HIPTRY(hipMemcpyAsync(loglik, f->loglik.p, (size_t)Np * sizeof(double), hipMemcpyDeviceToHost, s));
    std::vector<int32_t> nm(Np), oc(Np);
    HIPTRY(hipMemcpyAsync(nm.data(), f->nmerged.p, (size_t)Np * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIPTRY(hipMemcpyAsync(oc.data(), f->outcnt.p, (size_t)Np * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    if (nsurvivors)
        HIPTRY(hipMemcpyAsync(nsurvivors, f->nsurv.p, (size_t)Np * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIPTRY(hipStreamSynchronize(s));
    for (int p = 0; p < Np; ++p) {
        if (nmerged) nmerged[p] = nm[p];
        if (nnondetect) nnondetect[p] = oc[p] - nm[p];
    }
    return SCPHD_OK;
}

int scphd_get_fit_debug(scphd_t* f, double* out) {
    if (int rc_live = check_live(f)) return rc_live;
    if (!out) return SCPHD_ERR_ARG;
    HIPTRY(hipSetDevice(f->dev));
    const int64_t n = f->stats.n_gaussians_in;
    for (int q = 0; q < FP_COUNT && n > 0; ++q)
        HIPTRY(hipMemcpyAsync(out + q * n, f->fit.p + q * n, (size_t)n * sizeof(double), hipMemcpyDeviceToHost,
                              f->stream));
    HIPTRY(hipStreamSynchronize(f->stream));
    return SCPHD_OK;
}

int scphd_get_ancestors(scphd_t* f, int32_t* ancestors) {
    if (int rc_live = check_live(f)) return rc_live;
    if (!ancestors) return SCPHD_ERR_ARG;
    HIPTRY(hipSetDevice(f->dev));
    HIPTRY(hipMemcpyAsync(ancestors, f->anc.p, (size_t)f->Np * sizeof(int32_t), hipMemcpyDeviceToHost, f->stream));
    HIPTRY(hipStreamSynchronize(f->stream));
    return SCPHD_OK;
}

int scphd_set_timing(scphd_t* f, int32_t enabled) {
    if (check_handle(f)) return SCPHD_ERR_ARG;
    f->timing = enabled != 0;
    return SCPHD_OK;
}

int scphd_get_phase_ms(scphd_t* f, double* ms8) {
    if (check_handle(f) || !ms8) return SCPHD_ERR_ARG;
    for (int i = 0; i < 8; ++i) ms8[i] = f->phase_ms[i];
    return SCPHD_OK;
}

int scphd_debug_set_limits(scphd_t* f, int32_t merge_lds_cap, int32_t eta_heavy_cap, int32_t eta_ncand_cap,
                           int32_t merge_force_rounds, int32_t merge_all_pairs) {
    if (check_handle(f)) return SCPHD_ERR_ARG;
    if (merge_lds_cap < -1 || eta_heavy_cap < -1 || eta_ncand_cap < -1 || eta_heavy_cap > ETA_HEAVY ||
        eta_ncand_cap > ETA_NCAND)
        return fail(f, SCPHD_ERR_ARG, "debug limits out of range");
    f->dbg_lds_cap = merge_lds_cap;
    f->dbg_heavy = eta_heavy_cap;
    f->dbg_ncand = eta_ncand_cap;
    f->dbg_rounds = merge_force_rounds != 0;
    f->dbg_all_pairs = merge_all_pairs != 0;
    return SCPHD_OK;
}

int scphd_debug_normals(uint64_t seed, uint32_t stream, uint64_t step, uint32_t parent, uint32_t gauss,
                        int32_t n_elem, double* out) {
    if (n_elem < 0 || n_elem > (1 << 24) || (n_elem > 0 && !out)) return SCPHD_ERR_ARG;
    if (n_elem == 0) return SCPHD_OK;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return SCPHD_ERR_NODEVICE;
    double* d = nullptr;
    const size_t bytes = (size_t)n_elem * 6 * sizeof(double);
    if (hipMalloc((void**)&d, bytes) != hipSuccess) return SCPHD_ERR_HIP;
    hipError_t e = launch_debug_normals(nullptr, NoiseKey::from_seed(seed), stream, step, parent, gauss, n_elem, d);
    if (e == hipSuccess) e = hipMemcpy(out, d, bytes, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    return e == hipSuccess ? SCPHD_OK : SCPHD_ERR_HIP;
}

int scphd_set_noise(scphd_t* f, int32_t mode, uint64_t seed, const double* injected, size_t n) {
    if (check_handle(f)) return SCPHD_ERR_ARG;
    if (f->ensemble) return fail(f, SCPHD_ERR_ARG, "an ensemble's keys are fixed by scphd_create_ensemble");
    if (mode != 0 || injected || n)
        return fail(f, SCPHD_ERR_ARG, "only mode 0 (Philox streams keyed by seed) is provided");
    f->prm.seed = seed;
    f->key = NoiseKey::from_seed(seed);
    return SCPHD_OK;
}

int scphd_get_estimate(scphd_t* f, double* mean12, double* cov144, double* cardinality) {
    if (int rc_live = check_live(f)) return rc_live;
    HIPTRY(hipSetDevice(f->dev));
    HIPTRY(launch_estimate(f->stream, f->ext.p, f->pw.p, f->goff.p, f->gw.p, f->Np, nullptr, 1,
                           (double*)((char*)f->dmap + MAP_EST)));
    HIPTRY(hipStreamSynchronize(f->stream));
    const double* h = (const double*)((const char*)f->hmap + MAP_EST);
    if (mean12) std::memcpy(mean12, h, 12 * sizeof(double));
    if (cov144) std::memcpy(cov144, h + 12, 144 * sizeof(double));
    if (cardinality) *cardinality = h[156];
    return SCPHD_OK;
}

int scphd_get_rig_estimates(scphd_t* f, double* mean12, double* cov144, double* cardinality) {
    if (int rc_live = check_live(f)) return rc_live;
    HIPTRY(hipSetDevice(f->dev));
    HIPTRY(launch_estimate(f->stream, f->ext.p, f->pw.p, f->goff.p, f->gw.p, f->Np,
                           f->ensemble ? f->dpoff.p : nullptr, f->n_rigs,
                           (double*)((char*)f->dmap + MAP_EST)));
    HIPTRY(hipStreamSynchronize(f->stream));
    const double* h = (const double*)((const char*)f->hmap + MAP_EST);
    for (int r = 0; r < f->n_rigs; ++r, h += EST_DOUBLES) {
        if (mean12) std::memcpy(mean12 + 12 * r, h, 12 * sizeof(double));
        if (cov144) std::memcpy(cov144 + 144 * r, h + 12, 144 * sizeof(double));
        if (cardinality) cardinality[r] = h[156];
    }
    return SCPHD_OK;
}

int scphd_get_rig_resample(scphd_t* f, int32_t* resampled, double* neff) {
    if (int rc_live = check_live(f)) return rc_live;
    if (f->n_resample == 0) return fail(f, SCPHD_ERR_STATE, "no resampling step yet");
    HIPTRY(hipSetDevice(f->dev));
    if (resampled)
        HIPTRY(hipMemcpyAsync(resampled, f->drflag.p, f->n_rigs * sizeof(int32_t), hipMemcpyDeviceToHost, f->stream));
    if (neff) HIPTRY(hipMemcpyAsync(neff, f->dneff.p, f->n_rigs * sizeof(double), hipMemcpyDeviceToHost, f->stream));
    HIPTRY(hipStreamSynchronize(f->stream));
    return SCPHD_OK;
}

}  /* extern "C" */

namespace {

/* the map's clouds into the spare generation, sized for the map alone, which
 * then becomes the current one */
template <typename T>
int compact_map(scphd_filter* f) {
    const int nxt = 1 - f->cur;
    HIPTRY(f->gen[nxt].ensure(f->Ng, 0, f->S, f->esz, f->stream, true));
    HIPTRY(launch_gather<T>(f->stream, f->gen[f->cur].planes<T>(), f->gen[nxt].planes<T>(), f->gblock.p, f->Ng,
                            f->S));
    HIPTRY(launch_iota(f->stream, f->gblock.p, f->Ng));
    f->cur = nxt;
    f->blocks_used = f->Ng;
    return SCPHD_OK;
}

}  // namespace

extern "C" {

int scphd_compact(scphd_t* f) {
    if (check_handle(f)) return SCPHD_ERR_ARG;
    if (int rc = between_frames(f, "scphd_compact")) return rc;
    HIPTRY(hipSetDevice(f->dev));
    hipStream_t s = f->stream;
    HIPTRY(hipStreamSynchronize(s));
    HIPTRY(hipStreamSynchronize(f->side));
    flush_retired(f);
    /* the spare generation first, so that the compacted map can use its memory */
    f->gen[1 - f->cur].release();
    flush_retired(f);
    if (f->Ng > 0) {
        int rc = f->esz == 8 ? compact_map<double>(f) : compact_map<float>(f);
        if (rc) return rc;
        HIPTRY(hipStreamSynchronize(s));
    }
    /* the generation the map came from, and the work arrays of an update
     * (each update sizes them again); the debug outputs of the last update
     * go with them */
    f->gen[1 - f->cur].release();
    for (auto* b : {&f->fit, &f->pre, &f->lik, &f->eta, &f->sw, &f->smu, &f->mw, &f->mmu, &f->mP}) b->release();
    for (auto* b : {&f->keep, &f->keepscan, &f->ndlist, &f->holes, &f->redo, &f->ndfill, &f->mcnt, &f->mncand, &f->mcand, &f->sg,
                    &f->sm, &f->glead, &f->gsta, &f->gend, &f->gmem, &f->gparent2, &f->gblock2, &f->nsurv,
                    &f->survoff, &f->nmerged, &f->mergeoff, &f->outcnt, &f->ngroups})
        b->release();
    f->gw2.release();
    f->gpd2.release();
    f->mscratch.release();
    f->last_M = 0;
    f->has_outputs = false;
    f->stats = scphd_update_stats{};
    flush_retired(f); /* nothing queued uses them */
    return SCPHD_OK;
}

int scphd_reserve(scphd_t* f, int64_t n_gaussians, int32_t max_meas) {
    if (check_handle(f) || n_gaussians < 0 || max_meas < 0) return SCPHD_ERR_ARG;
    HIPTRY(hipSetDevice(f->dev));
    const int64_t G = std::max<int64_t>(n_gaussians, f->Ng);
    const int64_t Np = f->Np, M = std::max(max_meas, 1);
    /* the current generation keeps every block (after resampling the table
     * may reference any of them); the other one holds nothing between calls */
    Generation& cur = f->gen[f->cur];
    HIPTRY(cur.ensure(G, cur.cap_blocks, f->S, f->esz, f->stream));
    HIPTRY(f->gen[1 - f->cur].ensure(G, 0, f->S, f->esz, f->stream));
    HIPTRY(f->fit.ensure((int64_t)FP_COUNT * G));
    HIPTRY(f->pre.ensure((int64_t)PP_COUNT * G));
    HIPTRY(f->lik.ensure((int64_t)LK_COUNT * (G + 3)));
    HIPTRY(f->keep.ensure(G));
    HIPTRY(f->keepscan.ensure(G + 1));
    HIPTRY(f->redo.ensure(G + 1));
    HIPTRY(f->ndfill.ensure(G + 1));
    HIPTRY(f->ndlist.ensure(G));
    HIPTRY(f->holes.ensure(G));
    HIPTRY(f->scantmp.ensure(scan_tmp_size(std::max<int64_t>(G, Np) + 1) + 16));
    /* the successor table, and the current one keeping its entries, so that
     * the table swaps of later updates and resamplings find both sized */
    HIPTRY(f->gparent2.ensure(G + 1));
    HIPTRY(f->gblock2.ensure(G + 1));
    HIPTRY(f->gw2.ensure(G + 1));
    HIPTRY(f->gpd2.ensure(G + 1));
    HIPTRY(grow_keep(f->gparent, G + 1, f->Ng, f->stream));
    HIPTRY(grow_keep(f->gblock, G + 1, f->Ng, f->stream));
    HIPTRY(grow_keep(f->gw, G + 1, f->Ng, f->stream));
    HIPTRY(grow_keep(f->gpd, G + 1, f->Ng, f->stream));
    /* survivors: about one per Gaussian in steady state */
    HIPTRY(f->sw.ensure(G + 1));
    HIPTRY(f->sg.ensure(G + 1));
    HIPTRY(f->sm.ensure(G + 1));
    const int64_t scap = f->sw.n;
    HIPTRY(f->smu.ensure(6 * scap));
    HIPTRY(f->mw.ensure(scap));
    HIPTRY(f->mmu.ensure(6 * scap));
    HIPTRY(f->mP.ensure(21 * scap));
    for (DBuf<int32_t>* b : {&f->glead, &f->gsta, &f->gend, &f->gmem}) HIPTRY(b->ensure(G + 1));
    /* the global-scratch merge workspace, sized like the survivor lists (a
     * growth between frames would otherwise allocate in the update) */
    HIPTRY(f->mscratch.ensure(merge_scratch_bytes(G + 1, (int)Np)));
    /* per (parent, measurement) and the measurement sets */
    HIPTRY(f->eta.ensure(Np * M));
    HIPTRY(f->mcnt.ensure(Np * M));
    HIPTRY(f->mncand.ensure(Np * M));
    HIPTRY(f->mcand.ensure(Np * M * ETA_NCAND));
    HIPTRY(f->du.ensure((int64_t)f->n_rigs * M));
    HIPTRY(f->dv.ensure((int64_t)f->n_rigs * M));
    HIPTRY(hipStreamSynchronize(f->stream));
    flush_retired(f);
    return SCPHD_OK;
}

int scphd_get_memory(scphd_t* f, int64_t* bytes) {
    if (check_handle(f) || !bytes) return SCPHD_ERR_ARG;
    int64_t b = 0;
    for_each_buf(f, [&b](auto& d) { b += (int64_t)d.bytes; });
    for (const auto& g : f->gen)
        for (size_t x : g.pbytes) b += (int64_t)x;
    *bytes = b;
    return SCPHD_OK;
}

}  /* extern "C" */
