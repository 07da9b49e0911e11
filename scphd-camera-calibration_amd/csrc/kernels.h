/*
 * kernels.h -- launch interface of the gfx950 kernels of the SC-PHD filter
 * step.  Every launcher enqueues on the given stream and returns the HIP
 * status of the launch.
 *
 * Device data layout (all arrays structure-of-arrays, HBM resident):
 *   parents   ext[Np*12] (parent-major Extrinsics), pw[Np] linear weights,
 *             pose[Np*12] (R row-major 9, t 3) of the camera used by the
 *             current update.
 *   map table one entry per Gaussian in canonical order (parent ascending):
 *             goff[Np+1] offsets, gparent[Ng], gblock[Ng] sample block,
 *             gw[Ng] Gaussian weight (FP64).
 *   samples   a generation buffer of 7 planes (x y z vx vy vz w) of T
 *             (float or double); block b holds samples [b*S, (b+1)*S).
 *             Blocks may be shared by several table entries after
 *             resampling; kernels never write a block they read.
 */
#ifndef SCPHD_KERNELS_H
#define SCPHD_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device.h"

namespace scphd {

/* Noise keys: one rig (keys == nullptr: every parent uses `key`), or an
 * ensemble of rigs, rig r owning parents [poff[r], poff[r + 1]) and
 * prig[p] the rig of parent p; a parent draws with its rig's key and its
 * index inside the rig, so that every rig draws exactly the streams of a
 * filter of its own. */
/* Step-counter classes of the noise streams: the camera and landmark
 * prediction draw with the predict count, posterior sampling with the update
 * count, resampling with the resample count. */
enum StepCounter { SC_PREDICT = 0, SC_UPDATE = 1, SC_RESAMPLE = 2, SC_COUNT = 3 };

struct RigKeys {
    NoiseKey key;
    const NoiseKey* keys = nullptr;
    const int32_t* prig = nullptr;
    const int32_t* poff = nullptr;
    /* ensemble: per rig, SC_COUNT step counts of the handle when the rig
     * joined (null: every rig joined at creation); the launch draws with
     * counter class `which` */
    const uint32_t* sbase = nullptr;
    int which = SC_PREDICT;
    __device__ LaneKey of(int p, int& pl) const {
        if (!keys) {
            pl = p;
            return LaneKey{key.k0, key.k1, 0u};
        }
        const int r = prig[p];
        pl = p - poff[r];
        const NoiseKey k = keys[r];
        return LaneKey{k.k0, k.k1, sbase ? sbase[SC_COUNT * r + which] : 0u};
    }
};

/* per-Gaussian fit outputs, stride Ng (SoA planes) */
enum FitPlane { FP_MU = 0, FP_P = 6, FP_PD = 27, FP_COUNT = 28 };
/* per-Gaussian detection terms, stride Ng */
enum PrePlane {
    PP_MK = 0,  /* per Gaussian, PP_MK_LEN consecutive doubles: the fitted mean (6), then the gain (12) */
    PP_MK_LEN = 18,
    PP_PN = 18, /* updated covariance: from plane 18 on, 21 packed doubles per Gaussian, contiguous */
    PP_COUNT = 39
};
/* per-Gaussian likelihood record, LK_COUNT consecutive doubles (64 B):
 * predicted measurement, S^-1 entries (s0, s1 + s2, s3), log(pd_g w_g),
 * log(det S) / 2 and 1 when the updated covariance is PD, else 0 */
enum LikField { LK_MUU = 0, LK_MUV, LK_S0, LK_S12, LK_S3, LK_C, LK_H, LK_OK, LK_COUNT = 8 };

template <typename T>
struct Planes {
    T* p[7];
};

struct FilterModels {
    DisparityMeasurementModel meas;
    OrientedLinearCVMotionModel3D cam;
    LinearCVMotionModel3D feat;
};

struct BirthModel {
    double mean[6]; /* u, v filled per measurement; d0 vu0 vv0 vd0 */
    double P[21];   /* packed diagonal */
    double logw0;
    int pd_ok;
};

/* ---- prediction ---- */
hipError_t launch_parent_predict(hipStream_t s, double* ext, int Np, OrientedLinearCVMotionModel3D cam,
                                 double dt, RigKeys key, uint64_t step);
hipError_t launch_scale(hipStream_t s, double* x, int64_t n, double a);
/* one rig's Np parents (weights 1 / Np) */
hipError_t launch_prior(hipStream_t s, double* ext, double* pw, int Np, const double* a12, const double* b12,
                        NoiseKey key);
/* rfixed (ensembles, may be null): per rig, the fixed camera for that rig
 * (rig of parent p: prig[p]); else `fixed` for every parent */
hipError_t launch_pose(hipStream_t s, const double* ext, double* pose, int Np, int fixed,
                       const int32_t* prig = nullptr, const int32_t* rfixed = nullptr);

/* Apply a deferred sample prediction, writing every table entry to its own
 * block (block id = table index) of the destination generation. */
template <typename T>
hipError_t launch_materialize(hipStream_t s, Planes<T> src, Planes<T> dst, const int32_t* gblock,
                              const int32_t* gparent, const int32_t* goff, int64_t Ng, int S,
                              LinearCVMotionModel3D feat, double dt, RigKeys key, uint64_t step,
                              bool arith32 = false);

/* ---- update ---- */
struct FitArgs {
    const int32_t* gblock;
    const int32_t* gparent;
    const int32_t* goff;
    const double* gw;
    const double* pose;
    int64_t Ng;
    int S;
    DisparityMeasurementModel meas;
    const DisparityMeasurementModel* rig_meas = nullptr; /* ensemble: per rig (rigs may differ), else null */
    LinearCVMotionModel3D feat;
    int pending;
    double dt;
    RigKeys key;
    uint64_t pstep;
    double min_weight;
    double* fit;     /* FP_COUNT planes */
    int32_t* keep;   /* non-detection term kept */
    int32_t* redo;     /* Gaussians whose sample weights were all zero, refitted by k_fit_fix */
    int32_t* redo_n;   /* their number; zero before k_fit */
    int32_t* ndfill;   /* kept terms whose clouds k_nd_fill writes (decided after the sweep) */
    int32_t* ndfill_n; /* their number; zero before k_fit */
    int64_t nd_base; /* destination block of the non-detection cloud of entry g: nd_base + g */
    const float* gpd = nullptr; /* per entry: pd_g of its source term at the last update (1: posterior term) */
    int off32;       /* every source and destination plane is below 4 GiB: 32-bit byte offsets */
    int arith32 = 0; /* per-sample arithmetic in single precision (FP32 storage only) */
};

/* k_fit: the streaming sweep; k_fit_fix: the rare cases it lists in redo[];
 * k_nd_fill: the non-detection clouds it lists in ndfill[] */
template <typename T>
hipError_t launch_fit(hipStream_t s, const FitArgs& a, Planes<T> src, Planes<T> dst);
template <typename T>
hipError_t launch_fit_fix(hipStream_t s, const FitArgs& a, Planes<T> src, Planes<T> dst);
template <typename T>
hipError_t launch_nd_fill(hipStream_t s, const FitArgs& a, Planes<T> src, Planes<T> dst);

hipError_t launch_precompute(hipStream_t s, const double* fit, const double* gw, int64_t Ng, double var_u,
                             double var_v, double* pre, double* lik);

struct EtaArgs {
    const int32_t* goff;
    const double* gw;
    const double* pre; /* PP_COUNT planes */
    const double* lik; /* LK_COUNT doubles per Gaussian */
    int64_t Ng;        /* plane stride */
    const double* zu;
    const double* zv;
    int M;             /* the measurement count; in an ensemble the largest, the stride of eta & co. */
    int Np;
    double bpc;        /* log(M w0 + kappa) */
    const int32_t* moff = nullptr; /* ensemble: rig r's measurements at [moff[r], moff[r + 1]); null for one rig */
    const double* bpcs = nullptr;  /* ensemble: bpc per rig */
    const int32_t* prig = nullptr; /* ensemble: the rig of each parent */
    double min_weight;
    const BirthModel* births; /* per rig (one for a plain filter), device memory */
    double* eta;       /* Np*M */
    double* loglik;    /* Np   */
    int32_t* nsurv;    /* Np   */
    int32_t* mcnt;     /* Np*M detection terms kept per (parent, measurement) */
    int32_t* mncand;   /* Np*M candidates per (parent, measurement): terms that may be kept */
    int32_t* mcand;    /* Np*M*ETA_NCAND their map indices (ascending), when at most ETA_NCAND */
    double lw_gate;    /* log(min_weight) - 1e-3: below it exp() < min_weight for certain */
    int heavy_cap;     /* at most ETA_HEAVY: more Gaussians above bpc than this -> whole-map max pass */
    int ncand_cap;     /* at most ETA_NCAND: more candidates than this -> whole-map sweeps */
};
static constexpr int ETA_NCAND = 8;
static constexpr int ETA_HEAVY = 2048; /* LDS list of the Gaussians that can exceed bpc */
hipError_t launch_eta(hipStream_t s, const EtaArgs& a);

struct SurvArgs {
    EtaArgs e;
    const double* fit;
    const int32_t* survoff; /* Np+1 */
    double* sw;
    int32_t* sg;            /* table index, or -1 - rig for a birth term */
    int32_t* sm;
    double* smu;            /* 6 consecutive doubles per survivor */
    int64_t cap;
    int32_t* err;           /* set to 1 when a parent's count differs from nsurv */
};
hipError_t launch_survivors(hipStream_t s, const SurvArgs& a);

struct MergeArgs {
    const int32_t* survoff;
    const double* sw;
    const int32_t* sg;
    const int32_t* sm;
    const double* smu;
    int64_t scap;
    const double* pre; /* updated covariances of the table entries (PP_PN records) */
    int64_t Ng;
    const BirthModel* births; /* per rig, device memory: birth terms' covariance */
    double min_separation;
    int Np;
    int32_t lds_cap;   /* parents with more survivors use global scratch */
    int32_t force_rounds; /* nonzero: every parent takes the round loop (verification of that path) */
    int32_t all_pairs;    /* nonzero: candidates from all pairs, as for ill-conditioned survivors (verification) */
    int32_t big_lds_up;   /* big parents with next_pow2(U) <= this sort in the launch's LDS (12 B per slot) */
    int32_t big_lds_bytes; /* the big launch's LDS: the sort keys, then the walk's per-survivor arrays (<= 64 KB) */
    char* scratch;     /* global scratch for big parents: merge_scratch_bytes(Stot, Np) */
    double* mw;        /* merged outputs at survoff[p] + k */
    double* mmu;       /* 6 planes, stride scap */
    double* mP;        /* 21 planes, stride scap */
    int32_t* nmerged;  /* Np: merged count */
    int32_t* ngroups;  /* Np: groups formed by the walk, -1 when the round loop is needed */
    int32_t* glead;    /* per survivor slot: group leader (survivor index) */
    int32_t* gsta;     /* per survivor slot: group member range [gsta, gend) in gmem */
    int32_t* gend;
    int32_t* gmem;     /* per survivor slot: group members (survivor indices) */
    int64_t Stot;
};
/* Greedy merge of every parent's survivors: groups (sort, neighbour lists,
 * walk), moments per group, and the round loop for parents whose
 * neighbour lists overflow the workspace. */
hipError_t launch_merge(hipStream_t s, const MergeArgs& a, int lds_bytes);
int merge_lds_bytes(int cap);
int64_t merge_scratch_bytes(int64_t Stot, int Np);
hipError_t launch_iota(hipStream_t s, int32_t* x, int64_t n, int32_t start = 0);
/* Removal of a rig's parents [P0, P0 + Nr) and their Gaussians [G0, G0 + Gr)
 * from an ensemble's map table: the entries after them move down, their
 * parent indices by Nr (gpd and gpd_out may be null) */
hipError_t launch_table_remove(hipStream_t s, const int32_t* gparent, const int32_t* gblock, const double* gw,
                               const float* gpd, int64_t G0, int64_t Gr, int64_t Ng_new, int P1, int Nr,
                               int32_t* gparent_out, int32_t* gblock_out, double* gw_out, float* gpd_out);
/* goff without the rig's parents: goff_out[p] = goff[p] below P0, goff[p + Nr] - Gr from P0 (Np_new + 1 entries) */
hipError_t launch_goff_remove(hipStream_t s, const int32_t* goff, int P0, int Nr, int32_t Gr, int Np_new,
                              int32_t* goff_out);
/* x[i] -= Nr where x[i] >= P1 (ancestor indices after a removal), n entries */
hipError_t launch_index_shift(hipStream_t s, int32_t* x, int64_t n, int P1, int Nr);
/* out[0..4] (device-visible host memory): off[n], the largest segment of
 * off, the number of segments above limit, *x0, *x1 */
hipError_t launch_readout(hipStream_t s, const int32_t* off, int n, int32_t limit, const int32_t* x0,
                          const int32_t* x1, int32_t* out);
/* *out = the number of NaN entries of w[0, n) */
hipError_t launch_nan_count(hipStream_t s, const double* w, int n, int32_t* out);
/* *out = sum over rigs r of (goff[poff[r+1]] - goff[poff[r]]) (moff[r+1] - moff[r]) */
hipError_t launch_pair_count(hipStream_t s, const int32_t* goff, const int32_t* moff, const int32_t* poff,
                             int nrigs, double* out);

/* Per-rig launches: one workgroup per rig, rig r's parents [poff[r],
 * poff[r + 1]); poff == nullptr: one rig of Np parents. */
/* rfixed (ensembles, may be null): rigs whose frame used the fixed camera
 * keep their weights */
hipError_t launch_parent_weights(hipStream_t s, double* pw, const double* loglik, int Np, const int32_t* poff,
                                 int nrigs, const int32_t* rfixed = nullptr);

/* out[6 n]: six normals of every element 0..n-1 of one noise stream */
hipError_t launch_debug_normals(hipStream_t s, NoiseKey key, uint32_t stream, uint64_t step, uint32_t parent,
                                uint32_t gauss, int n, double* out);

/* per rig, EST_DOUBLES at out + EST_DOUBLES r: weighted parent mean (12),
 * covariance (144, row-major), summed Gaussian weight of the first
 * highest-weight parent */
static constexpr int EST_DOUBLES = 157;
hipError_t launch_estimate(hipStream_t s, const double* ext, const double* pw, const int32_t* goff,
                           const double* gw, int Np, const int32_t* poff, int nrigs, double* out);

struct SampleArgs {
    const int32_t* mergeoff; /* Np+1 */
    const int32_t* survoff;
    const double* mmu;
    const double* mP;
    int64_t scap;
    const double* pose;
    int Np;
    int64_t Kt;
    int S;
    DisparityMeasurementModel meas;
    const DisparityMeasurementModel* rig_meas = nullptr; /* ensemble: per rig, else null */
    RigKeys key;
    uint64_t ustep;
    int64_t post_base;
    /* merged Gaussian gm's block: holes[gm] while gm < nholes (the blocks
     * of dropped non-detection terms), else post_base + gm - nholes */
    const int32_t* holes = nullptr;
    int64_t nholes = 0;
    int off32; /* every destination offset is below 4 GiB: 32-bit byte offsets */
    int arith32 = 0; /* draws and disparity -> world in single precision (FP32 storage only) */
};
template <typename T>
hipError_t launch_sample(hipStream_t s, const SampleArgs& a, Planes<T> dst);

hipError_t launch_outcount(hipStream_t s, const int32_t* nmerged, const int32_t* goff, const int32_t* keepscan,
                           int Np, int32_t* outcnt);
/* ndlist[keepscan[g]] = g for the kept entries; with holes, also
 * holes[g - keepscan[g]] = g for the dropped ones */
hipError_t launch_ndlist(hipStream_t s, const int32_t* keep, const int32_t* keepscan, int64_t Ng, int32_t* ndlist,
                         int32_t* holes = nullptr);

struct TableArgs {
    const int32_t* outoff;   /* Np+1 */
    const int32_t* mergeoff; /* Np+1 */
    const int32_t* survoff;  /* Np+1 */
    const int32_t* goff;     /* input table offsets */
    const int32_t* keepscan; /* Ng+1 */
    const int32_t* ndlist;
    const double* mw;
    const double* gw;        /* input table weights */
    const double* fit;       /* FP_PD plane */
    int64_t Ng_in;
    int Np;
    int64_t Nout;
    int64_t post_base;
    const int32_t* holes = nullptr; /* posterior blocks as in SampleArgs */
    int64_t nholes = 0;
    int64_t nd_base;
    int32_t* gparent;
    int32_t* gblock;
    double* gwo;
    float* gpd;     /* per entry: the source term's pd_g (non-detection entries), 1 (posterior entries) */
};
hipError_t launch_table(hipStream_t s, const TableArgs& a);

/* ---- resampling ---- */
struct ResampleArgs {
    double* pw;
    double* ext;
    double* ext_tmp;
    double* cum;
    int32_t* anc;
    int32_t* newcnt;
    const int32_t* goff;
    int Np;                /* parents of the one rig when poff is null */
    const int32_t* poff = nullptr; /* ensemble: rig r's parents [poff[r], poff[r + 1]) */
    int nrigs = 1;         /* one workgroup per rig */
    double min_neff;
    NoiseKey key;          /* one rig */
    const NoiseKey* keys = nullptr; /* ensemble: per rig, or nullptr */
    const uint32_t* sbase = nullptr; /* ensemble: per rig, the handle's step counts at joining (SC_COUNT each) */
    uint64_t rstep;
    int32_t* flag;  /* out: 1 if any rig resampled (zero before the launch) */
    int32_t* rflag = nullptr; /* out: per rig, 1 if resampled (may be null) */
    double* neff;   /* out: per rig */
    int defer = 0;  /* nonzero: leave ext and pw unchanged (k_resample_commit applies the gather) */
};
hipError_t launch_resample(hipStream_t s, const ResampleArgs& a);
/* Apply a deferred resampling of one rig's N parents: anc_out <- anc, and
 * when go, ext <- ext_tmp and pw <- 1/N. */
hipError_t launch_resample_commit(hipStream_t s, double* ext, const double* ext_tmp, double* pw, int32_t* anc_out,
                                  const int32_t* anc, int N, int go);
hipError_t launch_fill_i32(hipStream_t s, int32_t* x, int32_t v, int64_t n);
hipError_t launch_table_resample(hipStream_t s, const int32_t* anc, const int32_t* goff_old,
                                 const int32_t* goff_new, const int32_t* gblock_old, const double* gw_old,
                                 int Np, int64_t Nnew, int32_t* gparent, int32_t* gblock, double* gw,
                                 const float* gpd_old, float* gpd);

/* ---- read-back ---- */
template <typename T>
hipError_t launch_extract(hipStream_t s, Planes<T> src, const int32_t* gblock, int64_t Ng, int S, double* out);
/* the clouds of the Ng map entries into blocks 0 .. Ng-1 of dst */
template <typename T>
hipError_t launch_gather(hipStream_t s, Planes<T> src, Planes<T> dst, const int32_t* gblock, int64_t Ng, int S);
template <typename T>
hipError_t launch_inject(hipStream_t s, const double* in, int64_t Ng, int S, Planes<T> dst);

/* ---- scan ---- */
/* out[0..n] = exclusive prefix sums of in[0..n), out[n] = total.
 * tmp needs scan_tmp_size(n) ints. */
int64_t scan_tmp_size(int64_t n);
hipError_t launch_scan(hipStream_t s, const int32_t* in, int32_t* out, int64_t n, int32_t* tmp);

}  // namespace scphd

#endif
