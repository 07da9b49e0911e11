/*
 * kernels.hip -- gfx950 kernels of the SC-PHD camera calibration step.
 *
 * Reference: cheesinglee/scphd-camera-calibration.  Each kernel names the
 * reference code whose semantics it implements; none is a translation of it:
 * the reference launches ~40 thrust/CUDA kernels per step over dense
 * N_g x M intermediates, this file fuses them into per-Gaussian wavefront
 * kernels and per-parent workgroup kernels that never materialise the
 * dense (Gaussian, measurement) array.
 */
#include "kernels.h"

#include <float.h>

#include <algorithm>

#include <type_traits>

namespace scphd {

static constexpr int NT = 256; /* threads per workgroup for per-parent kernels */
static constexpr int WPB = 4;  /* wavefronts (Gaussians) per workgroup for per-Gaussian kernels */

/* ------------------------------------------------------------ reductions */
/* DPP lane moves on the two 32-bit halves of a double (full EXEC assumed). */
template <int CTRL>
__device__ inline double dpp_f64(double x) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(x), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(x), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

template <int CTRL>
__device__ inline float dpp_f32(float x) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), CTRL, 0xF, 0xF, false));
}

__device__ inline double readlane_f64(double x, int l) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(x), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(x), l);
    return __hiloint2double(hi, lo);
}

/* Sum over a full wavefront: four DPP steps inside each 16-lane row (swap
 * neighbours, swap pairs, half-row mirror, row mirror) leave every lane
 * with its row's sum; the four row sums are read into scalar registers and
 * added.  Every lane gets the result.  All 64 lanes must be active. */
__device__ inline double wave_sum_dpp(double x) {
    x += dpp_f64<0xB1>(x);  /* quad_perm [1,0,3,2] */
    x += dpp_f64<0x4E>(x);  /* quad_perm [2,3,0,1] */
    x += dpp_f64<0x141>(x); /* row_half_mirror */
    x += dpp_f64<0x140>(x); /* row_mirror */
    return (readlane_f64(x, 0) + readlane_f64(x, 16)) + (readlane_f64(x, 32) + readlane_f64(x, 48));
}

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ inline double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

/* Sum over a workgroup of NT threads; every thread gets the result. */
__device__ inline double block_sum(double v, double* sh) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double r = 0.0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) r += sh[i];
    return r;
}

__device__ inline double block_max(double v, double* sh) {
    v = wave_max(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double r = -INFINITY;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) r = fmax(r, sh[i]);
    return r;
}

/* Exclusive scan over a workgroup of NT threads; *total gets the sum. */
__device__ inline int block_exscan(int v, int* sh, int* total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    __syncthreads();
    if (lane == 63) sh[w] = x;
    __syncthreads();
    int base = 0, tot = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) {
        if (i < w) base += sh[i];
        tot += sh[i];
    }
    *total = tot;
    return base + x - v;
}

__device__ inline int upper_parent(const int32_t* off, int n, int64_t x) {
    /* largest p in [0, n) with off[p] <= x, for nondecreasing off[0..n] with off[n] > x */
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if ((int64_t)off[mid] <= x)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

/* ------------------------------------------------------------ prediction */
struct Bounds12 {
    double a[12], b[12];
};

/* Uniform prior over [a, b] per extrinsic component, weights 1/N
 * (src/scphdcameracalibration.cu:835-980). */
__global__ void k_prior(double* ext, double* pw, int Np, Bounds12 ab, NoiseKey key) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= Np) return;
    for (int c = 0; c < 12; ++c) {
        uint32_t x[4];
        noise_block(key, NS_PRIOR, 0, (uint32_t)p, 0, 0, (uint32_t)(c / 2), x);
        const double r = (c % 2 == 0) ? uniform_co(x[0], x[1]) : uniform_co(x[2], x[3]);
        ext[(int64_t)p * 12 + c] = r * (ab.b[c] - ab.a[c]) + ab.a[c];
    }
    pw[p] = 1.0 / Np;
}

/* Parent extrinsics through OrientedLinearCVMotionModel3D::computeNoisyMotion
 * with N(0, std^2) accelerations (predict_camera functor,
 * src/scphdcameracalibration.cu:129-152, :1062-1100). */
__global__ void k_parent_predict(double* ext, int Np, OrientedLinearCVMotionModel3D cam, double dt, RigKeys rk,
                                 uint64_t step) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= Np) return;
    double a[6];
    if (!rk.keys) { /* one rig: the key stays in scalar registers */
        normals<6>(rk.key, NS_CAMERA, step, (uint32_t)p, 0, 0, a);
    } else {
        int pl;
        const LaneKey key = rk.of(p, pl);
        normals<6>(key, NS_CAMERA, step, (uint32_t)pl, 0, 0, a);
    }
    double* e = ext + (int64_t)p * 12;
    EuclideanPoint c{e[0], e[1], e[2], e[3], e[4], e[5]};
    EuclideanPoint g{e[6], e[7], e[8], e[9], e[10], e[11]};
    cam.computeNoisyMotion(c, g, dt, a[0] * cam.std_ax_cartesian(), a[1] * cam.std_ay_cartesian(),
                           a[2] * cam.std_az_cartesian(), a[3] * cam.std_ax_angular(), a[4] * cam.std_ay_angular(),
                           a[5] * cam.std_az_angular());
    e[0] = c.x; e[1] = c.y; e[2] = c.z; e[3] = c.vx; e[4] = c.vy; e[5] = c.vz;
    e[6] = g.x; e[7] = g.y; e[8] = g.z; e[9] = g.vx; e[10] = g.vy; e[11] = g.vz;
}

__global__ void k_scale(double* x, int64_t n, double a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] *= a;
}

/* Camera used by an update: the parent's extrinsics, or all-zero extrinsics
 * for the fixed camera (src/scphdcameracalibration.cu:1167-1176). */
__global__ void k_pose(const double* ext, double* pose, int Np, int fixed, const int32_t* prig,
                       const int32_t* rfixed) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= Np) return;
    if (rfixed) fixed = rfixed[prig[p]];
    CameraPose c;
    if (fixed) {
        c = CameraPose::identity();
    } else {
        const double* e = ext + (int64_t)p * 12;
        Extrinsics x{{e[0], e[1], e[2], e[3], e[4], e[5]}, {e[6], e[7], e[8], e[9], e[10], e[11]}};
        c = CameraPose::fromExtrinsics(x);
    }
    double* o = pose + (int64_t)p * 12;
    for (int i = 0; i < 9; ++i) o[i] = c.R[i];
    for (int i = 0; i < 3; ++i) o[9 + i] = c.t[i];
}

__device__ inline CameraPose load_pose(const double* pose, int p) {
    CameraPose c;
    const double* o = pose + (int64_t)p * 12;
#pragma unroll
    for (int i = 0; i < 9; ++i) c.R[i] = o[i];
#pragma unroll
    for (int i = 0; i < 3; ++i) c.t[i] = o[9 + i];
    return c;
}

/* Landmark sample prediction (predict_features functor,
 * src/scphdcameracalibration.cu:154-178): the constant-velocity step with
 * acceleration draws a[i] * std_i (LinearCVMotionModel3D::computeNoisyMotion),
 * x + v dt + a std dt^2 / 2 and v + a std dt, the products std dt^2 / 2 and
 * std dt formed once per kernel (loop invariants) instead of per sample. */
__device__ inline EuclideanPoint noisy_motion(EuclideanPoint x, const LinearCVMotionModel3D& feat, double dt,
                                              const double a[3]) {
    const double h = 0.5 * dt * dt;
    x.x = fma(a[0], feat.std_ax() * h, fma(x.vx, dt, x.x));
    x.y = fma(a[1], feat.std_ay() * h, fma(x.vy, dt, x.y));
    x.z = fma(a[2], feat.std_az() * h, fma(x.vz, dt, x.z));
    x.vx = fma(a[0], feat.std_ax() * dt, x.vx);
    x.vy = fma(a[1], feat.std_ay() * dt, x.vy);
    x.vz = fma(a[2], feat.std_az() * dt, x.vz);
    return x;
}

/* the prediction of the sample (p, k, s) */
template <typename Key>
__device__ inline EuclideanPoint predict_sample(EuclideanPoint x, const LinearCVMotionModel3D& feat, double dt,
                                                Key key, uint64_t step, uint32_t p, uint32_t k, uint32_t s) {
    double a[3];
    feature_normals(key, step, p, k, s, a);
    return noisy_motion(x, feat, dt, a);
}

/* the same for the members of a noise group in order (phase = member) */
template <typename Key>
__device__ inline EuclideanPoint predict_sample_seq(EuclideanPoint x, const LinearCVMotionModel3D& feat, double dt,
                                                    Key key, uint64_t step, uint32_t p, uint32_t k, uint32_t s,
                                                    uint32_t phase, FeatureCarry& c) {
    double a[3];
    feature_normals_seq(key, step, p, k, s, phase, c, a);
    return noisy_motion(x, feat, dt, a);
}

template <typename T>
__global__ void k_materialize(Planes<T> src, Planes<T> dst, const int32_t* gblock, const int32_t* gparent,
                              const int32_t* goff, int64_t Ng, int S, LinearCVMotionModel3D feat, double dt,
                              RigKeys rk, uint64_t step, int arith32) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Ng * S) return;
    const int64_t g = i / S;
    const int s = (int)(i - g * S);
    const int p = gparent[g];
    const int k = (int)(g - goff[p]);
    const int64_t si = (int64_t)gblock[g] * S + s;
    const int64_t di = g * S + s;
    if (arith32) { /* the single-precision prediction of the FP32-arithmetic variant */
        float x[6], n[3];
#pragma unroll
        for (int c = 0; c < 6; ++c) x[c] = (float)src.p[c][si];
        if (!rk.keys) {
            feature_normals(rk.key, step, (uint32_t)p, (uint32_t)k, (uint32_t)s, n);
        } else {
            int pl;
            const LaneKey key = rk.of(p, pl);
            feature_normals(key, step, (uint32_t)pl, (uint32_t)k, (uint32_t)s, n);
        }
        predict_f(x, motion_f(feat, dt), n);
#pragma unroll
        for (int c = 0; c < 6; ++c) dst.p[c][di] = (T)x[c];
        dst.p[6][di] = src.p[6][si];
        return;
    }
    EuclideanPoint x{(double)src.p[0][si], (double)src.p[1][si], (double)src.p[2][si],
                     (double)src.p[3][si], (double)src.p[4][si], (double)src.p[5][si]};
    if (!rk.keys) { /* one rig: the key stays in scalar registers */
        x = predict_sample(x, feat, dt, rk.key, step, (uint32_t)p, (uint32_t)k, (uint32_t)s);
    } else {
        int pl;
        const LaneKey key = rk.of(p, pl);
        x = predict_sample(x, feat, dt, key, step, (uint32_t)pl, (uint32_t)k, (uint32_t)s);
    }
    dst.p[0][di] = (T)x.x; dst.p[1][di] = (T)x.y; dst.p[2][di] = (T)x.z;
    dst.p[3][di] = (T)x.vx; dst.p[4][di] = (T)x.vy; dst.p[5][di] = (T)x.vz;
    dst.p[6][di] = src.p[6][si];
}

/* ------------------------------------------------------- reproject + fit */
/* Sum over the 16 lanes of a row (four DPP steps: swap neighbours, swap
 * pairs, half-row mirror, row mirror); every lane of the row gets the sum.
 * The row must be fully active; rows are independent. */
__device__ inline double row_sum16(double x) {
    x += dpp_f64<0xB1>(x);
    x += dpp_f64<0x4E>(x);
    x += dpp_f64<0x141>(x);
    x += dpp_f64<0x140>(x);
    return x;
}
__device__ inline float row_sum16(float x) {
    x += dpp_f32<0xB1>(x);
    x += dpp_f32<0x4E>(x);
    x += dpp_f32<0x141>(x);
    x += dpp_f32<0x140>(x);
    return x;
}

/* One sample of a Gaussian's cloud: the deferred prediction applied, its
 * disparity-space point, its field-of-view flag and stored weight. */
/* Plane element at byte offset `off` from a uniform plane base.  With a
 * 32-bit offset the address is a scalar base plus one per-lane offset
 * register (the global_load saddr form); 64-bit offsets serve planes of
 * 4 GiB and more. */
template <typename T, typename Off>
__device__ inline T& at(T* base, Off off) {
    return *(T*)((char*)base + off);
}

template <typename T, typename Off>
__device__ inline void load_sample_off(const Planes<T>& src, Off off, T r[7]) {
#pragma unroll
    for (int c = 0; c < 7; ++c) r[c] = at(src.p[c], off);
}

template <typename T, typename Off>
__device__ inline void write_nd_off(const DisparityMeasurementModel& meas, Planes<T>& dst, Off off,
                                    const EuclideanPoint& x, bool in) {
    at(dst.p[0], off) = (T)x.x;
    at(dst.p[1], off) = (T)x.y;
    at(dst.p[2], off) = (T)x.z;
    at(dst.p[3], off) = (T)x.vx;
    at(dst.p[4], off) = (T)x.vy;
    at(dst.p[5], off) = (T)x.vz;
    at(dst.p[6], off) = (T)(1.0 - (in ? meas.pd() : 0.0));
}

template <typename T>
__device__ inline void load_sample(const Planes<T>& src, int64_t si, T r[7]) {
#pragma unroll
    for (int c = 0; c < 7; ++c) r[c] = src.p[c][si];
}

template <typename T, typename Key>
__device__ inline void eval_sample(const FitArgs& a, const DisparityMeasurementModel& meas, const T r[7], Key kk,
                                   int pl, int k, int s, const CameraPose& pose, EuclideanPoint& x, DisparityPoint& d,
                                   bool& in, double& w, FeatureCarry* fc = nullptr, uint32_t phase = 0) {
    x = EuclideanPoint{(double)r[0], (double)r[1], (double)r[2], (double)r[3], (double)r[4], (double)r[5]};
    w = (double)r[6];
    if (a.pending) {
        if (fc)
            x = predict_sample_seq(x, a.feat, a.dt, kk, a.pstep, (uint32_t)pl, (uint32_t)k, (uint32_t)s, phase, *fc);
        else
            x = predict_sample(x, a.feat, a.dt, kk, a.pstep, (uint32_t)pl, (uint32_t)k, (uint32_t)s);
    }
    d = disparity_fast(meas, pose, x);
    in = meas.inRange(d);
}

template <typename T, typename Key>
__device__ inline void fit_sample(const FitArgs& a, const DisparityMeasurementModel& meas, const Planes<T>& src,
                                  int64_t si, Key kk, int pl, int k, int s, const CameraPose& pose,
                                  EuclideanPoint& x, DisparityPoint& d, bool& in, double& w) {
    T r[7];
    load_sample(src, si, r);
    eval_sample(a, meas, r, kk, pl, k, s, pose, x, d, in, w);
}

/* The FP32-arithmetic form of eval_sample: the sample's world point x (after
 * the deferred prediction), its disparity-space point q and field-of-view
 * flag, and its stored weight, all in single precision. */
struct SampleCtxF {
    MeasF m;
    PoseF pose;
    MotionF mo;
};

template <typename T, typename Key>
__device__ inline void eval_sample_f(const FitArgs& a, const SampleCtxF& cx, const T r[7], Key kk, int pl, int k,
                                     int s, float x[6], float q[6], bool& in, float& w, FeatureCarry* fc = nullptr,
                                     uint32_t phase = 0) {
#pragma unroll
    for (int c = 0; c < 6; ++c) x[c] = (float)r[c];
    w = (float)r[6];
    if (a.pending) {
        float n[3];
        if (fc)
            feature_normals_seq(kk, a.pstep, (uint32_t)pl, (uint32_t)k, (uint32_t)s, phase, *fc, n);
        else
            feature_normals(kk, a.pstep, (uint32_t)pl, (uint32_t)k, (uint32_t)s, n);
        predict_f(x, cx.mo, n);
    }
    disparity_f(cx.m, cx.pose, x, q);
    in = in_range_f(cx.m, q);
}

template <typename T, typename Off>
__device__ inline void write_nd_off_f(const DisparityMeasurementModel& meas, Planes<T>& dst, Off off, const float x[6],
                                      bool in) {
#pragma unroll
    for (int c = 0; c < 6; ++c) at(dst.p[c], off) = (T)x[c];
    at(dst.p[6], off) = (T)(1.0 - (in ? meas.pd() : 0.0));
}

__device__ inline void write_fit(const FitArgs& a, int64_t g, int lane, const double* mu, const double* P, double pdg) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int q = lane + 16 * h;
        double v = 0.0;
#pragma unroll
        for (int c = 0; c < 6; ++c)
            if (q == FP_MU + c) v = mu[c];
#pragma unroll
        for (int e = 0; e < 21; ++e)
            if (q == FP_P + e) v = P[e];
        if (q == FP_PD) v = pdg;
        if (q < FP_COUNT) a.fit[(int64_t)q * a.Ng + g] = v;
    }
}

/* Sixteen lanes (one row of a wavefront) per Gaussian, four Gaussians per
 * wavefront: deferred prediction, world -> disparity with the parent's
 * camera and the field-of-view Pd (worldToDisparityKernel,
 * src/kernels.cu:6-47), per-Gaussian Pd (computePdKernel :82-99), sample
 * reweighting by pd and normalisation (src/scphdcameracalibration.cu:1602-1608,
 * normalizeWeightsKernel src/kernels.cu:996-1018), and the weighted moment
 * fit (fitGaussiansKernel :102-438, covariance normaliser (W^2 + sum w^2) / W).
 *
 * Samples stream through once: each lane accumulates, with weights
 * w' = w pd, the sums of w', w'^2, w' (q - K) and w' (q - K)(q - K)^T about
 * a shift K, from which the normalised mean and centred covariance follow;
 * 30 row sums by DPP.  K is the first sample with w' > 0 (the first sample
 * when there is none).
 *
 * The non-detection cloud (kept when w_g (1 - pd_g) >= min_weight,
 * computeNonDetections src/scphdcameracalibration.cu:1290-1420; sample
 * weights 1 - pd_i; the samples themselves are the predicted cloud, which
 * the reference's world -> disparity -> world round trip reproduces to
 * rounding) is written during the sweep when the term is kept for certain,
 * w_g (1 - Pd) >= min_weight (pd_g <= Pd).  A term that may go either way
 * is decided after the sweep and, when kept, listed for k_nd_fill, which
 * writes its cloud; most such terms are dropped, and their clouds are never
 * written.  A cloud whose weights are all zero, which the reference
 * normalises to weight 1 per sample, is listed in redo[] and refitted by
 * k_fit_fix when pd_g > 0; with pd_g = 0 its fit cannot reach any output and
 * is the first sample with zero covariance. */
/* The noise key of parent p and p's index in its rig: the kernel argument
 * (in scalar registers) for one rig, the rig's key per lane in an ensemble. */
template <bool RG>
__device__ inline std::conditional_t<RG, LaneKey, NoiseKey> row_key(const RigKeys& rk, int p, int& pl) {
    if constexpr (RG) {
        return rk.of(p, pl);
    } else {
        pl = p;
        return rk.key;
    }
}

/* The measurement model of parent p's rig, `m`: the kernel argument itself
 * for one rig (RG false, the model stays in scalar registers), else a copy
 * of the rig's entry of per_rig. */
template <bool RG>
struct RowMeas;
template <>
struct RowMeas<false> {
    const DisparityMeasurementModel& m;
    __device__ RowMeas(const DisparityMeasurementModel& one, const DisparityMeasurementModel*, const RigKeys&, int)
        : m(one) {}
};
template <>
struct RowMeas<true> {
    DisparityMeasurementModel m;
    __device__ RowMeas(const DisparityMeasurementModel&, const DisparityMeasurementModel* per_rig, const RigKeys& rk,
                       int p)
        : m(per_rig[rk.prig[p]]) {}
};

/* waves per SIMD the fit kernel is compiled for (register budget 512 / n) */
#ifndef SCPHD_FIT_WAVES
#define SCPHD_FIT_WAVES 2
#endif

/* A, the arithmetic type of the per-sample work: double, or float for the
 * FP32-arithmetic variant (FP32 storage), whose moment sums are single
 * precision up to the row sums and whose in-view count is an integer. */
template <typename T, typename Off, bool RG, typename A = double>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(SCPHD_FIT_WAVES, SCPHD_FIT_WAVES))) void k_fit(
    FitArgs a, Planes<T> src, Planes<T> dst) {
    constexpr bool F32 = std::is_same_v<A, float>;
    static_assert(!F32 || std::is_same_v<T, float>, "single-precision arithmetic needs FP32 storage");
    /* the rows' camera poses are staged through LDS once and then held in
     * registers for the whole sweep */
    __shared__ double s_pose[16 * 12];
    const int lane = threadIdx.x & 15, row = threadIdx.x >> 4;
    const int64_t g = (int64_t)blockIdx.x * 16 + row;
    const bool act = g < a.Ng;
    const int p = act ? a.gparent[g] : 0;
    if (act && lane < 12) s_pose[row * 12 + lane] = a.pose[(int64_t)p * 12 + lane];
    __syncthreads();
    if (!act) return; /* whole rows; rows never exchange data */
    CameraPose pose;
#pragma unroll
    for (int i = 0; i < 9; ++i) pose.R[i] = s_pose[row * 12 + i];
#pragma unroll
    for (int i = 0; i < 3; ++i) pose.t[i] = s_pose[row * 12 + 9 + i];
    const int k = (int)(g - a.goff[p]);
    int pl;
    const auto kk = row_key<RG>(a.key, p, pl); /* the rig's key, p's index in its rig */
    const RowMeas<RG> rm(a.meas, a.rig_meas, a.key, p);
    const DisparityMeasurementModel& meas = rm.m;
    const int64_t b0 = (int64_t)a.gblock[g] * a.S;
    const int S = a.S;
    const double wg = a.gw[g];
    const double Pd = meas.pd();
    const bool nd_sure = wg * (1.0 - Pd) >= a.min_weight; /* kept for certain */
    /* the cloud is written in the sweep when the term is kept for certain, or
     * when it would be kept at the pd_g of its source term at the last update
     * (a non-detection term's cloud moves little in one step): a guess that
     * costs only the write of a cloud nothing references if it proves wrong */
    const bool nd_write = nd_sure || (a.gpd && wg * (1.0 - (double)a.gpd[g]) >= a.min_weight);
    const int64_t d0 = (a.nd_base + g) * (int64_t)S;
    SampleCtxF cx;
    if constexpr (F32) cx = SampleCtxF{meas_f(meas), pose_f(pose), motion_f(a.feat, a.dt)};
    const A PdA = (A)Pd;

    [[maybe_unused]] double spd = 0.0; /* sum of pd_i (double arithmetic) */
    [[maybe_unused]] float nin = 0.0f; /* in-view samples (single-precision arithmetic) */
    A sw = 0, sw2 = 0, m1[6], m2[21], K[6];
    bool kw = false; /* K is a weighted sample (uniform along the row) */
    FeatureCarry fc; /* prediction normals carried between the steps of a noise group */
#pragma unroll
    for (int c = 0; c < 6; ++c) m1[c] = 0.0;
#pragma unroll
    for (int e = 0; e < 21; ++e) m2[e] = 0.0;
    /* three sample buffers: a step consumes one and loads the step two
     * ahead into the one its predecessor consumed.  Full steps go in groups
     * of three with the buffers in place, so no register copy waits on a
     * load in flight; the last steps (fewer than three, or a partial one)
     * rotate the buffers.  Full steps carry no per-lane sample guard. */
    T bA[7], bB[7], bC[7];
    Off sof = (Off)((b0 + lane) * (int64_t)sizeof(T)); /* byte offsets, advanced 16 samples a step */
    Off dof = (Off)((d0 + lane) * (int64_t)sizeof(T));
    const Off last = (Off)((b0 + S - 1) * (int64_t)sizeof(T)); /* the cloud's last sample */
    if (lane < S) load_sample_off(src, sof, bA);
    sof += (Off)(16 * sizeof(T));
    if (lane + 16 < S) load_sample_off(src, sof, bB);
    /* one step of 16 samples, s0 .. s0 + 15, from `cur`, loading the step
     * two ahead into `fill`; FULL: all of them exist.  That load is
     * unconditional (no branch in the loop): past the end of the cloud it
     * reads the last sample again. */
    auto step = [&](auto full, const T(&cur)[7], T(&fill)[7], int s0) {
        constexpr bool FULL = decltype(full)::value;
        const int s = s0 + lane;
        const bool valid = FULL || s < S;
        sof += (Off)(16 * sizeof(T));
        load_sample_off(src, s + 32 < S ? sof : last, fill);
        A q[6] = {0, 0, 0, 0, 0, 0};
        bool in = false;
        A w = 0;
        [[maybe_unused]] EuclideanPoint x;
        [[maybe_unused]] float xf[6];
        if constexpr (F32) {
            if (valid) eval_sample_f(a, cx, cur, kk, pl, k, s, xf, q, in, w, &fc, (uint32_t)(s0 >> 4) & 3u);
        } else {
            DisparityPoint d;
            if (valid) {
                eval_sample(a, meas, cur, kk, pl, k, s, pose, x, d, in, w, &fc, (uint32_t)(s0 >> 4) & 3u);
                q[0] = d.u; q[1] = d.v; q[2] = d.d; q[3] = d.vu; q[4] = d.vv; q[5] = d.vd;
            }
        }
        const A pd = in ? PdA : (A)0;
        const A wp = valid ? w * pd : (A)0;
        if (s0 == 0) {
            /* shift: the Gaussian's first sample, broadcast along the row
             * (the fit of a cloud with no weighted sample) */
#pragma unroll
            for (int c = 0; c < 6; ++c) K[c] = __shfl(q[c], 0, 16);
        }
        if (__ballot(!kw)) {
            /* rows still without a weighted sample take the first one
             * that appears as the shift: every sum is exactly zero until
             * then, and a shift inside the weighted cloud keeps the
             * one-pass moments exact to rounding (a single weighted
             * sample gives zero covariance exactly) */
            const uint32_t rb = (uint32_t)(__ballot(wp > (A)0) >> (threadIdx.x & 48)) & 0xFFFFu;
            if (!kw && rb != 0) {
                const int l = __builtin_ctz(rb);
#pragma unroll
                for (int c = 0; c < 6; ++c) K[c] = __shfl(q[c], l, 16);
                kw = true;
            }
        }
        if constexpr (F32)
            nin += in ? 1.0f : 0.0f;
        else
            spd += pd;
        sw += wp;
        sw2 += wp * wp;
        A dv[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) dv[c] = valid ? q[c] - K[c] : (A)0;
#pragma unroll
        for (int c = 0; c < 6; ++c) m1[c] += wp * dv[c];
#pragma unroll
        for (int r = 0; r < 6; ++r)
#pragma unroll
            for (int c = 0; c <= r; ++c) m2[pidx(r, c)] += wp * dv[r] * dv[c];
        if (nd_write && valid) {
            if constexpr (F32)
                write_nd_off_f(meas, dst, dof, xf, in);
            else
                write_nd_off(meas, dst, dof, x, in);
        }
        dof += (Off)(16 * sizeof(T));
    };
    const std::integral_constant<bool, true> full;
    const std::integral_constant<bool, false> guarded;
    int s0 = 0;
    for (; s0 + 48 <= S; s0 += 48) { /* three full steps, buffers in place */
        step(full, bA, bC, s0);
        step(full, bB, bA, s0 + 16);
        step(full, bC, bB, s0 + 32);
    }
    for (; s0 < S; s0 += 16) { /* the last steps: buffers rotate */
        step(guarded, bA, bC, s0);
#pragma unroll
        for (int c = 0; c < 7; ++c) {
            bA[c] = bB[c];
            bB[c] = bC[c];
        }
    }
    /* the row sums, then the rest in double */
    if constexpr (F32)
        spd = (double)row_sum16(nin) * Pd;
    else
        spd = row_sum16(spd);
    const double sw_r = (double)row_sum16(sw);
    const double sw2_r = (double)row_sum16(sw2);
    double m1r[6], m2r[21];
#pragma unroll
    for (int c = 0; c < 6; ++c) m1r[c] = (double)row_sum16(m1[c]);
#pragma unroll
    for (int e = 0; e < 21; ++e) m2r[e] = (double)row_sum16(m2[e]);
    const double pdg = spd / S;

    /* normalised weights w'_i / sum w': mean K + sum w (q - K); centred
     * covariance sum w (q - K)(q - K)^T - (2 - W) dm dm^T with W = sum w */
    const double r = sw_r > 0 ? 1.0 / sw_r : 0.0;
    const double W = sw_r * r, W2 = sw2_r * r * r;
    double mu[6], dm[6], P[21];
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        dm[c] = m1r[c] * r;
        mu[c] = (double)K[c] + dm[c];
    }
    /* W = 0: with pd_g = 0 the fit is the first sample with zero covariance
     * (its detection terms have weight 0); with pd_g > 0 (possible only when
     * Pd = 1) k_fit_fix refits it with unit weights, as the reference does */
    const double scale = W > 0 ? (W * W + W2) / W : 0.0;
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = 0; j <= i; ++j) P[pidx(i, j)] = (m2r[pidx(i, j)] * r - (2.0 - W) * dm[i] * dm[j]) * scale;
    write_fit(a, g, lane, mu, P, pdg);

    const bool keep = wg * (1 - pdg) >= a.min_weight;
    if (lane == 0) {
        a.keep[g] = keep ? 1 : 0;
        if (sw_r == 0 && spd != 0.0) a.redo[atomicAdd(a.redo_n, 1)] = (int32_t)g;
        if (keep && !nd_write) a.ndfill[atomicAdd(a.ndfill_n, 1)] = (int32_t)g;
    }
}

/* The non-detection clouds of the terms k_fit decided only after its sweep
 * and kept, listed in ndfill[0, *ndfill_n): the predicted samples and the
 * weights 1 - pd_i, computed by the same code as in the sweep.  One 16-lane
 * row per listed term; workgroups stride over the list (its length is read
 * on the device). */
template <typename T, typename Off, bool RG, typename A = double>
__global__ __launch_bounds__(256) void k_nd_fill(FitArgs a, Planes<T> src, Planes<T> dst) {
    constexpr bool F32 = std::is_same_v<A, float>;
    const int lane = threadIdx.x & 15;
    const int n = *a.ndfill_n;
    const int stride = gridDim.x * 16;
    for (int i = blockIdx.x * 16 + (threadIdx.x >> 4); i < n; i += stride) {
        const int64_t g = a.ndfill[i];
        const int p = a.gparent[g];
        const int k = (int)(g - a.goff[p]);
        int pl;
        const auto kk = row_key<RG>(a.key, p, pl);
        const RowMeas<RG> rm(a.meas, a.rig_meas, a.key, p);
        const DisparityMeasurementModel& meas = rm.m;
        const int64_t b0 = (int64_t)a.gblock[g] * a.S;
        const int64_t d0 = (a.nd_base + g) * (int64_t)a.S;
        const CameraPose pose = load_pose(a.pose, p);
        SampleCtxF cx;
        if constexpr (F32) cx = SampleCtxF{meas_f(meas), pose_f(pose), motion_f(a.feat, a.dt)};
        /* four steps' loads issued before their arithmetic; the four steps
         * are one noise group (phase j) */
        FeatureCarry fc;
        for (int s0 = 0; s0 < a.S; s0 += 64) {
            T r[4][7];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int s = s0 + 16 * j + lane;
                if (s < a.S) load_sample_off(src, (Off)((b0 + s) * (int64_t)sizeof(T)), r[j]);
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int s = s0 + 16 * j + lane;
                if (s < a.S) {
                    bool in;
                    if constexpr (F32) {
                        float x[6], q[6], w;
                        eval_sample_f(a, cx, r[j], kk, pl, k, s, x, q, in, w, &fc, (uint32_t)j);
                        write_nd_off_f(meas, dst, (Off)((d0 + s) * (int64_t)sizeof(T)), x, in);
                    } else {
                        EuclideanPoint x;
                        DisparityPoint d;
                        double w;
                        eval_sample(a, meas, r[j], kk, pl, k, s, pose, x, d, in, w, &fc, (uint32_t)j);
                        write_nd_off(meas, dst, (Off)((d0 + s) * (int64_t)sizeof(T)), x, in);
                    }
                }
            }
        }
    }
}

/* The rare case of k_fit, one row per listed Gaussian (redo[0, *redo_n)):
 * every weight zero.  Refitted by two sweeps with unit weights (mean = sum
 * of the samples, as the reference's normalisation of an all-zero cloud
 * gives).  Workgroups stride over the list (its length is read on the
 * device). */
template <typename T, bool RG, typename A = double>
__global__ __launch_bounds__(256) void k_fit_fix(FitArgs a, Planes<T> src, Planes<T> dst) {
    constexpr bool F32 = std::is_same_v<A, float>;
    const int lane = threadIdx.x & 15;
    const int nredo = *a.redo_n;
    const int S = a.S;
    for (int i = blockIdx.x * 16 + (threadIdx.x >> 4); i < nredo; i += gridDim.x * 16) {
        const int64_t g = a.redo[i];
        const int p = a.gparent[g];
        const int k = (int)(g - a.goff[p]);
        int pl;
        const auto kk = row_key<RG>(a.key, p, pl);
        const RowMeas<RG> rm(a.meas, a.rig_meas, a.key, p);
        const DisparityMeasurementModel& meas = rm.m;
        const int64_t b0 = (int64_t)a.gblock[g] * S;
        const CameraPose pose = load_pose(a.pose, p);
        SampleCtxF cx;
        if constexpr (F32) cx = SampleCtxF{meas_f(meas), pose_f(pose), motion_f(a.feat, a.dt)};
        /* the sample's disparity-space point in the arithmetic type */
        auto point = [&](int s, A q[6]) {
            if constexpr (F32) {
                T r[7];
                load_sample(src, b0 + s, r);
                float x[6], w;
                bool in;
                eval_sample_f(a, cx, r, kk, pl, k, s, x, q, in, w);
            } else {
                EuclideanPoint x;
                DisparityPoint d;
                bool in;
                double w;
                fit_sample(a, meas, src, b0 + s, kk, pl, k, s, pose, x, d, in, w);
                q[0] = d.u; q[1] = d.v; q[2] = d.d; q[3] = d.vu; q[4] = d.vv; q[5] = d.vd;
            }
        };
        A sq[6] = {0, 0, 0, 0, 0, 0};
        for (int s0 = 0; s0 < S; s0 += 16) {
            const int s = s0 + lane;
            if (s < S) {
                A q[6];
                point(s, q);
#pragma unroll
                for (int c = 0; c < 6; ++c) sq[c] += q[c];
            }
        }
        A muA[6], cc[21];
        double mu[6], P[21];
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            muA[c] = row_sum16(sq[c]);
            mu[c] = (double)muA[c];
        }
#pragma unroll
        for (int e = 0; e < 21; ++e) cc[e] = 0;
        for (int s0 = 0; s0 < S; s0 += 16) {
            const int s = s0 + lane;
            if (s < S) {
                A q[6], dv[6];
                point(s, q);
#pragma unroll
                for (int c = 0; c < 6; ++c) dv[c] = q[c] - muA[c];
#pragma unroll
                for (int r = 0; r < 6; ++r)
#pragma unroll
                    for (int c = 0; c <= r; ++c) cc[pidx(r, c)] += dv[r] * dv[c];
            }
        }
        const double W = (double)S, W2 = (double)S;
        const double scale = (W * W + W2) / W;
#pragma unroll
        for (int e = 0; e < 21; ++e) P[e] = (double)row_sum16(cc[e]) * scale;
        const double pdg = a.fit[(int64_t)FP_PD * a.Ng + g];
        write_fit(a, g, lane, mu, P, pdg);
    }
}

/* Measurement-independent detection terms of each fitted Gaussian
 * (computeDetectionsKernel, src/kernels.cu:492-598): innovation covariance,
 * gain, Joseph covariance and its PD test, log(pd) + log(w), log det. */
__global__ __launch_bounds__(128) void k_precompute(const double* fit, const double* gw, int64_t Ng, double var_u,
                                                    double var_v,
                             double* pre, double* lik) {
    /* each Gaussian's records are formed by its thread, then written by the
     * workgroup through LDS as one contiguous stretch (records of
     * consecutive Gaussians are adjacent), so that the stores coalesce */
    __shared__ double stage[128 * 21];
    const int t = threadIdx.x;
    const int64_t g0 = (int64_t)blockIdx.x * 128;
    const int64_t g = g0 + t;
    const int nb = (int)min((int64_t)128, Ng - g0);
    const bool act = g < Ng;
    DetectionTerms r;
    bool ok = false;
    double pdg = 0.0, mu[6];
    if (act) {
        double P[21];
#pragma unroll
        for (int e = 0; e < 21; ++e) P[e] = fit[(FP_P + e) * Ng + g];
        detection_terms(P, var_u, var_v, r);
        ok = is_pd6(r.Pn);
        pdg = fit[FP_PD * Ng + g];
#pragma unroll
        for (int c = 0; c < 6; ++c) mu[c] = fit[(FP_MU + c) * Ng + g];
    }
    /* out[g0 L ..) <- the block's L-double records */
    auto put = [&](double* out, int L) {
        __syncthreads();
        for (int e = t; e < nb * L; e += 128) out[g0 * L + e] = stage[e];
    };
    if (act) {
        double* rec = stage + t * LK_COUNT;
        rec[LK_MUU] = mu[0];
        rec[LK_MUV] = mu[1];
        rec[LK_S0] = r.s0;
        rec[LK_S12] = r.s1 + r.s2;
        rec[LK_S3] = r.s3;
        rec[LK_C] = safe_log(pdg) + safe_log(gw[g]);
        rec[LK_H] = 0.5 * safe_log(r.det);
        rec[LK_OK] = ok ? 1.0 : 0.0;
    }
    put(lik, LK_COUNT);
    __syncthreads();
    if (act) {
        double* mk = stage + t * PP_MK_LEN;
#pragma unroll
        for (int c = 0; c < 6; ++c) mk[c] = mu[c];
#pragma unroll
        for (int k = 0; k < 12; ++k) mk[6 + k] = r.K[k];
    }
    put(pre + PP_MK * Ng, PP_MK_LEN);
    __syncthreads();
    if (act) {
#pragma unroll
        for (int e = 0; e < 21; ++e) stage[t * 21 + e] = r.Pn[e];
    }
    put(pre + PP_PN * Ng, 21);
}

/* ------------------------------------------- normalisers and survivors */
/* Likelihood records read through the constant address space: their
 * address is the same in every lane (one Gaussian, all measurements), so
 * the reads are scalar loads into SGPRs, with no LDS staging and no
 * per-lane copies.  Scalar loads return out of order, so every use waits
 * for all of them: records come four at a time, one wait per four (the
 * record buffer holds three records of padding past the last Gaussian). */
typedef __attribute__((address_space(4))) const double lik_t;

struct Lik4 {
    double v[4][LK_COUNT];
};

__device__ inline Lik4 load_lik4(const double* lik, int64_t g) {
    lik_t* r = (lik_t*)(lik + g * LK_COUNT);
    Lik4 x;
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int q = 0; q < LK_COUNT; ++q) x.v[k][q] = r[k * LK_COUNT + q];
    return x;
}

__device__ inline double det_lw(const double* r, double zu, double zv);

/* One record through the scalar path (its index the same in every lane). */
__device__ inline double det_lw(lik_t* r, double zu, double zv) {
    double v[LK_COUNT];
#pragma unroll
    for (int q = 0; q < LK_COUNT; ++q) v[q] = r[q];
    return det_lw((const double*)v, zu, zv);
}

__device__ inline lik_t* lik_rec(const double* lik, int64_t g) { return (lik_t*)(lik + g * LK_COUNT); }

/* log weight of the detection term of Gaussian record r and measurement
 * (zu, zv): log pd + log w - dist/2 - log 2 pi - log det / 2
 * (src/kernels.cu:527-539); log(0) when the updated covariance is not PD
 * (:597-598). */
__device__ inline double det_lw(const double* r, double zu, double zv) {
    const double i0 = zu - r[LK_MUU], i1 = zv - r[LK_MUV];
    const double dist = i0 * i0 * r[LK_S0] + i0 * i1 * r[LK_S12] + i1 * i1 * r[LK_S3];
    const double lw = r[LK_C] - 0.5 * dist - 1.8378770664093453 - r[LK_H];
    return r[LK_OK] != 0.0 ? lw : SCPHD_LOG0;
}

/* Per parent: eta_{p,m} = log( sum_g exp(lw_{g,m}) + M w0 + kappa ) for
 * every measurement (normaliser reduce_by_key + log_sum_exp,
 * src/scphdcameracalibration.cu:1729-1768, exact per-parent segments), the
 * single-cluster log-likelihood sum_m eta + sum_g w_g (:1813-1849) and the
 * number of detection/birth terms whose normalised weight reaches
 * min_weight (updateKernel's prune flag, src/kernels.cu:608-659).
 * One workgroup per parent, one thread per measurement, the Gaussians'
 * likelihood records read by scalar loads. */

/* e^x for -746 < x <= 0, the normaliser's terms (within 1 ulp of a
 * correctly rounded exp): x = k ln2 + r with |r| <= ln2/2 by Cody and
 * Waite's two-part reduction (ln2 split so that k ln2_hi is exact), e^r by
 * its Taylor polynomial to degree 13 (truncation below 5e-18 relative) in
 * Horner form, scaled by 2^k (gradual underflow below 2^-1022).  Cheaper
 * than the general exp, which also handles overflow and special values. */
__device__ inline double exp_nonpos(double x) {
    const double k = rint(x * 1.4426950408889634);
    const double r = fma(k, -1.90821492927058770002e-10, fma(k, -6.93147180369123816490e-01, x));
    double p = 1.6059043836821613e-10; /* 1/13! */
    p = fma(p, r, 2.08767569878681e-09);
    p = fma(p, r, 2.505210838544172e-08);
    p = fma(p, r, 2.755731922398589e-07);
    p = fma(p, r, 2.7557319223985893e-06);
    p = fma(p, r, 2.48015873015873e-05);
    p = fma(p, r, 0.0001984126984126984);
    p = fma(p, r, 0.001388888888888889);
    p = fma(p, r, 0.008333333333333333);
    p = fma(p, r, 0.041666666666666664);
    p = fma(p, r, 0.16666666666666666);
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return ldexp(p, (int)k);
}

__global__ __launch_bounds__(NT) void k_eta(EtaArgs a) {
    __shared__ double red[16];
    __shared__ int sh[8];
    __shared__ int s_heavy[ETA_HEAVY];
    const int p = blockIdx.x;
    const int t = threadIdx.x;
    const int64_t g0 = a.goff[p];
    const int G = a.goff[p + 1] - a.goff[p];
    double eta_sum = 0.0;
    int cnt = 0;
    /* this parent's measurement set: its rig's in an ensemble */
    const int rig = a.moff ? a.prig[p] : 0;
    const int M = a.moff ? a.moff[rig + 1] - a.moff[rig] : a.M;
    const int64_t zb = a.moff ? a.moff[rig] : 0;
    const double bpc = a.bpcs ? a.bpcs[rig] : a.bpc;
    const double logw0 = a.births[rig].logw0;
    const bool birth_ok = a.births[rig].pd_ok != 0;
    const int rounds = (M + NT - 1) / NT;
    /* Only a Gaussian with l > bpc can move the maximum off bpc, and
     * l <= C - log 2 pi - H (the value at dist = 0; the computed quadratic
     * form is never below -1e-9 of its scale), so the max pass visits only
     * those whose bound exceeds bpc (with a 1e-6 relative margin), in
     * ascending order: the same first maximum.  More than heavy_cap (at most
     * ETA_HEAVY) of them: the whole map. */
    int nheavy = 0;
    for (int j0 = 0; j0 < G; j0 += NT) {
        const int j = j0 + t;
        bool h = false;
        if (j < G) {
            const double* rec = a.lik + (g0 + j) * LK_COUNT;
            const double ub = rec[LK_C] - 1.8378770664093453 - rec[LK_H];
            h = rec[LK_OK] != 0.0 && ub + 1e-6 * (1.0 + fabs(ub)) > bpc;
        }
        int tot;
        const int pos = nheavy + block_exscan(h ? 1 : 0, sh, &tot);
        if (h && pos < ETA_HEAVY) s_heavy[pos] = j;
        nheavy += tot;
    }
    __syncthreads();
    const bool listed = nheavy <= a.heavy_cap;
    for (int r = 0; r < rounds; ++r) {
        const int m = r * NT + t;
        const bool valid = m < M;
        const double zu = valid ? a.zu[zb + m] : 0.0, zv = valid ? a.zv[zb + m] : 0.0;
        /* eta = mx + log1p(sum of the other terms' exp(l - mx)), mx the
         * first maximum of (bpc, l_0, ...), terms summed g ascending then
         * bpc: the normaliser definition the oracle uses bit for bit */
        double mx = bpc;
        int am = -1;
        double T = 0.0;
        int mcnt = 0;
        /* candidates: terms with l - mx above the gate (less 1e-6), the only
         * ones that can pass it against eta >= mx; the first ETA_NCAND kept */
        int nc = 0, cc[ETA_NCAND];
#pragma unroll
        for (int q = 0; q < ETA_NCAND; ++q) cc[q] = 0;
        const double cgate = a.lw_gate - 1e-6;
        if (valid) {
            if (listed) {
                for (int q = 0; q < nheavy; ++q) {
                    const int j = __builtin_amdgcn_readfirstlane(s_heavy[q]);
                    const double l = det_lw(lik_rec(a.lik, g0 + j), zu, zv);
                    if (l > mx) {
                        mx = l;
                        am = j;
                    }
                }
            } else {
                for (int j0 = 0; j0 < G; j0 += 4) {
                    const Lik4 R = load_lik4(a.lik, g0 + j0);
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const double l = det_lw(R.v[q], zu, zv);
                        if (j0 + q < G && l > mx) {
                            mx = l;
                            am = j0 + q;
                        }
                    }
                }
            }
            /* terms that cannot change T are skipped, which changes no bit:
             * exp(x) is +0 for x <= -746, and for T in [2^(e-1), 2^e) a term
             * below 2^(e-55), a quarter ulp, leaves T + exp(x) == T */
            double thr = -746.0;
            for (int j0 = 0; j0 < G; j0 += 4) {
                const Lik4 R = load_lik4(a.lik, g0 + j0);
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const double x = det_lw(R.v[q], zu, zv) - mx;
                    if (j0 + q < G && j0 + q != am && x > thr) {
                        T += exp_nonpos(x);
                        /* the bound holds only once T > 0: a term that
                         * underflowed to +0 leaves the threshold at -746 */
                        thr = T > 0.0 ? fmax(-746.0, (double)(__builtin_amdgcn_frexp_exp(T) - 55) * 0.6931471805599453)
                                      : -746.0;
                    }
                    if (j0 + q < G && x > cgate && R.v[q][LK_OK] != 0.0) {
                        const int j = j0 + q;
#pragma unroll
                        for (int e = 0; e < ETA_NCAND; ++e) cc[e] = nc == e ? j : cc[e];
                        ++nc;
                    }
                }
            }
        }
        if (am >= 0) T += exp(bpc - mx);
        const double eta = mx + log1p(T);
        if (valid) {
            const int64_t pm = (int64_t)p * a.M + m;
            a.eta[pm] = eta;
            eta_sum += eta;
            if (nc <= a.ncand_cap) { /* the candidates alone */
#pragma unroll
                for (int k = 0; k < ETA_NCAND; ++k) {
                    if (k >= nc) break;
                    const int j = cc[k];
                    const double* rec = a.lik + (g0 + j) * LK_COUNT;
                    const double x = det_lw(rec, zu, zv) - eta;
                    if (x > a.lw_gate && rec[LK_OK] != 0.0) mcnt += exp(x) >= a.min_weight ? 1 : 0;
                    a.mcand[pm * ETA_NCAND + k] = j;
                }
            } else {
                for (int j0 = 0; j0 < G; j0 += 4) {
                    const Lik4 R = load_lik4(a.lik, g0 + j0);
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const double x = det_lw(R.v[q], zu, zv) - eta;
                        if (j0 + q < G && x > a.lw_gate && R.v[q][LK_OK] != 0.0)
                            mcnt += exp(x) >= a.min_weight ? 1 : 0;
                    }
                }
            }
            a.mncand[pm] = nc;
            a.mcnt[pm] = mcnt;
            cnt += mcnt;
            const double w = exp(logw0 - eta);
            cnt += (w >= a.min_weight && birth_ok) ? 1 : 0;
        }
    }
    double swg = 0.0;
    for (int j = t; j < G; j += NT) swg += a.gw[g0 + j];
    const double se = block_sum(eta_sum, red);
    const double sg = block_sum(swg, red);
    const double sc = block_sum((double)cnt, red);
    if (t == 0) {
        a.loglik[p] = se + sg;
        a.nsurv[p] = (int32_t)sc;
    }
}

/* Writes the surviving detection terms (Kalman-updated mean, weight
 * exp(lw - eta)) and birth terms of every parent, compacted in the
 * reference's pack order (updateKernel, src/kernels.cu:608-659: detection
 * terms measurement-major, then births). */
__global__ __launch_bounds__(NT) void k_survivors(SurvArgs a) {
    __shared__ int sh[8];
    const EtaArgs& e = a.e;
    const int p = blockIdx.x;
    const int t = threadIdx.x;
    const int64_t g0 = e.goff[p];
    const int G = e.goff[p + 1] - e.goff[p];
    const int64_t base = a.survoff[p];
    const int cap = a.survoff[p + 1] - a.survoff[p];
    int run = 0;
    const int rig = e.moff ? e.prig[p] : 0;
    const BirthModel& birth = e.births[rig];
    const int M = e.moff ? e.moff[rig + 1] - e.moff[rig] : e.M;
    const int64_t zb = e.moff ? e.moff[rig] : 0;
    const int rounds = (M + NT - 1) / NT;
    for (int r = 0; r < rounds; ++r) {
        const int m = r * NT + t;
        const bool valid = m < M;
        const double zu = valid ? e.zu[zb + m] : 0.0, zv = valid ? e.zv[zb + m] : 0.0;
        const double eta = valid ? e.eta[(int64_t)p * e.M + m] : 0.0;
        const int cnt = valid ? e.mcnt[(int64_t)p * e.M + m] : 0; /* detection terms kept, from k_eta */
        int tot;
        int o = run + block_exscan(cnt, sh, &tot);
        if (valid && cnt > 0) {
            const int64_t pm = (int64_t)p * e.M + m;
            /* detection term of map entry j: kept when its normalised weight
             * reaches min_weight (the gate first: exp(x) < min_weight below it) */
            auto term = [&](int j, const double* rec) {
                const double x = det_lw(rec, zu, zv) - eta;
                if (!(x > e.lw_gate && rec[LK_OK] != 0.0)) return;
                const double w = exp(x);
                if (w < e.min_weight) return;
                if (o < cap) {
                    const int64_t gi = g0 + j;
                    const int64_t so = base + o;
                    const double i0 = zu - rec[LK_MUU], i1 = zv - rec[LK_MUV];
                    a.sw[so] = w;
                    a.sg[so] = (int32_t)gi;
                    a.sm[so] = m;
                    const double* mk = e.pre + PP_MK * e.Ng + gi * PP_MK_LEN;
#pragma unroll
                    for (int c = 0; c < 6; ++c) a.smu[(so) * 6 + c] = mk[c] + i0 * mk[6 + c] + i1 * mk[12 + c];
                }
                ++o;
            };
            const int nc = e.mncand[pm];
            if (nc <= e.ncand_cap) { /* the candidates k_eta listed, ascending */
                for (int k = 0; k < nc; ++k) {
                    const int j = e.mcand[pm * ETA_NCAND + k];
                    term(j, e.lik + (g0 + j) * LK_COUNT);
                }
            } else {
                for (int j0 = 0; j0 < G; j0 += 4) {
                    const Lik4 R = load_lik4(e.lik, g0 + j0);
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        if (j0 + q < G) term(j0 + q, R.v[q]);
                }
            }
        }
        run += tot;
    }
    for (int r = 0; r < rounds; ++r) {
        const int m = r * NT + t;
        const bool valid = m < M;
        const double eta = valid ? e.eta[(int64_t)p * e.M + m] : 0.0;
        const double w = valid ? exp(birth.logw0 - eta) : 0.0;
        const bool ok = valid && w >= e.min_weight && birth.pd_ok;
        int tot;
        const int o = run + block_exscan(ok ? 1 : 0, sh, &tot);
        if (ok && o < cap) {
            const int64_t so = base + o;
            a.sw[so] = w;
            a.sg[so] = -1 - rig;
            a.sm[so] = m;
            a.smu[(so) * 6 + 0] = e.zu[zb + m];
            a.smu[(so) * 6 + 1] = e.zv[zb + m];
#pragma unroll
            for (int c = 2; c < 6; ++c) a.smu[(so) * 6 + c] = birth.mean[c];
        }
        run += tot;
    }
    if (t == 0 && run != cap) *a.err = 1;
}

/* -------------------------------------------------------------- merging */
/* Workspace of one parent's merge, U survivor slots (Up = next power of two):
 *   key_w double[Up]  sort keys;            ord   int[Up]   (weight desc, index asc) order
 *   bu, bv double[U]  mean u, v;            bs    double[3U] (u,v) covariance block uu uv vv
 *   flag  uint8[U]    merged;               selfm uint8[U]  member of its own group
 *   nbo   int[U+1]    neighbour offsets;    cnt   int[U]    neighbour counts / fill cursors
 *   mem   int[U]      group members;
 *   nbr   int[ecap]   neighbour lists, ecap = 8 U + 64
 *   cand  uint32[ccap] candidate pairs i << 16 | j (i < j), ccap = 16 U + 64
 *   res   uint8[ccap] their membership
 * While the candidates are enumerated, ord holds the (u, index) order,
 * selfm the ill-conditioned flags and mem the list of those survivors. */
__host__ __device__ inline int next_pow2(int x) {
    int v = 1;
    while (v < x) v <<= 1;
    return v;
}

struct MergeLayout {
    int64_t ord, bu, bv, bs, flag, selfm, nbo, cnt, mem, nbr, ecap, cand, ccap, res, bytes;
    __host__ __device__ explicit MergeLayout(int U) {
        const int64_t Up = next_pow2(U > 0 ? U : 1);
        ord = Up * 8;
        bu = (ord + Up * 4 + 7) & ~(int64_t)7;
        bv = bu + (int64_t)U * 8;
        bs = bv + (int64_t)U * 8;
        flag = bs + (int64_t)U * 24;
        selfm = flag + U;
        nbo = (selfm + U + 7) & ~(int64_t)7;
        cnt = nbo + ((int64_t)U + 1) * 4;
        mem = cnt + (int64_t)U * 4;
        nbr = mem + (int64_t)U * 4;
        ecap = 8 * (int64_t)U + 64;
        cand = nbr + ecap * 4;
        ccap = 16 * (int64_t)U + 64;
        res = cand + ccap * 4;
        bytes = res + ccap;
    }
};

/* global scratch for parents that do not fit in LDS: MergeLayout(U).bytes
 * <= 190 U + 600 since Up < 2 U */
static constexpr int64_t MERGE_WS_PER_SLOT = 240, MERGE_WS_PER_PARENT = 1024;

int merge_lds_bytes(int cap) { return (int)MergeLayout(cap).bytes; }

int64_t merge_scratch_bytes(int64_t Stot, int Np) {
    return Stot * MERGE_WS_PER_SLOT + (int64_t)Np * MERGE_WS_PER_PARENT + 256;
}

/* Packed covariance entry e of a survivor's term: the updated covariance of
 * table entry g, or (g = -1 - r < 0) the birth covariance of rig r.  Both
 * live in global memory, so the choice is one global load from a chosen
 * address. */
__device__ inline double term_P(const MergeArgs& a, int g, int e) {
    return g >= 0 ? a.pre[(int64_t)PP_PN * a.Ng + (int64_t)g * 21 + e] : a.births[-1 - g].P[e];
}

__device__ inline void survivor_P(const MergeArgs& a, int64_t si, double* P) {
    const int g = a.sg[si];
#pragma unroll
    for (int e = 0; e < 21; ++e) P[e] = term_P(a, g, e);
}

__device__ inline void survivor_mu(const MergeArgs& a, int64_t si, double* mu) {
#pragma unroll
    for (int c = 0; c < 6; ++c) mu[c] = a.smu[(si) * 6 + c];
}

/* Workspace pointers of one parent's merge. */
struct MergeWs {
    double* key_w;
    int* ord;
    double* bu;
    double* bv;
    double* bs;
    uint8_t* flag;
    uint8_t* selfm;
    int* nbo;
    int* cnt;
    int* mem;
    int* nbr;
    uint32_t* cand;
    uint8_t* res;
    int64_t ecap, ccap;
    __device__ MergeWs(char* ws, int U) {
        const MergeLayout ly(U);
        key_w = (double*)ws;
        ord = (int*)(ws + ly.ord);
        bu = (double*)(ws + ly.bu);
        bv = (double*)(ws + ly.bv);
        bs = (double*)(ws + ly.bs);
        flag = (uint8_t*)(ws + ly.flag);
        selfm = (uint8_t*)(ws + ly.selfm);
        nbo = (int*)(ws + ly.nbo);
        cnt = (int*)(ws + ly.cnt);
        mem = (int*)(ws + ly.mem);
        nbr = (int*)(ws + ly.nbr);
        cand = (uint32_t*)(ws + ly.cand);
        res = (uint8_t*)(ws + ly.res);
        ecap = ly.ecap;
        ccap = ly.ccap;
    }
};

/* Can survivors i and j be merge partners?  The (u, v) marginal of the
 * squared Mahalanobis distance with Sigma = P_i + P_j is a lower bound of
 * the full 6-D distance for positive-definite Sigma (and a non-PD Sigma
 * never merges), so a pair is rejected when the marginal exceeds the
 * threshold by 0.1 %; an ill-conditioned (u, v) block is never rejected. */
__device__ inline bool merge_bound(const MergeWs& w, int i, int j, double thr) {
    const double du = w.bu[j] - w.bu[i], dv = w.bv[j] - w.bv[i];
    const double suu = w.bs[3 * i] + w.bs[3 * j];
    const double suv = w.bs[3 * i + 1] + w.bs[3 * j + 1];
    const double svv = w.bs[3 * i + 2] + w.bs[3 * j + 2];
    const double det = suu * svv - suv * suv;
    if (!(det > 1e-6 * suu * svv)) return true;
    const double q = svv * du * du - 2.0 * suv * du * dv + suu * dv * dv;
    return !(q > 1.001 * thr * det);
}

/* Merge membership of survivors i and j: squared Mahalanobis distance with
 * Sigma = P_i + P_j below min_separation (computeMahalDist,
 * src/device_math.cu:374-453).  The computed value is symmetric in i and j
 * to the last bit (the sum P_i + P_j commutes and negating the difference
 * negates every forward-substitution term). */
__device__ inline bool merge_full(const MergeArgs& a, int64_t sbase, int i, int j, double thr) {
    double mi[6], mj[6], Pi[21], Pj[21];
    survivor_mu(a, sbase + i, mi);
    survivor_mu(a, sbase + j, mj);
    survivor_P(a, sbase + i, Pi);
    survivor_P(a, sbase + j, Pj);
    return mahal6(mi, Pi, mj, Pj) < thr;
}

__device__ inline bool merge_pair(const MergeArgs& a, int64_t sbase, const MergeWs& w, int i, int j, double thr) {
    return merge_bound(w, i, j, thr) && merge_full(a, sbase, i, j, thr);
}

/* Bitonic sort of ord[0, Up) by key_w: ascending (key, index) when ASC,
 * else descending key with ascending index on ties.  Padding slots carry
 * keys that sort last.  Ends with a barrier. */
template <bool ASC>
__device__ inline void merge_bitonic(double* key, int* ord, int Up) {
    const int t = threadIdx.x;
    __syncthreads();
    for (int k = 2; k <= Up; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = t; i < Up; i += NT) {
                const int l = i ^ j;
                if (l > i) {
                    const double wi = key[i], wl = key[l];
                    const int ii = ord[i], il = ord[l];
                    const bool l_before_i = ASC ? ((wl < wi) || (wl == wi && il < ii))
                                                : ((wl > wi) || (wl == wi && il < ii));
                    const bool asc = (i & k) == 0;
                    if (asc == l_before_i) {
                        key[i] = wl;
                        key[l] = wi;
                        ord[i] = il;
                        ord[l] = ii;
                    }
                }
            }
            __syncthreads();
        }
    }
}

/* Load one parent's survivors' (u, v) marginals into the workspace, clear
 * the merged flags, and sort them by (u, index): key[r] is the u of rank r
 * (+inf for a u that is not finite, which sorts last), ord[r] its survivor.
 * key and ord hold next_pow2(U) entries. */
__device__ inline void merge_load_sort_u(const MergeArgs& a, int64_t sbase, int U, MergeWs& w, double* key,
                                         int* ord) {
    const int t = threadIdx.x;
    const int Up = next_pow2(U);
    for (int i = t; i < Up; i += NT) {
        if (i < U) {
            w.bu[i] = a.smu[(sbase + i) * 6 + 0];
            w.bv[i] = a.smu[(sbase + i) * 6 + 1];
            const int g = a.sg[sbase + i];
            w.bs[3 * i] = term_P(a, g, pidx(0, 0));
            w.bs[3 * i + 1] = term_P(a, g, pidx(1, 0));
            w.bs[3 * i + 2] = term_P(a, g, pidx(1, 1));
            w.flag[i] = 0;
            key[i] = isfinite(w.bu[i]) ? w.bu[i] : INFINITY;
            ord[i] = i;
        } else {
            key[i] = INFINITY;
            ord[i] = 0x7fffffff;
        }
    }
    merge_bitonic<true>(key, ord, Up);
}

/* Sort one parent's survivors by (weight desc, index asc) into ord, the
 * order in which the reference's argmax loop picks leaders (ties: lowest
 * pack index). */
__device__ inline void merge_sort_w(const MergeArgs& a, int64_t sbase, int U, double* key, int* ord) {
    const int t = threadIdx.x;
    const int Up = next_pow2(U);
    __syncthreads();
    for (int i = t; i < Up; i += NT) {
        key[i] = i < U ? a.sw[sbase + i] : -INFINITY;
        ord[i] = i < U ? i : 0x7fffffff;
    }
    merge_bitonic<false>(key, ord, Up);
}

/* Load and sort by weight (the round loop). */
__device__ inline void merge_load_sort(const MergeArgs& a, int64_t sbase, int U, MergeWs& w) {
    merge_load_sort_u(a, sbase, U, w, w.key_w, w.ord);
    merge_sort_w(a, sbase, U, w.key_w, w.ord);
}

__device__ inline bool merge_self(const MergeArgs& a, int64_t si, double thr) {
    double m[6], P[21];
    survivor_mu(a, si, m);
    survivor_P(a, si, P);
    return mahal6(m, P, m, P) < thr;
}

/* Step 2 of k_merge_groups: the merge candidates, from the survivors
 * sorted by (u, index) (key[r] = u of rank r, ord[r] its survivor; in LDS
 * when they fit, else in the workspace). */
__device__ inline void merge_candidates(const MergeArgs& a, int U, MergeWs& w, double thr, const double* key,
                                        const int* ord, int* s_nc, int* s_nill, double* s_red) {
    const int t = threadIdx.x;
    /* 2. candidates: the pairs that pass merge_bound, enumerated without
     * visiting all U (U - 1) / 2 pairs.
     *   A survivor is ill-conditioned when its (u, v) block is not PD with
     *   det > 2e-6 puu pvv (or holds a non-finite value).  For two that are
     *   not, the block of their sum has det > 1e-6 suu svv (the correlation
     *   of a sum of PSD 2x2 blocks is at most the larger of theirs:
     *   |auv + buv| <= r (sqrt(auu avv) + sqrt(buu bvv)) <= r sqrt(suu svv)
     *   by Cauchy-Schwarz), so merge_bound decides by q <= 1.001 thr det,
     *   and q suu - du^2 det = (suv du - suu dv)^2 >= 0 makes that imply
     *   du^2 <= 1.001 thr suu.  Such pairs lie within
     *   du^2 <= 1.002 thr (puu_i + max puu) of each other in u: each
     *   survivor scans its successors in (u, index) order up to that
     *   window.  (Every survivor of this filter has puu <= std_u^2: a
     *   Kalman update with H = [I2 0] never raises the measured variance
     *   above R, and births carry R; the windows span a few pixels.)
     *   Pairs with an ill-conditioned survivor are tested against all
     *   others.  The candidate set is the one the full U^2 sweep finds. */
    for (int i = t; i < U; i += NT) w.cnt[i] = 0;
    if (t == 0) *s_nill = 0;
    __syncthreads();
    double smax = 0.0;
    for (int i = t; i < U; i += NT) {
        const double suu = w.bs[3 * i], suv = w.bs[3 * i + 1], svv = w.bs[3 * i + 2];
        const bool good = !a.all_pairs && isfinite(w.bu[i]) && isfinite(w.bv[i]) && isfinite(suu) && isfinite(suv) &&
                          isfinite(svv) && suu > 0.0 && svv > 0.0 && suu * svv - suv * suv > 2e-6 * suu * svv;
        w.selfm[i] = good ? 0 : 1;
        if (good) {
            smax = fmax(smax, suu);
        } else {
            w.mem[atomicAdd(s_nill, 1)] = i;
        }
    }
    smax = block_max(smax, s_red); /* ends with a barrier after the flags */
    {
        auto add = [&](int i, int j) {
            const int lo = min(i, j), hi = max(i, j);
            if (merge_bound(w, lo, hi, thr)) {
                const int c = atomicAdd(s_nc, 1);
                if (c < w.ccap) w.cand[c] = ((uint32_t)lo << 16) | (uint32_t)hi;
            }
        };
        for (int r = t; r < U; r += NT) {
            const int i = ord[r];
            if (w.selfm[i]) continue;
            const double ui = key[r];
            const double lim2 = 1.002 * thr * (w.bs[3 * i] + smax);
            for (int r2 = r + 1; r2 < U; ++r2) {
                const int j = ord[r2];
                const double du = key[r2] - ui;
                if (!(du * du <= lim2)) break;
                if (!w.selfm[j]) add(i, j);
            }
        }
        const int nill = *s_nill;
        const int64_t nx = (int64_t)nill * U;
        for (int64_t e = t; e < nx; e += NT) {
            const int i = w.mem[e / U], j = (int)(e % U);
            if (j == i || (w.selfm[j] && j < i)) continue; /* two ill survivors: once */
            add(i, j);
        }
    }
}

/* Greedy Gaussian-mixture merge (phdUpdateMergeKernel, src/kernels.cu:662-874):
 * repeatedly take the heaviest unmerged term (ties: lowest pack index),
 * merge every unmerged term within squared Mahalanobis distance
 * min_separation of it by moment matching, emit the merge if PD else the
 * leader, stop when nothing is left or a group's weight is zero.
 *
 * Evaluated in phases instead of the reference's barrier-bound rounds.
 * k_merge_groups (one workgroup per parent):
 *   1. sort by (u, index);
 *   2. candidate pairs i < j that pass merge_bound, found by scanning a
 *      window in u around each survivor (and every pair of an
 *      ill-conditioned survivor), compacted in the workspace; then sort by
 *      (weight desc, index asc);
 *   3. the full 6x6 test, one lane per candidate pair (and per self pair),
 *      so that lanes stay busy however sparse the candidates are;
 *   4. neighbour lists from the members, each sorted ascending;
 *   5. one lane walks the sorted order forming each leader's group from its
 *      unmerged neighbours -- the same groups as the round loop, since
 *      membership depends only on the pair.
 * k_merge_moments (one thread per group): moments with members summed in
 * ascending index order.  k_merge_rounds: the round loop, for parents whose
 * candidate or neighbour lists overflow the workspace. */
template <bool BIG>
__global__ __launch_bounds__(NT) void k_merge_groups(MergeArgs a) {
    extern __shared__ char lds[];
    __shared__ int s_nc, s_tot;
    __shared__ int sh[8];
    const int p = blockIdx.x;
    const int t = threadIdx.x;
    const int64_t sbase = a.survoff[p];
    const int U = a.survoff[p + 1] - a.survoff[p];
    if (BIG != (U > a.lds_cap)) return; /* the other launch handles this parent (uniform) */
    if (U == 0) {
        if (t == 0) {
            a.ngroups[p] = 0;
            a.nmerged[p] = 0;
        }
        return;
    }
    if (U >= 65536 || a.force_rounds) { /* candidate pairs pack two 16-bit indices: the round loop takes it */
        if (t == 0) {
            a.ngroups[p] = -1;
            a.nmerged[p] = 0;
        }
        return;
    }
    /* a compile-time choice of workspace, so that the LDS instantiation
     * addresses it with ds_* instructions rather than flat ones */
    char* ws = BIG ? (a.scratch + sbase * MERGE_WS_PER_SLOT + (int64_t)p * MERGE_WS_PER_PARENT) : lds;
    MergeWs w(ws, U);
    const double thr = a.min_separation;
    if (t == 0) s_nc = 0;
    /* 1-2. sort by (u, index), candidates, sort by weight: the sort keys
     * in LDS for a big parent whose keys fit the launch's LDS, else in the
     * workspace */
    __shared__ int s_nill;
    __shared__ double s_red[8];
    const int Up = next_pow2(U);
    const bool keys_lds = BIG && Up <= a.big_lds_up;
    /* once the keys are sorted, the LDS they held takes the walk's most
     * accessed per-survivor arrays (global scratch otherwise): neighbour
     * counts and offsets always, and when the launch's LDS allows (28 B per
     * survivor) the component labels, bucket offsets and cursors and the
     * per-rank group records */
    int* lds_lab = nullptr;
    int* lds_boff = nullptr;
    int* lds_rinfo = nullptr;
    if (keys_lds) {
        double* key = (double*)lds;
        int* ord = (int*)(lds + (size_t)8 * Up);
        merge_load_sort_u(a, sbase, U, w, key, ord); /* ends with a barrier */
        merge_candidates(a, U, w, thr, key, ord, &s_nc, &s_nill, s_red);
        __syncthreads();
        if (s_nc <= w.ccap) {
            merge_sort_w(a, sbase, U, key, ord);
            for (int i = t; i < U; i += NT) w.ord[i] = ord[i];
            __syncthreads(); /* the sort keys are read */
            /* U + (U + 1) ints within the keys' 12 Up bytes */
            w.cnt = (int*)lds;
            w.nbo = (int*)lds + U;
            if ((int64_t)(7 * U + 2) * 4 <= a.big_lds_bytes) {
                lds_lab = (int*)lds + 2 * U + 1;
                lds_boff = lds_lab + U;
                lds_rinfo = lds_boff + 2 * U + 1;
            }
            for (int i = t; i < U; i += NT) w.cnt[i] = 0;
        }
    } else {
        merge_load_sort_u(a, sbase, U, w, w.key_w, w.ord);
        merge_candidates(a, U, w, thr, w.key_w, w.ord, &s_nc, &s_nill, s_red);
        __syncthreads();
        if (s_nc <= w.ccap) merge_sort_w(a, sbase, U, w.key_w, w.ord);
    }

    __syncthreads();
    const int nc = s_nc;
    if (nc > w.ccap) {
        if (t == 0) {
            a.ngroups[p] = -1;
            a.nmerged[p] = 0;
        }
        return;
    }
    /* 3. full tests: candidates, then every survivor against itself */
    for (int c = t; c < nc + U; c += NT) {
        if (c < nc) {
            const int i = (int)(w.cand[c] >> 16), j = (int)(w.cand[c] & 0xFFFFu);
            const bool m = merge_full(a, sbase, i, j, thr);
            w.res[c] = m ? 1 : 0;
            if (m) {
                atomicAdd(&w.cnt[i], 1);
                atomicAdd(&w.cnt[j], 1);
            }
        } else {
            const int i = c - nc;
            w.selfm[i] = merge_self(a, sbase + i, thr) ? 1 : 0;
        }
    }
    __syncthreads();
    /* 4. neighbour lists: offsets (rows [r0, r1) per thread), fill, sort */
    const int rc = (U + NT - 1) / NT;
    const int r0 = min(U, t * rc), r1 = min(U, r0 + rc);
    int mine = 0;
    for (int i = r0; i < r1; ++i) mine += w.cnt[i];
    int tot;
    int base = block_exscan(mine, sh, &tot);
    for (int i = r0; i < r1; ++i) {
        w.nbo[i] = base;
        base += w.cnt[i];
        w.cnt[i] = 0;
    }
    if (t == 0) {
        w.nbo[U] = tot;
        s_tot = tot;
    }
    __syncthreads();
    if (s_tot > w.ecap) {
        if (t == 0) {
            a.ngroups[p] = -1;
            a.nmerged[p] = 0;
        }
        return;
    }
    for (int c = t; c < nc; c += NT)
        if (w.res[c]) {
            const int i = (int)(w.cand[c] >> 16), j = (int)(w.cand[c] & 0xFFFFu);
            w.nbr[w.nbo[i] + atomicAdd(&w.cnt[i], 1)] = j;
            w.nbr[w.nbo[j] + atomicAdd(&w.cnt[j], 1)] = i;
        }
    __syncthreads();
    for (int i = t; i < U; i += NT) { /* insertion sort: lists are short */
        const int e0 = w.nbo[i], e1 = w.nbo[i + 1];
        for (int e = e0 + 1; e < e1; ++e) {
            const int x = w.nbr[e];
            int f = e - 1;
            while (f >= e0 && w.nbr[f] > x) {
                w.nbr[f + 1] = w.nbr[f];
                --f;
            }
            w.nbr[f + 1] = x;
        }
    }
    __syncthreads();
    /* 5. the greedy walk.  A survivor with an empty neighbour list is
     * isolated: no leader takes it, so when the walk reaches it, it forms
     * the group {itself}, or stops the walk when it is not its own member.
     * The sequential walk therefore visits only the connected survivors, in
     * rank order; every rank's outcome (group size, stop) is then placed in
     * rank order by two scans.  Per rank r: rinfo[2 r] = group size | STOP,
     * rinfo[2 r + 1] = start of the members in tmp (connected leaders). */
    constexpr int STOP = 1 << 30;
    int* rinfo = lds_rinfo ? lds_rinfo : (int*)w.bu; /* bu is free after the candidates: 8 bytes per slot */
    int* nl = w.cnt;         /* connected ranks, ascending; free after the lists */
    int* tmp = (int*)w.cand; /* members of connected groups; ccap >= U */
    {
        int mine = 0;
        for (int r = r0; r < r1; ++r) mine += w.nbo[w.ord[r] + 1] != w.nbo[w.ord[r]];
        int nn;
        int k = block_exscan(mine, sh, &nn);
        for (int r = r0; r < r1; ++r) {
            const int L = w.ord[r];
            if (w.nbo[L + 1] != w.nbo[L]) {
                nl[k++] = r;
                rinfo[2 * r] = 0;
            } else {
                rinfo[2 * r] = w.selfm[L] ? 1 : STOP;
            }
        }
        if (t == 0) s_tot = nn;
    }
    __syncthreads();
    /* The walk per connected component.  A leader takes only unmerged
     * neighbours, so the groups formed in one component of the neighbour
     * graph depend only on that component's ranks and flags: each
     * component is walked on its own, in rank order, by one lane.  (The
     * reference's loop ends at the first leader that is not its own member;
     * a component walked past that rank forms groups the rank-ordered
     * placement below discards, exactly as if the walk had stopped.)
     *   labels: the smallest survivor index of the component, by min
     *           propagation along the neighbour lists;
     *   buckets: the connected ranks grouped by label (counts, scan,
     *           scatter), each bucket sorted by rank by its lane;
     *   members of a group go to tmp inside its bucket's range, ascending
     *           by survivor index with the leader in its place. */
    {
        const int nn = s_tot;
        int* lab = lds_lab ? lds_lab : (int*)w.bv; /* label per survivor (bv is free after the candidates) */
        int* gb = (int*)w.bs;            /* bs is free after the candidates */
        int* boff = lds_boff ? lds_boff : gb; /* bucket offsets [U + 1] */
        int* bcur = boff + (U + 1);      /* scatter cursors [U] */
        int* bkt = gb + 2 * U + 1;       /* connected ranks by bucket [nn] */
        int* seg = (int*)w.key_w;        /* non-empty buckets (key_w is free after the sorts) */
        for (int i = t; i < U; i += NT) {
            lab[i] = i;
            boff[i] = 0;
        }
        if (t == 0) {
            boff[U] = 0;
            s_nc = 0;
        }
        __syncthreads();
        for (;;) {
            int changed = 0;
            for (int k = t; k < nn; k += NT) {
                const int i = w.ord[nl[k]];
                int m = lab[i];
                for (int e = w.nbo[i]; e < w.nbo[i + 1]; ++e) m = min(m, lab[w.nbr[e]]);
                if (m < lab[i]) {
                    lab[i] = m;
                    changed = 1;
                }
            }
            if (!__syncthreads_or(changed)) break;
        }
        for (int k = t; k < nn; k += NT) atomicAdd(&boff[lab[w.ord[nl[k]]]], 1);
        __syncthreads();
        {
            int mine = 0;
            for (int i = r0; i < r1; ++i) mine += boff[i];
            int tot;
            int base = block_exscan(mine, sh, &tot);
            for (int i = r0; i < r1; ++i) {
                const int c = boff[i];
                boff[i] = base;
                bcur[i] = base;
                if (c > 0) seg[atomicAdd(&s_nc, 1)] = i;
                base += c;
            }
            if (t == 0) boff[U] = tot;
        }
        __syncthreads();
        for (int k = t; k < nn; k += NT) {
            const int r = nl[k];
            bkt[atomicAdd(&bcur[lab[w.ord[r]]], 1)] = r;
        }
        __syncthreads();
        const int nseg = s_nc;
        for (int q = t; q < nseg; q += NT) {
            const int b0 = seg[q];
            const int k0 = boff[b0], k1 = bcur[b0]; /* the cursor ends at the bucket's end */
            for (int k = k0 + 1; k < k1; ++k) { /* ranks ascending: buckets are small */
                const int x = bkt[k];
                int f = k - 1;
                while (f >= k0 && bkt[f] > x) {
                    bkt[f + 1] = bkt[f];
                    --f;
                }
                bkt[f + 1] = x;
            }
            int tp = k0;
            for (int k = k0; k < k1; ++k) {
                const int r = bkt[k], L = w.ord[r];
                if (w.flag[L]) continue; /* taken by an earlier leader */
                const bool sm = w.selfm[L] != 0;
                const int st = tp;
                bool placed = !sm;
                for (int e = w.nbo[L]; e < w.nbo[L + 1]; ++e) {
                    const int j = w.nbr[e];
                    if (w.flag[j]) continue;
                    if (!placed && j > L) {
                        tmp[tp++] = L;
                        placed = true;
                    }
                    tmp[tp++] = j;
                    w.flag[j] = 1;
                }
                if (!placed) tmp[tp++] = L;
                if (sm) w.flag[L] = 1;
                const int gsz = tp - st;
                /* empty group: zero weight, the reference stops; a leader
                 * outside its group is taken again, finds nothing and stops */
                rinfo[2 * r] = gsz | (sm ? 0 : STOP);
                rinfo[2 * r + 1] = st;
                if (!sm || gsz == 0) break;
            }
        }
    }
    __syncthreads();
    if (t == 0) s_nc = U; /* first stopping rank */
    __syncthreads();
    for (int r = r0; r < r1; ++r)
        if (rinfo[2 * r] & STOP) {
            atomicMin(&s_nc, r);
            break;
        }
    __syncthreads();
    {
        const int rstop = s_nc; /* ranks after it form no group */
        int mg = 0, mm = 0;
        for (int r = r0; r < min(r1, rstop + 1); ++r) {
            const int n = rinfo[2 * r] & (STOP - 1);
            mg += n > 0;
            mm += n;
        }
        int ng;
        int gi = block_exscan(mg, sh, &ng);
        int mtot;
        int mp = block_exscan(mm, sh, &mtot);
        for (int r = r0; r < min(r1, rstop + 1); ++r) {
            const int n = rinfo[2 * r] & (STOP - 1);
            if (n == 0) continue;
            const int L = w.ord[r];
            a.glead[sbase + gi] = L;
            a.gsta[sbase + gi] = mp;
            a.gend[sbase + gi] = mp + n;
            if (w.nbo[L + 1] == w.nbo[L]) {
                w.mem[mp] = L;
            } else {
                const int s0 = rinfo[2 * r + 1];
                for (int q = 0; q < n; ++q) w.mem[mp + q] = tmp[s0 + q];
            }
            ++gi;
            mp += n;
        }
        if (t == 0) {
            a.ngroups[p] = ng;
            a.nmerged[p] = ng;
        }
    }
    __syncthreads();
    for (int i = t; i < U; i += NT) a.gmem[sbase + i] = w.mem[i];
}

__global__ __launch_bounds__(128) void k_merge_moments(MergeArgs a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.Stot) return;
    const int p = upper_parent(a.survoff, a.Np, i);
    const int64_t sbase = a.survoff[p];
    const int gi = (int)(i - sbase);
    if (gi >= a.ngroups[p]) return; /* also every slot of a parent left to the round loop */
    const int st = a.gsta[sbase + gi], en = a.gend[sbase + gi];
    double W = 0.0;
    for (int e = st; e < en; ++e) W += a.sw[sbase + a.gmem[sbase + e]];
    if (W == 0) atomicMin(&a.nmerged[p], gi);
    double mu[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) mu[c] = 0.0;
    for (int e = st; e < en; ++e) {
        const int64_t si = sbase + a.gmem[sbase + e];
        const double w = a.sw[si];
#pragma unroll
        for (int c = 0; c < 6; ++c) mu[c] += w * a.smu[(si) * 6 + c];
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) mu[c] = mu[c] / W;
    double P[21];
#pragma unroll
    for (int q = 0; q < 21; ++q) P[q] = 0.0;
    for (int e = st; e < en; ++e) {
        const int64_t si = sbase + a.gmem[sbase + e];
        const double w = a.sw[si];
        const int g = a.sg[si];
        double dl[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) dl[c] = mu[c] - a.smu[(si) * 6 + c];
#pragma unroll
        for (int r = 0; r < 6; ++r)
#pragma unroll
            for (int c = 0; c <= r; ++c) {
                const double Pe = term_P(a, g, pidx(r, c));
                P[pidx(r, c)] += w * (Pe + dl[c] * dl[r]);
            }
    }
#pragma unroll
    for (int q = 0; q < 21; ++q) P[q] = P[q] / W;
    const int64_t o = i;
    if (is_pd6(P)) {
        a.mw[o] = W;
#pragma unroll
        for (int c = 0; c < 6; ++c) a.mmu[c * a.scap + o] = mu[c];
#pragma unroll
        for (int q = 0; q < 21; ++q) a.mP[q * a.scap + o] = P[q];
    } else {
        const int64_t li = sbase + a.glead[sbase + gi];
        const int g = a.sg[li];
        a.mw[o] = a.sw[li];
#pragma unroll
        for (int c = 0; c < 6; ++c) a.mmu[c * a.scap + o] = a.smu[(li) * 6 + c];
#pragma unroll
        for (int q = 0; q < 21; ++q)
            a.mP[q * a.scap + o] = term_P(a, g, q);
    }
}

template <bool BIG>
__global__ __launch_bounds__(NT) void k_merge_rounds(MergeArgs a) {
    extern __shared__ char lds[];
    __shared__ int s_lead, s_pos, s_nm;
    __shared__ double s_W, s_lw, s_lmu[6], s_lP[21], s_mu[6], s_P[21];
    const int p = blockIdx.x;
    const int t = threadIdx.x;
    if (a.ngroups[p] != -1) return; /* uniform */
    const int64_t sbase = a.survoff[p];
    const int U = a.survoff[p + 1] - a.survoff[p];
    if (BIG != (U > a.lds_cap)) return;
    char* ws = BIG ? (a.scratch + sbase * MERGE_WS_PER_SLOT + (int64_t)p * MERGE_WS_PER_PARENT) : lds;
    MergeWs w(ws, U);
    const double thr = a.min_separation;
    merge_load_sort(a, sbase, U, w);
    for (int i = t; i < U; i += NT) w.selfm[i] = merge_self(a, sbase + i, thr) ? 1 : 0;
    if (t == 0) s_pos = 0;
    __syncthreads();
    const int lane = t & 63, wave = t >> 6;
    const int64_t mbase = sbase;
    int K = 0;
    uint8_t* memb = (uint8_t*)w.nbr; /* round membership bits; the neighbour-list area is free here */
    for (;;) {
        if (wave == 0) {
            int found = -1;
            for (int r0w = s_pos; r0w < U; r0w += 64) {
                const int r = r0w + lane;
                const bool c = r < U && !w.flag[w.ord[r]];
                const uint64_t b = __ballot(c);
                if (b) {
                    found = r0w + __ffsll((unsigned long long)b) - 1;
                    break;
                }
            }
            if (lane == 0) {
                /* restart at the leader's rank: a leader outside its own
                 * group stays selectable and the next round stops */
                s_lead = found >= 0 ? w.ord[found] : -1;
                s_pos = found >= 0 ? found : U;
            }
        }
        __syncthreads();
        const int lead = s_lead;
        if (lead < 0) break;
        if (t < 6) s_lmu[t] = a.smu[(sbase + lead) * 6 + t];
        if (t == 0) s_lw = a.sw[sbase + lead];
        if (t < 21) {
            const int g = a.sg[sbase + lead];
            s_lP[t] = term_P(a, g, t);
        }
        __syncthreads();
        for (int i = t; i < U; i += NT) {
            uint8_t mb = 0;
            if (!w.flag[i])
                mb = (i == lead ? w.selfm[i] != 0 : merge_pair(a, sbase, w, lead, i, thr)) ? 1 : 0;
            memb[i] = mb;
        }
        __syncthreads();
        if (wave == 0) {
            int nm = 0;
            for (int r0w = 0; r0w < U; r0w += 64) {
                const int i = r0w + lane;
                const bool mbit = i < U && memb[i];
                const uint64_t b = __ballot(mbit);
                if (mbit) w.mem[nm + __popcll(b & ((1ull << lane) - 1ull))] = i;
                nm += __popcll(b);
            }
            if (lane == 0) s_nm = nm;
        }
        __syncthreads();
        const int nm = s_nm;
        if (t == 0) {
            double W = 0.0;
            for (int j = 0; j < nm; ++j) W += a.sw[sbase + w.mem[j]];
            s_W = W;
        }
        __syncthreads();
        const double W = s_W;
        if (W == 0) break;
        if (t < 6) {
            double acc = 0.0;
            for (int j = 0; j < nm; ++j) {
                const int64_t si = sbase + w.mem[j];
                acc += a.sw[si] * a.smu[(si) * 6 + t];
            }
            s_mu[t] = acc / W;
        }
        __syncthreads();
        if (t < 21) {
            int r = 0;
            while ((r + 1) * (r + 2) / 2 <= t) ++r;
            const int c = t - r * (r + 1) / 2;
            double acc = 0.0;
            for (int j = 0; j < nm; ++j) {
                const int64_t si = sbase + w.mem[j];
                const int g = a.sg[si];
                const double Pe = term_P(a, g, t);
                const double dc = s_mu[c] - a.smu[(si) * 6 + c];
                const double dr = s_mu[r] - a.smu[(si) * 6 + r];
                acc += a.sw[si] * (Pe + dc * dr);
            }
            s_P[t] = acc / W;
        }
        __syncthreads();
        if (t == 0) {
            const bool ok = is_pd6(s_P);
            const int64_t o = mbase + K;
            a.mw[o] = ok ? W : s_lw;
            for (int c = 0; c < 6; ++c) a.mmu[c * a.scap + o] = ok ? s_mu[c] : s_lmu[c];
            for (int e = 0; e < 21; ++e) a.mP[e * a.scap + o] = ok ? s_P[e] : s_lP[e];
        }
        for (int j = t; j < nm; j += NT) w.flag[w.mem[j]] = 1;
        ++K;
        __syncthreads();
    }
    if (t == 0) a.nmerged[p] = K;
}

/* ------------------------------------------------------ parent weights */
/* w_p <- w_p exp(loglik_p), normalised, in the log domain with a max shift
 * (src/scphdcameracalibration.cu:1851-1875 multiplies in the linear
 * domain).  One workgroup. */
__global__ __launch_bounds__(1024) void k_parent_weights(double* pw, const double* loglik, int Np,
                                                         const int32_t* poff, const int32_t* rfixed) {
    __shared__ double red[16];
    if (rfixed && rfixed[blockIdx.x]) return; /* this rig's frame used the fixed camera */
    if (poff) { /* one workgroup per rig */
        pw += poff[blockIdx.x];
        loglik += poff[blockIdx.x];
        Np = poff[blockIdx.x + 1] - poff[blockIdx.x];
    }
    double mx = -INFINITY;
    for (int p = threadIdx.x; p < Np; p += blockDim.x) mx = fmax(mx, safe_log(pw[p]) + loglik[p]);
    mx = block_max(mx, red);
    double s = 0.0;
    for (int p = threadIdx.x; p < Np; p += blockDim.x) s += exp(safe_log(pw[p]) + loglik[p] - mx);
    s = block_sum(s, red);
    __syncthreads();
    for (int p = threadIdx.x; p < Np; p += blockDim.x) pw[p] = exp(safe_log(pw[p]) + loglik[p] - mx) / s;
}

/* ------------------------------------------------------ noise check */
/* Six normals per element elem = 0..n-1 of one counter stream, exactly as
 * the filter kernels draw them, for comparison with the CPU restatement. */
__global__ void k_debug_normals(NoiseKey key, uint32_t stream, uint64_t step, uint32_t parent, uint32_t gauss, int n,
                                double* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double z[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (stream == NS_FEATURE)
        feature_normals(key, step, parent, gauss, (uint32_t)i, z);
    else if (stream == NS_POSTERIOR)
        posterior_normals(key, step, parent, gauss, (uint32_t)i, z);
    else
        normals<6>(key, stream, step, parent, gauss, (uint32_t)i, z);
#pragma unroll
    for (int c = 0; c < 6; ++c) out[(int64_t)i * 6 + c] = z[c];
}

/* --------------------------------------------------------- estimate */
/* Sum each of N per-thread values over a 256-thread block; every thread
 * receives the totals (fixed order: wave sums, then waves ascending). */
template <int N>
__device__ inline void block_sum_n(double (&v)[N], double* sh) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const double t = wave_sum(v[i]);
        if (lane == 0) sh[w * N + i] = t;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = ((sh[i] + sh[N + i]) + sh[2 * N + i]) + sh[3 * N + i];
    __syncthreads();
}

/* One 256-thread workgroup per rig of Np parents: weighted mean and
 * covariance of the parent extrinsics and the summed Gaussian weight of the
 * first highest-weight parent's map (exp_traj, cov_traj and cardinality of
 * load_disparity_results.m:27-34).  out per rig: mean[12], cov[144]
 * row-major, cardinality (EST_DOUBLES). */
__global__ __launch_bounds__(256) void k_estimate(const double* ext, const double* pw, const int32_t* goff,
                                                   const double* gw, int Np, const int32_t* poff, double* out) {
    __shared__ double sh[4 * 78];
    if (poff) { /* one workgroup per rig */
        const int64_t P0 = poff[blockIdx.x];
        Np = poff[blockIdx.x + 1] - poff[blockIdx.x];
        ext += P0 * 12;
        pw += P0;
        goff += P0;
        out += (int64_t)blockIdx.x * EST_DOUBLES;
    }
    __shared__ double bws[4];
    __shared__ int bis[4];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    double s1[13];
#pragma unroll
    for (int c = 0; c < 13; ++c) s1[c] = 0.0;
    double bw = -1.0;
    int bi = 0x7fffffff;
    for (int p = t; p < Np; p += 256) {
        const double w = pw[p];
        s1[12] += w;
#pragma unroll
        for (int c = 0; c < 12; ++c) s1[c] += w * ext[(int64_t)p * 12 + c];
        if (w > bw) { /* p ascending within a thread: strict > keeps the first */
            bw = w;
            bi = p;
        }
    }
    /* first maximum: larger weight, then smaller index */
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double ow = __shfl_xor(bw, o, 64);
        const int oi = __shfl_xor(bi, o, 64);
        if (ow > bw || (ow == bw && oi < bi)) {
            bw = ow;
            bi = oi;
        }
    }
    if (lane == 0) {
        bws[wv] = bw;
        bis[wv] = bi;
    }
    block_sum_n<13>(s1, sh);
    int best = bis[0];
    double bestw = bws[0];
    for (int k = 1; k < 4; ++k)
        if (bws[k] > bestw || (bws[k] == bestw && bis[k] < best)) {
            bestw = bws[k];
            best = bis[k];
        }
    const double W = s1[12];
    const double r = W > 0 ? 1.0 / W : 0.0;
    double mu[12];
#pragma unroll
    for (int c = 0; c < 12; ++c) mu[c] = s1[c] * r;
    /* covariance one row per pass (12 running sums per thread) */
    for (int i = 0; i < 12; ++i) {
        double row[12];
#pragma unroll
        for (int j = 0; j < 12; ++j) row[j] = 0.0;
        for (int p = t; p < Np; p += 256) {
            const double* x = ext + (int64_t)p * 12;
            const double wdi = pw[p] * (x[i] - mu[i]);
#pragma unroll
            for (int j = 0; j < 12; ++j) row[j] += wdi * (x[j] - mu[j]);
        }
        block_sum_n<12>(row, sh);
        if (t < 12) out[12 + i * 12 + t] = row[t] * r;
    }
    double card[1] = {0.0};
    if (best < Np)
        for (int64_t g = goff[best] + t; g < goff[best + 1]; g += 256) card[0] += gw[g];
    block_sum_n<1>(card, sh);
    if (t < 12) out[t] = mu[t];
    if (t == 0) {
        out[156] = card[0];
    }
}

/* ---------------------------------------------------- posterior samples */
/* the block of merged Gaussian gm: a dropped non-detection term's block
 * while there are any, else the next one after them */
__device__ inline int64_t posterior_block(const int32_t* holes, int64_t nholes, int64_t post_base, int64_t gm) {
    return gm < nholes ? (int64_t)holes[gm] : post_base + (gm - nholes);
}

/* Sixteen lanes (one row of a wavefront) per merged Gaussian, four per
 * wavefront: S draws x = L z + mu with L the Cholesky factor of the merged
 * covariance (sample_disparity_gaussian, src/scphdcameracalibration.cu:640-690),
 * factored once per lane and used for every sample the lane draws, mapped
 * to the world by the parent's camera (disparityToWorldKernel,
 * src/kernels.cu:49-80), sample weight 1/S (:2009-2010).  Merged Gaussian
 * gm's cloud goes to the block of a dropped non-detection term while there
 * are any (holes), else after them (posterior_block).  Destination
 * addresses are byte offsets from the plane bases (32-bit while every
 * destination lies below 4 GiB). */
template <typename T, typename Off, bool RG, typename A = double>
__global__ __launch_bounds__(256) void k_sample(SampleArgs a, Planes<T> dst) {
    constexpr bool F32 = std::is_same_v<A, float>;
    const int lane = threadIdx.x & 15;
    const int64_t gm = (int64_t)blockIdx.x * 16 + (threadIdx.x >> 4);
    if (gm >= a.Kt) return; /* whole rows */
    const int p = upper_parent(a.mergeoff, a.Np, gm);
    const int k = (int)(gm - a.mergeoff[p]);
    int pl;
    const auto key = row_key<RG>(a.key, p, pl);
    const RowMeas<RG> rm(a.meas, a.rig_meas, a.key, p);
    const DisparityMeasurementModel& meas = rm.m;
    const int64_t rec = (int64_t)a.survoff[p] + k;
    double mu[6], P[21], L[21];
#pragma unroll
    for (int c = 0; c < 6; ++c) mu[c] = a.mmu[c * a.scap + rec];
#pragma unroll
    for (int e = 0; e < 21; ++e) P[e] = a.mP[e * a.scap + rec];
    chol6(P, L);
    const CameraPose pose = load_pose(a.pose, p);
    const T w = (T)(1.0 / a.S);
    Off off = (Off)((posterior_block(a.holes, a.nholes, a.post_base, gm) * (int64_t)a.S + lane) * (int64_t)sizeof(T));
    PosteriorCarry pc; /* normals carried between the two steps of a noise group */
    if constexpr (F32) {
        /* the FP32-arithmetic variant: the factor and mean rounded to
         * single precision once, each draw and its unprojection in single
         * precision */
        const MeasF mf = meas_f(meas);
        const PoseF pf = pose_f(pose);
        float Lf[21], muf[6];
#pragma unroll
        for (int e = 0; e < 21; ++e) Lf[e] = (float)L[e];
#pragma unroll
        for (int c = 0; c < 6; ++c) muf[c] = (float)mu[c];
        for (int s = lane; s < a.S; s += 16) {
            float z[6], q[6], x[6];
            posterior_normals_seq(key, a.ustep, (uint32_t)pl, (uint32_t)k, (uint32_t)s, (uint32_t)(s >> 4) & 1u, pc,
                                  z);
#pragma unroll
            for (int r = 0; r < 6; ++r) {
                float acc = muf[r];
#pragma unroll
                for (int c = 0; c <= r; ++c) acc = __builtin_fmaf(Lf[pidx(r, c)], z[c], acc);
                q[r] = acc;
            }
            world_from_disparity_f(mf, pf, q, x);
#pragma unroll
            for (int c = 0; c < 6; ++c) at(dst.p[c], off) = (T)x[c];
            at(dst.p[6], off) = w;
            off += (Off)(16 * sizeof(T));
        }
        return;
    }
    for (int s = lane; s < a.S; s += 16) {
        double z[6];
        posterior_normals_seq(key, a.ustep, (uint32_t)pl, (uint32_t)k, (uint32_t)s, (uint32_t)(s >> 4) & 1u, pc, z);
        double q[6];
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            double acc = 0.0;
#pragma unroll
            for (int c = 0; c <= r; ++c) acc += L[pidx(r, c)] * z[c];
            q[r] = acc + mu[r];
        }
        const DisparityPoint d{q[0], q[1], q[2], q[3], q[4], q[5]};
        const EuclideanPoint x = world_from_disparity_fast(meas, pose, d);
        at(dst.p[0], off) = (T)x.x;
        at(dst.p[1], off) = (T)x.y;
        at(dst.p[2], off) = (T)x.z;
        at(dst.p[3], off) = (T)x.vx;
        at(dst.p[4], off) = (T)x.vy;
        at(dst.p[5], off) = (T)x.vz;
        at(dst.p[6], off) = w;
        off += (Off)(16 * sizeof(T));
    }
}

/* -------------------------------------------------------- map assembly */
__global__ void k_outcount(const int32_t* nmerged, const int32_t* goff, const int32_t* keepscan, int Np,
                           int32_t* outcnt) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= Np) return;
    outcnt[p] = nmerged[p] + (keepscan[goff[p + 1]] - keepscan[goff[p]]);
}

__global__ void k_ndlist(const int32_t* keep, const int32_t* keepscan, int64_t Ng, int32_t* ndlist,
                         int32_t* holes) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= Ng) return;
    if (keep[g])
        ndlist[keepscan[g]] = (int32_t)g;
    else if (holes)
        holes[g - keepscan[g]] = (int32_t)g;
}


/* New map table per parent: [merged posterior Gaussians | kept
 * non-detection Gaussians in input order] (recombineNonDetections,
 * src/scphdcameracalibration.cu:1422-1574, weights w (1 - pd) for the
 * non-detection terms, :1314-1318). */
__global__ void k_table(TableArgs a) {
    const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= a.Nout) return;
    const int p = upper_parent(a.outoff, a.Np, o);
    const int j = (int)(o - a.outoff[p]);
    const int K = a.mergeoff[p + 1] - a.mergeoff[p];
    if (j < K) {
        a.gblock[o] = (int32_t)posterior_block(a.holes, a.nholes, a.post_base, (int64_t)a.mergeoff[p] + j);
        a.gwo[o] = a.mw[(int64_t)a.survoff[p] + j];
        a.gpd[o] = 1.0f;
    } else {
        const int g = a.ndlist[a.keepscan[a.goff[p]] + (j - K)];
        const double pdg = a.fit[(int64_t)FP_PD * a.Ng_in + g];
        a.gblock[o] = (int32_t)(a.nd_base + g);
        a.gwo[o] = a.gw[g] * (1 - pdg);
        a.gpd[o] = (float)pdg;
    }
    a.gparent[o] = p;
}

/* ---------------------------------------------------------- resampling */
/* n_eff = 1 / (N sum w^2) (computeNeff :2641-2649); when below min_neff,
 * systematic resampling with points u + j/N against the sequential
 * cumulative sum of the weights (resample :2651-2702), the overrun fallback
 * to the first maximum of exp(w), gather of parent states, and weights 1/N.
 * One workgroup.  The cumulative sum keeps the reference's left-to-right
 * rounding: thread 0 accumulates chunks staged in LDS. */
static constexpr int RS_CH = 2048;

__global__ __launch_bounds__(1024) void k_resample(ResampleArgs a) {
    __shared__ double red[16];
    __shared__ double s_u;
    __shared__ double chunk[RS_CH];
    __shared__ double s_em[16];
    __shared__ int s_am[16];
    /* one workgroup per rig: its Np parents from P0, ancestors within it */
    const int rig = blockIdx.x;
    const int N = a.poff ? a.poff[rig + 1] - a.poff[rig] : a.Np;
    const int64_t P0 = a.poff ? a.poff[rig] : 0;
    a.pw += P0;
    a.ext += P0 * 12;
    a.ext_tmp += P0 * 12;
    a.cum += P0;
    a.anc += P0;
    a.newcnt += P0;
    a.goff += P0;
    const NoiseKey key = a.keys ? a.keys[rig] : a.key;
    /* the rig's own resample count: the handle's less its count at joining */
    const uint64_t rstep = a.rstep - (a.sbase ? a.sbase[SC_COUNT * rig + SC_RESAMPLE] : 0u);
    const int t = threadIdx.x;
    double ss = 0.0;
    for (int p = t; p < N; p += blockDim.x) ss += a.pw[p] * a.pw[p];
    ss = block_sum(ss, red);
    const double neff = 1.0 / (N * ss);
    const bool go = !(neff >= a.min_neff);
    if (t == 0) {
        a.neff[rig] = neff;
        if (a.rflag) a.rflag[rig] = go ? 1 : 0;
        if (go) atomicOr(a.flag, 1); /* any rig resampled; zero before the launch */
    }
    if (!go) {
        for (int p = t; p < N; p += blockDim.x) {
            a.anc[p] = (int32_t)(P0 + p);
            a.newcnt[p] = a.goff[p + 1] - a.goff[p];
        }
        return;
    }
    /* first index of the maximum of exp(w) */
    double em = -1.0;
    int am = -1;
    for (int i = t; i < N; i += blockDim.x) {
        const double e = exp(a.pw[i]);
        if (e > em) {
            em = e;
            am = i;
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double e2 = __shfl_xor(em, o, 64);
        const int a2 = __shfl_xor(am, o, 64);
        if (e2 > em || (e2 == em && a2 >= 0 && (am < 0 || a2 < am))) {
            em = e2;
            am = a2;
        }
    }
    if ((t & 63) == 0) {
        s_em[t >> 6] = em;
        s_am[t >> 6] = am;
    }
    if (t == 0) {
        uint32_t x[4];
        noise_block(key, NS_RESAMPLE, rstep, 0, 0, 0, 0, x);
        s_u = uniform_co(x[0], x[1]) * (1.0 / N);
    }
    /* sequential cumulative sum */
    double c = 0.0;
    for (int c0 = 0; c0 < N; c0 += RS_CH) {
        const int n = min(RS_CH, N - c0);
        __syncthreads();
        for (int i = t; i < n; i += blockDim.x) chunk[i] = a.pw[c0 + i];
        __syncthreads();
        if (t == 0) {
            int i = 0;
            if (c0 == 0) {
                c = chunk[0];
                a.cum[0] = c;
                i = 1;
            }
            for (; i < n; ++i) {
                c += chunk[i];
                a.cum[c0 + i] = c;
            }
        }
    }
    __syncthreads();
    int amax = -1;
    {
        double bm = -1.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
            if (s_am[w] >= 0 && (s_em[w] > bm || (s_em[w] == bm && s_am[w] < amax))) {
                bm = s_em[w];
                amax = s_am[w];
            }
        }
    }
    const double interval = 1.0 / N;
    for (int j = t; j < N; j += blockDim.x) {
        const double r = s_u + j * interval;
        int lo = 0, hi = N;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (!(r > a.cum[mid]))
                hi = mid;
            else
                lo = mid + 1;
        }
        const int anc = lo < N ? lo : amax;
        a.anc[j] = (int32_t)(P0 + anc);
        for (int q = 0; q < 12; ++q) a.ext_tmp[(int64_t)j * 12 + q] = a.ext[(int64_t)anc * 12 + q];
        a.newcnt[j] = a.goff[anc + 1] - a.goff[anc];
    }
    if (a.defer) return; /* k_resample_commit applies the gather */
    __syncthreads();
    for (int j = t; j < N; j += blockDim.x) {
        for (int q = 0; q < 12; ++q) a.ext[(int64_t)j * 12 + q] = a.ext_tmp[(int64_t)j * 12 + q];
        a.pw[j] = 1.0 / N;
    }
}

__global__ void k_resample_commit(double* ext, const double* ext_tmp, double* pw, int32_t* anc_out,
                                  const int32_t* anc, int N, int go) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    anc_out[j] = anc[j];
    if (go) {
        for (int q = 0; q < 12; ++q) ext[(int64_t)j * 12 + q] = ext_tmp[(int64_t)j * 12 + q];
        pw[j] = 1.0 / N;
    }
}

__global__ void k_fill_i32(int32_t* x, int32_t v, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}

__global__ void k_iota(int32_t* x, int64_t n, int32_t start) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = start + (int32_t)i;
}

/* ---------------------------------------------- ensemble membership */
/* The map table without one rig's Gaussians [G0, G0 + Gr): entry i of the
 * result is entry i below G0, else entry i + Gr, whose parent index (at or
 * after the removed rig's parents, which end at P1) moves down by Nr. */
__global__ void k_table_remove(const int32_t* gparent, const int32_t* gblock, const double* gw, const float* gpd,
                               int64_t G0, int64_t Gr, int64_t Ng_new, int P1, int Nr, int32_t* gparent_out,
                               int32_t* gblock_out, double* gw_out, float* gpd_out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Ng_new) return;
    const int64_t j = i < G0 ? i : i + Gr;
    const int32_t p = gparent[j];
    gparent_out[i] = p >= P1 ? p - Nr : p;
    gblock_out[i] = gblock[j];
    gw_out[i] = gw[j];
    if (gpd_out) gpd_out[i] = gpd[j];
}

__global__ void k_goff_remove(const int32_t* goff, int P0, int Nr, int32_t Gr, int Np_new, int32_t* goff_out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > Np_new) return;
    goff_out[p] = p < P0 ? goff[p] : goff[p + Nr] - Gr;
}

__global__ void k_index_shift(int32_t* x, int64_t n, int P1, int Nr) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && x[i] >= P1) x[i] -= Nr;
}

__global__ void k_table_resample(const int32_t* anc, const int32_t* goff_old, const int32_t* goff_new,
                                 const int32_t* gblock_old, const double* gw_old, int Np, int64_t Nnew,
                                 int32_t* gparent, int32_t* gblock, double* gw, const float* gpd_old,
                                 float* gpd) {
    const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= Nnew) return;
    const int p = upper_parent(goff_new, Np, o);
    const int64_t src = (int64_t)goff_old[anc[p]] + (o - goff_new[p]);
    gparent[o] = p;
    gblock[o] = gblock_old[src];
    gw[o] = gw_old[src];
    if (gpd) gpd[o] = gpd_old[src];
}

/* ------------------------------------------------------------ read-back */
/* The map's clouds into consecutive blocks: block g of dst <- block
 * gblock[g] of src, every plane. */
template <typename T>
__global__ void k_gather(Planes<T> src, Planes<T> dst, const int32_t* gblock, int64_t Ng, int S) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Ng * S) return;
    const int64_t g = i / S;
    const int64_t si = (int64_t)gblock[g] * S + (i - g * S);
#pragma unroll
    for (int c = 0; c < 7; ++c) dst.p[c][i] = src.p[c][si];
}

template <typename T>
__global__ void k_extract(Planes<T> src, const int32_t* gblock, int64_t Ng, int S, double* out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = Ng * S;
    if (i >= n) return;
    const int64_t g = i / S;
    const int64_t si = (int64_t)gblock[g] * S + (i - g * S);
#pragma unroll
    for (int c = 0; c < 7; ++c) out[c * n + i] = (double)src.p[c][si];
}

template <typename T>
__global__ void k_inject(const double* in, int64_t Ng, int S, Planes<T> dst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = Ng * S;
    if (i >= n) return;
#pragma unroll
    for (int c = 0; c < 7; ++c) dst.p[c][i] = (T)in[c * n + i];
}

/* ----------------------------------------------------------------- scan */
static constexpr int SCAN_T = 256, SCAN_I = 8, SCAN_TILE = SCAN_T * SCAN_I;

__global__ __launch_bounds__(SCAN_T) void k_scan_tiles(const int32_t* in, int32_t* out, int64_t n, int32_t* sums) {
    __shared__ int sh[8];
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_I;
    int v[SCAN_I];
    int s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_I; ++i) {
        v[i] = (base + i < n) ? in[base + i] : 0;
        s += v[i];
    }
    int tot;
    int off = block_exscan(s, sh, &tot);
#pragma unroll
    for (int i = 0; i < SCAN_I; ++i) {
        if (base + i < n) out[base + i] = off;
        off += v[i];
    }
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(SCAN_T) void k_scan_sums(int32_t* sums, int64_t nb, int32_t* total_out) {
    __shared__ int sh[8];
    int carry = 0;
    for (int64_t c0 = 0; c0 < nb; c0 += SCAN_T) {
        const int64_t i = c0 + threadIdx.x;
        const int v = i < nb ? sums[i] : 0;
        int tot;
        const int o = block_exscan(v, sh, &tot);
        if (i < nb) sums[i] = carry + o;
        carry += tot;
    }
    if (threadIdx.x == 0) *total_out = carry;
}

__global__ void k_scan_add(int32_t* out, int64_t n, const int32_t* sums) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += sums[i / SCAN_TILE];
}

/* ============================================================ launchers */
static inline unsigned nblocks(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

#define LAUNCH_CHECK() return hipGetLastError()

hipError_t launch_prior(hipStream_t s, double* ext, double* pw, int Np, const double* a12, const double* b12,
                        NoiseKey key) {
    Bounds12 ab;
    for (int i = 0; i < 12; ++i) {
        ab.a[i] = a12[i];
        ab.b[i] = b12[i];
    }
    k_prior<<<nblocks(Np, 256), 256, 0, s>>>(ext, pw, Np, ab, key);
    LAUNCH_CHECK();
}

hipError_t launch_parent_predict(hipStream_t s, double* ext, int Np, OrientedLinearCVMotionModel3D cam, double dt,
                                 RigKeys key, uint64_t step) {
    k_parent_predict<<<nblocks(Np, 256), 256, 0, s>>>(ext, Np, cam, dt, key, step);
    LAUNCH_CHECK();
}

hipError_t launch_scale(hipStream_t s, double* x, int64_t n, double a) {
    if (n <= 0) return hipSuccess;
    k_scale<<<nblocks(n, 256), 256, 0, s>>>(x, n, a);
    LAUNCH_CHECK();
}

hipError_t launch_pose(hipStream_t s, const double* ext, double* pose, int Np, int fixed, const int32_t* prig,
                       const int32_t* rfixed) {
    if (rfixed && !prig) return hipErrorInvalidValue;
    k_pose<<<nblocks(Np, 256), 256, 0, s>>>(ext, pose, Np, fixed, prig, rfixed);
    LAUNCH_CHECK();
}

template <typename T>
hipError_t launch_materialize(hipStream_t s, Planes<T> src, Planes<T> dst, const int32_t* gblock,
                              const int32_t* gparent, const int32_t* goff, int64_t Ng, int S,
                              LinearCVMotionModel3D feat, double dt, RigKeys key, uint64_t step, bool arith32) {
    if (Ng <= 0) return hipSuccess;
    if (arith32 && !std::is_same_v<T, float>) return hipErrorInvalidValue;
    k_materialize<T><<<nblocks(Ng * S, 256), 256, 0, s>>>(src, dst, gblock, gparent, goff, Ng, S, feat, dt, key,
                                                          step, arith32 ? 1 : 0);
    LAUNCH_CHECK();
}

template <typename T>
hipError_t launch_fit(hipStream_t s, const FitArgs& a, Planes<T> src, Planes<T> dst) {
    if (a.Ng <= 0) return hipSuccess;
    const unsigned nb = nblocks(a.Ng, 16);
    if constexpr (std::is_same_v<T, float>) {
        if (a.arith32) {
            if (a.rig_meas) {
                if (a.off32)
                    k_fit<T, uint32_t, true, float><<<nb, 256, 0, s>>>(a, src, dst);
                else
                    k_fit<T, uint64_t, true, float><<<nb, 256, 0, s>>>(a, src, dst);
            } else if (a.off32)
                k_fit<T, uint32_t, false, float><<<nb, 256, 0, s>>>(a, src, dst);
            else
                k_fit<T, uint64_t, false, float><<<nb, 256, 0, s>>>(a, src, dst);
            LAUNCH_CHECK();
        }
    } else if (a.arith32) {
        return hipErrorInvalidValue;
    }
    if (a.rig_meas) {
        if (a.off32)
            k_fit<T, uint32_t, true><<<nb, 256, 0, s>>>(a, src, dst);
        else
            k_fit<T, uint64_t, true><<<nb, 256, 0, s>>>(a, src, dst);
    } else if (a.off32)
        k_fit<T, uint32_t, false><<<nb, 256, 0, s>>>(a, src, dst);
    else
        k_fit<T, uint64_t, false><<<nb, 256, 0, s>>>(a, src, dst);
    LAUNCH_CHECK();
}

template <typename T>
hipError_t launch_fit_fix(hipStream_t s, const FitArgs& a, Planes<T> src, Planes<T> dst) {
    if (a.Ng <= 0) return hipSuccess;
    const unsigned nb = (unsigned)std::min<int64_t>(nblocks(a.Ng, 16), 256);
    if constexpr (std::is_same_v<T, float>) {
        if (a.arith32) {
            if (a.rig_meas)
                k_fit_fix<T, true, float><<<nb, 256, 0, s>>>(a, src, dst);
            else
                k_fit_fix<T, false, float><<<nb, 256, 0, s>>>(a, src, dst);
            LAUNCH_CHECK();
        }
    } else if (a.arith32) {
        return hipErrorInvalidValue;
    }
    if (a.rig_meas)
        k_fit_fix<T, true><<<nb, 256, 0, s>>>(a, src, dst);
    else
        k_fit_fix<T, false><<<nb, 256, 0, s>>>(a, src, dst);
    LAUNCH_CHECK();
}

template <typename T>
hipError_t launch_nd_fill(hipStream_t s, const FitArgs& a, Planes<T> src, Planes<T> dst) {
    if (a.Ng <= 0) return hipSuccess;
    const int64_t blocks = std::min<int64_t>(nblocks(a.Ng, 16), 2048); /* the list is at most Ng long */
    if constexpr (std::is_same_v<T, float>) {
        if (a.arith32) {
            if (a.rig_meas) {
                if (a.off32)
                    k_nd_fill<T, uint32_t, true, float><<<blocks, 256, 0, s>>>(a, src, dst);
                else
                    k_nd_fill<T, uint64_t, true, float><<<blocks, 256, 0, s>>>(a, src, dst);
            } else if (a.off32)
                k_nd_fill<T, uint32_t, false, float><<<blocks, 256, 0, s>>>(a, src, dst);
            else
                k_nd_fill<T, uint64_t, false, float><<<blocks, 256, 0, s>>>(a, src, dst);
            LAUNCH_CHECK();
        }
    } else if (a.arith32) {
        return hipErrorInvalidValue;
    }
    if (a.rig_meas) {
        if (a.off32)
            k_nd_fill<T, uint32_t, true><<<blocks, 256, 0, s>>>(a, src, dst);
        else
            k_nd_fill<T, uint64_t, true><<<blocks, 256, 0, s>>>(a, src, dst);
    } else if (a.off32)
        k_nd_fill<T, uint32_t, false><<<blocks, 256, 0, s>>>(a, src, dst);
    else
        k_nd_fill<T, uint64_t, false><<<blocks, 256, 0, s>>>(a, src, dst);
    LAUNCH_CHECK();
}

hipError_t launch_precompute(hipStream_t s, const double* fit, const double* gw, int64_t Ng, double var_u,
                             double var_v, double* pre, double* lik) {
    if (Ng <= 0) return hipSuccess;
    k_precompute<<<nblocks(Ng, 128), 128, 0, s>>>(fit, gw, Ng, var_u, var_v, pre, lik);
    LAUNCH_CHECK();
}

hipError_t launch_eta(hipStream_t s, const EtaArgs& a) {
    k_eta<<<a.Np, NT, 0, s>>>(a);
    LAUNCH_CHECK();
}

hipError_t launch_survivors(hipStream_t s, const SurvArgs& a) {
    k_survivors<<<a.e.Np, NT, 0, s>>>(a);
    LAUNCH_CHECK();
}

hipError_t launch_merge(hipStream_t s, const MergeArgs& a, int lds_bytes) {
    k_merge_groups<false><<<a.Np, NT, lds_bytes, s>>>(a);
    if (a.scratch) k_merge_groups<true><<<a.Np, NT, a.big_lds_bytes, s>>>(a);
    if (a.Stot > 0) k_merge_moments<<<nblocks(a.Stot, 128), 128, 0, s>>>(a);
    k_merge_rounds<false><<<a.Np, NT, lds_bytes, s>>>(a);
    if (a.scratch) k_merge_rounds<true><<<a.Np, NT, 0, s>>>(a);
    LAUNCH_CHECK();
}

hipError_t launch_parent_weights(hipStream_t s, double* pw, const double* loglik, int Np, const int32_t* poff,
                                 int nrigs, const int32_t* rfixed) {
    if (rfixed && !poff) return hipErrorInvalidValue;
    k_parent_weights<<<poff ? nrigs : 1, 1024, 0, s>>>(pw, loglik, Np, poff, rfixed);
    LAUNCH_CHECK();
}

hipError_t launch_debug_normals(hipStream_t s, NoiseKey key, uint32_t stream, uint64_t step, uint32_t parent,
                                uint32_t gauss, int n, double* out) {
    if (n <= 0) return hipSuccess;
    k_debug_normals<<<nblocks(n, 256), 256, 0, s>>>(key, stream, step, parent, gauss, n, out);
    LAUNCH_CHECK();
}

hipError_t launch_estimate(hipStream_t s, const double* ext, const double* pw, const int32_t* goff,
                           const double* gw, int Np, const int32_t* poff, int nrigs, double* out) {
    if (poff ? nrigs <= 0 : Np <= 0) return hipErrorInvalidValue;
    k_estimate<<<poff ? nrigs : 1, 256, 0, s>>>(ext, pw, goff, gw, Np, poff, out);
    LAUNCH_CHECK();
}

template <typename T>
hipError_t launch_sample(hipStream_t s, const SampleArgs& a, Planes<T> dst) {
    if (a.Kt <= 0) return hipSuccess;
    const unsigned nb = nblocks(a.Kt, 16);
    if constexpr (std::is_same_v<T, float>) {
        if (a.arith32) {
            if (a.rig_meas) {
                if (a.off32)
                    k_sample<T, uint32_t, true, float><<<nb, 256, 0, s>>>(a, dst);
                else
                    k_sample<T, uint64_t, true, float><<<nb, 256, 0, s>>>(a, dst);
            } else if (a.off32)
                k_sample<T, uint32_t, false, float><<<nb, 256, 0, s>>>(a, dst);
            else
                k_sample<T, uint64_t, false, float><<<nb, 256, 0, s>>>(a, dst);
            LAUNCH_CHECK();
        }
    } else if (a.arith32) {
        return hipErrorInvalidValue;
    }
    if (a.rig_meas) {
        if (a.off32)
            k_sample<T, uint32_t, true><<<nb, 256, 0, s>>>(a, dst);
        else
            k_sample<T, uint64_t, true><<<nb, 256, 0, s>>>(a, dst);
    } else if (a.off32)
        k_sample<T, uint32_t, false><<<nb, 256, 0, s>>>(a, dst);
    else
        k_sample<T, uint64_t, false><<<nb, 256, 0, s>>>(a, dst);
    LAUNCH_CHECK();
}

hipError_t launch_outcount(hipStream_t s, const int32_t* nmerged, const int32_t* goff, const int32_t* keepscan,
                           int Np, int32_t* outcnt) {
    k_outcount<<<nblocks(Np, 256), 256, 0, s>>>(nmerged, goff, keepscan, Np, outcnt);
    LAUNCH_CHECK();
}

hipError_t launch_ndlist(hipStream_t s, const int32_t* keep, const int32_t* keepscan, int64_t Ng, int32_t* ndlist,
                         int32_t* holes) {
    if (Ng <= 0) return hipSuccess;
    k_ndlist<<<nblocks(Ng, 256), 256, 0, s>>>(keep, keepscan, Ng, ndlist, holes);
    LAUNCH_CHECK();
}

hipError_t launch_table(hipStream_t s, const TableArgs& a) {
    if (a.Nout <= 0) return hipSuccess;
    k_table<<<nblocks(a.Nout, 256), 256, 0, s>>>(a);
    LAUNCH_CHECK();
}

hipError_t launch_resample(hipStream_t s, const ResampleArgs& a) {
    k_resample<<<a.poff ? a.nrigs : 1, 1024, 0, s>>>(a);
    LAUNCH_CHECK();
}

/* Scalars the host needs after a scan, written by the device straight into
 * mapped host memory, so that a read-back costs no copy-engine transfer:
 * out[0] = off[n], out[1] = the largest segment off[p + 1] - off[p],
 * out[2] = the number of segments above `limit` (0 when limit <= 0),
 * out[3] = *x0 and out[4] = *x1 (0 for a null pointer).  One workgroup. */
__global__ __launch_bounds__(1024) void k_readout(const int32_t* off, int n, int32_t limit, const int32_t* x0,
                                                  const int32_t* x1, volatile int32_t* out) {
    __shared__ int smax[16], scnt[16];
    int mx = 0, c = 0;
    for (int p = threadIdx.x; p < n; p += blockDim.x) {
        const int seg = off[p + 1] - off[p];
        mx = max(mx, seg);
        c += (limit > 0 && seg > limit) ? 1 : 0;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        mx = max(mx, __shfl_xor(mx, o, 64));
        c += __shfl_xor(c, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        smax[threadIdx.x >> 6] = mx;
        scnt[threadIdx.x >> 6] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
            mx = max(mx, smax[w]);
            c += scnt[w];
        }
        out[0] = off[n];
        out[1] = mx;
        out[2] = c;
        out[3] = x0 ? *x0 : 0;
        out[4] = x1 ? *x1 : 0;
    }
}

/* out = the number of NaN entries of w[0, n) (checkStuff,
 * src/scphdcameracalibration.cu:812-821, without the host copy).  One
 * workgroup. */
__global__ __launch_bounds__(1024) void k_nan_count(const double* w, int n, volatile int32_t* out) {
    __shared__ int sc[16];
    int c = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) c += w[i] != w[i] ? 1 : 0;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
    if ((threadIdx.x & 63) == 0) sc[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < (int)(blockDim.x >> 6); ++k) c += sc[k];
        *out = c;
    }
}

/* Detection pairs of an ensemble update: the sum over rigs of the rig's
 * map size times its measurement count (exact in FP64 below 2^53).  One
 * workgroup. */
__global__ __launch_bounds__(256) void k_pair_count(const int32_t* goff, const int32_t* moff, const int32_t* poff,
                                                    int nrigs, volatile double* out) {
    __shared__ double red[16];
    double c = 0.0;
    for (int r = threadIdx.x; r < nrigs; r += blockDim.x)
        c += (double)(goff[poff[r + 1]] - goff[poff[r]]) * (double)(moff[r + 1] - moff[r]);
    c = block_sum(c, red);
    if (threadIdx.x == 0) *out = c;
}

hipError_t launch_pair_count(hipStream_t s, const int32_t* goff, const int32_t* moff, const int32_t* poff,
                             int nrigs, double* out) {
    k_pair_count<<<1, 256, 0, s>>>(goff, moff, poff, nrigs, out);
    LAUNCH_CHECK();
}

hipError_t launch_nan_count(hipStream_t s, const double* w, int n, int32_t* out) {
    k_nan_count<<<1, 1024, 0, s>>>(w, n, out);
    LAUNCH_CHECK();
}

hipError_t launch_readout(hipStream_t s, const int32_t* off, int n, int32_t limit, const int32_t* x0,
                          const int32_t* x1, int32_t* out) {
    k_readout<<<1, 1024, 0, s>>>(off, n, limit, x0, x1, out);
    LAUNCH_CHECK();
}

hipError_t launch_resample_commit(hipStream_t s, double* ext, const double* ext_tmp, double* pw, int32_t* anc_out,
                                  const int32_t* anc, int N, int go) {
    if (N <= 0) return hipSuccess;
    k_resample_commit<<<nblocks(N, 256), 256, 0, s>>>(ext, ext_tmp, pw, anc_out, anc, N, go);
    LAUNCH_CHECK();
}

hipError_t launch_fill_i32(hipStream_t s, int32_t* x, int32_t v, int64_t n) {
    if (n <= 0) return hipSuccess;
    k_fill_i32<<<nblocks(n, 256), 256, 0, s>>>(x, v, n);
    LAUNCH_CHECK();
}

hipError_t launch_iota(hipStream_t s, int32_t* x, int64_t n, int32_t start) {
    if (n <= 0) return hipSuccess;
    k_iota<<<nblocks(n, 256), 256, 0, s>>>(x, n, start);
    LAUNCH_CHECK();
}

hipError_t launch_table_remove(hipStream_t s, const int32_t* gparent, const int32_t* gblock, const double* gw,
                               const float* gpd, int64_t G0, int64_t Gr, int64_t Ng_new, int P1, int Nr,
                               int32_t* gparent_out, int32_t* gblock_out, double* gw_out, float* gpd_out) {
    if (Ng_new <= 0) return hipSuccess;
    if ((gpd_out != nullptr) != (gpd != nullptr)) return hipErrorInvalidValue;
    k_table_remove<<<nblocks(Ng_new, 256), 256, 0, s>>>(gparent, gblock, gw, gpd, G0, Gr, Ng_new, P1, Nr,
                                                        gparent_out, gblock_out, gw_out, gpd_out);
    LAUNCH_CHECK();
}

hipError_t launch_goff_remove(hipStream_t s, const int32_t* goff, int P0, int Nr, int32_t Gr, int Np_new,
                              int32_t* goff_out) {
    k_goff_remove<<<nblocks((int64_t)Np_new + 1, 256), 256, 0, s>>>(goff, P0, Nr, Gr, Np_new, goff_out);
    LAUNCH_CHECK();
}

hipError_t launch_index_shift(hipStream_t s, int32_t* x, int64_t n, int P1, int Nr) {
    if (n <= 0) return hipSuccess;
    k_index_shift<<<nblocks(n, 256), 256, 0, s>>>(x, n, P1, Nr);
    LAUNCH_CHECK();
}

hipError_t launch_table_resample(hipStream_t s, const int32_t* anc, const int32_t* goff_old,
                                 const int32_t* goff_new, const int32_t* gblock_old, const double* gw_old, int Np,
                                 int64_t Nnew, int32_t* gparent, int32_t* gblock, double* gw,
                                 const float* gpd_old, float* gpd) {
    if (Nnew <= 0) return hipSuccess;
    k_table_resample<<<nblocks(Nnew, 256), 256, 0, s>>>(anc, goff_old, goff_new, gblock_old, gw_old, Np, Nnew,
                                                         gparent, gblock, gw, gpd_old, gpd);
    LAUNCH_CHECK();
}

template <typename T>
hipError_t launch_gather(hipStream_t s, Planes<T> src, Planes<T> dst, const int32_t* gblock, int64_t Ng, int S) {
    if (Ng <= 0) return hipSuccess;
    k_gather<T><<<nblocks(Ng * S, 256), 256, 0, s>>>(src, dst, gblock, Ng, S);
    LAUNCH_CHECK();
}

template <typename T>
hipError_t launch_extract(hipStream_t s, Planes<T> src, const int32_t* gblock, int64_t Ng, int S, double* out) {
    if (Ng <= 0) return hipSuccess;
    k_extract<T><<<nblocks(Ng * S, 256), 256, 0, s>>>(src, gblock, Ng, S, out);
    LAUNCH_CHECK();
}

template <typename T>
hipError_t launch_inject(hipStream_t s, const double* in, int64_t Ng, int S, Planes<T> dst) {
    if (Ng <= 0) return hipSuccess;
    k_inject<T><<<nblocks(Ng * S, 256), 256, 0, s>>>(in, Ng, S, dst);
    LAUNCH_CHECK();
}

int64_t scan_tmp_size(int64_t n) { return (n + SCAN_TILE - 1) / SCAN_TILE + 1; }

hipError_t launch_scan(hipStream_t s, const int32_t* in, int32_t* out, int64_t n, int32_t* tmp) {
    const int64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (nb > 0) k_scan_tiles<<<(unsigned)nb, SCAN_T, 0, s>>>(in, out, n, tmp);
    k_scan_sums<<<1, SCAN_T, 0, s>>>(tmp, nb, out + n);
    if (nb > 1) k_scan_add<<<nblocks(n, 256), 256, 0, s>>>(out, n, tmp);
    LAUNCH_CHECK();
}

/* explicit instantiations */
template hipError_t launch_materialize<float>(hipStream_t, Planes<float>, Planes<float>, const int32_t*,
                                              const int32_t*, const int32_t*, int64_t, int, LinearCVMotionModel3D,
                                              double, RigKeys, uint64_t, bool);
template hipError_t launch_materialize<double>(hipStream_t, Planes<double>, Planes<double>, const int32_t*,
                                               const int32_t*, const int32_t*, int64_t, int, LinearCVMotionModel3D,
                                               double, RigKeys, uint64_t, bool);
template hipError_t launch_fit<float>(hipStream_t, const FitArgs&, Planes<float>, Planes<float>);
template hipError_t launch_fit<double>(hipStream_t, const FitArgs&, Planes<double>, Planes<double>);
template hipError_t launch_fit_fix<float>(hipStream_t, const FitArgs&, Planes<float>, Planes<float>);
template hipError_t launch_fit_fix<double>(hipStream_t, const FitArgs&, Planes<double>, Planes<double>);
template hipError_t launch_nd_fill<float>(hipStream_t, const FitArgs&, Planes<float>, Planes<float>);
template hipError_t launch_nd_fill<double>(hipStream_t, const FitArgs&, Planes<double>, Planes<double>);
template hipError_t launch_sample<float>(hipStream_t, const SampleArgs&, Planes<float>);
template hipError_t launch_sample<double>(hipStream_t, const SampleArgs&, Planes<double>);
template hipError_t launch_gather<float>(hipStream_t, Planes<float>, Planes<float>, const int32_t*, int64_t, int);
template hipError_t launch_gather<double>(hipStream_t, Planes<double>, Planes<double>, const int32_t*, int64_t,
                                          int);
template hipError_t launch_extract<float>(hipStream_t, Planes<float>, const int32_t*, int64_t, int, double*);
template hipError_t launch_extract<double>(hipStream_t, Planes<double>, const int32_t*, int64_t, int, double*);
template hipError_t launch_inject<float>(hipStream_t, const double*, int64_t, int, Planes<float>);
template hipError_t launch_inject<double>(hipStream_t, const double*, int64_t, int, Planes<double>);

}  // namespace scphd
