/*
 * device.h -- numerics shared by the gfx950 kernels of the SC-PHD filter:
 * counter-based noise, safe logarithm, packed 6x6 symmetric algebra.
 *
 * Covariances are stored packed, lower triangle, row by row:
 *   idx(i, j) = i (i + 1) / 2 + j   for i >= j   (21 values).
 */
#ifndef SCPHD_DEVICE_H
#define SCPHD_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <float.h>

#include "scphd_models.h"

#define SCPHD_LOG0 (-(double)FLT_MAX) /* log(0) value, src/device_math.cu:8 */

namespace scphd {

/* ----------------------------------------------------------------- noise */
/* x ^ y ^ k.  A key word k that is the same in every lane (LANE false) is
 * read from a scalar register: one three-input bitwise instruction, truth
 * table 0x96, on the device.  A key word that may differ between lanes
 * (LANE true: the rigs of an ensemble) stays a plain expression; forcing it
 * into a scalar register would make the compiler loop over its distinct
 * values. */
template <bool LANE>
SCPHD_HD inline uint32_t xor3_key(uint32_t x, uint32_t y, uint32_t k) {
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (!LANE) {
        uint32_t r;
        asm("v_bitop3_b32 %0, %1, %2, %3 bitop3:0x96" : "=v"(r) : "v"(x), "v"(y), "s"(k));
        return r;
    } else {
        return x ^ y ^ k;
    }
#else
    return x ^ y ^ k;
#endif
}

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
 * easy as 1, 2, 3", SC'11): 10 rounds of two 32x32->64 multiplies, key
 * bumped by the Weyl constants between rounds. */
template <bool LANE = false>
SCPHD_HD inline void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                   uint32_t k1, uint32_t out[4]) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t a = (uint64_t)0xD2511F53u * c0;
        const uint64_t b = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = xor3_key<LANE>((uint32_t)(b >> 32), c1, k0);
        const uint32_t n2 = xor3_key<LANE>((uint32_t)(a >> 32), c3, k1);
        c1 = (uint32_t)b;
        c3 = (uint32_t)a;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

/* Noise streams.  Counter words: c0 = element << 2 | block,
 * c1 = Gaussian index, c2 = parent, c3 = step << 4 | stream; key = seed. */
enum NoiseStream : uint32_t {
    NS_PRIOR = 0,
    NS_CAMERA = 1,
    NS_FEATURE = 2,
    NS_POSTERIOR = 3,
    NS_RESAMPLE = 4,
};

struct NoiseKey {
    uint32_t k0, k1;
    SCPHD_HD static NoiseKey from_seed(uint64_t s) {
        NoiseKey k;
        k.k0 = (uint32_t)s;
        k.k1 = (uint32_t)(s >> 32);
        return k;
    }
};

/* a key that may differ between the lanes of a wavefront (ensembles: each
 * rig has its own), with the count sb of the handle's steps that preceded
 * the rig's joining: the rig draws step s of a stream with counter s - sb,
 * the counter a filter of its own would use.  sb = 0 gives the same streams
 * as the NoiseKey of the same words. */
struct LaneKey {
    uint32_t k0, k1;
    uint32_t sb = 0;
};

SCPHD_HD inline void noise_block(NoiseKey key, uint32_t stream, uint64_t step, uint32_t parent,
                                 uint32_t gauss, uint32_t elem, uint32_t block, uint32_t out[4]) {
    philox4x32_10<false>((elem << 2) | block, gauss, parent, (uint32_t)((step << 4) | stream), key.k0, key.k1, out);
}
SCPHD_HD inline void noise_block(LaneKey key, uint32_t stream, uint64_t step, uint32_t parent,
                                 uint32_t gauss, uint32_t elem, uint32_t block, uint32_t out[4]) {
    philox4x32_10<true>((elem << 2) | block, gauss, parent, (((uint32_t)step - key.sb) << 4) | stream, key.k0,
                        key.k1, out);
}

/* 53-bit uniforms from two 32-bit words. */
SCPHD_HD inline double uniform_co(uint32_t lo, uint32_t hi) { /* [0, 1) */
    return (double)((((uint64_t)hi << 32) | lo) >> 11) * 0x1.0p-53;
}
SCPHD_HD inline double uniform_oc(uint32_t lo, uint32_t hi) { /* (0, 1] */
    return (double)(((((uint64_t)hi << 32) | lo) >> 11) + 1) * 0x1.0p-53;
}

/* One Box-Muller pair from two Philox words (Box and Muller 1958),
 * evaluated with single-precision adds, multiplies, fused multiply-adds and
 * the correctly rounded square root only, so that any IEEE-754
 * implementation returns the same bits:
 *   a = (x0 + 1) 2^-32 in (0, 1], its leading 24 bits kept:
 *       a = f 2^e with f in [sqrt(1/2), sqrt(2)), log a = e ln2 + s P(s),
 *       s = f - 1, P the degree-8 Chebyshev fit of log(1 + s)/s on that
 *       range (|error| < 2e-8);
 *   c = (x1 >> 8) 2^-24 in [0, 1): the octant of 2 pi c and the offset t in
 *       [0, 1] inside it come from the integer bits, sin(pi/4 t) = t S(t^2)
 *       and cos(pi/4 t) = C(t^2) are Taylor polynomials to t^9 and t^8
 *       (|error| < 4e-8);
 *   r = sqrt(-2 log a); z0 = r cos 2 pi c, z1 = r sin 2 pi c (single
 *   precision; box_muller widens them to double). */
/* The correctly rounded single-precision square root of x = +-0 or x in
 * [2^-96, 2^128): the hardware root (within 1 ulp for such x) moved by one
 * ulp down or up when the residual x - s'^2 of that neighbour shows it is
 * the nearer root, the same correction as the compiler's IEEE sqrt, without
 * its scaling of tiny inputs and its special cases (+-0 pass through the
 * correction unchanged: the neighbour below is a NaN, the one above gives a
 * zero residual). */
__device__ inline float sqrt_rn_normal(float x) {
    const float s = __builtin_amdgcn_sqrtf(x);
    const float dn = __uint_as_float(__float_as_uint(s) - 1u);
    const float up = __uint_as_float(__float_as_uint(s) + 1u);
    float r = __builtin_fmaf(-dn, s, x) <= 0.0f ? dn : s;
    r = __builtin_fmaf(-up, s, x) > 0.0f ? up : r;
    return r;
}

__device__ inline void box_muller_f(uint32_t x0, uint32_t x1, float& z0, float& z1) {
#pragma clang fp contract(off)
    const uint32_t n = x0 + 1u; /* 0 stands for a = 1 (log a = 0) */
    const uint32_t nn = n != 0u ? n : 1u;
    const int lz = __clz((int)nn);
    const uint32_t m = nn << lz; /* [2^31, 2^32) */
    float f = __uint_as_float(0x3F800000u | ((m >> 8) & 0x7FFFFFu)); /* (m >> 8) 2^-23, in [1, 2) */
    int e = -1 - lz;
    if (f > 0x1.6a09e6p+0f) {
        f *= 0.5f;
        e += 1;
    }
    const float sx = f - 1.0f;
    float pl = 0x1.6626eap-4f;
    pl = __builtin_fmaf(pl, sx, -0x1.26729ep-3f);
    pl = __builtin_fmaf(pl, sx, 0x1.322850p-3f);
    pl = __builtin_fmaf(pl, sx, -0x1.5329bep-3f);
    pl = __builtin_fmaf(pl, sx, 0x1.98b80ap-3f);
    pl = __builtin_fmaf(pl, sx, -0x1.0005a6p-2f);
    pl = __builtin_fmaf(pl, sx, 0x1.555790p-2f);
    pl = __builtin_fmaf(pl, sx, -0x1.fffff8p-2f);
    pl = __builtin_fmaf(pl, sx, 1.0f);
    const float lf = sx * pl;
    const float L = n != 0u ? __builtin_fmaf((float)e, 0x1.62e430p-1f, lf) : 0.0f;
    /* -2 L is -0 (a = 1) or at least 2^-23 (a <= 1 - 2^-24): correctly
     * rounded without the tiny-input path */
    const float r = sqrt_rn_normal(-2.0f * L);
    const uint32_t c = x1 >> 8;
    const uint32_t oct = c >> 21;
    const uint32_t fr = c & 0x1FFFFFu;
    const float t = (float)((oct & 1u) ? (0x200000u - fr) : fr) * 0x1.0p-21f;
    const float t2 = t * t;
    float ps = 0x1.507834p-22f;
    ps = __builtin_fmaf(ps, t2, -0x1.32d2ccp-15f);
    ps = __builtin_fmaf(ps, t2, 0x1.466bc6p-9f);
    ps = __builtin_fmaf(ps, t2, -0x1.4abbcep-4f);
    ps = __builtin_fmaf(ps, t2, 0x1.921fb6p-1f);
    const float sv = t * ps;
    float pc = 0x1.e1f506p-19f;
    pc = __builtin_fmaf(pc, t2, -0x1.55d3c8p-12f);
    pc = __builtin_fmaf(pc, t2, 0x1.03c1f0p-6f);
    pc = __builtin_fmaf(pc, t2, -0x1.3bd3ccp-2f);
    pc = __builtin_fmaf(pc, t2, 1.0f);
    /* quadrant angle phi = 2 pi c - q pi/2: odd octants measure t from pi/2
     * (sin and cos exchanged), odd quadrants exchange them again, so one
     * exchange when exactly one applies; the quadrant's signs are sign-bit
     * flips: -cos for q = 1, 2 and -sin for q = 2, 3 */
    const uint32_t q = oct >> 1;
    const bool swap = ((oct ^ q) & 1u) != 0u;
    const float sn0 = swap ? pc : sv, cs0 = swap ? sv : pc;
    const float sn = __uint_as_float(__float_as_uint(sn0) ^ ((q & 2u) << 30));
    const float cs = __uint_as_float(__float_as_uint(cs0) ^ (((q + 1u) & 2u) << 30));
    z0 = r * cs;
    z1 = r * sn;
}

__device__ inline void box_muller(uint32_t x0, uint32_t x1, double& z0, double& z1) {
    float a, b;
    box_muller_f(x0, x1, a, b);
    z0 = (double)a;
    z1 = (double)b;
}

/* N (<= 8) standard normals, four per Philox block: words (x0, x1) give
 * (z0, z1), words (x2, x3) give (z2, z3) (box_muller). */
template <int N, typename Key>
__device__ inline void normals(Key key, uint32_t stream, uint64_t step, uint32_t parent, uint32_t gauss,
                               uint32_t elem, double z[N]) {
#pragma unroll
    for (int b = 0; 4 * b < N; ++b) {
        uint32_t x[4];
        noise_block(key, stream, step, parent, gauss, elem, (uint32_t)b, x);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int i = 4 * b + 2 * h;
            if (i < N) {
                double z0, z1;
                box_muller(x[2 * h], x[2 * h + 1], z0, z1);
                z[i] = z0;
                if (i + 1 < N) z[i + 1] = z1;
            }
        }
    }
}

/* Landmark-prediction normals: three per sample, four samples to three
 * Philox blocks.  Sample s is member i = (s >> 4) & 3 of the group whose
 * element is feature_elem(s) = (s >> 6) << 4 | (s & 15): the four samples
 * that one lane of a 16-lane row visits in four consecutive steps of 16.
 * The group's blocks b = 0, 1, 2 give twelve normals, n[4 b + 2 h] and
 * n[4 b + 2 h + 1] from box_muller of block b's words 2 h and 2 h + 1, and
 * member i takes n[3 i], n[3 i + 1], n[3 i + 2]. */
SCPHD_HD constexpr uint32_t feature_elem(uint32_t s) { return ((s >> 6) << 4) | (s & 15u); }

/* any one sample (one or two blocks) */
template <typename Key, typename A>
__device__ inline void feature_normals(Key key, uint64_t step, uint32_t p, uint32_t k, uint32_t s, A a[3]) {
    const uint32_t e = feature_elem(s), i = (s >> 4) & 3u;
    uint32_t x[4];
    float n0, n1, n2, n3;
    noise_block(key, NS_FEATURE, step, p, k, e, i == 0u ? 0u : i - 1u, x); /* the block of n[3 i] */
    box_muller_f(x[0], x[1], n0, n1);
    box_muller_f(x[2], x[3], n2, n3);
    if (i == 0u) {
        a[0] = n0; a[1] = n1; a[2] = n2;
    } else if (i == 3u) {
        a[0] = n1; a[1] = n2; a[2] = n3;
    } else {
        uint32_t y[4];
        float m0, m1;
        noise_block(key, NS_FEATURE, step, p, k, e, i, y);
        box_muller_f(y[0], y[1], m0, m1);
        if (i == 1u) {
            a[0] = n3; a[1] = m0; a[2] = m1;
        } else {
            a[0] = n2; a[1] = n3; a[2] = m0;
        }
    }
}

/* A group's members in order (i = 0, 1, 2, 3, the same in every lane that
 * calls): one block for each of the first three, the normals the next
 * member needs carried in c. */
struct FeatureCarry {
    float c0 = 0.0f, c1 = 0.0f, c2 = 0.0f;
};

template <typename Key, typename A>
__device__ inline void feature_normals_seq(Key key, uint64_t step, uint32_t p, uint32_t k, uint32_t s, uint32_t i,
                                           FeatureCarry& c, A a[3]) {
    float n0 = 0.0f, n1 = 0.0f, n2 = 0.0f, n3 = 0.0f;
    if (i != 3u) {
        uint32_t x[4];
        noise_block(key, NS_FEATURE, step, p, k, feature_elem(s), i, x);
        box_muller_f(x[0], x[1], n0, n1);
        box_muller_f(x[2], x[3], n2, n3);
    }
    if (i == 0u) {
        a[0] = n0; a[1] = n1; a[2] = n2;
        c.c0 = n3;
    } else if (i == 1u) {
        a[0] = c.c0; a[1] = n0; a[2] = n1;
        c.c0 = n2; c.c1 = n3;
    } else if (i == 2u) {
        a[0] = c.c0; a[1] = c.c1; a[2] = n0;
        c.c0 = n1; c.c1 = n2; c.c2 = n3;
    } else {
        a[0] = c.c0; a[1] = c.c1; a[2] = c.c2;
    }
}

/* Posterior-sampling normals: six per sample, two samples to three Philox
 * blocks.  Sample s is member i = (s >> 4) & 1 of the group whose element
 * is posterior_elem(s) = (s >> 5) << 4 | (s & 15) (the two samples one lane
 * visits in consecutive steps of 16); the group's twelve normals are laid
 * out as for the landmark prediction, and member i takes n[6 i .. 6 i + 5]. */
SCPHD_HD constexpr uint32_t posterior_elem(uint32_t s) { return ((s >> 5) << 4) | (s & 15u); }

template <typename Key>
__device__ inline void posterior_normals(Key key, uint64_t step, uint32_t p, uint32_t k, uint32_t s, double z[6]) {
    const uint32_t e = posterior_elem(s), i = (s >> 4) & 1u;
    uint32_t x[4], y[4];
    float n[6];
    noise_block(key, NS_POSTERIOR, step, p, k, e, i == 0u ? 0u : 1u, x);
    noise_block(key, NS_POSTERIOR, step, p, k, e, i == 0u ? 1u : 2u, y);
    if (i == 0u) { /* n0..n3 from block 0, n4 n5 from block 1's first pair */
        box_muller_f(x[0], x[1], n[0], n[1]);
        box_muller_f(x[2], x[3], n[2], n[3]);
        box_muller_f(y[0], y[1], n[4], n[5]);
    } else { /* n6 n7 from block 1's second pair, n8..n11 from block 2 */
        box_muller_f(x[2], x[3], n[0], n[1]);
        box_muller_f(y[0], y[1], n[2], n[3]);
        box_muller_f(y[2], y[3], n[4], n[5]);
    }
#pragma unroll
    for (int j = 0; j < 6; ++j) z[j] = (double)n[j];
}

/* The pair in order (i = 0 then 1, the same in every lane that calls): two
 * blocks for the first, one for the second, block 1's second pair carried. */
struct PosteriorCarry {
    float c0 = 0.0f, c1 = 0.0f;
};

template <typename Key, typename A>
__device__ inline void posterior_normals_seq(Key key, uint64_t step, uint32_t p, uint32_t k, uint32_t s, uint32_t i,
                                             PosteriorCarry& c, A z[6]) {
    const uint32_t e = posterior_elem(s);
    float n[6];
    uint32_t x[4];
    if (i == 0u) {
        noise_block(key, NS_POSTERIOR, step, p, k, e, 0u, x);
        box_muller_f(x[0], x[1], n[0], n[1]);
        box_muller_f(x[2], x[3], n[2], n[3]);
        noise_block(key, NS_POSTERIOR, step, p, k, e, 1u, x);
        box_muller_f(x[0], x[1], n[4], n[5]);
        box_muller_f(x[2], x[3], c.c0, c.c1);
    } else {
        n[0] = c.c0;
        n[1] = c.c1;
        noise_block(key, NS_POSTERIOR, step, p, k, e, 2u, x);
        box_muller_f(x[0], x[1], n[2], n[3]);
        box_muller_f(x[2], x[3], n[4], n[5]);
    }
#pragma unroll
    for (int j = 0; j < 6; ++j) z[j] = (A)n[j];
}

SCPHD_HD inline double safe_log(double x) { return x <= 0 ? SCPHD_LOG0 : log(x); }

SCPHD_HD constexpr int pidx(int i, int j) { return i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i; }

/* Lower Cholesky factor (packed) of a packed symmetric 6x6, row by row
 * (Cholesky-Banachiewicz), no PD check: a non-PD input leaves NaN. */
SCPHD_HD inline void chol6(const double* A, double* L) {
#pragma unroll
    for (int i = 0; i < 6; ++i) {
#pragma unroll
        for (int j = 0; j <= i; ++j) {
            double t = A[pidx(i, j)];
            if (i == j) {
#pragma unroll
                for (int k = 0; k < j; ++k) t -= L[pidx(j, k)] * L[pidx(j, k)];
                L[pidx(i, j)] = sqrt(t);
            } else {
#pragma unroll
                for (int k = 0; k < j; ++k) t -= L[pidx(i, k)] * L[pidx(j, k)];
                L[pidx(i, j)] = t / L[pidx(j, j)];
            }
        }
    }
}

/* Positive-definiteness test (isPosDef, src/device_math.cu:800-812, which
 * checks the Cholesky factor for NaN).  Here every Cholesky pivot (the value
 * under the square root) must exceed PD_REL times its diagonal entry: the
 * same answer as the NaN test for every matrix whose smallest relative pivot
 * is above 1e-9, and a deterministic "not PD" for numerically singular ones
 * (rank-deficient refits of clouds with a few in-range samples sit at 1e-13
 * and below, where the NaN test is decided by rounding). */
static constexpr double PD_REL = 1e-9;

SCPHD_HD inline bool is_pd6(const double* A) {
    double L[21];
    bool ok = true;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
#pragma unroll
        for (int j = 0; j <= i; ++j) {
            double t = A[pidx(i, j)];
            if (i == j) {
#pragma unroll
                for (int k = 0; k < j; ++k) t -= L[pidx(j, k)] * L[pidx(j, k)];
                ok = ok && (t > PD_REL * A[pidx(i, i)]);
                L[pidx(i, j)] = sqrt(t);
            } else {
#pragma unroll
                for (int k = 0; k < j; ++k) t -= L[pidx(i, k)] * L[pidx(j, k)];
                L[pidx(i, j)] = t / L[pidx(j, j)];
            }
        }
    }
    return ok;
}

/* Inverse of a packed lower-triangular 6x6 by forward substitution. */
SCPHD_HD inline void trinv6(const double* L, double* Li) {
#pragma unroll
    for (int c = 0; c < 6; ++c) {
#pragma unroll
        for (int i = c; i < 6; ++i) {
            double v = (i == c) ? 1.0 : 0.0;
#pragma unroll
            for (int j = c; j < i; ++j) v -= L[pidx(i, j)] * Li[pidx(j, c)];
            Li[pidx(i, c)] = v / L[pidx(i, i)];
        }
    }
}

/* Squared Mahalanobis distance with Sigma = Pa + Pb (packed covariances). */
SCPHD_HD inline double mahal6(const double* ma, const double* Pa, const double* mb, const double* Pb) {
    double d[6], S[21], L[21], Li[21];
#pragma unroll
    for (int i = 0; i < 6; ++i) d[i] = ma[i] - mb[i];
#pragma unroll
    for (int i = 0; i < 21; ++i) S[i] = Pa[i] + Pb[i];
    chol6(S, L);
    trinv6(L, Li);
    double dist = 0.0;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j <= i; ++j) s += d[j] * Li[pidx(i, j)];
        dist += s * s;
    }
    return dist;
}

/* 1/x within about one ulp: the hardware reciprocal estimate refined by two
 * Newton steps (finite, normal x). */
__device__ inline double rcp_nr(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-x, r, 1.0);
    return __builtin_fma(r, e, r);
}

/* DisparityMeasurementModel::computeMeasurement of a world point seen by
 * camera c, with the three quotients fx/z, fy/z, 1/z^2 taken from one
 * reciprocal 1/z (within a few ulps of the divisions). */
__device__ inline DisparityPoint disparity_fast(const DisparityMeasurementModel& m, const CameraPose& c,
                                                const EuclideanPoint& pw) {
    const EuclideanPoint p = c.worldToCamera(pw);
    const double iz = rcp_nr(p.z);
    const double fz = m.fx() * iz, gz = m.fy() * iz, iz2 = iz * iz;
    DisparityPoint q;
    q.u = m.u0() - fz * p.x;
    q.v = m.v0() - gz * p.y;
    q.d = -fz;
    q.vu = -m.fx() * (p.vx * p.z - p.vz * p.x) * iz2;
    q.vv = -m.fy() * (p.vy * p.z - p.vz * p.y) * iz2;
    q.vd = m.fx() * p.vz * iz2;
    return q;
}

/* DisparityMeasurementModel::invertMeasurement into the world frame of
 * camera c, with 1/d and 1/d^2 taken from one reciprocal. */
__device__ inline EuclideanPoint world_from_disparity_fast(const DisparityMeasurementModel& m, const CameraPose& c,
                                                           const DisparityPoint& q) {
    const double id = rcp_nr(q.d), id2 = id * id, r = m.fx() / m.fy();
    const double du = q.u - m.u0(), dv = q.v - m.v0();
    EuclideanPoint e;
    e.x = du * id;
    e.y = r * dv * id;
    e.z = -m.fx() * id;
    e.vx = (q.vu * q.d - q.vd * du) * id2;
    e.vy = r * (q.vv * q.d - q.vd * dv) * id2;
    e.vz = m.fx() * q.vd * id2;
    return c.cameraToWorld(e);
}

/* Measurement-independent part of the 2-D Kalman detection update with
 * H = [I2 0], R = diag(var_u, var_v) (computeDetectionsKernel,
 * src/kernels.cu:492-595): symmetrised innovation covariance and its
 * inverse, gain K = P H^T S^-1 (6x2, column 0 then column 1) and the
 * Joseph-form covariance (I-KH) P (I-KH)^T + K R K^T. */
struct DetectionTerms {
    double s0, s1, s2, s3; /* S^-1, column-major */
    double det;
    double K[12];
    double Pn[21];
};

SCPHD_HD inline void detection_terms(const double* P, double var_u, double var_v, DetectionTerms& r) {
    double g0 = P[pidx(0, 0)] + var_u;
    double g1 = P[pidx(1, 0)];
    double g2 = P[pidx(0, 1)];
    double g3 = P[pidx(1, 1)] + var_v;
    g1 = (g1 + g2) / 2;
    g2 = g1;
    const double det = g0 * g3 - g1 * g2;
    r.det = det;
    r.s0 = g3 / det;
    r.s1 = -g1 / det;
    r.s2 = -g2 / det;
    r.s3 = g0 / det;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        r.K[i] = P[pidx(i, 0)] * r.s0 + P[pidx(i, 1)] * r.s1;
        r.K[6 + i] = P[pidx(i, 0)] * r.s2 + P[pidx(i, 1)] * r.s3;
    }
    /* A = I - K H: columns 0, 1 are e_c - K(:, c); columns 2..5 are e_c. */
    double AP[36]; /* row-major A P */
#pragma unroll
    for (int i = 0; i < 6; ++i) {
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                const double a = ((i == k) ? 1.0 : 0.0) - (k == 0 ? r.K[i] : (k == 1 ? r.K[6 + i] : 0.0));
                s += a * P[pidx(k, j)];
            }
            AP[i * 6 + j] = s;
        }
    }
#pragma unroll
    for (int i = 0; i < 6; ++i) {
#pragma unroll
        for (int j = 0; j <= i; ++j) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                const double a = ((j == k) ? 1.0 : 0.0) - (k == 0 ? r.K[j] : (k == 1 ? r.K[6 + j] : 0.0));
                s += AP[i * 6 + k] * a;
            }
            s += r.K[i] * var_u * r.K[j] + r.K[6 + i] * var_v * r.K[6 + j];
            r.Pn[pidx(i, j)] = s;
        }
    }
}

/* ------------------------------------------ single-precision sample work */
/* The FP32-arithmetic variant evaluates the per-sample work -- landmark
 * prediction, world -> disparity, the field-of-view test, posterior draws
 * and disparity -> world -- in single precision.  Every multiply-add below
 * is an explicit fused one and contraction is off, so each kernel that
 * evaluates a sample gets the same bits.  Reciprocals are the hardware's
 * single-precision estimate (v_rcp_f32, 1 ulp). */
struct PoseF {
    float R[9], t[3];
};
struct MeasF {
    float fx, fy, u0, v0, w, h, rxy; /* rxy = fx / fy */
};
/* constant-velocity prediction: dt, dt^2 / 2 and the acceleration stds */
struct MotionF {
    float dt, h, sx, sy, sz;
};

__device__ inline PoseF pose_f(const CameraPose& c) {
    PoseF f;
#pragma unroll
    for (int i = 0; i < 9; ++i) f.R[i] = (float)c.R[i];
#pragma unroll
    for (int i = 0; i < 3; ++i) f.t[i] = (float)c.t[i];
    return f;
}
__device__ inline MeasF meas_f(const DisparityMeasurementModel& m) {
    return MeasF{(float)m.fx(), (float)m.fy(), (float)m.u0(), (float)m.v0(), (float)m.img_width(),
                 (float)m.img_height(), (float)(m.fx() / m.fy())};
}
__device__ inline MotionF motion_f(const LinearCVMotionModel3D& m, double dt) {
    return MotionF{(float)dt, (float)(0.5 * dt * dt), (float)m.std_ax(), (float)m.std_ay(), (float)m.std_az()};
}

/* computeNoisyMotion with acceleration draws n * std:
 * p + v dt + a dt^2 / 2, v + a dt */
__device__ inline void predict_f(float x[6], const MotionF& m, const float n[3]) {
#pragma clang fp contract(off)
    const float ax = n[0] * m.sx, ay = n[1] * m.sy, az = n[2] * m.sz;
    x[0] = __builtin_fmaf(ax, m.h, __builtin_fmaf(x[3], m.dt, x[0]));
    x[1] = __builtin_fmaf(ay, m.h, __builtin_fmaf(x[4], m.dt, x[1]));
    x[2] = __builtin_fmaf(az, m.h, __builtin_fmaf(x[5], m.dt, x[2]));
    x[3] = __builtin_fmaf(ax, m.dt, x[3]);
    x[4] = __builtin_fmaf(ay, m.dt, x[4]);
    x[5] = __builtin_fmaf(az, m.dt, x[5]);
}

/* DisparityMeasurementModel::computeMeasurement of world point x seen by
 * camera c: p = R^T (x - t), u = u0 - fx p_x / p_z, v = v0 - fy p_y / p_z,
 * d = -fx / p_z and the velocities' images */
__device__ inline void disparity_f(const MeasF& m, const PoseF& c, const float x[6], float q[6]) {
#pragma clang fp contract(off)
    const float dx = x[0] - c.t[0], dy = x[1] - c.t[1], dz = x[2] - c.t[2];
    const float px = __builtin_fmaf(c.R[0], dx, __builtin_fmaf(c.R[3], dy, c.R[6] * dz));
    const float py = __builtin_fmaf(c.R[1], dx, __builtin_fmaf(c.R[4], dy, c.R[7] * dz));
    const float pz = __builtin_fmaf(c.R[2], dx, __builtin_fmaf(c.R[5], dy, c.R[8] * dz));
    const float vx = __builtin_fmaf(c.R[0], x[3], __builtin_fmaf(c.R[3], x[4], c.R[6] * x[5]));
    const float vy = __builtin_fmaf(c.R[1], x[3], __builtin_fmaf(c.R[4], x[4], c.R[7] * x[5]));
    const float vz = __builtin_fmaf(c.R[2], x[3], __builtin_fmaf(c.R[5], x[4], c.R[8] * x[5]));
    const float iz = __builtin_amdgcn_rcpf(pz);
    const float fz = m.fx * iz, gz = m.fy * iz, iz2 = iz * iz;
    q[0] = __builtin_fmaf(-fz, px, m.u0);
    q[1] = __builtin_fmaf(-gz, py, m.v0);
    q[2] = -fz;
    q[3] = -m.fx * __builtin_fmaf(vx, pz, -(vz * px)) * iz2;
    q[4] = -m.fy * __builtin_fmaf(vy, pz, -(vz * py)) * iz2;
    q[5] = m.fx * vz * iz2;
}

/* DisparityMeasurementModel::inRange */
__device__ inline bool in_range_f(const MeasF& m, const float q[6]) {
    return q[0] >= 0.0f && q[0] <= m.w && q[1] >= 0.0f && q[1] <= m.h && q[2] >= 0.0f;
}

/* DisparityMeasurementModel::invertMeasurement into the world frame of
 * camera c */
__device__ inline void world_from_disparity_f(const MeasF& m, const PoseF& c, const float q[6], float x[6]) {
#pragma clang fp contract(off)
    const float id = __builtin_amdgcn_rcpf(q[2]), id2 = id * id;
    const float du = q[0] - m.u0, dv = q[1] - m.v0;
    const float ex = du * id, ey = m.rxy * dv * id, ez = -m.fx * id;
    const float evx = __builtin_fmaf(q[3], q[2], -(q[5] * du)) * id2;
    const float evy = m.rxy * __builtin_fmaf(q[4], q[2], -(q[5] * dv)) * id2;
    const float evz = m.fx * q[5] * id2;
    x[0] = __builtin_fmaf(c.R[0], ex, __builtin_fmaf(c.R[1], ey, __builtin_fmaf(c.R[2], ez, c.t[0])));
    x[1] = __builtin_fmaf(c.R[3], ex, __builtin_fmaf(c.R[4], ey, __builtin_fmaf(c.R[5], ez, c.t[1])));
    x[2] = __builtin_fmaf(c.R[6], ex, __builtin_fmaf(c.R[7], ey, __builtin_fmaf(c.R[8], ez, c.t[2])));
    x[3] = __builtin_fmaf(c.R[0], evx, __builtin_fmaf(c.R[1], evy, c.R[2] * evz));
    x[4] = __builtin_fmaf(c.R[3], evx, __builtin_fmaf(c.R[4], evy, c.R[5] * evz));
    x[5] = __builtin_fmaf(c.R[6], evx, __builtin_fmaf(c.R[7], evy, c.R[8] * evz));
}

}  // namespace scphd

#endif /* SCPHD_DEVICE_H */
