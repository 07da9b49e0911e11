/*
 * scphd_oracle.cpp -- FP64 CPU restatement of the reference SC-PHD camera
 * calibration step (cheesinglee/scphd-camera-calibration).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The GPU
 * product never links, loads or falls back to it.
 *
 * Every function cites the reference file:line whose arithmetic it restates.
 * Deliberate, documented deviations (the same ones the GPU path follows):
 *   D1  noise: Philox4x32-10 counter streams + Box-Muller (box_muller,
 *       single-precision polynomial log / sin / cos) instead of thrust's
 *       time-seeded default_random_engine (reference :92-127, :1062-1118,
 *       :640-690, :2663-2666, prior :912-979).  The landmark-prediction
 *       normals of four samples share three Philox blocks, the posterior
 *       normals of two samples share three (feature_elem, posterior_elem).
 *   D2  exact per-parent segmentation of normalisers, cardinalities and map
 *       offsets (reference reduce_by_key misaligns when a parent map is empty
 *       or M == 1, :1750-1756, :1824-1837, :2566-2572).
 *   D3  parent weights normalised in the log domain with a max shift
 *       (reference multiplies linear weights, :1852-1875, which underflows).
 *   D4  merge tie-break: weight descending, pack index ascending.
 *   D5  sums that the reference reduces in parallel trees or thrust calls are
 *       sequential in ascending index order.
 *   D6  systematic resampling points r_j = u + j/N (one rounding each)
 *       instead of the running r += 1/N (:2700).
 *   D7  world->camera as R^T (p - t) (reference expands R^T p - R^T t,
 *       src/disparitymeasurementmodel.cu:25-42).
 *   D8  the Kalman covariance update is the Joseph form
 *       (I-KH) P (I-KH)^T + K R K^T evaluated on the lower triangle and
 *       mirrored; the reference writes 36 expanded expressions
 *       (src/kernels.cu:560-595) that equal it to rounding.
 *   D10 the per-(parent, measurement) normaliser is summed in a fixed order
 *       with log1p (see log_normaliser); the reference's reduce_by_key order
 *       is unspecified.
 *   D11 normalised sample weights as w / sum (normalise_weights).
 *   D9  positive-definiteness: relative Cholesky pivot > 1e-9 instead of
 *       "no NaN in the factor" (isPosDef, src/device_math.cu:800-812); see
 *       is_pos_def6.
 *   D13 the fit of a cloud with pd_g = 0 (entirely out of view) is its first
 *       sample with zero covariance; the reference refits it with unit
 *       weights, but its detection terms have weight 0 either way.
 * With storage32 set, the sample cloud (positions, velocities, weights) is
 * rounded to float at the end of every update, as the GPU's FP32 storage
 * does; all arithmetic stays FP64.
 */
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

/* ---------------------------------------------------------------- params */
/* Mirror of scphd_params (include/scphd.h); layout checked by a test. */
struct Params {
    int32_t n_particles, particles_per_feature, max_gaussians, storage;
    int32_t verbosity, slam, max_steps, save_all_maps;
    double min_weight, min_separation, min_neff, ps;
    double prior_a[12], prior_b[12];
    double cam_accel_std[6];
    double feat_accel_std[3];
    double fx, fy, u0, v0, std_u, std_v, pd, lambda;
    double birth_w0, birth_d0, birth_vu0, birth_vv0, birth_vd0;
    double birth_var_d0, birth_var_vu0, birth_var_vv0, birth_var_vd0;
    double max_velocity, min_range;
    uint64_t seed;
};

/* ------------------------------------------------------------ basic math */
static const double LOG0 = -FLT_MAX;  /* src/device_math.cu:8 */

/* src/device_math.cu:22-29 */
static inline double safe_log(double x) { return x <= 0 ? LOG0 : std::log(x); }

/* src/device_math.cu:308-317 */
static inline double wrap_angle(double a) {
    double r = std::fmod(a, 2.0 * M_PI);
    if (r > M_PI)
        r -= 2.0 * M_PI;
    else if (r < -M_PI)
        r += 2.0 * M_PI;
    return r;
}

/* Lower Cholesky factor of a 6x6 column-major matrix (element (i,j) at
 * i + 6 j), reading the lower triangle only; no PD check.
 * src/device_math.cu:123-151 */
static void cholesky6(const double* A, double* L) {
    for (int i = 0; i < 36; ++i) L[i] = 0.0;
    for (int i = 0; i < 6; ++i) {
        for (int j = 0; j <= i; ++j) {
            double t = A[i + 6 * j];
            if (i == j) {
                for (int k = 0; k < j; ++k) t -= L[j + 6 * k] * L[j + 6 * k];
                L[i + 6 * j] = std::sqrt(t);
            } else {
                for (int k = 0; k < j; ++k) t -= L[i + 6 * k] * L[j + 6 * k];
                L[i + 6 * j] = t / L[j + 6 * j];
            }
        }
    }
}

/* isPosDef, src/device_math.cu:800-812, calls a matrix PD when its Cholesky
 * factor holds no NaN.  D9: PD here means every Cholesky pivot (the value
 * under the square root) exceeds PD_REL times its diagonal entry.  The two
 * tests agree on every matrix whose smallest relative pivot is above 1e-9;
 * below that the matrix is numerically singular (rank-deficient fits of
 * clouds with a handful of in-range samples land at 1e-13 and below) and
 * the NaN test's outcome is decided by rounding. */
static const double PD_REL = 1e-9;

static bool is_pos_def6(const double* A) {
    double L[36];
    for (int i = 0; i < 36; ++i) L[i] = 0.0;
    for (int i = 0; i < 6; ++i) {
        for (int j = 0; j <= i; ++j) {
            double t = A[i + 6 * j];
            if (i == j) {
                for (int k = 0; k < j; ++k) t -= L[j + 6 * k] * L[j + 6 * k];
                if (!(t > PD_REL * A[i + 6 * i])) return false;
                L[i + 6 * j] = std::sqrt(t);
            } else {
                for (int k = 0; k < j; ++k) t -= L[i + 6 * k] * L[j + 6 * k];
                L[i + 6 * j] = t / L[j + 6 * j];
            }
        }
    }
    return true;
}

/* Inverse of a lower-triangular 6x6 by forward substitution, column by
 * column.  src/device_math.cu:153-175 */
static void tri_inverse6(const double* L, double* Li) {
    for (int c = 0; c < 6; ++c) {
        for (int i = 0; i < 6; ++i) {
            if (i >= c) {
                double v = (i == c) ? 1.0 : 0.0;
                for (int j = 0; j < i; ++j) v -= L[i + 6 * j] * Li[j + 6 * c];
                Li[i + 6 * c] = v / L[i + 6 * i];
            } else {
                Li[i + 6 * c] = 0.0;
            }
        }
    }
}

/* Squared Mahalanobis distance of two Gaussians with Sigma = Pa + Pb.
 * computeMahalDist<6,36>, src/device_math.cu:374-453 */
static double mahal6(const double* ma, const double* Pa, const double* mb, const double* Pb) {
    double d[6], S[36], L[36], Li[36];
    for (int i = 0; i < 6; ++i) d[i] = ma[i] - mb[i];
    for (int i = 0; i < 36; ++i) S[i] = Pa[i] + Pb[i];
    cholesky6(S, L);
    tri_inverse6(L, Li);
    double dist = 0.0;
    for (int i = 0; i < 6; ++i) {
        double s = 0.0;
        for (int j = 0; j <= i; ++j) s += d[j] * Li[i + 6 * j];
        dist += s * s;
    }
    return dist;
}

/* force_symmetric_covariance, src/device_math.cu:722-737 */
static void force_symmetric6(double* P) {
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < i; ++j) {
            double a = (P[i + 6 * j] + P[j + 6 * i]) / 2;
            P[i + 6 * j] = a;
            P[j + 6 * i] = a;
        }
}

/* log-sum-exp functor, src/scphdcameracalibration.cu:506-517 */
static inline double lse2(double a, double b) {
    if (a > b) return a + safe_log(1 + std::exp(b - a));
    return b + safe_log(1 + std::exp(a - b));
}

/* D10: normaliser eta = log(exp(bpc) + sum_g exp(l_g)), which the reference
 * folds with thrust::reduce_by_key and the pairwise log_sum_exp functor in
 * an unspecified association order (:1750-1762).  Defined here as
 *   mx = first maximum of (bpc, l_0, ..., l_{G-1});
 *   T  = sum over the other terms, g ascending then bpc, of exp(term - mx);
 *   eta = mx + log1p(T).
 * log1p keeps T's relative precision when the Gaussian terms are ~1e-16 of
 * the clutter term, so equal inputs give equal normalisers to the last bit
 * on CPU and GPU (birth terms of clutter measurements tie exactly). */
static double log_normaliser(const double* l, int G, double bpc) {
    double mx = bpc;
    int am = -1;
    for (int g = 0; g < G; ++g)
        if (l[g] > mx) {
            mx = l[g];
            am = g;
        }
    double T = 0.0;
    for (int g = 0; g < G; ++g)
        if (g != am) T += std::exp(l[g] - mx);
    if (am >= 0) T += std::exp(bpc - mx);
    return mx + std::log1p(T);
}

/* normalizeWeightsKernel, src/kernels.cu:996-1018: w_i <- exp(log w_i -
 * log sum w), with log 0 = -FLT_MAX so that an all-zero cloud gets weight 1
 * everywhere.  D11: evaluated as w_i / sum (equal to the reference's form
 * to rounding), 0 for w_i = 0, and 1 everywhere when the sum is 0. */
static void normalise_weights(double* w, int S) {
    double sum = 0.0;
    for (int s = 0; s < S; ++s) sum += w[s];
    for (int s = 0; s < S; ++s) w[s] = sum > 0 ? (w[s] > 0 ? w[s] / sum : 0.0) : 1.0;
}

/* ------------------------------------------------------------- geometry */
/* Rotation camera->world R = Rx(theta) Ry(phi) Rz(psi), row-major, from the
 * coefficients of transformCameraToWorld, src/disparitymeasurementmodel.cu:82-144 */
static void rotation(const double* e, double R[3][3]) {
    const double ct = std::cos(e[6]), st = std::sin(e[6]);
    const double cp = std::cos(e[7]), sp = std::sin(e[7]);
    const double cs = std::cos(e[8]), ss = std::sin(e[8]);
    R[0][0] = cp * cs;                R[0][1] = -cp * ss;               R[0][2] = sp;
    R[1][0] = ct * ss + st * sp * cs; R[1][1] = ct * cs - st * sp * ss; R[1][2] = -st * cp;
    R[2][0] = st * ss - ct * sp * cs; R[2][1] = st * cs + ct * sp * ss; R[2][2] = ct * cp;
}

/* transformWorldToCamera, src/disparitymeasurementmodel.cu:3-80 (D7) */
static void world_to_camera(const double R[3][3], const double* t, const double* pw, double* pc) {
    const double dx = pw[0] - t[0], dy = pw[1] - t[1], dz = pw[2] - t[2];
    for (int i = 0; i < 3; ++i) {
        pc[i] = R[0][i] * dx + R[1][i] * dy + R[2][i] * dz;
        pc[3 + i] = R[0][i] * pw[3] + R[1][i] * pw[4] + R[2][i] * pw[5];
    }
}

/* transformCameraToWorld, src/disparitymeasurementmodel.cu:82-144 */
static void camera_to_world(const double R[3][3], const double* t, const double* pc, double* pw) {
    for (int i = 0; i < 3; ++i) {
        pw[i] = R[i][0] * pc[0] + R[i][1] * pc[1] + R[i][2] * pc[2] + t[i];
        pw[3 + i] = R[i][0] * pc[3] + R[i][1] * pc[4] + R[i][2] * pc[5];
    }
}

/* DisparityMeasurementModel::computeMeasurement(EuclideanPoint),
 * src/disparitymeasurementmodel.cu:165-196 */
static void project(const Params& P, const double* pc, double* q) {
    const double x = pc[0], y = pc[1], z = pc[2], vx = pc[3], vy = pc[4], vz = pc[5];
    q[0] = P.u0 - P.fx / z * x;
    q[1] = P.v0 - P.fy / z * y;
    q[2] = -P.fx / z;
    q[3] = -P.fx * (vx * z - vz * x) / (z * z);
    q[4] = -P.fy * (vy * z - vz * y) / (z * z);
    q[5] = P.fx * vz / (z * z);
}

/* DisparityMeasurementModel::invertMeasurement (camera part),
 * src/disparitymeasurementmodel.cu:198-217 */
static void unproject(const Params& P, const double* q, double* pc) {
    const double u = q[0], v = q[1], d = q[2], vu = q[3], vv = q[4], vd = q[5];
    pc[0] = (u - P.u0) / d;
    pc[1] = (P.fx / P.fy) * (v - P.v0) / d;
    pc[2] = -P.fx / d;
    pc[3] = (vu * d - vd * (u - P.u0)) / (d * d);
    pc[4] = (P.fx / P.fy) * (vv * d - vd * (v - P.v0)) / (d * d);
    pc[5] = P.fx * vd / (d * d);
}

/* Field of view test of worldToDisparityKernel, src/kernels.cu:37-45;
 * image = 2 u0 x 2 v0 (src/disparitymeasurementmodel.cu:159-160). */
static inline bool in_range(const Params& P, const double* q) {
    return q[0] >= 0 && q[0] <= 2 * P.u0 && q[1] >= 0 && q[1] <= 2 * P.v0 && q[2] >= 0;
}

/* ------------------------------------------------------------------ RNG */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11). */
static inline void philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

enum Stream { ST_PRIOR = 0, ST_CAMERA = 1, ST_FEATURE = 2, ST_POSTERIOR = 3, ST_RESAMPLE = 4 };

/* Counter layout shared with the GPU library:
 *   c0 = element << 2 | block, c1 = Gaussian index, c2 = parent, c3 = step << 4 | stream
 *   key = seed (low word, high word). */
static inline void draw(uint64_t seed, uint32_t stream, uint64_t step, uint32_t parent,
                        uint32_t gauss, uint32_t elem, uint32_t block, uint32_t out[4]) {
    const uint32_t ctr[4] = {(elem << 2) | block, gauss, parent, (uint32_t)((step << 4) | stream)};
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    philox4x32_10(ctr, key, out);
}

static inline double u01(uint32_t lo, uint32_t hi) {  /* [0, 1) */
    const uint64_t b = (((uint64_t)hi << 32) | lo) >> 11;
    return (double)b * 0x1.0p-53;
}

/* Box-Muller pair (Box and Muller 1958) from two Philox words, in single
 * precision with explicit fused multiply-adds (bit-identical on any
 * IEEE-754 machine):
 *   a = (x0 + 1) 2^-32, truncated to 24 significant bits, written f 2^e with
 *   f in [sqrt(1/2), sqrt(2)); log a = e ln2 + s P(s), s = f - 1, P the
 *   degree-8 Chebyshev fit of log(1 + s)/s;
 *   c = (x1 >> 8) 2^-24: octant o = top 3 bits, offset t in [0, 1] (from
 *   the low 21 bits, mirrored in odd octants); sin(pi/4 t) = t S(t^2),
 *   cos(pi/4 t) = C(t^2), Taylor to t^9 / t^8;
 *   r = sqrt(-2 log a); z0 = r cos 2 pi c, z1 = r sin 2 pi c, all in single
 *   precision (correctly rounded sqrt), widened to double. */
static void box_muller(uint32_t x0, uint32_t x1, double& z0, double& z1) {
    static const float PLOG[9] = {1.0f,          -0x1.fffff8p-2f, 0x1.555790p-2f, -0x1.0005a6p-2f, 0x1.98b80ap-3f,
                                  -0x1.5329bep-3f, 0x1.322850p-3f, -0x1.26729ep-3f, 0x1.6626eap-4f};
    static const float PSIN[5] = {0x1.921fb6p-1f, -0x1.4abbcep-4f, 0x1.466bc6p-9f, -0x1.32d2ccp-15f, 0x1.507834p-22f};
    static const float PCOS[5] = {1.0f, -0x1.3bd3ccp-2f, 0x1.03c1f0p-6f, -0x1.55d3c8p-12f, 0x1.e1f506p-19f};
    float la = 0.0f;
    const uint64_t n = (uint64_t)x0 + 1;  /* in [1, 2^32] */
    if (n < (1ull << 32)) {
        int k = 0;  /* n = m 2^(k - 31) with m in [2^31, 2^32) */
        uint64_t m = n;
        while (m < (1ull << 31)) {
            m <<= 1;
            --k;
        }
        float f = std::ldexp((float)(uint32_t)(m >> 8), -23);  /* [1, 2) */
        int e = k - 1;  /* a = n 2^-32 = (m 2^-31) 2^(k - 1) */
        if (f > 0x1.6a09e6p+0f) {
            f = f / 2.0f;
            e += 1;
        }
        const float s = f - 1.0f;
        float p = PLOG[8];
        for (int i = 7; i >= 0; --i) p = std::fmaf(p, s, PLOG[i]);
        la = std::fmaf((float)e, 0x1.62e430p-1f, s * p);
    }
    const float r = std::sqrt(-2.0f * la);  /* float overload: correctly rounded */
    const uint32_t c = x1 >> 8;
    const uint32_t oct = c >> 21, off = c & 0x1FFFFF;
    const bool odd = (oct & 1) != 0;
    const float t = std::ldexp((float)(odd ? 0x200000 - off : off), -21);
    const float t2 = t * t;
    float ps = PSIN[4], pc = PCOS[4];
    for (int i = 3; i >= 0; --i) {
        ps = std::fmaf(ps, t2, PSIN[i]);
        pc = std::fmaf(pc, t2, PCOS[i]);
    }
    const float st = t * ps;
    /* angle within the quadrant, then the quadrant's rotation */
    const float sq = odd ? pc : st, cq = odd ? st : pc;
    float sn, cs;
    switch (oct >> 1) {
        case 0: sn = sq; cs = cq; break;
        case 1: sn = cq; cs = -sq; break;
        case 2: sn = -sq; cs = -cq; break;
        default: sn = -cq; cs = sq; break;
    }
    z0 = (double)(r * cs);
    z1 = (double)(r * sn);
}

/* n (<= 8) standard normals, four per Philox block: words (x0, x1) give the
 * first pair, (x2, x3) the second. */
static void normals(uint64_t seed, uint32_t stream, uint64_t step, uint32_t parent, uint32_t gauss,
                    uint32_t elem, int n, double* z) {
    for (int b = 0; 4 * b < n; ++b) {
        uint32_t x[4];
        draw(seed, stream, step, parent, gauss, elem, (uint32_t)b, x);
        for (int h = 0; h < 2; ++h) {
            const int i = 4 * b + 2 * h;
            if (i >= n) break;
            double z0, z1;
            box_muller(x[2 * h], x[2 * h + 1], z0, z1);
            z[i] = z0;
            if (i + 1 < n) z[i + 1] = z1;
        }
    }
}

/* The normals of a group of samples that share Philox blocks: blocks
 * b = 0 .. nb - 1 of element elem give n[4 b + 2 h], n[4 b + 2 h + 1] =
 * box_muller(words 2 h, 2 h + 1 of block b). */
static void group_normals(uint64_t seed, uint32_t stream, uint64_t step, uint32_t parent, uint32_t gauss,
                          uint32_t elem, int nb, double* n) {
    for (int b = 0; b < nb; ++b) {
        uint32_t x[4];
        draw(seed, stream, step, parent, gauss, elem, (uint32_t)b, x);
        box_muller(x[0], x[1], n[4 * b], n[4 * b + 1]);
        box_muller(x[2], x[3], n[4 * b + 2], n[4 * b + 3]);
    }
}

/* D1, landmark prediction: three normals per sample, four samples to three
 * blocks.  Sample s is member i = (s >> 4) & 3 of the group with element
 * (s >> 6) << 4 | (s & 15) -- the samples s, s + 16, s + 32, s + 48 of one
 * 64-sample stretch -- and takes the group's normals n[3 i .. 3 i + 2]. */
static inline uint32_t feature_elem(uint32_t s) { return ((s >> 6) << 4) | (s & 15u); }

/* D1, posterior sampling: six normals per sample, two samples to three
 * blocks.  Sample s is member i = (s >> 4) & 1 of the group with element
 * (s >> 5) << 4 | (s & 15) and takes n[6 i .. 6 i + 5]. */
static inline uint32_t posterior_elem(uint32_t s) { return ((s >> 5) << 4) | (s & 15u); }

/* the normals of one sample of either stream (the unit tests' view) */
static void stream_normals(uint64_t seed, uint32_t stream, uint64_t step, uint32_t parent, uint32_t gauss,
                           uint32_t s, double* z) {
    double n[12];
    if (stream == ST_FEATURE) {
        const int i = (int)((s >> 4) & 3u);
        group_normals(seed, stream, step, parent, gauss, feature_elem(s), 3, n);
        for (int c = 0; c < 3; ++c) z[c] = n[3 * i + c];
    } else {
        const int i = (int)((s >> 4) & 1u);
        group_normals(seed, stream, step, parent, gauss, posterior_elem(s), 3, n);
        for (int c = 0; c < 6; ++c) z[c] = n[6 * i + c];
    }
}

/* exported for the unit tests: one Box-Muller pair from raw words */
extern "C" void orc_box_muller(uint32_t x0, uint32_t x1, double* z) { box_muller(x0, x1, z[0], z[1]); }

/* ---------------------------------------------------------------- state */
struct Comp {  /* one 6-D Gaussian: weight, mean, full column-major covariance */
    double w;
    double m[6];
    double P[36];
};

struct State {
    Params prm;
    int Np = 0, S = 0;
    std::vector<double> ext;      /* Np*12 */
    std::vector<double> pw;       /* Np    */
    std::vector<int32_t> gcount;  /* Np    */
    std::vector<double> wg;       /* Ng    */
    std::vector<double> smp[7];   /* Ng*S each: x y z vx vy vz w */
    int64_t n_predict = 0, n_update = 0, n_resample = 0;
    /* intermediates of the last update / resample */
    int64_t last_M = 0;
    std::vector<double> eta, loglik;
    std::vector<int32_t> nmerged, nnd, nsurv, anc;
    std::vector<double> lastfit; /* 28 planes x Ng of the last update: mean, packed cov, pd */
    int64_t lastfit_n = 0;
    int32_t last_resampled = 0;
    double last_neff = 0.0;
};

static std::vector<int64_t> offsets(const std::vector<int32_t>& cnt) {
    std::vector<int64_t> off(cnt.size() + 1, 0);
    for (size_t i = 0; i < cnt.size(); ++i) off[i + 1] = off[i] + cnt[i];
    return off;
}

/* Prior: uniform [a, b] per extrinsic component, weights 1/N.
 * src/scphdcameracalibration.cu:835-980 (D1: Philox stream instead of an
 * unseeded default_random_engine). */
static void init_prior(State& st) {
    const Params& P = st.prm;
    st.Np = P.n_particles;
    st.S = P.particles_per_feature;
    st.ext.assign((size_t)st.Np * 12, 0.0);
    st.pw.assign(st.Np, 1.0 / st.Np);
    st.gcount.assign(st.Np, 0);
    st.wg.clear();
    for (auto& v : st.smp) v.clear();
    for (int p = 0; p < st.Np; ++p) {
        for (int c = 0; c < 12; ++c) {
            uint32_t x[4];
            draw(P.seed, ST_PRIOR, 0, (uint32_t)p, 0, 0, (uint32_t)(c / 2), x);
            const double r = (c % 2 == 0) ? u01(x[0], x[1]) : u01(x[2], x[3]);
            st.ext[(size_t)p * 12 + c] = r * (P.prior_b[c] - P.prior_a[c]) + P.prior_a[c];
        }
    }
    st.anc.resize(st.Np);
    for (int p = 0; p < st.Np; ++p) st.anc[p] = p;
}

/* ------------------------------------------------------------- predict */
/* LinearCVMotionModel3D::computeNoisyMotion, src/linearcvmotionmodel3d.cuh:28-40 */
static inline void cv_noisy(double* x, double* v, double a, double dt) {
    *x = *x + *v * dt;
    *x += 0.5 * a * dt * dt;
    *v += a * dt;
}

/* predict(dt), src/scphdcameracalibration.cu:1034-1151: parents through
 * OrientedLinearCVMotionModel3D::computeNoisyMotion
 * (src/OrientedLinearCVMotionModel3D.cuh:40-54), samples through
 * predict_features (:154-178), Gaussian weights * ps (:1144-1150). */
static void predict(State& st, double dt) {
    const Params& P = st.prm;
    const uint64_t step = (uint64_t)st.n_predict;
#pragma omp parallel for schedule(static)
    for (int p = 0; p < st.Np; ++p) {
        double a[6];
        normals(P.seed, ST_CAMERA, step, (uint32_t)p, 0, 0, 6, a);
        double* e = &st.ext[(size_t)p * 12];
        for (int c = 0; c < 3; ++c) cv_noisy(&e[c], &e[3 + c], a[c] * P.cam_accel_std[c], dt);
        for (int c = 0; c < 3; ++c)
            cv_noisy(&e[6 + c], &e[9 + c], a[3 + c] * P.cam_accel_std[3 + c], dt);
        for (int c = 0; c < 3; ++c) e[6 + c] = wrap_angle(e[6 + c]);
    }
    const std::vector<int64_t> goff = offsets(st.gcount);
    const int S = st.S;
#pragma omp parallel for schedule(dynamic, 1)
    for (int p = 0; p < st.Np; ++p) {
        for (int64_t g = goff[p]; g < goff[p + 1]; ++g) {
            const uint32_t k = (uint32_t)(g - goff[p]);
            /* groups of samples s0, s0 + 16, s0 + 32, s0 + 48 (D1): the
             * blocks their members need, then their motion */
            for (int s0 = 0; s0 < S; s0 += ((s0 & 15) == 15 ? 49 : 1)) {
                int nm = 0;
                while (nm < 4 && s0 + 16 * nm < S) ++nm;
                double n[12];
                group_normals(P.seed, ST_FEATURE, step, (uint32_t)p, k, feature_elem((uint32_t)s0), (3 * nm + 3) / 4, n);
                for (int m = 0; m < nm; ++m) {
                    const size_t i = (size_t)g * S + s0 + 16 * m;
                    for (int c = 0; c < 3; ++c) {
                        const double ac = n[3 * m + c] * P.feat_accel_std[c];
                        const double x = st.smp[c][i], v = st.smp[3 + c][i];
                        st.smp[c][i] = x + v * dt + 0.5 * ac * dt * dt;
                        st.smp[3 + c][i] = v + ac * dt;
                    }
                }
            }
            st.wg[g] *= P.ps;
        }
    }
    st.n_predict++;
}

/* ---------------------------------------------------------------- update */
/* Joseph-form Kalman covariance with H = [I2 0], R = diag(vu, vv) (D8). */
static void joseph(const double* P, const double* K /*6x2 col-major*/, double vu, double vv, double* Pn) {
    double A[36];
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) A[i + 6 * j] = (i == j) ? 1.0 : 0.0;
    for (int i = 0; i < 6; ++i) {
        A[i + 0] -= K[i];
        A[i + 6] -= K[6 + i];
    }
    double AP[36];
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) {
            double s = 0.0;
            for (int k = 0; k < 6; ++k) s += A[i + 6 * k] * P[k + 6 * j];
            AP[i + 6 * j] = s;
        }
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j <= i; ++j) {
            double s = 0.0;
            for (int k = 0; k < 6; ++k) s += AP[i + 6 * k] * A[j + 6 * k];
            s += K[i] * vu * K[j] + K[6 + i] * vv * K[6 + j];
            Pn[i + 6 * j] = s;
            Pn[j + 6 * i] = s;
        }
}

/* Per-Gaussian terms of the detection update that do not depend on the
 * measurement (computeDetectionsKernel, src/kernels.cu:470-603). */
struct Pre {
    double s[4];    /* inverse innovation covariance, column-major */
    double det;
    double K[12];   /* gain, 6x2 column-major */
    double Pn[36];  /* updated covariance */
    bool pd_ok;     /* isPosDef(updated) */
};

static void precompute(const Params& P, const Comp& f, Pre& r) {
    const double vu = P.std_u * P.std_u, vv = P.std_v * P.std_v;
    const double* p = f.P;
    double sg[4] = {p[0] + vu, p[1], p[6], p[7] + vv};
    sg[1] = (sg[1] + sg[2]) / 2;
    sg[2] = sg[1];
    const double det = sg[0] * sg[3] - sg[1] * sg[2];
    r.det = det;
    r.s[0] = sg[3] / det;
    r.s[1] = -sg[1] / det;
    r.s[2] = -sg[2] / det;
    r.s[3] = sg[0] / det;
    for (int i = 0; i < 6; ++i) {
        r.K[i] = p[i] * r.s[0] + p[6 + i] * r.s[1];
        r.K[6 + i] = p[i] * r.s[2] + p[6 + i] * r.s[3];
    }
    joseph(p, r.K, vu, vv, r.Pn);
    r.pd_ok = is_pos_def6(r.Pn);
}

/* log weight of the detection term (src/kernels.cu:527-539) */
static inline double det_logw(const Pre& r, double logpd, double logw, double zu, double zv, const double* m) {
    const double i0 = zu - m[0], i1 = zv - m[1];
    const double dist = i0 * i0 * r.s[0] + i0 * i1 * (r.s[1] + r.s[2]) + i1 * i1 * r.s[3];
    double lw = logpd + logw - 0.5 * dist - safe_log(2 * M_PI) - 0.5 * safe_log(r.det);
    if (!r.pd_ok) lw = LOG0;
    return lw;
}

/* Greedy merge of one parent's survivors (in pack order) --
 * phdUpdateMergeKernel, src/kernels.cu:662-874 (D4, D5). */
static void merge(const Params& P, const std::vector<Comp>& in, std::vector<Comp>& out) {
    const size_t U = in.size();
    std::vector<char> used(U, 0);
    std::vector<size_t> mem;
    for (;;) {
        long js = -1;
        for (size_t i = 0; i < U; ++i)
            if (!used[i] && (js < 0 || in[i].w > in[(size_t)js].w)) js = (long)i;
        if (js < 0) break;
        const Comp& mx = in[(size_t)js];
        mem.clear();
        for (size_t i = 0; i < U; ++i) {
            if (used[i]) continue;
            const double d = mahal6(mx.m, mx.P, in[i].m, in[i].P);
            if (d < P.min_separation) mem.push_back(i);
        }
        double W = 0.0;
        for (size_t i : mem) W += in[i].w;
        if (W == 0) break;
        Comp g;
        g.w = W;
        double acc[6] = {0, 0, 0, 0, 0, 0};
        for (size_t i : mem)
            for (int c = 0; c < 6; ++c) acc[c] += in[i].w * in[i].m[c];
        for (int c = 0; c < 6; ++c) g.m[c] = acc[c] / W;
        double C[36];
        for (int c = 0; c < 36; ++c) C[c] = 0.0;
        for (size_t i : mem) {
            double dl[6];
            for (int c = 0; c < 6; ++c) dl[c] = g.m[c] - in[i].m[c];
            for (int a = 0; a < 6; ++a)
                for (int b = 0; b < 6; ++b) C[a * 6 + b] += in[i].w * (in[i].P[a * 6 + b] + dl[a] * dl[b]);
            used[i] = 1;
        }
        for (int c = 0; c < 36; ++c) g.P[c] = C[c] / W;
        force_symmetric6(g.P);
        if (is_pos_def6(g.P))
            out.push_back(g);
        else
            out.push_back(mx);
    }
}

/* update(u, v, fixed_camera), src/scphdcameracalibration.cu:1576-2094 */
static int update(State& st, const double* zu, const double* zv, int M, int fixed) {
    const Params& P = st.prm;
    const int Np = st.Np, S = st.S;
    const std::vector<int64_t> goff = offsets(st.gcount);
    const int64_t Ng = goff[Np];
    const double vu = P.std_u * P.std_u, vv = P.std_v * P.std_v;

    /* births, computeBirths :2589-2638 */
    std::vector<Comp> births((size_t)M);
    for (int m = 0; m < M; ++m) {
        Comp& b = births[m];
        b.m[0] = zu[m]; b.m[1] = zv[m]; b.m[2] = P.birth_d0;
        b.m[3] = P.birth_vu0; b.m[4] = P.birth_vv0; b.m[5] = P.birth_vd0;
        for (int c = 0; c < 36; ++c) b.P[c] = 0.0;
        b.P[0] = vu; b.P[7] = vv; b.P[14] = P.birth_var_d0;
        b.P[21] = P.birth_var_vu0; b.P[28] = P.birth_var_vv0; b.P[35] = P.birth_var_vd0;
        b.w = safe_log(P.birth_w0);
    }
    const bool births_pd = M > 0 ? is_pos_def6(births[0].P) : true;
    /* birth + clutter term, :1744-1745 */
    const double bpc = safe_log(M * P.birth_w0 + P.lambda / ((2 * P.v0) * (2 * P.v0)));

    st.eta.assign((size_t)Np * M, 0.0);
    st.loglik.assign(Np, 0.0);
    st.nmerged.assign(Np, 0);
    st.nnd.assign(Np, 0);
    st.nsurv.assign(Np, 0);
    st.last_M = M;
    st.lastfit.assign((size_t)28 * Ng, 0.0);
    st.lastfit_n = Ng;

    /* new maps, assembled per parent then concatenated */
    std::vector<std::vector<double>> new_wg(Np);
    std::vector<std::vector<double>> new_smp[7];
    for (auto& v : new_smp) v.assign(Np, std::vector<double>());

#pragma omp parallel for schedule(dynamic, 1)
    for (int p = 0; p < Np; ++p) {
        const int G = st.gcount[p];
        double zero[12] = {0};
        const double* e = fixed ? zero : &st.ext[(size_t)p * 12];
        double R[3][3];
        rotation(e, R);

        std::vector<Comp> pred((size_t)G);
        std::vector<double> pdg((size_t)G);
        std::vector<double> dsp((size_t)G * S * 6), pdi((size_t)G * S);
        std::vector<double> w((size_t)S);
        for (int k = 0; k < G; ++k) {
            const int64_t g = goff[p] + k;
            /* computeDisparityParticles: K1 + K3, :1155-1220, src/kernels.cu:6-99 */
            double pdsum = 0.0;
            for (int s = 0; s < S; ++s) {
                const size_t i = (size_t)g * S + s;
                double pwd[6] = {st.smp[0][i], st.smp[1][i], st.smp[2][i],
                                 st.smp[3][i], st.smp[4][i], st.smp[5][i]};
                double pc[6], q[6];
                world_to_camera(R, e, pwd, pc);
                project(P, pc, q);
                for (int c = 0; c < 6; ++c) dsp[((size_t)k * S + s) * 6 + c] = q[c];
                const double pd = in_range(P, q) ? P.pd : 0.0;
                pdi[(size_t)k * S + s] = pd;
                pdsum += pd;
                /* w_i *= pd_i, :1602-1608 */
                w[s] = st.smp[6][i] * pd;
            }
            pdg[k] = pdsum / S;
            Comp& f = pred[k];
            f.w = st.wg[g];
            const double* q0 = dsp.data() + (size_t)k * S * 6;
            if (pdsum == 0.0) {
                /* D13: a cloud entirely out of view (pd_g = 0) has detection
                 * terms of weight 0, so its fit cannot reach any output; it
                 * is defined as the first sample with zero covariance */
                for (int c = 0; c < 6; ++c) f.m[c] = q0[c];
                for (int c = 0; c < 36; ++c) f.P[c] = 0.0;
                continue;
            }
            normalise_weights(w.data(), S);
            /* fitGaussiansKernel, src/kernels.cu:102-438 */
            for (int c = 0; c < 6; ++c) {
                double a = 0.0;
                for (int s = 0; s < S; ++s) a += q0[s * 6 + c] * w[s];
                f.m[c] = a;
            }
            double W = 0.0, W2 = 0.0;
            for (int s = 0; s < S; ++s) W += w[s];
            for (int s = 0; s < S; ++s) W2 += w[s] * w[s];
            const double cn = W / (W * W + W2);
            for (int a = 0; a < 6; ++a)
                for (int b = 0; b <= a; ++b) {
                    double acc = 0.0;
                    for (int s = 0; s < S; ++s)
                        acc += (q0[s * 6 + a] - f.m[a]) * (q0[s * 6 + b] - f.m[b]) * w[s];
                    f.P[a + 6 * b] = acc / cn;
                    f.P[b + 6 * a] = acc / cn;
                }
        }

        for (int k = 0; k < G; ++k) {
            const size_t g = (size_t)(goff[p] + k);
            for (int c = 0; c < 6; ++c) st.lastfit[c * (size_t)Ng + g] = pred[k].m[c];
            int e = 0;
            for (int a = 0; a < 6; ++a)
                for (int b = 0; b <= a; ++b, ++e) st.lastfit[(6 + e) * (size_t)Ng + g] = pred[k].P[a + 6 * b];
            st.lastfit[27 * (size_t)Ng + g] = pdg[k];
        }

        /* detection terms and normalisers, K6 src/kernels.cu:470-603,
         * :1729-1768 (D2, D5) */
        std::vector<Pre> pre((size_t)G);
        std::vector<double> lw((size_t)G * M);
        for (int k = 0; k < G; ++k) {
            precompute(P, pred[k], pre[k]);
            const double lpd = safe_log(pdg[k]), lwg = safe_log(pred[k].w);
            for (int m = 0; m < M; ++m) lw[(size_t)m * G + k] = det_logw(pre[k], lpd, lwg, zu[m], zv[m], pred[k].m);
        }
        /* (data() + offset: with M = 0 or G = 0 the vectors are empty) */
        double* eta = st.eta.data() + (size_t)p * M;
        for (int m = 0; m < M; ++m) eta[m] = log_normaliser(lw.data() + (size_t)m * G, G, bpc);

        /* single-cluster parent log-likelihood, :1813-1849 */
        double ls = 0.0;
        for (int m = 0; m < M; ++m) ls += eta[m];
        double card = 0.0;
        for (int k = 0; k < G; ++k) card += pred[k].w;
        st.loglik[p] = ls + card;

        /* pack + prune flags, updateKernel src/kernels.cu:608-659 */
        std::vector<Comp> surv;
        for (int m = 0; m < M; ++m)
            for (int k = 0; k < G; ++k) {
                const double wt = std::exp(lw[(size_t)m * G + k] - eta[m]);
                if (wt < P.min_weight || !pre[k].pd_ok) continue;
                Comp c;
                c.w = wt;
                const double i0 = zu[m] - pred[k].m[0], i1 = zv[m] - pred[k].m[1];
                for (int r = 0; r < 6; ++r) c.m[r] = pred[k].m[r] + i0 * pre[k].K[r] + i1 * pre[k].K[6 + r];
                std::memcpy(c.P, pre[k].Pn, sizeof c.P);
                surv.push_back(c);
            }
        for (int m = 0; m < M; ++m) {
            const double wt = std::exp(births[m].w - eta[m]);
            if (wt < P.min_weight || !births_pd) continue;
            Comp c = births[m];
            c.w = wt;
            surv.push_back(c);
        }
        st.nsurv[p] = (int32_t)surv.size();

        /* GM merge, K8 */
        std::vector<Comp> mg;
        merge(P, surv, mg);
        const int K = (int)mg.size();
        st.nmerged[p] = K;

        /* non-detection terms kept, computeNonDetections :1290-1420 */
        std::vector<int> keep;
        for (int k = 0; k < G; ++k)
            if (pred[k].w * (1 - pdg[k]) >= P.min_weight) keep.push_back(k);
        st.nnd[p] = (int32_t)keep.size();

        /* posterior sampling (sample_disparity_gaussian :640-690, :1972-2010),
         * recombination [posterior | non-detection] (:1422-1574) and
         * disparity->world (K2, src/kernels.cu:49-80) */
        const size_t nG = (size_t)K + keep.size();
        new_wg[p].resize(nG);
        for (int c = 0; c < 7; ++c) new_smp[c][p].resize(nG * S);
        for (int k = 0; k < K; ++k) {
            new_wg[p][k] = mg[k].w;
            double L[36];
            cholesky6(mg[k].P, L);
            /* D1: samples s0 and s0 + 16 of a 32-sample stretch share three blocks */
            for (int s0 = 0; s0 < S; s0 += ((s0 & 15) == 15 ? 17 : 1)) {
              const int nm = s0 + 16 < S ? 2 : 1;
              double gn[12];
              group_normals(P.seed, ST_POSTERIOR, (uint64_t)st.n_update, (uint32_t)p, (uint32_t)k,
                            posterior_elem((uint32_t)s0), nm == 2 ? 3 : 2, gn);
              for (int m = 0; m < nm; ++m) {
                const int s = s0 + 16 * m;
                const double* z = gn + 6 * m;
                double q[6];
                for (int r = 0; r < 6; ++r) {
                    double a = 0.0;
                    for (int c = 0; c <= r; ++c) a += L[r + 6 * c] * z[c];
                    q[r] = a + mg[k].m[r];
                }
                double pc[6], pwd[6];
                unproject(P, q, pc);
                camera_to_world(R, e, pc, pwd);
                const size_t o = (size_t)k * S + s;
                for (int c = 0; c < 6; ++c) new_smp[c][p][o] = pwd[c];
                new_smp[6][p][o] = 1.0 / S;
              }
            }
        }
        for (size_t j = 0; j < keep.size(); ++j) {
            const int k = keep[j];
            new_wg[p][K + j] = pred[k].w * (1 - pdg[k]);
            for (int s = 0; s < S; ++s) {
                const double* q = &dsp[((size_t)k * S + s) * 6];
                double pc[6], pwd[6];
                unproject(P, q, pc);
                camera_to_world(R, e, pc, pwd);
                const size_t o = (K + j) * (size_t)S + s;
                for (int c = 0; c < 6; ++c) new_smp[c][p][o] = pwd[c];
                new_smp[6][p][o] = 1.0 - pdi[(size_t)k * S + s];
            }
        }
    }

    /* parent reweighting (D3), :1813-1876 */
    if (!fixed) {
        std::vector<double> lwp(Np);
        double mx = -INFINITY;
        for (int p = 0; p < Np; ++p) {
            lwp[p] = safe_log(st.pw[p]) + st.loglik[p];
            if (lwp[p] > mx) mx = lwp[p];
        }
        double sum = 0.0;
        for (int p = 0; p < Np; ++p) {
            lwp[p] = std::exp(lwp[p] - mx);
            sum += lwp[p];
        }
        for (int p = 0; p < Np; ++p) st.pw[p] = lwp[p] / sum;
    }

    /* concatenate new maps */
    int64_t nNg = 0;
    for (int p = 0; p < Np; ++p) nNg += (int64_t)new_wg[p].size();
    if (P.max_gaussians > 0)
        for (int p = 0; p < Np; ++p)
            if ((int64_t)new_wg[p].size() > P.max_gaussians) return -4;
    st.wg.resize(nNg);
    for (auto& v : st.smp) v.resize((size_t)nNg * S);
    int64_t o = 0;
    for (int p = 0; p < Np; ++p) {
        const int64_t n = (int64_t)new_wg[p].size();
        st.gcount[p] = (int32_t)n;
        std::copy(new_wg[p].begin(), new_wg[p].end(), st.wg.begin() + o);
        for (int c = 0; c < 7; ++c)
            std::copy(new_smp[c][p].begin(), new_smp[c][p].end(), st.smp[c].begin() + o * S);
        o += n;
    }
    if (P.storage == 0)
        for (int c = 0; c < 7; ++c)
            for (auto& x : st.smp[c]) x = (double)(float)x;
    (void)Ng;
    st.n_update++;
    return 0;
}

/* ---------------------------------------------------------------- resample */
/* Systematic resampling (Kitagawa 1996) as in resample(),
 * src/scphdcameracalibration.cu:2651-2702 (D6).  Returns ancestors. */
static void systematic(const double* w, int N, double u, int32_t* a) {
    const double interval = 1.0 / N;
    double c = w[0];
    int i = 0;
    for (int j = 0; j < N; ++j) {
        const double r = u + j * interval;
        while (r > c) {
            ++i;
            if (i >= N) {
                double mw = -1;
                int mi = -1;
                for (int k = 0; k < N; ++k)
                    if (std::exp(w[k]) > mw) {
                        mw = std::exp(w[k]);
                        mi = k;
                    }
                i = mi;
                c = 2;
                break;
            }
            c += w[i];
        }
        a[j] = i;
    }
}

static int resample(State& st, int32_t* resampled) {
    const int N = st.Np;
    double ss = 0.0;
    for (int p = 0; p < N; ++p) ss += st.pw[p] * st.pw[p];
    const double neff = 1.0 / (N * ss); /* computeNeff :2641-2649 */
    st.last_neff = neff;
    const uint64_t step = (uint64_t)st.n_resample;
    st.n_resample++;
    for (int p = 0; p < N; ++p) st.anc[p] = p;
    st.last_resampled = 0;
    if (resampled) *resampled = 0;
    if (neff >= st.prm.min_neff) return 0;
    uint32_t x[4];
    draw(st.prm.seed, ST_RESAMPLE, step, 0, 0, 0, 0, x);
    const double u = u01(x[0], x[1]) * (1.0 / N);
    systematic(st.pw.data(), N, u, st.anc.data());
    /* gather parents and copy ancestor maps (:2705-2811) */
    const std::vector<int64_t> goff = offsets(st.gcount);
    const int S = st.S;
    std::vector<double> next((size_t)N * 12);
    std::vector<int32_t> ng(N);
    int64_t tot = 0;
    for (int j = 0; j < N; ++j) {
        const int a = st.anc[j];
        std::copy(&st.ext[(size_t)a * 12], &st.ext[(size_t)a * 12] + 12, &next[(size_t)j * 12]);
        ng[j] = st.gcount[a];
        tot += ng[j];
    }
    std::vector<double> nwg((size_t)tot);
    std::vector<double> nsmp[7];
    for (auto& v : nsmp) v.resize((size_t)tot * S);
    int64_t o = 0;
    for (int j = 0; j < N; ++j) {
        const int a = st.anc[j];
        const int64_t b = goff[a], n = ng[j];
        std::copy(st.wg.begin() + b, st.wg.begin() + b + n, nwg.begin() + o);
        for (int c = 0; c < 7; ++c)
            std::copy(st.smp[c].begin() + b * S, st.smp[c].begin() + (b + n) * S, nsmp[c].begin() + o * S);
        o += n;
    }
    st.ext.swap(next);
    st.gcount.swap(ng);
    st.wg.swap(nwg);
    for (int c = 0; c < 7; ++c) st.smp[c].swap(nsmp[c]);
    for (int p = 0; p < N; ++p) st.pw[p] = 1.0 / N;
    st.last_resampled = 1;
    if (resampled) *resampled = 1;
    return 0;
}

}  // namespace orc

/* =================================================================== C API */
using namespace orc;

extern "C" {

int orc_params_size(void) { return (int)sizeof(Params); }

void* orc_create(const Params* p) {
    if (!p || p->n_particles <= 0 || p->particles_per_feature <= 0) return nullptr;
    State* st = new State();
    st->prm = *p;
    init_prior(*st);
    return st;
}

void orc_destroy(void* h) { delete (State*)h; }

int orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

int orc_predict(void* h, double dt) {
    predict(*(State*)h, dt);
    return 0;
}

int orc_update(void* h, const double* u, const double* v, int M, int fixed) {
    if (M < 0) return -1;
    return update(*(State*)h, u, v, M, fixed);
}

int orc_resample(void* h, int32_t* resampled) { return resample(*(State*)h, resampled); }

int orc_check(void* h) {
    State* st = (State*)h;
    for (double w : st->pw)
        if (w != w) return -3;
    return 0;
}

void orc_get_counts(void* h, int64_t* out /* Np S Ng Ns npred nupd nres M */) {
    State* st = (State*)h;
    int64_t Ng = 0;
    for (int32_t c : st->gcount) Ng += c;
    out[0] = st->Np; out[1] = st->S; out[2] = Ng; out[3] = Ng * st->S;
    out[4] = st->n_predict; out[5] = st->n_update; out[6] = st->n_resample; out[7] = st->last_M;
}

void orc_get_parents(void* h, double* states, double* weights) {
    State* st = (State*)h;
    if (states) std::copy(st->ext.begin(), st->ext.end(), states);
    if (weights) std::copy(st->pw.begin(), st->pw.end(), weights);
}

void orc_get_maps(void* h, int32_t* gcount, double* wg, double* samples) {
    State* st = (State*)h;
    if (gcount) std::copy(st->gcount.begin(), st->gcount.end(), gcount);
    if (wg) std::copy(st->wg.begin(), st->wg.end(), wg);
    if (samples) {
        const size_t n = st->smp[0].size();
        for (int c = 0; c < 7; ++c) std::copy(st->smp[c].begin(), st->smp[c].end(), samples + c * n);
    }
}

int orc_set_state(void* h, const double* states, const double* weights, const int32_t* gcount, int64_t Ng,
                  const double* wg, const double* samples, int64_t np, int64_t nu, int64_t nr) {
    State* st = (State*)h;
    int64_t tot = 0;
    for (int p = 0; p < st->Np; ++p) tot += gcount[p];
    if (tot != Ng) return -1;
    st->ext.assign(states, states + (size_t)st->Np * 12);
    st->pw.assign(weights, weights + st->Np);
    st->gcount.assign(gcount, gcount + st->Np);
    st->wg.assign(wg, wg + Ng);
    const size_t n = (size_t)Ng * st->S;
    for (int c = 0; c < 7; ++c) st->smp[c].assign(samples + c * n, samples + (c + 1) * n);
    st->n_predict = np; st->n_update = nu; st->n_resample = nr;
    return 0;
}

void orc_get_update_debug(void* h, double* eta, double* loglik, int32_t* nmerged, int32_t* nnd, int32_t* nsurv) {
    State* st = (State*)h;
    if (eta) std::copy(st->eta.begin(), st->eta.end(), eta);
    if (loglik) std::copy(st->loglik.begin(), st->loglik.end(), loglik);
    if (nmerged) std::copy(st->nmerged.begin(), st->nmerged.end(), nmerged);
    if (nnd) std::copy(st->nnd.begin(), st->nnd.end(), nnd);
    if (nsurv) std::copy(st->nsurv.begin(), st->nsurv.end(), nsurv);
}

void orc_get_fit_debug(void* h, double* out) {
    State* st = (State*)h;
    std::copy(st->lastfit.begin(), st->lastfit.end(), out);
}

void orc_get_ancestors(void* h, int32_t* anc, double* neff) {
    State* st = (State*)h;
    if (anc) std::copy(st->anc.begin(), st->anc.end(), anc);
    if (neff) *neff = st->last_neff;
}

/* ---- unit-level entry points (known-answer and invariant tests) ---- */
void orc_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out) { philox4x32_10(ctr, key, out); }

void orc_normals(uint64_t seed, uint32_t stream, uint64_t step, uint32_t parent, uint32_t gauss,
                 uint32_t elem, int n, double* z) {
    normals(seed, stream, step, parent, gauss, elem, n, z);
}

/* the landmark-prediction (stream 2: 3 values) or posterior (stream 3: 6
 * values) normals of sample s */
void orc_stream_normals(uint64_t seed, uint32_t stream, uint64_t step, uint32_t parent, uint32_t gauss, uint32_t s,
                        double* z) {
    stream_normals(seed, stream, step, parent, gauss, s, z);
}

double orc_wrap_angle(double a) { return wrap_angle(a); }
double orc_safe_log(double x) { return safe_log(x); }
double orc_lse2(double a, double b) { return lse2(a, b); }

void orc_rotation(const double* e, double* R9) {
    double R[3][3];
    rotation(e, R);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) R9[3 * i + j] = R[i][j];
}

/* world point -> disparity point for extrinsics e (K1 per sample); returns pd */
double orc_world_to_disparity(const Params* P, const double* e, const double* pw, double* q) {
    double R[3][3], pc[6];
    rotation(e, R);
    world_to_camera(R, e, pw, pc);
    project(*P, pc, q);
    return in_range(*P, q) ? P->pd : 0.0;
}

void orc_disparity_to_world(const Params* P, const double* e, const double* q, double* pw) {
    double R[3][3], pc[6];
    rotation(e, R);
    unproject(*P, q, pc);
    camera_to_world(R, e, pc, pw);
}

void orc_cholesky(const double* A, double* L) { cholesky6(A, L); }
int orc_is_pos_def(const double* A) { return is_pos_def6(A) ? 1 : 0; }
double orc_mahal(const double* ma, const double* Pa, const double* mb, const double* Pb) {
    return mahal6(ma, Pa, mb, Pb);
}

/* Kalman detection term of one Gaussian and one measurement. out: logw, mean[6], cov[36] */
int orc_kalman(const Params* P, double w, double pd, const double* m, const double* Pc, double zu, double zv,
               double* out) {
    Comp f;
    f.w = w;
    std::memcpy(f.m, m, sizeof f.m);
    std::memcpy(f.P, Pc, sizeof f.P);
    Pre r;
    precompute(*P, f, r);
    out[0] = det_logw(r, safe_log(pd), safe_log(w), zu, zv, m);
    const double i0 = zu - m[0], i1 = zv - m[1];
    for (int c = 0; c < 6; ++c) out[1 + c] = m[c] + i0 * r.K[c] + i1 * r.K[6 + c];
    std::memcpy(out + 7, r.Pn, sizeof r.Pn);
    return r.pd_ok ? 1 : 0;
}

/* Greedy merge of n components (w[n], m[n*6], P[n*36]); writes at most n
 * outputs, returns the count. */
int orc_merge(const Params* P, int n, const double* w, const double* m, const double* Pc, double* ow, double* om,
              double* oP) {
    std::vector<Comp> in((size_t)n), out;
    for (int i = 0; i < n; ++i) {
        in[i].w = w[i];
        std::memcpy(in[i].m, m + 6 * i, sizeof in[i].m);
        std::memcpy(in[i].P, Pc + 36 * i, sizeof in[i].P);
    }
    merge(*P, in, out);
    for (size_t k = 0; k < out.size(); ++k) {
        ow[k] = out[k].w;
        std::memcpy(om + 6 * k, out[k].m, sizeof out[k].m);
        std::memcpy(oP + 36 * k, out[k].P, sizeof out[k].P);
    }
    return (int)out.size();
}

void orc_systematic(const double* w, int N, double u, int32_t* a) { systematic(w, N, u, a); }

/* Fit of one sample cloud given disparity samples q[S*6] and raw weights
 * w[S] (already multiplied by pd): normalise + moment fit.  out: mean[6],
 * cov[36]. */
void orc_fit(int S, const double* q, const double* w_in, double* mean, double* cov) {
    std::vector<double> w(w_in, w_in + S);
    normalise_weights(w.data(), S);
    for (int c = 0; c < 6; ++c) {
        double a = 0.0;
        for (int s = 0; s < S; ++s) a += q[s * 6 + c] * w[s];
        mean[c] = a;
    }
    double W = 0.0, W2 = 0.0;
    for (int s = 0; s < S; ++s) W += w[s];
    for (int s = 0; s < S; ++s) W2 += w[s] * w[s];
    const double cn = W / (W * W + W2);
    for (int a = 0; a < 6; ++a)
        for (int b = 0; b <= a; ++b) {
            double acc = 0.0;
            for (int s = 0; s < S; ++s) acc += (q[s * 6 + a] - mean[a]) * (q[s * 6 + b] - mean[b]) * w[s];
            cov[a + 6 * b] = acc / cn;
            cov[b + 6 * a] = acc / cn;
        }
}

}  /* extern "C" */
