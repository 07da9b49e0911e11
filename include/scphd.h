/*
 * scphd.h -- C ABI of the MI355X (gfx950) single-cluster PHD stereo camera
 * calibration filter.
 *
 * Every entry point below replaces one operation of the reference class
 * SCPHDCameraCalibration (cheesinglee/scphd-camera-calibration,
 * src/scphdcameracalibration.cuh:29-122).  The citation next to each
 * declaration names the reference member (file:line) it stands in for.
 *
 * Conventions
 *   - Return codes are int: SCPHD_OK (0) or a negative SCPHD_ERR_* value.
 *     scphd_last_error() returns a message for the last failing call on a
 *     handle.  No call throws or exits the process.
 *   - Caller buffers (host pointers unless the name says _device) are only
 *     read or written during the call.
 *   - A handle owns all device state and one HIP stream.  Distinct handles
 *     may be used from distinct threads; one handle is not reentrant.
 *   - Arrays of parent extrinsics are parent-major, 12 doubles per parent:
 *     cartesian x y z vx vy vz, angular x y z vx vy vz (the reference
 *     Extrinsics struct, src/types.h:137-140).
 *   - Sample clouds cross the ABI as 7 structure-of-arrays planes of doubles
 *     (world x y z vx vy vz, sample weight), Gaussian-major in canonical
 *     order: parent ascending, then Gaussian within parent, then sample.
 */
#ifndef SCPHD_H
#define SCPHD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SCPHD_OK 0
#define SCPHD_ERR_ARG (-1)       /* invalid argument or size                 */
#define SCPHD_ERR_HIP (-2)       /* HIP runtime failure                      */
#define SCPHD_ERR_NAN (-3)       /* NaN parent weight (checkStuff)           */
#define SCPHD_ERR_CAPACITY (-4)  /* max_gaussians per parent exceeded        */
#define SCPHD_ERR_CONFIG (-5)    /* configuration file missing or malformed  */
#define SCPHD_ERR_STATE (-6)     /* call not valid in the current state      */
#define SCPHD_ERR_NODEVICE (-7)  /* no usable GPU                            */

#define SCPHD_STORAGE_FP32 0     /* sample clouds stored as float  */
#define SCPHD_STORAGE_FP64 1     /* sample clouds stored as double */
#define SCPHD_STORAGE_FP32_ARITH 2 /* stored as float, and the per-sample work
                                      (prediction, reprojection, moment sums,
                                      posterior draws) computed in float     */

/* Filter parameters: the keys the reference reads from its libconfig file
 * (src/scphdcameracalibration.cu:827-1005, :1145, :1246-1247, :1343, :1744,
 * :1920, :2596-2610, :2655; src/main.cpp:74,77).  Layout is part of the ABI. */
typedef struct scphd_params {
    int32_t n_particles;            /* parents N_p                              */
    int32_t particles_per_feature;  /* samples per Gaussian S                   */
    int32_t max_gaussians;          /* capacity per parent, 0 = unbounded       */
    int32_t storage;                /* SCPHD_STORAGE_*                          */
    int32_t verbosity;
    int32_t slam;                   /* main loop: update with camera 2 only     */
    int32_t max_steps;
    int32_t save_all_maps;          /* accepted, no effect (map dumps)          */
    double min_weight;              /* prune threshold                          */
    double min_separation;          /* merge threshold (Mahalanobis, squared)   */
    double min_neff;                /* resample when n_eff < min_neff           */
    double ps;                      /* survival probability                     */
    double prior_a[12];             /* uniform prior lower bounds, Extrinsics order */
    double prior_b[12];             /* uniform prior upper bounds               */
    double cam_accel_std[6];        /* camera CV model: cartesian ax ay az, angular ax ay az */
    double feat_accel_std[3];       /* landmark CV model: ax ay az              */
    double fx, fy, u0, v0;          /* disparity measurement model              */
    double std_u, std_v, pd, lambda;
    double birth_w0, birth_d0, birth_vu0, birth_vv0, birth_vd0;
    double birth_var_d0, birth_var_vu0, birth_var_vv0, birth_var_vd0;
    double max_velocity, min_range; /* read by the reference, unused            */
    uint64_t seed;                  /* Philox4x32-10 key for every noise stream */
} scphd_params;

typedef struct scphd_filter scphd_t;

/* Fill *p with the defaults of the reference configuration file
 * (config/calibration.cfg) and seed 0. */
int scphd_params_init(scphd_params* p);

/* Parse a libconfig-syntax file into *p.  Accepts the keys the reference code
 * reads (prior.*.{a,b}) and the shipped file's prior.*.{mean,std}, which maps
 * to a = mean - sqrt(3) std, b = mean + sqrt(3) std.  Replaces the libconfig
 * lookups of SCPHDCameraCalibration::SCPHDCameraCalibration
 * (src/scphdcameracalibration.cu:823-1006). err may be NULL. */
int scphd_params_load(const char* cfg_path, scphd_params* p, char* err, size_t errlen);

/* Construct a filter: pick the GPU, allocate device state, draw the uniform
 * prior over parent extrinsics.  Replaces the constructor
 * src/scphdcameracalibration.cu:823-1006 and initializeCuda :1008-1032. */
int scphd_create(const scphd_params* p, scphd_t** out);

/* Free all device state.  Replaces the (implicit) destructor of the
 * reference class, src/scphdcameracalibration.cuh:29. */
void scphd_destroy(scphd_t* f);

/* Message for the last failing call on f (never NULL). */
const char* scphd_last_error(const scphd_t* f);

/* Constant-velocity prediction of parents and landmark samples, survival
 * reweighting.  Replaces predict(double dt), src/scphdcameracalibration.cu:1034-1151. */
int scphd_predict(scphd_t* f, double dt);

/* GM-PHD update of every daughter map with one camera's measurement set and,
 * when fixed_camera == 0, the single-cluster parent reweighting.  u and v are
 * host arrays of n_meas pixel coordinates.  Replaces
 * update(vector<double> u, vector<double> v, bool fixed_camera),
 * src/scphdcameracalibration.cu:1576-2094. */
int scphd_update(scphd_t* f, const double* u, const double* v, int32_t n_meas,
                 int32_t fixed_camera);

/* Same as scphd_update with u and v already resident in device memory
 * (no counterpart in the reference, whose measurements arrive as host
 * vectors at src/scphdcameracalibration.cu:1659-1660). */
int scphd_update_device(scphd_t* f, const double* u_device, const double* v_device,
                        int32_t n_meas, int32_t fixed_camera);

/* Systematic resampling of parents when n_eff < min_neff.  *resampled (may be
 * NULL) is set to 1 when parents were resampled.  Replaces resample(),
 * src/scphdcameracalibration.cu:2651-2817 and computeNeff :2641-2649. */
int scphd_resample(scphd_t* f, int32_t* resampled);

/* scphd_resample that allocates no device memory.  No counterpart in the
 * reference: for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * results as scphd_resample; its scratch (cumulative weights, gathered
 * extrinsics, ancestors, new map sizes, scan partials, the decision flag)
 * lives in a static 8 MiB device array that every call on one GPU shares,
 * so calls that overlap on one GPU race.  Plain filters only
 * (SCPHD_ERR_ARG for an ensemble).  SCPHD_ERR_CAPACITY, with the filter
 * unchanged, when the parents need more than the scratch holds (about
 * 77,000), when a deferred predict needs a larger sample generation, or
 * when the resampled map table exceeds the handle's spare table --
 * scphd_reserve sizes both. */
int scphd_resample_noalloc(scphd_t* f, int32_t* resampled);

/* NaN guard on the parent weights: SCPHD_ERR_NAN instead of exit(2).
 * Replaces checkStuff(), src/scphdcameracalibration.cu:812-821. */
int scphd_check(scphd_t* f);

/* Block until all work queued on the handle's stream has finished. */
int scphd_sync(scphd_t* f);

typedef struct scphd_counts {
    int64_t n_particles;
    int64_t particles_per_feature;
    int64_t n_gaussians;      /* N_g, Gaussians over all parents              */
    int64_t n_samples;        /* N_g * S                                      */
    int64_t n_predict;        /* predict() calls so far (noise step counters) */
    int64_t n_update;
    int64_t n_resample;
    int64_t pending_predict;  /* 1 when a predict is deferred into the next update */
} scphd_counts;

int scphd_get_counts(scphd_t* f, scphd_counts* c);

/* Parent states (N_p x 12, parent-major) and normalised parent weights.
 * Replaces the particle/weight read-back of writeMat,
 * src/scphdcameracalibration.cu:2200-2260. Either pointer may be NULL. */
int scphd_get_parents(scphd_t* f, double* states, double* weights);

/* Daughter maps in canonical order: gcount[N_p] Gaussians per parent,
 * wg[N_g] Gaussian (feature) weights, samples[7 * N_s] planes
 * x y z vx vy vz w.  Any pointer may be NULL.  Materialises a deferred
 * predict first.  Replaces the map read-back of writeMat,
 * src/scphdcameracalibration.cu:2262-2311. */
int scphd_get_maps(scphd_t* f, int32_t* gcount, double* wg, double* samples);

/* Replace the whole filter state (parents, weights, maps, noise counters).
 * Arrays as in scphd_get_parents / scphd_get_maps; n_gaussians must equal the
 * sum of gcount.  No counterpart in the reference (its state is private);
 * used for lock-step comparison and rig snapshot/resume. */
int scphd_set_state(scphd_t* f, const double* states, const double* weights,
                    const int32_t* gcount, int64_t n_gaussians, const double* wg,
                    const double* samples, int64_t n_predict, int64_t n_update,
                    int64_t n_resample);

/* Diagnostics of the last update (all counts are totals over parents). */
typedef struct scphd_update_stats {
    int64_t n_meas;            /* M                                            */
    int64_t n_gaussians_in;    /* N_g entering the update                      */
    int64_t n_pairs;           /* N_g * M component updates                    */
    int64_t n_survivors;       /* detection + birth terms kept for merging     */
    int64_t n_merged;          /* Gaussians out of the merge                   */
    int64_t n_nondetect;       /* non-detection Gaussians kept                 */
    int64_t n_gaussians_out;   /* n_merged + n_nondetect                       */
    int64_t max_survivors_per_parent;
    int64_t cap_exceeded;      /* parents whose map exceeded max_gaussians     */
} scphd_update_stats;

int scphd_get_update_stats(scphd_t* f, scphd_update_stats* s);

/* Per-parent intermediates of the last update, for verification: eta
 * [N_p * M] log normalisers (parent-major; in an ensemble M is the largest
 * rig's measurement count and a rig with fewer leaves the rest of its rows
 * unwritten), loglik[N_p] parent log-likelihood,
 * nmerged/nnondetect/nsurvivors[N_p] counts.  Any pointer may be NULL.  No
 * counterpart in the reference. */
int scphd_get_update_debug(scphd_t* f, double* eta, double* loglik, int32_t* nmerged,
                           int32_t* nnondetect, int32_t* nsurvivors);

/* Per-Gaussian fit of the last update, 28 planes of n_gaussians_in values:
 * disparity-space mean (6), covariance (21, packed lower triangle row by
 * row), detection probability pd_g.  For verification of the refit
 * (fitGaussiansKernel, src/kernels.cu:102-438, and computePdKernel
 * :82-99).  out must hold 28 * n_gaussians_in doubles. */
int scphd_get_fit_debug(scphd_t* f, double* out);

/* Ancestor index of every parent chosen by the last scphd_resample that
 * resampled (identity otherwise).  No counterpart in the reference (host
 * array idx_resample, src/scphdcameracalibration.cu:2659). */
int scphd_get_ancestors(scphd_t* f, int32_t* ancestors);

/* Wall time in milliseconds of the last update, measured with HIP events on
 * the handle's stream while timing is enabled (scphd_set_timing):
 *   [0] reproject + fit phase   [1] normalisers and survivors
 *   [2] merge and parent weights [3] posterior sampling
 *   [4] map assembly            [5] whole update
 *   [6] reproject/fit kernel alone   [7] posterior-sampling kernel alone
 * No counterpart in the reference (its NVTX ranges,
 * src/scphdcameracalibration.cu:1062-2636, carry no timings). */
int scphd_set_timing(scphd_t* f, int32_t enabled);
int scphd_get_phase_ms(scphd_t* f, double* ms8);

/* Verification limits: lower the capacities at which the update falls back
 * to its general code paths, so that small inputs exercise them.
 *   merge_lds_cap       parents with more survivors merge in global scratch
 *                       (k_merge_groups<true>); default: what fits 60 KB LDS
 *   eta_heavy_cap       0..2048: more Gaussians above the birth/clutter term
 *                       -> the normaliser max pass sweeps the whole map
 *   eta_ncand_cap       0..8: more candidate terms per measurement -> the
 *                       normaliser count and the survivor pass sweep the map
 *   merge_force_rounds  nonzero: every parent merges in the round loop
 *                       (k_merge_rounds), the path of candidate or
 *                       neighbour-list overflow
 *   merge_all_pairs     nonzero: merge candidates come from every pair of
 *                       survivors, the path of survivors whose (u, v)
 *                       covariance is ill-conditioned, instead of windows
 *                       in u
 * -1 (0 for the last two) restores the normal limit.  Every setting gives the
 * same results; no counterpart in the reference, whose merge has no
 * capacity limits (src/kernels.cu:662-874). */
int scphd_debug_set_limits(scphd_t* f, int32_t merge_lds_cap, int32_t eta_heavy_cap, int32_t eta_ncand_cap,
                           int32_t merge_force_rounds, int32_t merge_all_pairs);

/* Noise source of the handle.  mode 0: the counter-based Philox4x32-10
 * streams, keyed by seed from the next call on (the same key as
 * scphd_params.seed).  Injected noise buffers (the reference's parity
 * protocol for an oracle that cannot reproduce the device generator) are not
 * needed here -- the FP64 oracle reproduces these streams bit for bit -- and
 * any other mode, or a non-NULL injected, is SCPHD_ERR_ARG.  No counterpart
 * in the reference, whose engines are seeded from time()
 * (src/scphdcameracalibration.cu:1047-1118). */
int scphd_set_noise(scphd_t* f, int32_t mode, uint64_t seed, const double* injected, size_t n);

/* Standard normals of one noise stream (counter: elem << 2 | block, gauss,
 * parent, step << 4 | stream; key = seed), drawn on the device exactly as
 * the filter kernels draw them, six per element i = 0..n_elem-1 into
 * out[6 * n_elem]: for stream 2 (landmark prediction) the three normals of
 * sample i, then three zeros; for stream 3 (posterior sampling) the six
 * normals of sample i (samples share Philox blocks in groups, see the
 * kernels); for the other streams the six normals of element i.  For
 * verifying the device noise against a host restatement; no counterpart in
 * the reference. */
int scphd_debug_normals(uint64_t seed, uint32_t stream, uint64_t step, uint32_t parent, uint32_t gauss,
                        int32_t n_elem, double* out);

/* Calibration estimate from the current parents, computed on the device:
 * mean12 = sum_p w_p x_p (w normalised), cov144 = sum_p w_p (x_p - mean)
 * (x_p - mean)^T (12 x 12, row-major), *cardinality = sum of the Gaussian
 * weights of the highest-weight parent's map (the map writeMat writes).
 * Any pointer may be NULL.  Replaces the evaluation of the written results,
 * load_disparity_results.m:27-34 (exp_traj, cov_traj, cardinality). */
int scphd_get_estimate(scphd_t* f, double* mean12, double* cov144, double* cardinality);

/* ---- ensembles: many rigs in one filter ----
 * No counterpart in the reference, which runs one rig per process
 * (src/main.cpp:53-99).  An ensemble of n_rigs rigs, rig r configured by
 * rigs[r], is stepped by the same kernel launches as one filter of
 * sum_r rigs[r].n_particles parents: rig r draws every noise stream from
 * the key of rigs[r].seed with its parents numbered from 0, projects with
 * its own camera model, normalises and resamples its own parents, and
 * reads its own measurement set, so each rig evolves exactly as the plain
 * filter scphd_create(&rigs[r]) fed the same measurements.  Rigs may differ
 * in n_particles, the prior bounds, fx fy u0 v0, pd, lambda, the birth
 * model and the seed; every other field must equal rigs[0]'s
 * (SCPHD_ERR_ARG otherwise).  Rig r's parents follow those of rigs 0 .. r-1
 * in scphd_get_parents, scphd_get_maps and scphd_set_state.
 * scphd_predict, scphd_resample (*resampled: any rig), scphd_check and the
 * getters take an ensemble handle; scphd_update, scphd_update_device and
 * scphd_set_noise refuse it. */
int scphd_create_ensemble(const scphd_params* rigs, int32_t n_rigs, scphd_t** out);

/* Update every rig: rig r's measurements are u[moff[r] .. moff[r + 1]) and
 * v[...] (host arrays; moff has n_rigs + 1 entries, moff[0] = 0, never
 * decreasing). */
int scphd_update_ensemble(scphd_t* f, const double* u, const double* v, const int32_t* moff,
                          int32_t fixed_camera);

/* scphd_update_ensemble with a fixed-camera flag per rig (fixed_camera[r],
 * n_rigs entries): rigs at different points of their frame sequence step
 * together.  No counterpart in the reference (its driver updates one rig,
 * src/main.cpp:52-95). */
int scphd_update_ensemble_rigs(scphd_t* f, const double* u, const double* v, const int32_t* moff,
                               const int32_t* fixed_camera);

/* Ensemble membership, between frames (after an update or a resampling,
 * before the next predict; SCPHD_ERR_STATE otherwise).  No counterpart in
 * the reference: these serve many rigs from one set of launches.
 *
 * scphd_ensemble_add_rig appends a rig (*rig_out: its index, the last): its
 * parents follow the others', drawn from its prior as scphd_create draws
 * them, with an empty map.  It draws its noise with counters that start at
 * zero when it joins, so it evolves exactly as a plain filter created with
 * its parameters and fed the same frames.  Its parameters must agree with
 * the ensemble's in what rigs share (scphd_create_ensemble).
 *
 * scphd_ensemble_remove_rig drops rig `rig` with its parents and map; the
 * rigs after it move down one index.  An ensemble whose last rig has left
 * keeps its device buffers and takes new rigs; every other call on it
 * returns SCPHD_ERR_STATE until one joins.  Debug outputs of the last update
 * (scphd_get_update_debug, scphd_get_fit_debug) describe the membership of
 * that update. */
int scphd_ensemble_add_rig(scphd_t* f, const scphd_params* p, int32_t* rig_out);
int scphd_ensemble_remove_rig(scphd_t* f, int32_t rig);

/* Hand back device memory a handle holds beyond its map: the map's clouds
 * move into a sample generation of the map's own size, and the spare
 * generation and an update's work arrays are released (the next update
 * allocates them at the sizes it needs).  Between frames only.  For
 * long-lived handles whose maps shrink, such as ensembles whose rigs leave;
 * buffers otherwise keep their largest size.  Results do not change;
 * scphd_get_update_debug and scphd_get_fit_debug have nothing to report
 * until the next update.  No counterpart in the reference. */
int scphd_compact(scphd_t* f);

/* scphd_get_estimate of every rig over its own parents, one launch:
 * mean12[12 r ..], cov144[144 r ..], cardinality[r] for rig r (any may be
 * NULL; a plain filter is one rig). */
int scphd_get_rig_estimates(scphd_t* f, double* mean12, double* cov144, double* cardinality);

/* Per rig, from the last scphd_resample: resampled[r] = 1 when rig r was
 * resampled, neff[r] its effective sample size (either may be NULL;
 * n_rigs entries each; a plain filter is one rig). */
int scphd_get_rig_resample(scphd_t* f, int32_t* resampled, double* neff);

/* Grow the device buffers now to hold n_gaussians map entries over all
 * parents (and as many survivors) and max_meas measurements per rig, so that
 * steps up to those sizes allocate nothing; without it the buffers grow
 * geometrically as the maps grow, and the steps that grow them carry the
 * allocation time.  No counterpart in the reference. */
int scphd_reserve(scphd_t* f, int64_t n_gaussians, int32_t max_meas);

/* Device memory currently held by the handle, in bytes (buffers grow
 * geometrically and are kept between steps).  No counterpart in the
 * reference. */
int scphd_get_memory(scphd_t* f, int64_t* bytes);

/* Read a measurement sequence from a MAT-file (level 5): variable varname
 * is a cell array whose cell k is a 2 x n_k double matrix (rows u, v).
 * First call with counts == NULL: sets *n_steps and *n_total.  Second call
 * with *n_steps / *n_total holding the capacities of counts[n_steps] and
 * u, v[n_total]: fills them (frame k's measurements follow frame k-1's).
 * SCPHD_ERR_CONFIG for a missing variable or a malformed file.  Replaces
 * readMatfile, src/main.cpp:19-49. */
int scphd_mat_read_measurements(const char* path, const char* varname, int32_t* n_steps, int64_t* n_total,
                                int32_t* counts, double* u, double* v);

/* Write the parents and the highest-weight parent's map to a MAT-file
 * (level 5): particles (12 x N_p), weights (1 x N_p), map_weights (1 x G),
 * map_particles (6 x G*S), map_particle_weights (1 x G*S).  Replaces
 * writeMat, src/scphdcameracalibration.cu:2200-2373 (which writes MAT 7.3
 * through HDF5). */
int scphd_write_mat(scphd_t* f, const char* path);

/* ======================================================== batched / served
 * Many rigs (independent calibration instances, each with its own
 * parameters, parent count and measurement stream) share one GPU.  Each
 * admitted rig keeps its filter state resident in device memory; frames are
 * queued per rig and processed in order by a pool of host workers, each rig
 * on its own HIP stream, so that the device work of different rigs
 * overlaps.  A rig is admitted only while the sum of the reservations of the
 * admitted rigs stays within the HBM budget; a refused rig can be offered
 * again once others leave.  A rig's results are the same whether it runs
 * alone or with others.  No counterpart in the reference, which runs one
 * rig per process (src/main.cpp:67-96, batch_run.sh:15). */
typedef struct scphd_batch scphd_batch_t;

typedef struct scphd_batch_opts {
    int64_t hbm_budget_bytes;  /* admission budget; 0: 90 % of the free device memory */
    int32_t max_rigs;          /* rig slots; 0: 1024                               */
    int32_t n_workers;         /* frames in flight on the device; 0: 8              */
} scphd_batch_opts;

typedef struct scphd_frame_result {
    int32_t rig;
    int32_t status;            /* SCPHD_OK or the error of the failing call       */
    int64_t frame_id;          /* as submitted                                     */
    double t_submit;           /* seconds on the batch clock (scphd_batch_now)     */
    double t_start;
    double t_end;
    int64_t n_meas;
    int64_t n_pairs;           /* component updates of the frame (N_g * M)         */
    int64_t n_gaussians;       /* Gaussians after the frame                        */
    int32_t resampled;
    int32_t reserved0;
    double estimate[12];       /* weighted parent mean after the frame             */
    double cardinality;
} scphd_frame_result;

typedef struct scphd_batch_stats {
    int64_t rigs;              /* admitted rigs                                    */
    int64_t reserved_bytes;    /* sum of their reservations                        */
    int64_t budget_bytes;
    int64_t used_bytes;        /* device memory held by their handles              */
    int64_t queued_frames;     /* submitted, not yet finished                      */
    int64_t frames_done;       /* since creation                                   */
    int64_t pairs_done;        /* component updates since creation                 */
    int64_t refused;           /* admissions refused for lack of budget            */
} scphd_batch_stats;

int scphd_batch_opts_init(scphd_batch_opts* o);
int scphd_batch_create(const scphd_batch_opts* o, scphd_batch_t** out);
/* Waits for queued frames, then frees every rig. */
void scphd_batch_destroy(scphd_batch_t* b);
const char* scphd_batch_last_error(const scphd_batch_t* b);
/* Seconds since the batch was created (the clock of scphd_frame_result). */
double scphd_batch_now(const scphd_batch_t* b);

/* Device memory a rig needs with about gaussians_per_parent Gaussians per
 * parent (state, both sample generations and per-step work arrays, with the
 * growth slack of the buffers). */
int64_t scphd_rig_footprint(const scphd_params* p, int32_t gaussians_per_parent);

/* Admit a rig with reservation reserve_bytes (0: scphd_rig_footprint with
 * max_gaussians, or 1000 when that is 0).  SCPHD_ERR_CAPACITY when the
 * budget or the slots are exhausted. */
int scphd_batch_add_rig(scphd_batch_t* b, const scphd_params* p, int64_t reserve_bytes, int32_t* rig);
/* Wait for the rig's frames in flight, drop its queued frames, free it. */
int scphd_batch_remove_rig(scphd_batch_t* b, int32_t rig);

/* Write <dir>/<frame_id>.mat (scphd_write_mat) after each of the rig's
 * frames, as the reference main writes k.mat (src/main.cpp:91-93); NULL or
 * "" stops it.  Only while none of the rig's frames is queued or running. */
int scphd_batch_set_output(scphd_batch_t* b, int32_t rig, const char* dir);

/* Queue one frame of a rig: predict(dt), update(u, v, fixed_camera),
 * writeMat (when an output directory is set), checkStuff, resample -- one
 * iteration of src/main.cpp:78-97.  u and v are copied; the call returns at
 * once. */
int scphd_batch_submit(scphd_batch_t* b, int32_t rig, int64_t frame_id, double dt, const double* u,
                       const double* v, int32_t n_meas, int32_t fixed_camera);
/* Block until every submitted frame has finished; *frames_done (may be NULL)
 * receives the number finished since the previous call. */
int scphd_batch_step(scphd_batch_t* b, int32_t* frames_done);
/* Move up to cap finished-frame results into out (oldest first). */
int scphd_batch_collect(scphd_batch_t* b, scphd_frame_result* out, int32_t cap, int32_t* n_out);
int scphd_batch_get_stats(scphd_batch_t* b, scphd_batch_stats* s);
/* The rig's filter handle, for read-back (scphd_get_parents, scphd_get_maps,
 * ...) while none of its frames is queued or running; NULL for an unknown rig. */
scphd_t* scphd_batch_filter(scphd_batch_t* b, int32_t rig);

#ifdef __cplusplus
}
#endif

#endif /* SCPHD_H */
