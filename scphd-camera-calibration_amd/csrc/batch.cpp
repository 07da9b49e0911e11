/*
 * batch.cpp -- batched and served operation of many rigs on one GPU
 * (scphd_batch_* in include/scphd.h).
 *
 * Each admitted rig owns a filter handle (scphd_t): its parents, maps and
 * sample generations stay resident in device memory for the rig's lifetime,
 * and its kernels run on the handle's own HIP stream.  Submitted frames wait
 * in a per-rig FIFO.  A pool of worker threads takes frames, oldest
 * submission first, from rigs that have no frame in flight, and runs one
 * iteration of the reference main loop (src/main.cpp:78-97: predict,
 * update, writeMat when an output directory is set, checkStuff,
 * resample) plus the calibration estimate.  One frame
 * per rig is in flight at a time, so a rig's frames keep their order, while
 * the frames of different rigs overlap on the device (up to n_workers at
 * once).  A rig's computation never depends on the other rigs: its noise is
 * keyed by its own seed and step counters, so its results are the same
 * alone and in any mix.
 *
 * Admission: a rig is admitted when its reservation fits in the HBM budget
 * left by the reservations of the admitted rigs; otherwise the call returns
 * SCPHD_ERR_CAPACITY and the caller offers it again later.
 */
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "scphd.h"

namespace {

struct Frame {
    int64_t id;
    double dt;
    std::vector<double> u, v;
    int32_t fixed;
    double t_submit;
};

struct Rig {
    int32_t id = -1;
    scphd_t* f = nullptr;
    int64_t reserve = 0;
    std::deque<Frame> q;
    bool busy = false;     /* a frame is running (or the handle is being created) */
    bool removing = false; /* scphd_batch_remove_rig is waiting for it */
    int64_t mem = 0;       /* device bytes held, refreshed after each frame */
    std::string outdir;    /* when set: <outdir>/<frame_id>.mat after each frame */
};

using Clock = std::chrono::steady_clock;

}  // namespace

struct scphd_batch {
    scphd_batch_opts opts{};
    Clock::time_point t0;
    std::mutex mu;
    std::condition_variable cv_work; /* a frame became runnable, or stop */
    std::condition_variable cv_idle; /* a frame finished */
    std::vector<std::unique_ptr<Rig>> slots;
    std::vector<std::thread> workers;
    std::deque<scphd_frame_result> done;
    bool stop = false;
    int64_t reserved = 0;
    int64_t queued = 0; /* submitted and not finished */
    int64_t frames_done = 0, pairs_done = 0, refused = 0, done_since_step = 0;
    std::string err;

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

/* Run one frame on the rig's handle; fills r. */
void run_frame(Rig* rig, const Frame& fr, scphd_frame_result& r) {
    scphd_t* f = rig->f;
    int rc = scphd_predict(f, fr.dt);
    if (rc == SCPHD_OK)
        rc = scphd_update(f, fr.u.data(), fr.v.data(), (int32_t)fr.u.size(), fr.fixed);
    scphd_update_stats st{};
    if (rc == SCPHD_OK) rc = scphd_get_update_stats(f, &st);
    int32_t resampled = 0;
    if (rc == SCPHD_OK && !rig->outdir.empty())
        rc = scphd_write_mat(f, (rig->outdir + "/" + std::to_string(fr.id) + ".mat").c_str());
    if (rc == SCPHD_OK) rc = scphd_check(f);
    if (rc == SCPHD_OK) rc = scphd_resample(f, &resampled);
    if (rc == SCPHD_OK) rc = scphd_get_estimate(f, r.estimate, nullptr, &r.cardinality);
    scphd_counts c{};
    if (rc == SCPHD_OK) rc = scphd_get_counts(f, &c);
    r.status = rc;
    r.n_meas = (int64_t)fr.u.size();
    r.n_pairs = rc == SCPHD_OK ? st.n_pairs : 0;
    r.n_gaussians = c.n_gaussians;
    r.resampled = resampled;
}

/* The runnable rig whose next frame was submitted first (ties: lower id). */
Rig* pick(scphd_batch* b) {
    Rig* best = nullptr;
    for (auto& s : b->slots) {
        Rig* r = s.get();
        if (!r || r->busy || r->removing || r->q.empty()) continue;
        if (!best || r->q.front().t_submit < best->q.front().t_submit) best = r;
    }
    return best;
}

void worker(scphd_batch* b) {
    std::unique_lock<std::mutex> lk(b->mu);
    for (;;) {
        Rig* rig = nullptr;
        b->cv_work.wait(lk, [&] { return b->stop || (rig = pick(b)) != nullptr; });
        if (!rig) return; /* stop and nothing runnable */
        Frame fr = std::move(rig->q.front());
        rig->q.pop_front();
        rig->busy = true;
        lk.unlock();
        scphd_frame_result r{};
        r.rig = rig->id;
        r.frame_id = fr.id;
        r.t_submit = fr.t_submit;
        r.t_start = b->now();
        run_frame(rig, fr, r);
        r.t_end = b->now();
        int64_t mem = 0;
        (void)scphd_get_memory(rig->f, &mem);
        lk.lock();
        rig->busy = false;
        rig->mem = mem;
        b->done.push_back(r);
        b->queued--;
        b->frames_done++;
        b->done_since_step++;
        b->pairs_done += r.n_pairs;
        b->cv_idle.notify_all();
        b->cv_work.notify_one(); /* the rig may have more frames */
    }
}

Rig* find(scphd_batch* b, int32_t id) {
    if (id < 0 || id >= (int32_t)b->slots.size()) return nullptr;
    return b->slots[id].get();
}

}  // namespace

extern "C" {

int scphd_batch_opts_init(scphd_batch_opts* o) {
    if (!o) return SCPHD_ERR_ARG;
    std::memset(o, 0, sizeof *o);
    o->max_rigs = 1024;
    o->n_workers = 8;
    return SCPHD_OK;
}

int scphd_batch_create(const scphd_batch_opts* o, scphd_batch_t** out) {
    if (!out) return SCPHD_ERR_ARG;
    *out = nullptr;
    scphd_batch_opts opts;
    scphd_batch_opts_init(&opts);
    if (o) {
        if (o->hbm_budget_bytes < 0 || o->max_rigs < 0 || o->n_workers < 0 || o->max_rigs > (1 << 20) ||
            o->n_workers > 64)
            return SCPHD_ERR_ARG;
        opts.hbm_budget_bytes = o->hbm_budget_bytes;
        if (o->max_rigs) opts.max_rigs = o->max_rigs;
        if (o->n_workers) opts.n_workers = o->n_workers;
    }
    if (opts.hbm_budget_bytes == 0) {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return SCPHD_ERR_NODEVICE;
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) return SCPHD_ERR_HIP;
        opts.hbm_budget_bytes = (int64_t)(0.9 * (double)fr);
    }
    auto* b = new scphd_batch();
    b->opts = opts;
    b->t0 = Clock::now();
    b->slots.resize(opts.max_rigs);
    for (int i = 0; i < opts.n_workers; ++i) b->workers.emplace_back(worker, b);
    *out = b;
    return SCPHD_OK;
}

void scphd_batch_destroy(scphd_batch_t* b) {
    if (!b) return;
    {
        std::unique_lock<std::mutex> lk(b->mu);
        b->cv_idle.wait(lk, [&] { return b->queued == 0; });
        b->stop = true;
    }
    b->cv_work.notify_all();
    for (auto& t : b->workers) t.join();
    for (auto& s : b->slots)
        if (s) scphd_destroy(s->f);
    delete b;
}

const char* scphd_batch_last_error(const scphd_batch_t* b) { return b ? b->err.c_str() : "null batch"; }

double scphd_batch_now(const scphd_batch_t* b) { return b ? b->now() : 0.0; }

int64_t scphd_rig_footprint(const scphd_params* p, int32_t gpp) {
    if (!p || p->n_particles <= 0 || p->particles_per_feature <= 0 || gpp < 0) return 0;
    const double esz = p->storage == SCPHD_STORAGE_FP64 ? 8.0 : 4.0;
    const double Np = p->n_particles, S = p->particles_per_feature, G = std::max(gpp, 1);
    const double slack = 1.25; /* geometric growth of every buffer */
    /* two sample generations: (non-detection + posterior) blocks of the next
     * map, and the current map */
    const double samples = 2.0 * Np * G * S * 7.0 * esz;
    /* per Gaussian: fit 28 + detection terms 39 + likelihood record 8 FP64,
     * two tables, flags */
    const double per_gauss = Np * G * (75.0 * 8.0 + 2.0 * 16.0 + 4.0 * 4.0);
    /* survivors (at most about one per Gaussian in steady state): lists,
     * merged means and covariances, merge groups, the global-scratch merge
     * workspace (240 B a slot) */
    const double per_surv = Np * G * (64.0 + 28.0 * 8.0 + 16.0 + 240.0);
    const double parents = Np * 256.0;
    return (int64_t)(slack * (samples + per_gauss + per_surv + parents)) + (64LL << 20);
}

int scphd_batch_add_rig(scphd_batch_t* b, const scphd_params* p, int64_t reserve, int32_t* rig) {
    if (!b || !p || !rig || reserve < 0) return SCPHD_ERR_ARG;
    *rig = -1;
    if (reserve == 0) reserve = scphd_rig_footprint(p, p->max_gaussians > 0 ? p->max_gaussians : 1000);
    int32_t slot = -1;
    {
        std::lock_guard<std::mutex> lk(b->mu);
        if (b->reserved + reserve > b->opts.hbm_budget_bytes) {
            b->refused++;
            b->err = "rig reservation exceeds the remaining HBM budget";
            return SCPHD_ERR_CAPACITY;
        }
        for (int32_t i = 0; i < (int32_t)b->slots.size(); ++i)
            if (!b->slots[i]) {
                slot = i;
                break;
            }
        if (slot < 0) {
            b->refused++;
            b->err = "no free rig slot";
            return SCPHD_ERR_CAPACITY;
        }
        /* hold the slot and the reservation while the handle is created */
        b->slots[slot].reset(new Rig());
        b->slots[slot]->id = slot;
        b->slots[slot]->reserve = reserve;
        b->slots[slot]->busy = true;
        b->reserved += reserve;
    }
    scphd_t* f = nullptr;
    const int rc = scphd_create(p, &f);
    std::lock_guard<std::mutex> lk(b->mu);
    if (rc != SCPHD_OK) {
        b->reserved -= reserve;
        b->slots[slot].reset();
        b->err = "scphd_create failed";
        return rc;
    }
    b->slots[slot]->f = f;
    (void)scphd_get_memory(f, &b->slots[slot]->mem);
    b->slots[slot]->busy = false;
    *rig = slot;
    return SCPHD_OK;
}

int scphd_batch_remove_rig(scphd_batch_t* b, int32_t id) {
    if (!b) return SCPHD_ERR_ARG;
    std::unique_ptr<Rig> gone;
    {
        std::unique_lock<std::mutex> lk(b->mu);
        Rig* r = find(b, id);
        if (!r || !r->f || r->removing) return SCPHD_ERR_ARG;
        r->removing = true;
        b->queued -= (int64_t)r->q.size();
        r->q.clear();
        b->cv_idle.wait(lk, [&] { return !r->busy; });
        b->reserved -= r->reserve;
        gone = std::move(b->slots[id]);
        b->cv_idle.notify_all();
    }
    scphd_destroy(gone->f);
    return SCPHD_OK;
}

int scphd_batch_set_output(scphd_batch_t* b, int32_t id, const char* dir) {
    if (!b) return SCPHD_ERR_ARG;
    std::lock_guard<std::mutex> lk(b->mu);
    Rig* r = find(b, id);
    if (!r || !r->f || r->removing || r->busy || !r->q.empty()) {
        b->err = "unknown rig, or frames of it are queued or running";
        return SCPHD_ERR_ARG;
    }
    r->outdir = dir ? dir : "";
    return SCPHD_OK;
}

int scphd_batch_submit(scphd_batch_t* b, int32_t id, int64_t frame_id, double dt, const double* u, const double* v,
                       int32_t n, int32_t fixed) {
    if (!b || n < 0 || (n > 0 && (!u || !v))) return SCPHD_ERR_ARG;
    Frame fr;
    fr.id = frame_id;
    fr.dt = dt;
    fr.u.assign(u, u + n);
    fr.v.assign(v, v + n);
    fr.fixed = fixed ? 1 : 0;
    {
        std::lock_guard<std::mutex> lk(b->mu);
        Rig* r = find(b, id);
        if (!r || !r->f || r->removing) {
            b->err = "unknown rig";
            return SCPHD_ERR_ARG;
        }
        fr.t_submit = b->now();
        r->q.push_back(std::move(fr));
        b->queued++;
    }
    b->cv_work.notify_one();
    return SCPHD_OK;
}

int scphd_batch_step(scphd_batch_t* b, int32_t* frames_done) {
    if (!b) return SCPHD_ERR_ARG;
    std::unique_lock<std::mutex> lk(b->mu);
    b->cv_idle.wait(lk, [&] { return b->queued == 0; });
    if (frames_done) *frames_done = (int32_t)b->done_since_step;
    b->done_since_step = 0;
    return SCPHD_OK;
}

int scphd_batch_collect(scphd_batch_t* b, scphd_frame_result* out, int32_t cap, int32_t* n_out) {
    if (!b || !n_out || cap < 0 || (cap > 0 && !out)) return SCPHD_ERR_ARG;
    std::lock_guard<std::mutex> lk(b->mu);
    int32_t n = 0;
    while (n < cap && !b->done.empty()) {
        out[n++] = b->done.front();
        b->done.pop_front();
    }
    *n_out = n;
    return SCPHD_OK;
}

int scphd_batch_get_stats(scphd_batch_t* b, scphd_batch_stats* s) {
    if (!b || !s) return SCPHD_ERR_ARG;
    std::lock_guard<std::mutex> lk(b->mu);
    std::memset(s, 0, sizeof *s);
    for (auto& r : b->slots)
        if (r && r->f) {
            s->rigs++;
            s->used_bytes += r->mem;
        }
    s->reserved_bytes = b->reserved;
    s->budget_bytes = b->opts.hbm_budget_bytes;
    s->queued_frames = b->queued;
    s->frames_done = b->frames_done;
    s->pairs_done = b->pairs_done;
    s->refused = b->refused;
    return SCPHD_OK;
}

scphd_t* scphd_batch_filter(scphd_batch_t* b, int32_t id) {
    if (!b) return nullptr;
    std::lock_guard<std::mutex> lk(b->mu);
    Rig* r = find(b, id);
    return r && !r->removing ? r->f : nullptr;
}

}  /* extern "C" */
