// pm_serve.cpp -- serving form: request loop with continuous batching.
//
// One scheduler thread owns the device.  Each loop turn:
//   1. builds a step over every admitted job whose images are ready (a job
//      takes 1 + 2*iterations steps: start, then one half-pass per step,
//      writing its outputs in its last step) and launches it;
//   2. while that step runs, copies the outputs of jobs that finished in the
//      previous step back to their host buffers and releases their memory;
//   3. admits queued jobs (FIFO) into free slots whose footprint fits the
//      arena, uploading and converting their images on a second stream;
//   4. waits for the step.
// Device memory comes from one arena reserved at creation (the HBM budget)
// and is sub-allocated first-fit, so the loop never calls hipMalloc/hipFree.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

#include "pm_batch.h"
#include "pm_runtime.h"

using namespace hp3d;

namespace {

using Clock = std::chrono::steady_clock;

size_t a256(size_t n) { return (n + 255) & ~(size_t)255; }

// First-fit sub-allocator over [0, size).
class Arena {
public:
    void reset(size_t size) { free_.clear(); free_[0] = size; size_ = size; used_ = 0; peak_ = 0; }
    bool alloc(size_t bytes, size_t* off)
    {
        bytes = a256(bytes);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            *off = it->first;
            const size_t rest = it->second - bytes, start = it->first + bytes;
            free_.erase(it);
            if (rest) free_[start] = rest;
            used_ += bytes;
            peak_ = std::max(peak_, used_);
            return true;
        }
        return false;
    }
    void release(size_t off, size_t bytes)
    {
        bytes = a256(bytes);
        used_ -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    size_t size() const { return size_; }
    size_t peak() const { return peak_; }

private:
    std::map<size_t, size_t> free_;
    size_t size_ = 0, used_ = 0, peak_ = 0;
};

struct SJob {
    uint64_t id = 0;
    hp3d_job_desc d{};
    std::vector<uint8_t> host_img;       // (1+S) images, W*H each
    hp3d_job_status st{};
    size_t footprint = 0, off = 0;       // footprint: bytes the job needs once admitted
    bool holds_block = false;            // [off, off + footprint) is this job's arena block
    int slot = -1;
    int step = 0;                        // next step to run
    bool ready = false, finished = false, cancel = false;
    hipEvent_t ready_ev = nullptr;
    Image imgs[1 + kMaxSrc];
    DevJob dj{};
};

}  // namespace

struct hp3d_pm_server {
    hp3d_server_opts o{};
    int device = 0;
    hipStream_t compute = nullptr, upload = nullptr;
    char* arena_base = nullptr;
    Arena arena;
    DevJob* d_jobs = nullptr;            // slot table
    StepItem* d_items = nullptr;
    DevJob* h_jobs = nullptr;            // pinned staging
    StepItem* h_items = nullptr;
    hipEvent_t ev_a = nullptr, ev_b = nullptr;
    std::vector<int> free_slots;

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<SJob>> queue;
    std::map<uint64_t, std::shared_ptr<SJob>> jobs;   // submitted and not yet forgotten
    std::vector<std::shared_ptr<SJob>> active;
    uint64_t next_id = 0;
    bool stop = false;
    std::thread thr;
    hp3d_server_stats stats{};
    Clock::time_point t0;
    std::string last_error;
    int variant = 0;

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
    size_t footprint(const hp3d_job_desc& d) const;
    void layout(SJob& j);
    void loop();
    void finish_job(const std::shared_ptr<SJob>& j, int state, int error);
};

size_t hp3d_pm_server::footprint(const hp3d_job_desc& d) const
{
    int pitch;
    size_t plain_bytes, coef_bytes;
    image_plane_sizes(d.width, d.height, o.store, &pitch, &plain_bytes, &coef_bytes);
    const size_t n = 1 + d.num_src;
    const size_t mask = 1 | (d.out_normal ? 2 : 0) | (d.out_cost ? 4 : 0);
    return a256(n * (size_t)d.width * d.height) + n * (a256(plain_bytes) + a256(coef_bytes)) +
           a256(job_state_bytes(d.width, d.height, (int)mask));
}

// Place the job's staging bytes, image planes and state inside its arena block.
void hp3d_pm_server::layout(SJob& j)
{
    const hp3d_job_desc& d = j.d;
    char* p = arena_base + j.off;
    p += a256((size_t)(1 + d.num_src) * d.width * d.height);          // staging first
    for (int k = 0; k <= d.num_src; ++k) {
        Image& im = j.imgs[k];
        im.W = d.width;
        im.H = d.height;
        im.store = o.store;
        image_plane_sizes(d.width, d.height, o.store, &im.pitch_plain, &im.plain_bytes,
                          &im.coef_bytes);
        im.den_scale = 255.0f * 255.0f;
        im.plain = p;
        p += a256(im.plain_bytes);
        im.coef = im.coef_bytes ? p : nullptr;
        p += a256(im.coef_bytes);
    }
    hp3d_view_desc vd{};
    vd.num_src = d.num_src;
    std::memcpy(vd.K_ref_inv, d.K_ref_inv, sizeof vd.K_ref_inv);
    std::memcpy(vd.A, d.A, sizeof vd.A);
    std::memcpy(vd.b, d.b, sizeof vd.b);
    vd.depth_min = d.depth_min;
    vd.depth_max = d.depth_max;
    vd.radius = o.radius;
    vd.stride = o.stride;
    vd.iterations = d.iterations;
    vd.top_k = d.top_k;
    vd.seed = d.seed;
    const Image* srcs[kMaxSrc];
    for (int k = 0; k < d.num_src; ++k) srcs[k] = &j.imgs[1 + k];
    const int mask = 1 | (d.out_normal ? 2 : 0) | (d.out_cost ? 4 : 0);
    bind_job(j.dj, &vd, &j.imgs[0], srcs, p, mask);
}

void hp3d_pm_server::finish_job(const std::shared_ptr<SJob>& j, int state, int error)
{
    // caller holds mu
    j->st.state = state;
    j->st.error = error;
    j->st.t_done = now();
    if (j->slot >= 0) {
        free_slots.push_back(j->slot);
        j->slot = -1;
    }
    // only an admitted job holds an arena block: a job cancelled while
    // queued has a footprint but no block, and must release nothing
    if (j->holds_block) {
        arena.release(j->off, j->footprint);
        j->holds_block = false;
    }
    if (j->ready_ev) {
        (void)hipEventDestroy(j->ready_ev);
        j->ready_ev = nullptr;
    }
    j->host_img.clear();
    j->host_img.shrink_to_fit();
    if (state == HP3D_JOB_DONE) {
        stats.done++;
        const double n = (double)((2 * o.radius) / o.stride + 1);
        stats.taps += (double)j->d.width * j->d.height *
                      (j->d.num_src * n * n + j->d.iterations * 6.0 * j->d.num_src * n * n);
        stats.mpix_iters += (double)j->d.width * j->d.height * j->d.iterations / 1e6;
    } else if (state == HP3D_JOB_FAILED) {
        stats.failed++;
    } else if (state == HP3D_JOB_CANCELLED) {
        stats.cancelled++;
    }
    cv_done.notify_all();
}

void hp3d_pm_server::loop()
{
    (void)hipSetDevice(device);
    std::vector<std::shared_ptr<SJob>> finished_prev;
    for (;;) {
        std::vector<std::shared_ptr<SJob>> stepping;
        int n_items = 0, total_tiles = 0;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_work.wait(lk, [&] {
                return stop || !queue.empty() || !active.empty() || !finished_prev.empty();
            });
            if (stop && finished_prev.empty()) break;
            // cancelled or injected failures leave before the step
            for (auto it = active.begin(); it != active.end();) {
                auto& j = *it;
                if (j->cancel || (j->d.fail_at_step > 0 && j->step >= j->d.fail_at_step)) {
                    finish_job(j, j->cancel ? HP3D_JOB_CANCELLED : HP3D_JOB_FAILED,
                               j->cancel ? HP3D_ERR_CANCELLED : HP3D_ERR_DEVICE);
                    it = active.erase(it);
                } else {
                    ++it;
                }
            }
            // 1. step over every ready job
            for (auto& j : active) {
                if (!j->ready) j->ready = hipEventQuery(j->ready_ev) == hipSuccess;
                if (!j->ready) continue;
                StepItem it = make_item(j->slot, j->step, j->d.iterations, j->d.seed,
                                        j->d.width, j->d.height, variant);
                it.tile_begin = total_tiles;
                total_tiles += it.tiles;
                h_items[n_items++] = it;
                if (j->step == 0) h_jobs[j->slot] = j->dj;
                stepping.push_back(j);
            }
            stats.max_active = std::max(stats.max_active, n_items);
        }
        hipError_t e = hipSuccess;
        if (n_items) {
            for (auto& j : stepping)
                if (j->step == 0 && e == hipSuccess)
                    e = hipMemcpyAsync(d_jobs + j->slot, h_jobs + j->slot, sizeof(DevJob),
                                       hipMemcpyHostToDevice, compute);
            if (e == hipSuccess)
                e = hipMemcpyAsync(d_items, h_items, n_items * sizeof(StepItem),
                                   hipMemcpyHostToDevice, compute);
            if (e == hipSuccess) e = hipEventRecord(ev_a, compute);
            if (e == hipSuccess)
                e = launch_step(o.store, o.radius, o.stride, variant, d_jobs, d_items, n_items,
                                total_tiles, compute);
            if (e == hipSuccess) e = hipEventRecord(ev_b, compute);
        }
        // 2. outputs of the jobs that finished in the previous step
        for (auto& j : finished_prev) {
            const size_t px = (size_t)j->d.width * j->d.height;
            hipError_t ce = hipMemcpy(j->d.out_depth, j->dj.out_depth, px * 4, hipMemcpyDeviceToHost);
            if (ce == hipSuccess && j->d.out_normal)
                ce = hipMemcpy(j->d.out_normal, j->dj.out_normal, 3 * px * 4, hipMemcpyDeviceToHost);
            if (ce == hipSuccess && j->d.out_cost)
                ce = hipMemcpy(j->d.out_cost, j->dj.out_cost, px * 4, hipMemcpyDeviceToHost);
            std::lock_guard<std::mutex> lk(mu);
            finish_job(j, ce == hipSuccess ? HP3D_JOB_DONE : HP3D_JOB_FAILED,
                       ce == hipSuccess ? HP3D_OK : HP3D_ERR_DEVICE);
        }
        finished_prev.clear();
        // 3. admissions (FIFO: the head waits until it fits)
        {
            std::lock_guard<std::mutex> lk(mu);
            while (!queue.empty() && !free_slots.empty()) {
                auto j = queue.front();
                if (j->cancel) {
                    queue.pop_front();
                    finish_job(j, HP3D_JOB_CANCELLED, HP3D_ERR_CANCELLED);
                    continue;
                }
                size_t off = 0;
                if (!arena.alloc(j->footprint, &off)) break;
                queue.pop_front();
                j->off = off;
                j->holds_block = true;
                j->slot = free_slots.back();
                free_slots.pop_back();
                layout(*j);
                const int W = j->d.width, H = j->d.height;
                char* staging = arena_base + off;
                hipError_t ue = hipMemcpyAsync(staging, j->host_img.data(), j->host_img.size(),
                                               hipMemcpyHostToDevice, upload);
                for (int k = 0; k <= j->d.num_src && ue == hipSuccess; ++k)
                    ue = launch_img_prep((const uint8_t*)staging + (size_t)k * W * H, nullptr,
                                         (size_t)W, W, H, o.store, j->imgs[k].plain,
                                         j->imgs[k].coef, j->imgs[k].pitch_plain, upload);
                if (ue == hipSuccess) ue = hipEventCreateWithFlags(&j->ready_ev, hipEventDisableTiming);
                if (ue == hipSuccess) ue = hipEventRecord(j->ready_ev, upload);
                j->st.state = HP3D_JOB_RUNNING;
                j->st.t_admit = now();
                if (ue != hipSuccess) {
                    finish_job(j, HP3D_JOB_FAILED, HP3D_ERR_DEVICE);
                    continue;
                }
                active.push_back(j);
            }
        }
        // 4. wait for the step, then advance the jobs that took part
        if (n_items) {
            if (e == hipSuccess) e = hipEventSynchronize(ev_b);
            float ms = 0.f;
            if (e == hipSuccess && hipEventElapsedTime(&ms, ev_a, ev_b) == hipSuccess)
                stats.busy_ms += ms;
            std::lock_guard<std::mutex> lk(mu);
            stats.steps++;
            for (auto& j : stepping) {
                if (e != hipSuccess) {
                    last_error = hipGetErrorString(e);
                    finish_job(j, HP3D_JOB_FAILED, HP3D_ERR_DEVICE);
                    active.erase(std::find(active.begin(), active.end(), j));
                    continue;
                }
                j->st.steps++;
                if (j->step++ == 2 * j->d.iterations) {
                    active.erase(std::find(active.begin(), active.end(), j));
                    finished_prev.push_back(j);
                }
            }
        } else if (!active.empty()) {
            // only jobs whose images are still uploading: let the upload finish
            (void)hipStreamSynchronize(upload);
        }
    }
}

extern "C" {

int hp3d_server_create(const hp3d_server_opts* opts, hp3d_pm_server** out)
{
    if (!opts || !out) return HP3D_ERR_INVALID_ARG;
    *out = nullptr;
    if (opts->max_jobs < 1 || opts->max_jobs > 4096 || !window_supported(opts->radius, opts->stride) ||
        (opts->store != HP3D_STORE_F32 && opts->store != HP3D_STORE_F16) || opts->hbm_budget == 0)
        return HP3D_ERR_INVALID_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || opts->device < 0 || opts->device >= ndev)
        return HP3D_ERR_DEVICE;
    if (hipSetDevice(opts->device) != hipSuccess) return HP3D_ERR_DEVICE;
    auto s = std::make_unique<hp3d_pm_server>();
    s->o = *opts;
    s->device = opts->device;
    s->variant = !step_variant_tunable(opts->radius, opts->stride) ? 0
                 : opts->variant >= 0                              ? opts->variant
                                                                   : kDefaultStepVariant;
    if (s->variant >= kNumStepVariants) return HP3D_ERR_INVALID_ARG;
    bool ok = hipStreamCreateWithFlags(&s->compute, hipStreamNonBlocking) == hipSuccess &&
              hipStreamCreateWithFlags(&s->upload, hipStreamNonBlocking) == hipSuccess &&
              hipMalloc((void**)&s->arena_base, opts->hbm_budget) == hipSuccess &&
              hipMalloc((void**)&s->d_jobs, opts->max_jobs * sizeof(DevJob)) == hipSuccess &&
              hipMalloc((void**)&s->d_items, opts->max_jobs * sizeof(StepItem)) == hipSuccess &&
              hipHostMalloc((void**)&s->h_jobs, opts->max_jobs * sizeof(DevJob)) == hipSuccess &&
              hipHostMalloc((void**)&s->h_items, opts->max_jobs * sizeof(StepItem)) == hipSuccess &&
              hipEventCreate(&s->ev_a) == hipSuccess && hipEventCreate(&s->ev_b) == hipSuccess;
    if (!ok) {
        hp3d_server_destroy(s.release());
        return HP3D_ERR_DEVICE;
    }
    s->arena.reset(opts->hbm_budget);
    for (int k = opts->max_jobs - 1; k >= 0; --k) s->free_slots.push_back(k);
    s->t0 = Clock::now();
    s->thr = std::thread([p = s.get()] { p->loop(); });
    *out = s.release();
    return HP3D_OK;
}

int hp3d_server_submit(hp3d_pm_server* s, const hp3d_job_desc* d, uint64_t* id)
{
    if (!s || !d || !id) return HP3D_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    const int need = 2 * s->o.radius + 1;
    if (d->num_src < 1 || d->num_src > kMaxSrc || d->width < need || d->height < need ||
        !d->ref || !d->out_depth || d->iterations < 0 || d->iterations > kMaxIterations ||
        d->top_k < 1 || d->top_k > kMaxTopK || !(d->depth_min < d->depth_max) ||
        !std::isfinite(d->depth_min) || !std::isfinite(d->depth_max)) {
        s->last_error = "submit: invalid job description";
        s->stats.rejected++;
        return HP3D_ERR_INVALID_ARG;
    }
    for (int k = 0; k < d->num_src; ++k)
        if (!d->src[k]) {
            s->last_error = "submit: missing source image";
            s->stats.rejected++;
            return HP3D_ERR_INVALID_ARG;
        }
    auto j = std::make_shared<SJob>();
    j->d = *d;
    j->footprint = s->footprint(*d);
    // taps address an image plane with 32-bit buffer offsets and 24-bit row
    // offsets
    if (!image_dims_supported(d->width, d->height, s->o.store)) {
        s->last_error = "submit: image plane above 2 GiB - 64 KiB or row pitch of 2^24 elements";
        s->stats.rejected++;
        return HP3D_ERR_UNSUPPORTED;
    }
    if (j->footprint > s->arena.size()) {
        s->last_error = "submit: job footprint exceeds the server's HBM budget";
        s->stats.rejected++;
        return HP3D_ERR_OOM_BUDGET;
    }
    const size_t px = (size_t)d->width * d->height;
    j->host_img.resize((1 + d->num_src) * px);
    std::memcpy(j->host_img.data(), d->ref, px);
    for (int k = 0; k < d->num_src; ++k) std::memcpy(j->host_img.data() + (1 + k) * px, d->src[k], px);
    j->id = ++s->next_id;
    j->st.state = HP3D_JOB_QUEUED;
    j->st.t_submit = s->now();
    s->jobs[j->id] = j;
    s->queue.push_back(j);
    s->stats.submitted++;
    *id = j->id;
    s->cv_work.notify_all();
    return HP3D_OK;
}

int hp3d_server_poll(hp3d_pm_server* s, uint64_t id, hp3d_job_status* st)
{
    if (!s || !st) return HP3D_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return HP3D_ERR_NOT_FOUND;
    *st = it->second->st;
    return HP3D_OK;
}

int hp3d_server_wait(hp3d_pm_server* s, uint64_t id, double timeout_s, hp3d_job_status* st)
{
    if (!s) return HP3D_ERR_INVALID_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return HP3D_ERR_NOT_FOUND;
    auto j = it->second;
    auto done = [&] { return j->st.state >= HP3D_JOB_DONE; };
    if (timeout_s < 0) s->cv_done.wait(lk, done);
    else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), done))
        return HP3D_ERR_AGAIN;
    if (st) *st = j->st;
    return HP3D_OK;
}

int hp3d_server_cancel(hp3d_pm_server* s, uint64_t id)
{
    if (!s) return HP3D_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return HP3D_ERR_NOT_FOUND;
    if (it->second->st.state >= HP3D_JOB_DONE) return HP3D_OK;
    it->second->cancel = true;
    s->cv_work.notify_all();
    return HP3D_OK;
}

int hp3d_server_forget(hp3d_pm_server* s, uint64_t id)
{
    if (!s) return HP3D_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return HP3D_ERR_NOT_FOUND;
    // a queued or running job may still write the caller's output buffers
    if (it->second->st.state < HP3D_JOB_DONE) return HP3D_ERR_AGAIN;
    s->jobs.erase(it);
    return HP3D_OK;
}

int hp3d_server_get_stats(hp3d_pm_server* s, hp3d_server_stats* st)
{
    if (!s || !st) return HP3D_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    *st = s->stats;
    st->wall_s = s->now();
    st->arena_bytes = s->arena.size();
    st->arena_peak = s->arena.peak();
    return HP3D_OK;
}

const char* hp3d_server_last_error(hp3d_pm_server* s)
{
    return s ? s->last_error.c_str() : "NULL server";
}

int hp3d_server_destroy(hp3d_pm_server* s)
{
    if (!s) return HP3D_ERR_INVALID_ARG;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        for (auto& j : s->queue) j->cancel = true;
        for (auto& j : s->active) j->cancel = true;
        s->cv_work.notify_all();
    }
    if (s->thr.joinable()) s->thr.join();
    (void)hipSetDevice(s->device);
    if (s->compute) (void)hipStreamSynchronize(s->compute);
    if (s->upload) (void)hipStreamSynchronize(s->upload);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        for (auto& j : s->queue) s->finish_job(j, HP3D_JOB_CANCELLED, HP3D_ERR_CANCELLED);
        for (auto& j : s->active) s->finish_job(j, HP3D_JOB_CANCELLED, HP3D_ERR_CANCELLED);
        s->queue.clear();
        s->active.clear();
    }
    if (s->arena_base) (void)hipFree(s->arena_base);
    if (s->d_jobs) (void)hipFree(s->d_jobs);
    if (s->d_items) (void)hipFree(s->d_items);
    if (s->h_jobs) (void)hipHostFree(s->h_jobs);
    if (s->h_items) (void)hipHostFree(s->h_items);
    if (s->ev_a) (void)hipEventDestroy(s->ev_a);
    if (s->ev_b) (void)hipEventDestroy(s->ev_b);
    if (s->compute) (void)hipStreamDestroy(s->compute);
    if (s->upload) (void)hipStreamDestroy(s->upload);
    delete s;
    return HP3D_OK;
}

}  // extern "C"
