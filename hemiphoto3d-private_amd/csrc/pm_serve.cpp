// pm_serve.cpp -- serving form: request loop with continuous batching.
//
// Two host threads own the device work of a server:
//   * the scheduler admits queued jobs (FIFO) into free slots when their
//     footprint fits the arena, uploads their host images on the upload
//     stream, and launches one step at a time over every admitted job whose
//     images are ready (a job takes 1 + 2*iterations steps: start, then one
//     half-pass per step, writing its outputs in its last step).  Up to two
//     steps are queued on the compute stream, so the next step is built and
//     launched while the previous one runs; the scheduler blocks only on the
//     older of two queued steps.
//   * the completion thread waits for a job's last step, downloads its
//     outputs to the caller's buffers on a copy stream and releases the job.
// Resident images (hp3d_server_image_add) are uploaded once, with optional
// pyramid levels, and used by any number of jobs by handle.  All device
// memory -- job images, state, outputs and resident images -- comes from one
// arena reserved at creation (the HBM budget), sub-allocated first-fit, so
// the loop never calls hipMalloc/hipFree.
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

constexpr int kMaxLevels = 4;

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

// An image kept on the device for any number of jobs, with its pyramid.
struct ResImage {
    uint64_t id = 0;
    int levels = 0;
    size_t off = 0, bytes = 0;        // arena block of the planes
    Image lv[kMaxLevels];
    int refs = 0;                     // queued or running jobs using it
    bool removed = false;             // freed once refs drops to 0
};

struct SJob {
    uint64_t id = 0;
    hp3d_job_desc d{};
    std::vector<uint8_t> host_img;       // the images not resident, W*H each
    int n_host = 0;                      // how many
    hp3d_job_status st{};
    size_t footprint = 0, off = 0;       // footprint: bytes the job needs once admitted
    bool holds_block = false;            // [off, off + footprint) is this job's arena block
    int slot = -1;
    int step = 0;                        // next step to launch
    int inflight = 0;                    // launched steps not yet seen complete
    int completed = 0;                   // steps seen complete
    bool ready = false, cancel = false;
    bool dev_error = false;              // a step of the job failed on the device
    hipEvent_t ready_ev = nullptr;       // the job's host images are uploaded
    hipEvent_t last_ev = nullptr;        // recorded after the job's last step
    Image imgs[1 + kMaxSrc];
    std::shared_ptr<ResImage> res[1 + kMaxSrc];   // resident images used (null: uploaded)
    DevJob dj{};
};

// One launched step: the jobs in it and the buffers it uses.
struct Flight {
    std::vector<std::shared_ptr<SJob>> jobs;
};

}  // namespace

struct hp3d_pm_server {
    hp3d_server_opts o{};
    int device = 0;
    hipStream_t compute = nullptr, upload = nullptr, copy = nullptr, image = nullptr;
    char* arena_base = nullptr;
    Arena arena;
    DevJob* d_jobs = nullptr;            // slot table
    StepItem* d_items[2] = {nullptr, nullptr};
    DevJob* h_jobs = nullptr;            // pinned staging
    StepItem* h_items[2] = {nullptr, nullptr};
    hipEvent_t ev_a[2] = {nullptr, nullptr}, ev_b[2] = {nullptr, nullptr};
    Flight flight[2];
    std::vector<int> free_slots;

    std::mutex mu;
    std::condition_variable cv_work, cv_done, cv_comp;
    std::deque<std::shared_ptr<SJob>> queue;
    std::map<uint64_t, std::shared_ptr<SJob>> jobs;   // submitted and not yet forgotten
    std::vector<std::shared_ptr<SJob>> active;        // admitted, last step not yet launched
    std::deque<std::shared_ptr<SJob>> completing;     // last step launched
    std::map<uint64_t, std::shared_ptr<ResImage>> images;
    uint64_t next_id = 0, next_image = 0;
    bool stop = false, comp_stop = false;
    std::thread thr, comp_thr;
    hp3d_server_stats stats{};
    Clock::time_point t0;
    std::string last_error;
    int variant = 0;

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
    size_t footprint(const SJob& j) const;
    void layout(SJob& j);
    void loop();
    void complete_loop();
    void retire(int b);
    void finish_job(const std::shared_ptr<SJob>& j, int state, int error);
    void drop_image_ref(const std::shared_ptr<ResImage>& r);
};

// Bytes of a job's block: staging and planes of its non-resident images,
// state and staged outputs.
size_t hp3d_pm_server::footprint(const SJob& j) const
{
    const hp3d_job_desc& d = j.d;
    int pitch;
    size_t plain_bytes, coef_bytes;
    image_plane_sizes(d.width, d.height, o.store, &pitch, &plain_bytes, &coef_bytes);
    const size_t n = (size_t)j.n_host;
    const size_t mask = 1 | (d.out_normal ? 2 : 0) | (d.out_cost ? 4 : 0);
    return a256(n * (size_t)d.width * d.height) + n * (a256(plain_bytes) + a256(coef_bytes)) +
           a256(job_state_bytes(d.width, d.height, (int)mask));
}

// Place the job's staging bytes, non-resident image planes and state inside
// its arena block; resident images are bound to their level's planes.
void hp3d_pm_server::layout(SJob& j)
{
    const hp3d_job_desc& d = j.d;
    char* p = arena_base + j.off;
    p += a256((size_t)j.n_host * d.width * d.height);                // staging first
    for (int k = 0; k <= d.num_src; ++k) {
        Image& im = j.imgs[k];
        if (j.res[k]) {
            im = j.res[k]->lv[d.level];
            continue;
        }
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

void hp3d_pm_server::drop_image_ref(const std::shared_ptr<ResImage>& r)
{
    // caller holds mu
    if (--r->refs == 0 && r->removed) arena.release(r->off, r->bytes);
}

void hp3d_pm_server::finish_job(const std::shared_ptr<SJob>& j, int state, int error)
{
    // caller holds mu
    j->st.state = state;
    j->st.error = error;
    j->st.t_done = now();
    j->st.steps = state == HP3D_JOB_DONE ? 1 + 2 * j->d.iterations : j->completed;
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
    for (auto& r : j->res)
        if (r) {
            drop_image_ref(r);
            r.reset();
        }
    for (hipEvent_t* e : {&j->ready_ev, &j->last_ev})
        if (*e) {
            (void)hipEventDestroy(*e);
            *e = nullptr;
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

// Wait for the step in buffer b and account for it.
void hp3d_pm_server::retire(int b)
{
    const auto w0 = Clock::now();
    hipError_t e = hipEventSynchronize(ev_b[b]);
    const double waited = std::chrono::duration<double, std::milli>(Clock::now() - w0).count();
    float ms = 0.f;
    if (e == hipSuccess && hipEventElapsedTime(&ms, ev_a[b], ev_b[b]) != hipSuccess) ms = 0.f;
    std::lock_guard<std::mutex> lk(mu);
    stats.steps++;
    stats.busy_ms += ms;
    stats.wait_ms += waited;
    for (auto& j : flight[b].jobs) {
        j->inflight--;
        j->completed++;
        if (j->st.state < HP3D_JOB_DONE) j->st.steps = j->completed;
        if (e != hipSuccess && j->st.state < HP3D_JOB_DONE) {
            last_error = hipGetErrorString(e);
            j->dev_error = true;     // leaves at the next turn with inflight 0
        }
    }
    flight[b].jobs.clear();
    cv_work.notify_all();
}

void hp3d_pm_server::loop()
{
    (void)hipSetDevice(device);
    std::deque<int> order;      // step buffers in flight, oldest first
    int next_buf = 0;
    for (;;) {
        if (order.size() == 2) {
            retire(order.front());
            order.pop_front();
        }
        std::vector<std::shared_ptr<SJob>> stepping, fresh, last;
        int n_items = 0, total_tiles = 0;
        const int b = next_buf;
        bool uploading = false;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_work.wait(lk, [&] {
                return stop || !queue.empty() || !active.empty() || !order.empty();
            });
            if (stop && order.empty() && active.empty()) break;
            // cancelled, failed and (at shutdown) every job leave once no
            // step of theirs is in flight and its image uploads are done (its
            // arena block may go to a job with resident images only, whose
            // steps do not wait for the upload stream)
            for (auto it = active.begin(); it != active.end();) {
                auto& j = *it;
                const bool fail =
                    j->dev_error || (j->d.fail_at_step > 0 && j->step >= j->d.fail_at_step);
                const bool uploaded =
                    j->ready || !j->ready_ev || hipEventQuery(j->ready_ev) != hipErrorNotReady;
                if (j->inflight == 0 && uploaded && (j->cancel || fail || stop)) {
                    finish_job(j, fail ? HP3D_JOB_FAILED : HP3D_JOB_CANCELLED,
                               fail ? HP3D_ERR_DEVICE : HP3D_ERR_CANCELLED);
                    it = active.erase(it);
                } else {
                    ++it;
                }
            }
            // admissions (FIFO: the head waits until it fits)
            while (!stop && !queue.empty() && !free_slots.empty()) {
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
                j->st.state = HP3D_JOB_RUNNING;
                j->st.t_admit = now();
                hipError_t ue = hipEventCreateWithFlags(&j->last_ev, hipEventDisableTiming);
                if (ue == hipSuccess && j->n_host > 0) {
                    const int W = j->d.width, H = j->d.height;
                    char* staging = arena_base + off;
                    ue = hipMemcpyAsync(staging, j->host_img.data(), j->host_img.size(),
                                        hipMemcpyHostToDevice, upload);
                    int h = 0;
                    for (int k = 0; k <= j->d.num_src && ue == hipSuccess; ++k) {
                        if (j->res[k]) continue;
                        ue = launch_img_prep((const uint8_t*)staging + (size_t)h++ * W * H, nullptr,
                                             (size_t)W, W, H, o.store, j->imgs[k].plain,
                                             j->imgs[k].coef, j->imgs[k].pitch_plain, upload);
                    }
                    if (ue == hipSuccess)
                        ue = hipEventCreateWithFlags(&j->ready_ev, hipEventDisableTiming);
                    if (ue == hipSuccess) ue = hipEventRecord(j->ready_ev, upload);
                    stats.image_uploads += j->n_host;
                } else {
                    j->ready = true;
                }
                for (auto& r : j->res) stats.image_reuses += r ? 1 : 0;
                if (ue != hipSuccess) {
                    last_error = hipGetErrorString(ue);
                    // what was queued of its upload runs before the block is reused
                    (void)hipStreamSynchronize(upload);
                    finish_job(j, HP3D_JOB_FAILED, HP3D_ERR_DEVICE);
                    continue;
                }
                active.push_back(j);
            }
            // the next step: every ready job that has a step left
            for (auto it = active.begin(); it != active.end();) {
                auto& j = *it;
                if (!j->ready) {
                    j->ready = hipEventQuery(j->ready_ev) == hipSuccess;
                    if (!j->ready) {
                        uploading = true;
                        ++it;
                        continue;
                    }
                }
                if (stop || j->cancel || j->dev_error ||
                    (j->d.fail_at_step > 0 && j->step >= j->d.fail_at_step)) {
                    ++it;
                    continue;
                }
                StepItem si = make_item(j->slot, j->step, j->d.iterations, j->d.seed, j->d.width,
                                        j->d.height, variant);
                si.tile_begin = total_tiles;
                total_tiles += si.tiles;
                h_items[b][n_items++] = si;
                if (j->step == 0) {
                    h_jobs[j->slot] = j->dj;
                    fresh.push_back(j);
                }
                stepping.push_back(j);
                j->inflight++;
                if (j->step++ == 2 * j->d.iterations) {   // its last step
                    last.push_back(j);
                    it = active.erase(it);
                } else {
                    ++it;
                }
            }
            stats.max_active = std::max(stats.max_active, n_items);
        }
        if (n_items) {
            hipError_t e = hipSuccess;
            for (auto& j : fresh)
                if (e == hipSuccess)
                    e = hipMemcpyAsync(d_jobs + j->slot, h_jobs + j->slot, sizeof(DevJob),
                                       hipMemcpyHostToDevice, compute);
            if (e == hipSuccess)
                e = hipMemcpyAsync(d_items[b], h_items[b], n_items * sizeof(StepItem),
                                   hipMemcpyHostToDevice, compute);
            if (e == hipSuccess) e = hipEventRecord(ev_a[b], compute);
            if (e == hipSuccess)
                e = launch_step(o.store, o.radius, o.stride, variant, d_jobs, d_items[b], n_items,
                                total_tiles, compute);
            if (e == hipSuccess) e = hipEventRecord(ev_b[b], compute);
            // jobs whose last step this was go to the completion thread only
            // once the event it waits on is recorded
            for (auto& j : last)
                if (e == hipSuccess) e = hipEventRecord(j->last_ev, compute);
            {
                std::lock_guard<std::mutex> lk(mu);
                if (e != hipSuccess) {
                    last_error = hipGetErrorString(e);
                    for (auto& j : stepping) j->dev_error = true;
                    // the finishing jobs cannot complete: back to the active
                    // list, to leave as failed once their steps are retired
                    for (auto& j : last) active.push_back(j);
                } else {
                    for (auto& j : last) completing.push_back(j);
                }
                flight[b].jobs = std::move(stepping);
                cv_comp.notify_all();
            }
            order.push_back(b);
            next_buf ^= 1;
        } else if (!order.empty()) {
            retire(order.front());
            order.pop_front();
        } else if (uploading) {
            (void)hipStreamSynchronize(upload);
        }
    }
}

// Downloads the outputs of jobs whose last step was launched, in launch
// order, and completes them.
void hp3d_pm_server::complete_loop()
{
    (void)hipSetDevice(device);
    for (;;) {
        std::shared_ptr<SJob> j;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_comp.wait(lk, [&] { return comp_stop || !completing.empty(); });
            if (completing.empty()) break;
            j = completing.front();
            completing.pop_front();
        }
        hipError_t e = hipEventSynchronize(j->last_ev);
        const size_t px = (size_t)j->d.width * j->d.height;
        if (e == hipSuccess)
            e = hipMemcpyAsync(j->d.out_depth, j->dj.out_depth, px * 4, hipMemcpyDeviceToHost, copy);
        if (e == hipSuccess && j->d.out_normal)
            e = hipMemcpyAsync(j->d.out_normal, j->dj.out_normal, 3 * px * 4,
                               hipMemcpyDeviceToHost, copy);
        if (e == hipSuccess && j->d.out_cost)
            e = hipMemcpyAsync(j->d.out_cost, j->dj.out_cost, px * 4, hipMemcpyDeviceToHost, copy);
        if (e == hipSuccess) e = hipStreamSynchronize(copy);
        std::lock_guard<std::mutex> lk(mu);
        if (e != hipSuccess) last_error = hipGetErrorString(e);
        finish_job(j, e == hipSuccess ? HP3D_JOB_DONE : HP3D_JOB_FAILED,
                   e == hipSuccess ? HP3D_OK : HP3D_ERR_DEVICE);
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
              hipStreamCreateWithFlags(&s->copy, hipStreamNonBlocking) == hipSuccess &&
              hipStreamCreateWithFlags(&s->image, hipStreamNonBlocking) == hipSuccess &&
              hipMalloc((void**)&s->arena_base, opts->hbm_budget) == hipSuccess &&
              hipMalloc((void**)&s->d_jobs, opts->max_jobs * sizeof(DevJob)) == hipSuccess &&
              hipHostMalloc((void**)&s->h_jobs, opts->max_jobs * sizeof(DevJob)) == hipSuccess;
    for (int b = 0; b < 2 && ok; ++b)
        ok = hipMalloc((void**)&s->d_items[b], opts->max_jobs * sizeof(StepItem)) == hipSuccess &&
             hipHostMalloc((void**)&s->h_items[b], opts->max_jobs * sizeof(StepItem)) == hipSuccess &&
             hipEventCreate(&s->ev_a[b]) == hipSuccess && hipEventCreate(&s->ev_b[b]) == hipSuccess;
    if (!ok) {
        hp3d_server_destroy(s.release());
        return HP3D_ERR_DEVICE;
    }
    s->arena.reset(opts->hbm_budget);
    for (int k = opts->max_jobs - 1; k >= 0; --k) s->free_slots.push_back(k);
    s->t0 = Clock::now();
    s->thr = std::thread([p = s.get()] { p->loop(); });
    s->comp_thr = std::thread([p = s.get()] { p->complete_loop(); });
    *out = s.release();
    return HP3D_OK;
}

int hp3d_server_submit(hp3d_pm_server* s, const hp3d_job_desc* d, uint64_t* id)
{
    if (!s || !d || !id) return HP3D_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto reject = [&](int code, const char* msg) {
        s->last_error = msg;
        s->stats.rejected++;
        return code;
    };
    auto j = std::make_shared<SJob>();
    j->d = *d;
    // resident images: known, not removed, with the level asked for, of one size
    int rW = 0, rH = 0;
    for (int k = 0; k <= d->num_src && k <= kMaxSrc; ++k) {
        const uint64_t h = k == 0 ? d->ref_image : d->src_image[k - 1];
        if (!h) continue;
        auto it = s->images.find(h);
        if (it == s->images.end())
            return reject(HP3D_ERR_NOT_FOUND, "submit: unknown or removed resident image");
        if (d->level < 0 || d->level >= it->second->levels)
            return reject(HP3D_ERR_INVALID_ARG, "submit: resident image has no such level");
        const Image& lv = it->second->lv[d->level];
        if ((rW || rH) && (lv.W != rW || lv.H != rH))
            return reject(HP3D_ERR_INVALID_ARG, "submit: resident images differ in size");
        rW = lv.W;
        rH = lv.H;
        j->res[k] = it->second;
    }
    if (rW) {
        if (j->d.width == 0 && j->d.height == 0) {
            j->d.width = rW;
            j->d.height = rH;
        }
        if (j->d.width != rW || j->d.height != rH)
            return reject(HP3D_ERR_INVALID_ARG, "submit: width/height differ from the resident "
                                                "images at that level");
    }
    const int need = 2 * s->o.radius + 1;
    if (d->num_src < 1 || d->num_src > kMaxSrc || j->d.width < need || j->d.height < need ||
        (!j->res[0] && !d->ref) || !d->out_depth || d->iterations < 0 ||
        d->iterations > kMaxIterations || d->top_k < 1 || d->top_k > kMaxTopK ||
        !(d->depth_min < d->depth_max) || !std::isfinite(d->depth_min) ||
        !std::isfinite(d->depth_max) || (!rW && d->level != 0))
        return reject(HP3D_ERR_INVALID_ARG, "submit: invalid job description");
    for (int k = 0; k < d->num_src; ++k)
        if (!d->src[k] && !j->res[1 + k])
            return reject(HP3D_ERR_INVALID_ARG, "submit: missing source image");
    for (int i = 0; i < 9; ++i)
        if (!std::isfinite(d->K_ref_inv[i]))
            return reject(HP3D_ERR_INVALID_ARG, "submit: K_ref_inv is not finite");
    for (int k = 0; k < d->num_src; ++k)
        for (int i = 0; i < 12; ++i)
            if (!std::isfinite(i < 9 ? d->A[k][i] : d->b[k][i - 9]))
                return reject(HP3D_ERR_INVALID_ARG, "submit: a source's A or b is not finite");
    // taps address an image plane with 32-bit buffer offsets and 24-bit row
    // offsets
    if (!image_dims_supported(j->d.width, j->d.height, s->o.store))
        return reject(HP3D_ERR_UNSUPPORTED,
                      "submit: image plane above 2 GiB - 64 KiB or row pitch of 2^24 elements");
    for (int k = 0; k <= d->num_src; ++k) j->n_host += j->res[k] ? 0 : 1;
    j->footprint = s->footprint(*j);
    if (j->footprint > s->arena.size())
        return reject(HP3D_ERR_OOM_BUDGET, "submit: job footprint exceeds the server's HBM budget");
    const size_t px = (size_t)j->d.width * j->d.height;
    j->host_img.resize((size_t)j->n_host * px);
    int h = 0;
    for (int k = 0; k <= d->num_src; ++k)
        if (!j->res[k])
            std::memcpy(j->host_img.data() + (size_t)h++ * px, k == 0 ? d->ref : d->src[k - 1], px);
    for (auto& r : j->res)
        if (r) r->refs++;
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

int hp3d_server_image_add(hp3d_pm_server* s, const uint8_t* gray, int W, int H, size_t pitch,
                          int levels, uint64_t* image)
{
    if (!s || !gray || !image) return HP3D_ERR_INVALID_ARG;
    if (pitch == 0) pitch = (size_t)W;
    if (W < 2 || H < 2 || pitch < (size_t)W || levels < 1 || levels > kMaxLevels ||
        (W >> (levels - 1)) < 2 || (H >> (levels - 1)) < 2) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->last_error = "image_add: bad size, pitch or level count";
        return HP3D_ERR_INVALID_ARG;
    }
    if (!image_dims_supported(W, H, s->o.store)) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->last_error = "image_add: image plane above 2 GiB - 64 KiB or row pitch of 2^24 elements";
        return HP3D_ERR_UNSUPPORTED;
    }
    auto r = std::make_shared<ResImage>();
    r->levels = levels;
    const size_t stage = a256(pitch * (size_t)(H - 1) + W);
    size_t planes = 0;
    for (int l = 0; l < levels; ++l) {
        Image& im = r->lv[l];
        im.W = W >> l;
        im.H = H >> l;
        im.store = s->o.store;
        im.den_scale = 255.0f * 255.0f;
        image_plane_sizes(im.W, im.H, s->o.store, &im.pitch_plain, &im.plain_bytes, &im.coef_bytes);
        planes += a256(im.plain_bytes) + a256(im.coef_bytes);
    }
    size_t off = 0;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (!s->arena.alloc(stage + planes, &off)) {
            s->last_error = "image_add: the image does not fit the server's HBM budget now";
            return HP3D_ERR_OOM_BUDGET;
        }
    }
    (void)hipSetDevice(s->device);
    char* p = s->arena_base + off + stage;
    for (int l = 0; l < levels; ++l) {
        r->lv[l].plain = p;
        p += a256(r->lv[l].plain_bytes);
        r->lv[l].coef = p;
        p += a256(r->lv[l].coef_bytes);
    }
    hipError_t e = hipMemcpyAsync(s->arena_base + off, gray, pitch * (size_t)(H - 1) + W,
                                  hipMemcpyHostToDevice, s->image);
    if (e == hipSuccess)
        e = launch_img_prep((const uint8_t*)(s->arena_base + off), nullptr, pitch, W, H,
                            s->o.store, r->lv[0].plain, r->lv[0].coef, r->lv[0].pitch_plain,
                            s->image);
    for (int l = 1; l < levels && e == hipSuccess; ++l)
        e = launch_img_down(r->lv[l - 1].plain, r->lv[l - 1].pitch_plain, r->lv[l].W, r->lv[l].H,
                            s->o.store, r->lv[l].plain, r->lv[l].coef, r->lv[l].pitch_plain,
                            s->image);
    if (e == hipSuccess) e = hipStreamSynchronize(s->image);
    std::lock_guard<std::mutex> lk(s->mu);
    s->arena.release(off, stage);      // the upload staging in front of the planes
    r->off = off + stage;
    r->bytes = planes;
    if (e != hipSuccess) {
        s->arena.release(r->off, r->bytes);
        s->last_error = hipGetErrorString(e);
        return HP3D_ERR_DEVICE;
    }
    r->id = ++s->next_image;
    s->images[r->id] = r;
    s->stats.image_uploads++;
    *image = r->id;
    return HP3D_OK;
}

int hp3d_server_image_remove(hp3d_pm_server* s, uint64_t image)
{
    if (!s) return HP3D_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->images.find(image);
    if (it == s->images.end()) return HP3D_ERR_NOT_FOUND;
    auto r = it->second;
    s->images.erase(it);
    r->removed = true;
    if (r->refs == 0) s->arena.release(r->off, r->bytes);
    return HP3D_OK;
}

int hp3d_server_image_read(hp3d_pm_server* s, uint64_t image, int level, int* width, int* height,
                           float* out, size_t capacity)
{
    if (!s) return HP3D_ERR_INVALID_ARG;
    Image im;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->images.find(image);
        if (it == s->images.end()) return HP3D_ERR_NOT_FOUND;
        if (level < 0 || level >= it->second->levels) return HP3D_ERR_INVALID_ARG;
        im = it->second->lv[level];
    }
    if (width) *width = im.W;
    if (height) *height = im.H;
    if (!out) return HP3D_OK;
    if (capacity < (size_t)im.W * im.H) return HP3D_ERR_INVALID_ARG;
    const size_t esz = im.store == HP3D_STORE_F32 ? 4 : 2;
    std::vector<unsigned char> rows(im.plain_bytes);
    (void)hipSetDevice(s->device);
    hipError_t e = hipMemcpyAsync(rows.data(), im.plain, im.plain_bytes, hipMemcpyDeviceToHost,
                                  s->image);
    if (e == hipSuccess) e = hipStreamSynchronize(s->image);
    if (e != hipSuccess) return HP3D_ERR_DEVICE;
    for (int y = 0; y < im.H; ++y)
        for (int x = 0; x < im.W; ++x) {
            const size_t q = ((size_t)y * im.pitch_plain + x) * esz;
            float v;
            if (esz == 4) {
                std::memcpy(&v, &rows[q], 4);
            } else {
                _Float16 h;
                std::memcpy(&h, &rows[q], 2);
                v = (float)h;
            }
            out[(size_t)y * im.W + x] = v + kShiftDN;
        }
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
    it->second->cancel = true;   // a job whose last step is launched completes anyway
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
    st->resident_images = (long long)s->images.size();
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
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->comp_stop = true;
        s->cv_comp.notify_all();
    }
    if (s->comp_thr.joinable()) s->comp_thr.join();
    (void)hipSetDevice(s->device);
    for (hipStream_t st : {s->compute, s->upload, s->copy, s->image})
        if (st) (void)hipStreamSynchronize(st);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        for (auto& j : s->queue) s->finish_job(j, HP3D_JOB_CANCELLED, HP3D_ERR_CANCELLED);
        for (auto& j : s->active) s->finish_job(j, HP3D_JOB_CANCELLED, HP3D_ERR_CANCELLED);
        s->queue.clear();
        s->active.clear();
        s->images.clear();
    }
    if (s->arena_base) (void)hipFree(s->arena_base);
    if (s->d_jobs) (void)hipFree(s->d_jobs);
    if (s->h_jobs) (void)hipHostFree(s->h_jobs);
    for (int b = 0; b < 2; ++b) {
        if (s->d_items[b]) (void)hipFree(s->d_items[b]);
        if (s->h_items[b]) (void)hipHostFree(s->h_items[b]);
        if (s->ev_a[b]) (void)hipEventDestroy(s->ev_a[b]);
        if (s->ev_b[b]) (void)hipEventDestroy(s->ev_b[b]);
    }
    for (hipStream_t st : {s->compute, s->upload, s->copy, s->image})
        if (st) (void)hipStreamDestroy(st);
    delete s;
    return HP3D_OK;
}

}  // extern "C"
