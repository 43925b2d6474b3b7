// pm_batch.cpp -- batched form (hp3d_pm_batch_run) and the step-kernel
// autotuner (hp3d_pm_autotune).
//
// A batch replaces the reference's per-view host loop
// (src/dense_recon/patchmatch_mvs.cu:1076-1258, one view at a time, a device
// synchronisation after every launch) by a schedule of steps: step 0 starts
// every view (init + initial cost), step s >= 1 runs half-pass s-1 of every
// view that still has one, and each view writes its outputs in its last
// step.  The step launches are captured once in a hipGraph; a later batch
// with the same schedule shape replays the graph.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>

#include "pm_batch.h"
#include "pm_runtime.h"

namespace hp3d {

static size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

size_t job_state_bytes(int W, int H, int want_outputs_mask)
{
    const size_t plane = (size_t)H * ((W + 1) / 2);
    size_t n = align256(2 * plane * sizeof(float4)) + align256(2 * plane * sizeof(float)) +
               align256(2 * plane);
    const size_t px = (size_t)W * H * sizeof(float);
    if (want_outputs_mask & 1) n += align256(px);       // depth
    if (want_outputs_mask & 2) n += align256(3 * px);   // normal
    if (want_outputs_mask & 4) n += align256(px);       // cost
    if (want_outputs_mask & 8) n += align256(px);       // confidence
    return n;
}

void bind_job(DevJob& j, const hp3d_view_desc* d, const Image* ref, const Image* const* srcs,
              char* base, int outputs_mask)
{
    std::memset(&j, 0, sizeof j);
    fill_dev_view(j.v, d, ref, srcs);
    const size_t plane = (size_t)j.v.H * j.v.Wc;
    char* p = base;
    float4* hyp = reinterpret_cast<float4*>(p);
    p += align256(2 * plane * sizeof(float4));
    float* cost = reinterpret_cast<float*>(p);
    p += align256(2 * plane * sizeof(float));
    unsigned char* chg = reinterpret_cast<unsigned char*>(p);
    p += align256(2 * plane);
    j.v.hyp[0] = hyp;
    j.v.hyp[1] = hyp + plane;
    j.v.cost[0] = cost;
    j.v.cost[1] = cost + plane;
    j.v.chg[0] = chg;
    j.v.chg[1] = chg + plane;
    const size_t px = (size_t)j.v.W * j.v.H * sizeof(float);
    if (outputs_mask & 1) { j.out_depth = reinterpret_cast<float*>(p); p += align256(px); }
    if (outputs_mask & 2) { j.out_normal = reinterpret_cast<float*>(p); p += align256(3 * px); }
    if (outputs_mask & 4) { j.out_cost = reinterpret_cast<float*>(p); p += align256(px); }
    if (outputs_mask & 8) { j.out_conf = reinterpret_cast<float*>(p); p += align256(px); }
    j.seed = d->seed;
}

// The autotuner's choice is kept per (window radius, stride, storage, size
// bucket, mean source count): the size bucket is log2 of the batch's pixels
// in units of 2^18 (a quarter megapixel), 0..15, and the mean source count
// is rounded to 1..8 -- the quantities that move the best tile shape (tail
// length per step, work per tile).
int tune_key(int radius, int stride, int store, double pixels, double mean_src)
{
    int bucket = 0;
    while (bucket < 15 && pixels >= (double)(1ull << (19 + bucket))) ++bucket;
    const int S = std::max(1, std::min(8, (int)std::lround(mean_src)));
    return ((((radius * 4 + stride) * 2 + store) * 16 + bucket) * 9) + S;
}

int batch_tune_key(const hp3d_view_desc* descs, const std::vector<const Image*>& refs, int n)
{
    double px = 0.0, src = 0.0;
    for (int j = 0; j < n; ++j) {
        px += (double)refs[j]->W * refs[j]->H;
        src += descs[j].num_src;
    }
    return tune_key(descs[0].radius, descs[0].stride, refs[0]->store, px, src / n);
}

int chosen_variant(const hp3d_pm_ctx* ctx, int radius, int stride, int key, int requested)
{
    if (!step_variant_tunable(radius, stride)) return 0;
    if (requested >= 0) return requested;
    auto it = ctx->tuned.find(key);
    return it == ctx->tuned.end() ? kDefaultStepVariant : it->second;
}

StepItem make_item(int job, int step, int iterations, uint64_t seed, int W, int H, int variant)
{
    StepItem it{};
    it.job = job;
    const int tx = step_tile_tx(variant), ty = step_tile_ty(variant);
    const int Wc = (W + 1) / 2;
    it.tiles_x = (Wc + tx - 1) / tx;
    it.tiles = it.tiles_x * ((H + ty - 1) / ty);
    if (step == 0) {
        it.mode = kStepStart;
    } else {
        it.mode = kStepPropagate;
        it.iteration = (step - 1) / 2;
        it.parity = (step - 1) % 2;
        it.pass_seed = pass_seed(seed, it.iteration, it.parity);
    }
    it.emit = step == 2 * iterations;   // the view's last step
    it.step = step;
    return it;
}

namespace {

struct Plan {
    std::vector<StepItem> items;      // all steps, concatenated
    std::vector<int> offset, count, tiles;
};

// Steps [first, last) of the given views (iterations per view).
Plan build_plan(const std::vector<DevJob>& jobs, const hp3d_view_desc* descs, int n, int variant,
                int first_step, int last_step)
{
    Plan p;
    for (int s = first_step; s < last_step; ++s) {
        const int off = (int)p.items.size();
        int tiles = 0;
        for (int j = 0; j < n; ++j) {
            const int iters = descs[j].iterations;
            if (s > 2 * iters) continue;
            StepItem it = make_item(j, s, iters, descs[j].seed, jobs[j].v.W, jobs[j].v.H, variant);
            it.tile_begin = tiles;
            tiles += it.tiles;
            p.items.push_back(it);
        }
        p.offset.push_back(off);
        p.count.push_back((int)p.items.size() - off);
        p.tiles.push_back(tiles);
    }
    return p;
}

// Every step of the given views in one list for the persistent kernel:
// items in step order, tile_begin counting tiles across all steps, each
// view's tile counters at done_base.  *counters: tiles over all views.
Plan build_flat_plan(const std::vector<DevJob>& jobs, const hp3d_view_desc* descs, int n,
                     int variant, int max_steps, int* counters, int* units)
{
    Plan p;
    std::vector<int> base(n, 0);
    int total = 0;
    for (int j = 0; j < n; ++j) {
        base[j] = total;
        total += make_item(j, 0, descs[j].iterations, descs[j].seed, jobs[j].v.W, jobs[j].v.H,
                           variant).tiles;
    }
    int u = 0;
    for (int s = 0; s < max_steps; ++s)
        for (int j = 0; j < n; ++j) {
            const int iters = descs[j].iterations;
            if (s > 2 * iters) continue;
            StepItem it = make_item(j, s, iters, descs[j].seed, jobs[j].v.W, jobs[j].v.H, variant);
            it.tile_begin = u;
            it.done_base = base[j];
            u += it.tiles;
            p.items.push_back(it);
        }
    *counters = total;
    *units = u;
    return p;
}

int validate_batch(hp3d_pm_ctx* ctx, const hp3d_view_desc* descs, int n,
                   std::vector<const Image*>& refs, std::vector<std::vector<const Image*>>& srcs)
{
    if (!descs || n < 1) return ctx->fail(HP3D_ERR_INVALID_ARG, "batch: no views");
    refs.resize(n);
    srcs.assign(n, std::vector<const Image*>(kMaxSrc, nullptr));
    for (int j = 0; j < n; ++j) {
        int rc = validate_desc(ctx, &descs[j], &refs[j], srcs[j].data());
        if (rc) return ctx->fail(rc, "batch view %d: %s", j, ctx->last_error.c_str());
        if (descs[j].init_depth || descs[j].num_seeds > 0 || descs[j].first_iteration != 0)
            return ctx->fail(HP3D_ERR_INVALID_ARG, "batch view %d: caller hypotheses, SfM seeds "
                             "and resumed runs are only supported by hp3d_pm_run_view", j);
        if (descs[j].radius != descs[0].radius || descs[j].stride != descs[0].stride ||
            refs[j]->store != refs[0]->store)
            return ctx->fail(HP3D_ERR_INVALID_ARG, "batch view %d: radius, stride and storage "
                             "must match view 0", j);
    }
    return HP3D_OK;
}

// Lay out every view's state (and staged outputs) in the batch arena and
// upload the job table.
int stage_jobs(hp3d_pm_ctx* ctx, const hp3d_view_desc* descs, int n,
               const std::vector<const Image*>& refs,
               const std::vector<std::vector<const Image*>>& srcs, const hp3d_view_out* outs,
               std::vector<DevJob>& jobs, std::vector<size_t>& job_off)
{
    jobs.resize(n);
    job_off.resize(n);
    std::vector<int> masks(n, 0);
    size_t total = 0;
    for (int j = 0; j < n; ++j) {
        if (outs && !outs[j].on_device)
            masks[j] = (outs[j].depth ? 1 : 0) | (outs[j].normal ? 2 : 0) |
                       (outs[j].cost ? 4 : 0) | (outs[j].confidence ? 8 : 0);
        job_off[j] = total;
        total += job_state_bytes(refs[j]->W, refs[j]->H, masks[j]);
    }
    int rc = ctx->grow(ctx->batch_state, total, "batch state");
    if (rc) return rc;
    char* base = static_cast<char*>(ctx->batch_state.ptr);
    for (int j = 0; j < n; ++j) {
        bind_job(jobs[j], &descs[j], refs[j], srcs[j].data(), base + job_off[j], masks[j]);
        if (outs && outs[j].on_device) {
            jobs[j].out_depth = outs[j].depth;
            jobs[j].out_normal = outs[j].normal;
            jobs[j].out_cost = outs[j].cost;
            jobs[j].out_conf = outs[j].confidence;
        }
    }
    rc = ctx->grow(ctx->batch_jobs, n * sizeof(DevJob), "batch job table");
    if (rc) return rc;
    hipError_t e = hipMemcpyAsync(ctx->batch_jobs.ptr, jobs.data(), n * sizeof(DevJob),
                                  hipMemcpyHostToDevice, ctx->stream);
    return e == hipSuccess ? HP3D_OK : ctx->hip_fail(e, "job table upload");
}

int upload_items(hp3d_pm_ctx* ctx, const Plan& p)
{
    int rc = ctx->grow(ctx->batch_items, p.items.size() * sizeof(StepItem), "step items");
    if (rc) return rc;
    hipError_t e = hipMemcpyAsync(ctx->batch_items.ptr, p.items.data(),
                                  p.items.size() * sizeof(StepItem), hipMemcpyHostToDevice,
                                  ctx->stream);
    return e == hipSuccess ? HP3D_OK : ctx->hip_fail(e, "step item upload");
}

hipError_t launch_plan(hp3d_pm_ctx* ctx, const Plan& p, int store, int radius, int stride,
                       int variant)
{
    const DevJob* jobs = static_cast<const DevJob*>(ctx->batch_jobs.ptr);
    const StepItem* items = static_cast<const StepItem*>(ctx->batch_items.ptr);
    for (size_t s = 0; s < p.count.size(); ++s) {
        if (p.count[s] == 0) continue;
        hipError_t e = launch_step(store, radius, stride, variant, jobs, items + p.offset[s],
                                   p.count[s], p.tiles[s], ctx->stream);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace
}  // namespace hp3d

using namespace hp3d;

extern "C" {

int hp3d_pm_batch_run(hp3d_pm_ctx* ctx, const hp3d_view_desc* descs, int n,
                      const hp3d_view_out* outs, const hp3d_batch_opts* opts)
{
    if (!ctx) return HP3D_ERR_INVALID_ARG;
    if (!outs) return ctx->fail(HP3D_ERR_INVALID_ARG, "batch: outputs are NULL");
    std::vector<const Image*> refs;
    std::vector<std::vector<const Image*>> srcs;
    int rc = validate_batch(ctx, descs, n, refs, srcs);
    if (rc) return rc;
    for (int j = 0; j < n; ++j)
        if (!outs[j].depth) return ctx->fail(HP3D_ERR_INVALID_ARG, "batch view %d: no depth output", j);
    (void)hipSetDevice(ctx->device);
    const int radius = descs[0].radius, stride = descs[0].stride, store = refs[0]->store;
    const int variant = chosen_variant(ctx, radius, stride, batch_tune_key(descs, refs, n),
                                       opts ? opts->variant : -1);
    if (variant >= kNumStepVariants)
        return ctx->fail(HP3D_ERR_INVALID_ARG, "variant %d outside 0..%d", variant, kNumStepVariants - 1);
    const bool use_graph = opts ? opts->use_graph != 0 : true;

    std::vector<DevJob> jobs;
    std::vector<size_t> job_off;
    rc = stage_jobs(ctx, descs, n, refs, srcs, outs, jobs, job_off);
    if (rc) return rc;
    int max_steps = 0;
    double taps = 0.0;
    for (int j = 0; j < n; ++j) {
        max_steps = std::max(max_steps, 1 + 2 * descs[j].iterations);
        taps += algorithmic_taps(jobs[j].v.W, jobs[j].v.H, jobs[j].v.S, radius, stride,
                                 descs[j].iterations);
    }
    const bool persistent = opts ? opts->persistent != 0 : false;
    const int chunk = opts && opts->chunk > 0 ? opts->chunk : kDefaultPersistChunk;
    int counters = 0, units = 0;
    Plan plan;
    if (persistent) {
        // the persistent kernel addresses state planes with 31-bit byte offsets
        for (int j = 0; j < n; ++j)
            if ((size_t)jobs[j].v.H * jobs[j].v.Wc * sizeof(float4) >= 0x7fffffffull)
                return ctx->fail(HP3D_ERR_UNSUPPORTED, "batch view %d: %dx%d too large for the "
                                 "persistent kernel (state plane of 2 GiB)", j, jobs[j].v.W,
                                 jobs[j].v.H);
        plan = build_flat_plan(jobs, descs, n, variant, max_steps, &counters, &units);
        rc = ctx->grow(ctx->batch_done, (size_t)(counters + kCtlWords) * sizeof(int),
                       "persistent tile counters");
        if (rc) return rc;
    } else {
        plan = build_plan(jobs, descs, n, variant, 0, max_steps);
    }
    rc = upload_items(ctx, plan);
    if (rc) return rc;
    int* done = static_cast<int*>(ctx->batch_done.ptr);
    unsigned long long wait_ticks = 0;
    if (persistent) {
        int khz = 0;
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ctx->device) != hipSuccess ||
            khz <= 0)
            khz = 100000;
        wait_ticks = (unsigned long long)khz * 1000ull * kPersistWaitSeconds;
    }
    // the whole batch as launches on the context stream: the step kernels,
    // or the counter reset and the one persistent kernel
    auto launch_all = [&]() -> hipError_t {
        if (!persistent) return launch_plan(ctx, plan, store, radius, stride, variant);
        hipError_t le = zero_ints(done, counters + kCtlWords, ctx->stream);
        if (le == hipSuccess)
            le = launch_persistent(store, radius, stride, variant,
                                   static_cast<const DevJob*>(ctx->batch_jobs.ptr),
                                   static_cast<const StepItem*>(ctx->batch_items.ptr),
                                   (int)plan.items.size(), units, done, counters, chunk,
                                   wait_ticks, ctx->stream);
        return le;
    };

    ctx->stats = hp3d_pm_stats{};
    ctx->stats.taps = taps;
    ctx->mark(kPhaseStart);
    hipError_t e = hipSuccess;
    if (use_graph) {
        // the graph depends on the buffers it reads and on the schedule shape
        std::vector<unsigned char> key;
        auto put = [&key](const void* p, size_t b) {
            const unsigned char* c = static_cast<const unsigned char*>(p);
            key.insert(key.end(), c, c + b);
        };
        put(&ctx->batch_jobs.ptr, sizeof(void*));
        put(&ctx->batch_items.ptr, sizeof(void*));
        int meta[7] = {store, radius, stride, variant, persistent ? 1 : 0, chunk, counters};
        put(meta, sizeof meta);
        put(&ctx->batch_done.ptr, sizeof(void*));
        put(plan.offset.data(), plan.offset.size() * sizeof(int));
        put(plan.count.data(), plan.count.size() * sizeof(int));
        put(plan.tiles.data(), plan.tiles.size() * sizeof(int));
        if (!ctx->batch_graph || key != ctx->batch_graph_key) {
            if (ctx->batch_graph) {
                (void)hipGraphExecDestroy(ctx->batch_graph);
                ctx->batch_graph = nullptr;
            }
            hipGraph_t g = nullptr;
            e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
            if (e == hipSuccess) {
                hipError_t le = launch_all();
                e = hipStreamEndCapture(ctx->stream, &g);
                if (le != hipSuccess) e = le;
            }
            if (e == hipSuccess) e = hipGraphInstantiate(&ctx->batch_graph, g, nullptr, nullptr, 0);
            if (g) (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                ctx->batch_graph = nullptr;
                return ctx->hip_fail(e, "batch graph capture");
            }
            ctx->batch_graph_key = key;
        }
        e = hipGraphLaunch(ctx->batch_graph, ctx->stream);
    } else {
        e = launch_all();
    }
    if (e != hipSuccess) return ctx->hip_fail(e, "batch launch");
    ctx->mark(kPhasePropagate);
    for (int j = 0; j < n && e == hipSuccess; ++j) {
        if (outs[j].on_device) continue;
        const size_t px = (size_t)jobs[j].v.W * jobs[j].v.H;
        struct { float* dst; const float* src; size_t count; } cp[4] = {
            {outs[j].depth, jobs[j].out_depth, px}, {outs[j].normal, jobs[j].out_normal, 3 * px},
            {outs[j].cost, jobs[j].out_cost, px}, {outs[j].confidence, jobs[j].out_conf, px}};
        for (auto& c : cp)
            if (c.dst && e == hipSuccess)
                e = hipMemcpyAsync(c.dst, c.src, c.count * sizeof(float), hipMemcpyDeviceToHost,
                                   ctx->stream);
    }
    if (e != hipSuccess) return ctx->hip_fail(e, "batch output copy");
    ctx->mark(kPhaseFinalize);
    int ctl[kCtlWords] = {0, 0};
    if (persistent)
        e = hipMemcpyAsync(ctl, done + counters, sizeof ctl, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    ctx->collect_marks(max_steps - 1);
    if (e != hipSuccess) return ctx->hip_fail(e, "batch synchronize");
    if (ctl[kCtlAbort])
        return ctx->fail(HP3D_ERR_DEVICE, "persistent batch: a tile waited more than %d s for "
                         "its neighbours; outputs are incomplete", kPersistWaitSeconds);
    return HP3D_OK;
}

int hp3d_pm_autotune(hp3d_pm_ctx* ctx, const hp3d_view_desc* descs, int n, int* best_variant,
                     float* times_ms)
{
    if (!ctx) return HP3D_ERR_INVALID_ARG;
    std::vector<const Image*> refs;
    std::vector<std::vector<const Image*>> srcs;
    int rc = validate_batch(ctx, descs, n, refs, srcs);
    if (rc) return rc;
    (void)hipSetDevice(ctx->device);
    const int radius = descs[0].radius, stride = descs[0].stride, store = refs[0]->store;
    std::vector<DevJob> jobs;
    std::vector<size_t> job_off;
    rc = stage_jobs(ctx, descs, n, refs, srcs, nullptr, jobs, job_off);
    if (rc) return rc;
    // time a start step and two half-passes (views with at least one iteration)
    std::vector<hp3d_view_desc> tdesc(descs, descs + n);
    for (auto& d : tdesc) d.iterations = std::max(d.iterations, 1);
    hipEvent_t ev0, ev1;
    if (hipEventCreate(&ev0) != hipSuccess || hipEventCreate(&ev1) != hipSuccess)
        return ctx->fail(HP3D_ERR_DEVICE, "event create");
    int best = 0;
    float best_ms = 1e30f;
    for (int var = 0; var < kNumStepVariants; ++var) {
        if (times_ms) times_ms[var] = -1.0f;
        if (var > 0 && !step_variant_tunable(radius, stride)) continue;
        Plan plan = build_plan(jobs, tdesc.data(), n, var, 0, 3);
        rc = upload_items(ctx, plan);
        if (rc) break;
        hipError_t e = launch_plan(ctx, plan, store, radius, stride, var);   // warm-up
        if (e == hipSuccess) e = hipEventRecord(ev0, ctx->stream);
        if (e == hipSuccess) e = launch_plan(ctx, plan, store, radius, stride, var);
        if (e == hipSuccess) e = hipEventRecord(ev1, ctx->stream);
        if (e == hipSuccess) e = hipEventSynchronize(ev1);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev0, ev1);
        if (e != hipSuccess) { rc = ctx->hip_fail(e, "autotune run"); break; }
        if (times_ms) times_ms[var] = ms;
        if (ms < best_ms) { best_ms = ms; best = var; }
    }
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
    if (rc) return rc;
    ctx->tuned[batch_tune_key(descs, refs, n)] = best;
    if (best_variant) *best_variant = best;
    return HP3D_OK;
}

int hp3d_pm_variant_info(int variant, int* tx, int* ty, int* wpe)
{
    if (variant < 0 || variant >= kNumStepVariants) return HP3D_ERR_INVALID_ARG;
    if (tx) *tx = kStepVariants[variant].tx;
    if (ty) *ty = kStepVariants[variant].ty;
    if (wpe) *wpe = kStepVariants[variant].wpe;
    return HP3D_OK;
}

int hp3d_pm_tune_save(const hp3d_pm_ctx* ctx, const char* path)
{
    if (!ctx || !path) return HP3D_ERR_INVALID_ARG;
    std::ofstream f(path);
    if (!f) return HP3D_ERR_IO;
    f << "# hp3d-pm autotune: key variant\n";
    for (const auto& kv : ctx->tuned) f << kv.first << " " << kv.second << "\n";
    return f.good() ? HP3D_OK : HP3D_ERR_IO;
}

int hp3d_pm_tune_load(hp3d_pm_ctx* ctx, const char* path)
{
    if (!ctx || !path) return HP3D_ERR_INVALID_ARG;
    std::ifstream f(path);
    if (!f) return HP3D_ERR_IO;
    std::string line;
    while (std::getline(f, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream is(line);
        int k, v;
        if (!(is >> k >> v) || v < 0 || v >= kNumStepVariants) return HP3D_ERR_IO;
        ctx->tuned[k] = v;
    }
    return HP3D_OK;
}

}  // extern "C"
