// pm_solve.cpp -- per-view launch sequence (patchmatch_mvs.cu:1093-1258).
//
// init -> cost(red) -> cost(black) -> iterations x {half-pass red, black}
// -> finalize, all on the context stream, one synchronisation at the end.
#include <algorithm>
#include <cstring>

#include "pm_runtime.h"

namespace hp3d {

namespace {

// Bind the context's state planes to v, growing them if the view is larger.
int bind_state(hp3d_pm_ctx* ctx, DevView& v)
{
    const size_t plane = (size_t)v.H * v.Wc;
    int rc = ctx->grow(ctx->hyp, 2 * plane * sizeof(float4), "hypothesis planes");
    if (rc) return rc;
    rc = ctx->grow(ctx->cost, 2 * plane * sizeof(float), "cost planes");
    if (rc) return rc;
    rc = ctx->grow(ctx->chg, 2 * plane, "change flags");
    if (rc) return rc;
    float4* h = static_cast<float4*>(ctx->hyp.ptr);
    float* c = static_cast<float*>(ctx->cost.ptr);
    unsigned char* f = static_cast<unsigned char*>(ctx->chg.ptr);
    v.hyp[0] = h;
    v.hyp[1] = h + plane;
    v.cost[0] = c;
    v.cost[1] = c + plane;
    v.chg[0] = f;
    v.chg[1] = f + plane;
    return HP3D_OK;
}

// Upload caller hypotheses (host depth H*W, normal H*W*3) into the planes.
int import_hypotheses(hp3d_pm_ctx* ctx, const DevView& v, const float* depth, const float* normal)
{
    const size_t n = (size_t)v.W * v.H;
    int rc = ctx->grow(ctx->in_depth, n * sizeof(float), "input depth");
    if (rc) return rc;
    rc = ctx->grow(ctx->in_normal, 3 * n * sizeof(float), "input normal");
    if (rc) return rc;
    hipError_t e = hipMemcpyAsync(ctx->in_depth.ptr, depth, n * sizeof(float),
                                  hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(ctx->in_normal.ptr, normal, 3 * n * sizeof(float),
                           hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
        e = launch_import(v, (const float*)ctx->in_depth.ptr, (const float*)ctx->in_normal.ptr,
                          ctx->stream);
    return e == hipSuccess ? HP3D_OK : ctx->hip_fail(e, "hypothesis import");
}

// Paint the descriptor's SfM seeds over the current hypotheses.
int apply_seeds(hp3d_pm_ctx* ctx, const DevView& v, const hp3d_view_desc* d)
{
    if (d->num_seeds <= 0) return HP3D_OK;
    const size_t bytes = (size_t)d->num_seeds * sizeof(hp3d_seed_point);
    int rc = ctx->grow(ctx->seed_pts, bytes, "seed points");
    if (rc) return rc;
    rc = ctx->grow(ctx->seed_owner, (size_t)v.W * v.H * sizeof(int), "seed owner map");
    if (rc) return rc;
    hipError_t e = hipMemcpyAsync(ctx->seed_pts.ptr, d->seeds, bytes, hipMemcpyHostToDevice,
                                  ctx->stream);
    if (e == hipSuccess)
        e = launch_seed(v, static_cast<const hp3d_seed_point*>(ctx->seed_pts.ptr), d->num_seeds,
                        d->seed_radius, static_cast<int*>(ctx->seed_owner.ptr), ctx->stream);
    return e == hipSuccess ? HP3D_OK : ctx->hip_fail(e, "seeding");
}

// Finalize into device pointers, or into staging buffers followed by D2H copies.
int emit_outputs(hp3d_pm_ctx* ctx, const DevView& v, float* depth, float* normal, float* cost,
                 float* conf, bool on_device)
{
    const size_t n = (size_t)v.W * v.H;
    hipError_t e;
    if (on_device) {
        e = launch_finalize(v, depth, normal, cost, conf, ctx->stream);
        return e == hipSuccess ? HP3D_OK : ctx->hip_fail(e, "finalize");
    }
    int rc = HP3D_OK;
    float* dd = nullptr;
    float* dn = nullptr;
    float* dc = nullptr;
    if (depth) {
        rc = ctx->grow(ctx->out_depth, n * sizeof(float), "output depth");
        dd = static_cast<float*>(ctx->out_depth.ptr);
    }
    if (!rc && normal) {
        rc = ctx->grow(ctx->out_normal, 3 * n * sizeof(float), "output normal");
        dn = static_cast<float*>(ctx->out_normal.ptr);
    }
    if (!rc && (cost || conf)) {
        rc = ctx->grow(ctx->out_cost, 2 * n * sizeof(float), "output cost");
        dc = static_cast<float*>(ctx->out_cost.ptr);
    }
    if (rc) return rc;
    e = launch_finalize(v, dd, dn, cost ? dc : nullptr, conf ? dc + n : nullptr, ctx->stream);
    if (e == hipSuccess && depth)
        e = hipMemcpyAsync(depth, dd, n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess && normal)
        e = hipMemcpyAsync(normal, dn, 3 * n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess && cost)
        e = hipMemcpyAsync(cost, dc, n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess && conf)
        e = hipMemcpyAsync(conf, dc + n, n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream);
    return e == hipSuccess ? HP3D_OK : ctx->hip_fail(e, "output copy");
}

int finish(hp3d_pm_ctx* ctx, int n_prop)
{
    hipError_t e = hipStreamSynchronize(ctx->stream);
    ctx->collect_marks(n_prop);
    return e == hipSuccess ? HP3D_OK : ctx->hip_fail(e, "stream synchronize");
}

}  // namespace

namespace {

// Everything of one view up to its last half-pass, on the context stream:
// init (or caller hypotheses), seeds, initial cost, the iterations.
int enqueue_passes(hp3d_pm_ctx* ctx, const hp3d_view_desc* d, const Image* ref, DevView& v,
                   int* n_prop)
{
    int rc = HP3D_OK;
    hipError_t e = hipSuccess;
    if (d->init_depth) {
        rc = import_hypotheses(ctx, v, d->init_depth, d->init_normal);
        if (rc) return rc;
    } else {
        e = launch_init(v, d->seed, ctx->stream);
        if (e != hipSuccess) return ctx->hip_fail(e, "init launch");
    }
    rc = apply_seeds(ctx, v, d);
    if (rc) return rc;
    ctx->mark(kPhaseInit);
    for (int parity = 0; parity < 2 && e == hipSuccess; ++parity)
        e = launch_pass(v, ref->store, d->radius, d->stride, kPassCost, 0, parity, 0, ctx->stream);
    if (e != hipSuccess) return ctx->hip_fail(e, "cost launch");
    ctx->mark(kPhaseCost);
    for (int it = d->first_iteration; it < d->first_iteration + d->iterations; ++it)
        for (int parity = 0; parity < 2; ++parity) {
            e = launch_pass(v, ref->store, d->radius, d->stride, kPassPropagate, it, parity,
                            pass_seed(d->seed, it, parity), ctx->stream);
            if (e != hipSuccess) return ctx->hip_fail(e, "half-pass launch");
            ctx->mark(kPhasePropagate);
            ++*n_prop;
        }
    return HP3D_OK;
}

// Bytes of one view's staged host outputs: depth, normal, cost, confidence.
size_t staged_bytes(const hp3d_view_desc* d, const Image* ref, const hp3d_view_out* o)
{
    if (o->on_device) return 0;
    const size_t n = (size_t)ref->W * ref->H * sizeof(float);
    (void)d;
    return (o->depth ? n : 0) + (o->normal ? 3 * n : 0) + (o->cost ? n : 0) +
           (o->confidence ? n : 0);
}

}  // namespace

int solve_view(hp3d_pm_ctx* ctx, const hp3d_view_desc* d, const Image* ref,
               const Image* const* srcs, const hp3d_view_out* out)
{
    DevView v;
    fill_dev_view(v, d, ref, srcs);
    int rc = bind_state(ctx, v);
    if (rc) return rc;
    ctx->stats = hp3d_pm_stats{};
    ctx->stats.taps = algorithmic_taps(v.W, v.H, v.S, d->radius, d->stride, d->iterations);
    ctx->mark(kPhaseStart);
    int n_prop = 0;
    rc = enqueue_passes(ctx, d, ref, v, &n_prop);
    if (rc) return rc;
    rc = emit_outputs(ctx, v, out->depth, out->normal, out->cost, out->confidence,
                      out->on_device != 0);
    if (rc) return rc;
    ctx->mark(kPhaseFinalize);
    return finish(ctx, n_prop);
}

int solve_views(hp3d_pm_ctx* ctx, int n, const hp3d_view_desc* d, const Image* const* refs,
                const Image* const* const* srcs, const hp3d_view_out* out)
{
    // the pipeline's stream, events and two staging slots (sized for the
    // largest view before the loop, so no buffer is reallocated while a
    // download may still read it)
    hipError_t e = hipSuccess;
    if (!ctx->copy_stream)
        e = hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking);
    for (int k = 0; k < 2 && e == hipSuccess; ++k) {
        if (!ctx->ev_fin[k]) e = hipEventCreateWithFlags(&ctx->ev_fin[k], hipEventDisableTiming);
        if (e == hipSuccess && !ctx->ev_copied[k])
            e = hipEventCreateWithFlags(&ctx->ev_copied[k], hipEventDisableTiming);
    }
    if (e != hipSuccess) return ctx->hip_fail(e, "pipeline set-up");
    size_t slot_bytes = 0;
    for (int i = 0; i < n; ++i) slot_bytes = std::max(slot_bytes, staged_bytes(&d[i], refs[i], &out[i]));
    e = hipStreamSynchronize(ctx->copy_stream);
    if (e != hipSuccess) return ctx->hip_fail(e, "copy stream");
    for (int k = 0; k < 2 && slot_bytes; ++k) {
        int rc = ctx->grow(ctx->out_slot[k], slot_bytes, "output staging slot");
        if (rc) return rc;
    }
    // state planes for the largest view up front: growing them inside the
    // loop would free planes that queued work still uses
    {
        DevView big;
        std::memset(&big, 0, sizeof big);
        size_t best = 0;
        for (int i = 0; i < n; ++i) {
            const size_t plane = (size_t)refs[i]->H * ((refs[i]->W + 1) / 2);
            if (plane > best) {
                best = plane;
                big.H = refs[i]->H;
                big.Wc = (refs[i]->W + 1) / 2;
            }
        }
        int rc = bind_state(ctx, big);
        if (rc) return rc;
        // likewise the buffers that caller hypotheses and SfM seeds pass
        // through (a later view's import would otherwise free the buffer an
        // earlier view's queued import kernel reads)
        size_t in_px = 0, seed_bytes = 0, seed_px = 0;
        for (int i = 0; i < n; ++i) {
            const size_t px = (size_t)refs[i]->W * refs[i]->H;
            if (d[i].init_depth) in_px = std::max(in_px, px);
            if (d[i].num_seeds > 0) {
                seed_bytes = std::max(seed_bytes, (size_t)d[i].num_seeds * sizeof(hp3d_seed_point));
                seed_px = std::max(seed_px, px);
            }
        }
        if (in_px) rc = ctx->grow(ctx->in_depth, in_px * sizeof(float), "input depth");
        if (!rc && in_px) rc = ctx->grow(ctx->in_normal, 3 * in_px * sizeof(float), "input normal");
        if (!rc && seed_bytes) rc = ctx->grow(ctx->seed_pts, seed_bytes, "seed points");
        if (!rc && seed_px) rc = ctx->grow(ctx->seed_owner, seed_px * sizeof(int), "seed owner map");
        if (rc) return rc;
    }
    ctx->stats = hp3d_pm_stats{};
    ctx->mark(kPhaseStart);
    int n_prop = 0;
    bool slot_used[2] = {false, false};
    // staged output pointers of the view in each slot
    float* sd[2][4] = {};
    int rc = HP3D_OK;
    for (int i = 0; i <= n && rc == HP3D_OK; ++i) {
        if (i < n) {
            DevView v;
            fill_dev_view(v, &d[i], refs[i], srcs[i]);
            rc = bind_state(ctx, v);
            if (rc) break;
            ctx->stats.taps += algorithmic_taps(v.W, v.H, v.S, d[i].radius, d[i].stride,
                                                d[i].iterations);
            rc = enqueue_passes(ctx, &d[i], refs[i], v, &n_prop);
            if (rc) break;
            const hp3d_view_out& o = out[i];
            if (o.on_device) {
                e = launch_finalize(v, o.depth, o.normal, o.cost, o.confidence, ctx->stream);
            } else {
                const int k = i & 1;
                // the slot's previous view must be downloaded before it is overwritten
                if (slot_used[k]) e = hipStreamWaitEvent(ctx->stream, ctx->ev_copied[k], 0);
                float* p = static_cast<float*>(ctx->out_slot[k].ptr);
                const size_t px = (size_t)v.W * v.H;
                float* dst[4] = {o.depth, o.normal, o.cost, o.confidence};
                const size_t cnt[4] = {px, 3 * px, px, px};
                for (int q = 0; q < 4; ++q) {
                    sd[k][q] = dst[q] ? p : nullptr;
                    if (dst[q]) p += cnt[q];
                }
                if (e == hipSuccess)
                    e = launch_finalize(v, sd[k][0], sd[k][1], sd[k][2], sd[k][3], ctx->stream);
                if (e == hipSuccess) e = hipEventRecord(ctx->ev_fin[k], ctx->stream);
                slot_used[k] = true;
            }
            if (e != hipSuccess) { rc = ctx->hip_fail(e, "finalize"); break; }
            ctx->mark(kPhaseFinalize);
        }
        // download the previous view while this one computes; enqueued after
        // this view's launches, so a copy that blocks the host (pageable
        // destination) still overlaps them
        const int j = i - 1;
        if (j >= 0 && !out[j].on_device) {
            const int k = j & 1;
            const size_t px = (size_t)refs[j]->W * refs[j]->H;
            float* dst[4] = {out[j].depth, out[j].normal, out[j].cost, out[j].confidence};
            const size_t cnt[4] = {px, 3 * px, px, px};
            e = hipStreamWaitEvent(ctx->copy_stream, ctx->ev_fin[k], 0);
            for (int q = 0; q < 4 && e == hipSuccess; ++q)
                if (dst[q])
                    e = hipMemcpyAsync(dst[q], sd[k][q], cnt[q] * sizeof(float),
                                       hipMemcpyDeviceToHost, ctx->copy_stream);
            if (e == hipSuccess) e = hipEventRecord(ctx->ev_copied[k], ctx->copy_stream);
            if (e != hipSuccess) rc = ctx->hip_fail(e, "output download");
        }
    }
    // drain both streams whatever happened, so no copy outlives the call
    hipError_t e1 = hipStreamSynchronize(ctx->stream);
    hipError_t e2 = hipStreamSynchronize(ctx->copy_stream);
    ctx->collect_marks(n_prop);
    if (rc) return rc;
    if (e1 != hipSuccess) return ctx->hip_fail(e1, "stream synchronize");
    if (e2 != hipSuccess) return ctx->hip_fail(e2, "copy stream synchronize");
    return HP3D_OK;
}

int init_only(hp3d_pm_ctx* ctx, const hp3d_view_desc* d, const Image* ref,
              const Image* const* srcs, float* depth, float* normal)
{
    DevView v;
    fill_dev_view(v, d, ref, srcs);
    int rc = bind_state(ctx, v);
    if (rc) return rc;
    hipError_t e = launch_init(v, d->seed, ctx->stream);
    if (e != hipSuccess) return ctx->hip_fail(e, "init launch");
    rc = apply_seeds(ctx, v, d);
    if (rc) return rc;
    rc = emit_outputs(ctx, v, depth, normal, nullptr, nullptr, false);
    if (rc) return rc;
    return finish(ctx, 0);
}

int cost_only(hp3d_pm_ctx* ctx, const hp3d_view_desc* d, const Image* ref,
              const Image* const* srcs, const float* depth, const float* normal, float* cost)
{
    DevView v;
    fill_dev_view(v, d, ref, srcs);
    int rc = bind_state(ctx, v);
    if (rc) return rc;
    rc = import_hypotheses(ctx, v, depth, normal);
    if (rc) return rc;
    hipError_t e = hipSuccess;
    for (int parity = 0; parity < 2 && e == hipSuccess; ++parity)
        e = launch_pass(v, ref->store, d->radius, d->stride, kPassCost, 0, parity, 0, ctx->stream);
    if (e != hipSuccess) return ctx->hip_fail(e, "cost launch");
    rc = emit_outputs(ctx, v, nullptr, nullptr, cost, nullptr, false);
    if (rc) return rc;
    return finish(ctx, 0);
}

}  // namespace hp3d
