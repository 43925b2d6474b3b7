// pm_solve.cpp -- per-view launch sequence (patchmatch_mvs.cu:1093-1258).
//
// init -> cost(red) -> cost(black) -> iterations x {half-pass red, black}
// -> finalize, all on the context stream, one synchronisation at the end.
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
    int n_prop = 0;
    for (int it = d->first_iteration; it < d->first_iteration + d->iterations; ++it)
        for (int parity = 0; parity < 2; ++parity) {
            e = launch_pass(v, ref->store, d->radius, d->stride, kPassPropagate, it, parity,
                            pass_seed(d->seed, it, parity), ctx->stream);
            if (e != hipSuccess) return ctx->hip_fail(e, "half-pass launch");
            ctx->mark(kPhasePropagate);
            ++n_prop;
        }
    rc = emit_outputs(ctx, v, out->depth, out->normal, out->cost, out->confidence,
                      out->on_device != 0);
    if (rc) return rc;
    ctx->mark(kPhaseFinalize);
    return finish(ctx, n_prop);
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
