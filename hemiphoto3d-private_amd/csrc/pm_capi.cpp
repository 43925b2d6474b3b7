// pm_capi.cpp -- host runtime behind include/hp3d_pm.h.
//
// Replaces the per-view driver of runPatchMatchMVS
// (src/dense_recon/patchmatch_mvs.cu:1028-1258): images live in HBM as
// plain + row-difference planes, hypothesis state as two colour
// planes, and every launch of a view goes onto the context's stream with no
// host synchronisation until the outputs are needed (the reference
// synchronises after every launch, src/utils/cuda_utils.cuh:18-22).
#include "pm_runtime.h"

#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>

using namespace hp3d;

// ---------------------------------------------------------------------------
// Context helpers
// ---------------------------------------------------------------------------
int hp3d_pm_ctx::fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    last_error = buf;
    return code;
}

int hp3d_pm_ctx::hip_fail(hipError_t e, const char* what)
{
    return fail(HP3D_ERR_DEVICE, "%s: %s", what, hipGetErrorString(e));
}

int hp3d_pm_ctx::dev_alloc(void** p, size_t bytes, const char* what)
{
    *p = nullptr;
    if (bytes == 0) return HP3D_OK;
    if (hbm_in_use + bytes > hbm_budget)
        return fail(HP3D_ERR_OOM_BUDGET, "%s: %zu bytes would exceed the HBM budget (%zu of %zu in use)",
                    what, bytes, hbm_in_use, hbm_budget);
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return hip_fail(e, what);
    }
    hbm_in_use += bytes;
    return HP3D_OK;
}

void hp3d_pm_ctx::dev_free(void* p, size_t bytes)
{
    if (!p) return;
    (void)hipFree(p);
    hbm_in_use -= bytes;
}

int hp3d_pm_ctx::grow(DevBuf& b, size_t bytes, const char* what)
{
    if (b.bytes >= bytes) return HP3D_OK;
    dev_free(b.ptr, b.bytes);
    b.ptr = nullptr;
    b.bytes = 0;
    int rc = dev_alloc(&b.ptr, bytes, what);
    if (rc == HP3D_OK) b.bytes = bytes;
    return rc;
}

const Image* hp3d_pm_ctx::find(hp3d_img h) const
{
    auto it = images.find(h);
    return it == images.end() ? nullptr : &it->second;
}

// ---------------------------------------------------------------------------
// Validation and device-view assembly shared by every solve entry point
// ---------------------------------------------------------------------------
int hp3d::validate_desc(hp3d_pm_ctx* ctx, const hp3d_view_desc* d, const Image** ref,
                        const Image* srcs[kMaxSrc])
{
    if (!d) return ctx->fail(HP3D_ERR_INVALID_ARG, "view descriptor is NULL");
    if (d->num_src < 1 || d->num_src > kMaxSrc)
        return ctx->fail(HP3D_ERR_INVALID_ARG, "num_src %d outside 1..%d (a view without sources "
                         "is skipped by the caller, patchmatch_mvs.cu:1089-1092)", d->num_src, kMaxSrc);
    if (!window_supported(d->radius, d->stride))
        return ctx->fail(HP3D_ERR_UNSUPPORTED, "window radius %d stride %d not supported "
                         "(radius 1..5, stride 1 or 2)", d->radius, d->stride);
    if (d->top_k < 1 || d->top_k > kMaxTopK)
        return ctx->fail(HP3D_ERR_INVALID_ARG, "top_k %d outside 1..%d", d->top_k, kMaxTopK);
    if (d->iterations < 0 || d->iterations > kMaxIterations)
        return ctx->fail(HP3D_ERR_INVALID_ARG, "iterations %d outside 0..2^20", d->iterations);
    if (d->first_iteration < 0 || d->first_iteration > kMaxIterations)
        return ctx->fail(HP3D_ERR_INVALID_ARG, "first_iteration %d outside 0..2^20",
                         d->first_iteration);
    if (!(d->depth_min < d->depth_max) || !std::isfinite(d->depth_min) ||
        !std::isfinite(d->depth_max))
        return ctx->fail(HP3D_ERR_INVALID_ARG, "depth range [%g, %g] is empty", d->depth_min,
                         d->depth_max);
    if ((d->init_depth == nullptr) != (d->init_normal == nullptr))
        return ctx->fail(HP3D_ERR_INVALID_ARG, "init_depth and init_normal must be given together");
    if (d->num_seeds < 0 || (d->num_seeds > 0 && !d->seeds))
        return ctx->fail(HP3D_ERR_INVALID_ARG, "num_seeds %d without a seed array", d->num_seeds);
    if (d->num_seeds > 0 && (d->seed_radius < 0 || d->seed_radius > kMaxSeedRadius))
        return ctx->fail(HP3D_ERR_INVALID_ARG, "seed_radius %d outside 0..%d", d->seed_radius,
                         kMaxSeedRadius);
    for (int k = 0; k < d->num_seeds; ++k) {
        const hp3d_seed_point& s = d->seeds[k];
        if (s.x < -kMaxSeedCoord || s.x > kMaxSeedCoord || s.y < -kMaxSeedCoord ||
            s.y > kMaxSeedCoord || !std::isfinite(s.depth) || !std::isfinite(s.normal[0]) ||
            !std::isfinite(s.normal[1]) || !std::isfinite(s.normal[2]))
            return ctx->fail(HP3D_ERR_INVALID_ARG, "seed %d: pixel (%d, %d) out of range or "
                             "non-finite hypothesis", k, s.x, s.y);
    }
    *ref = ctx->find(d->ref);
    if (!*ref) return ctx->fail(HP3D_ERR_NOT_FOUND, "unknown reference image handle %llu",
                                (unsigned long long)d->ref);
    for (int s = 0; s < d->num_src; ++s) {
        srcs[s] = ctx->find(d->src[s]);
        if (!srcs[s])
            return ctx->fail(HP3D_ERR_NOT_FOUND, "unknown source image handle %llu",
                             (unsigned long long)d->src[s]);
        if (srcs[s]->W != (*ref)->W || srcs[s]->H != (*ref)->H)
            return ctx->fail(HP3D_ERR_INVALID_ARG, "source %d is %dx%d, reference is %dx%d "
                             "(all views of a job share one size)", s, srcs[s]->W, srcs[s]->H,
                             (*ref)->W, (*ref)->H);
        if (srcs[s]->store != (*ref)->store)
            return ctx->fail(HP3D_ERR_INVALID_ARG, "source %d storage differs from the reference", s);
    }
    const int need = 2 * d->radius + 1;
    if ((*ref)->W < need || (*ref)->H < need)
        return ctx->fail(HP3D_ERR_INVALID_ARG, "image %dx%d smaller than the %dx%d window",
                         (*ref)->W, (*ref)->H, need, need);
    for (int i = 0; i < 9; ++i)
        if (!std::isfinite(d->K_ref_inv[i]))
            return ctx->fail(HP3D_ERR_INVALID_ARG, "K_ref_inv is not finite");
    for (int s = 0; s < d->num_src; ++s)
        for (int i = 0; i < 12; ++i)
            if (!std::isfinite(i < 9 ? d->A[s][i] : d->b[s][i - 9]))
                return ctx->fail(HP3D_ERR_INVALID_ARG, "source %d: A or b is not finite", s);
    return HP3D_OK;
}

void hp3d::fill_dev_view(DevView& v, const hp3d_view_desc* d, const Image* ref,
                         const Image* const* srcs)
{
    std::memset(&v, 0, sizeof v);
    v.ref_plain = ref->plain;
    v.pitch_plain = ref->pitch_plain;
    v.plain_bytes = (int)ref->plain_bytes;
    v.coef_bytes = (int)ref->coef_bytes;
    for (int i = 0; i < 9; ++i) {
        v.kinv[i] = (float)d->K_ref_inv[i];
        v.kinv_d[i] = d->K_ref_inv[i];
    }
    for (int s = 0; s < d->num_src; ++s) {
        v.src_plain[s] = srcs[s]->plain;
        v.src_coef[s] = srcs[s]->coef;
        for (int i = 0; i < 9; ++i) v.src[s].A[i] = d->A[s][i];
        for (int i = 0; i < 3; ++i) v.src[s].b[i] = d->b[s][i];
    }
    v.W = ref->W;
    v.H = ref->H;
    v.S = d->num_src;
    v.Wc = (ref->W + 1) / 2;
    v.dmin = d->depth_min;
    v.dmax = d->depth_max;
    // the reference's den < 1e-6 guard is in gray^2 units (gray = DN/255,
    // patchmatch_mvs.cu:317); planes hold DN, so scale by 255^2
    v.den_eps = ref->den_scale * 1e-6f;
    v.top_k = d->top_k;
    v.first_iteration = d->first_iteration;
}

double hp3d::algorithmic_taps(int W, int H, int S, int radius, int stride, int iterations)
{
    const double n = (2 * radius) / stride + 1;
    const double k = n * n;
    return (double)W * H * (S * k + iterations * 6.0 * S * k);
}

uint64_t hp3d::pass_seed(uint64_t seed, int iteration, int parity)
{
    return seed + 100ull * (uint64_t)iteration + (parity ? 50ull : 0ull);
}

// ---------------------------------------------------------------------------
// Event timing
// ---------------------------------------------------------------------------
void hp3d_pm_ctx::mark(int phase)
{
    if (!(flags & HP3D_CTX_TIMING)) return;
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    if (hipEventRecord(e, stream) != hipSuccess) {
        (void)hipEventDestroy(e);
        return;
    }
    marks.push_back({phase, e});
}

void hp3d_pm_ctx::collect_marks(int n_prop)
{
    stats.n_propagate = n_prop;
    if (!(flags & HP3D_CTX_TIMING) || marks.size() < 2) {
        for (auto& m : marks) (void)hipEventDestroy(m.ev);
        marks.clear();
        return;
    }
    stats.ms_init = stats.ms_cost = stats.ms_propagate = stats.ms_finalize = 0.0;
    for (size_t i = 1; i < marks.size(); ++i) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, marks[i - 1].ev, marks[i].ev) != hipSuccess) continue;
        switch (marks[i].phase) {
            case kPhaseInit: stats.ms_init += ms; break;
            case kPhaseCost: stats.ms_cost += ms; break;
            case kPhasePropagate: stats.ms_propagate += ms; break;
            case kPhaseFinalize: stats.ms_finalize += ms; break;
            default: break;
        }
    }
    float total = 0.f;
    if (hipEventElapsedTime(&total, marks.front().ev, marks.back().ev) == hipSuccess)
        stats.ms_total = total;
    for (auto& m : marks) (void)hipEventDestroy(m.ev);
    marks.clear();
}

// ---------------------------------------------------------------------------
// Public API
// ---------------------------------------------------------------------------
extern "C" {

const char* hp3d_pm_version(void) { return "hp3d-pm 0.1 (gfx950)"; }

int hp3d_pm_create(const hp3d_pm_opts* opts, hp3d_pm_ctx** out)
{
    if (!out) return HP3D_ERR_INVALID_ARG;
    *out = nullptr;
    hp3d_pm_opts o{};
    if (opts) o = *opts;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return HP3D_ERR_DEVICE;
    if (o.device < 0 || o.device >= ndev) return HP3D_ERR_INVALID_ARG;
    if (hipSetDevice(o.device) != hipSuccess) return HP3D_ERR_DEVICE;
    auto* ctx = new hp3d_pm_ctx();
    ctx->device = o.device;
    ctx->flags = o.flags;
    if (o.stream) {
        ctx->stream = (hipStream_t)o.stream;
        ctx->own_stream = false;
    } else {
        if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
            delete ctx;
            return HP3D_ERR_DEVICE;
        }
        ctx->own_stream = true;
    }
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
        if (ctx->own_stream) (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return HP3D_ERR_DEVICE;
    }
    ctx->hbm_budget = o.hbm_budget ? o.hbm_budget : (size_t)(0.9 * (double)free_b);
    noalloc_setup(o.device);
    *out = ctx;
    return HP3D_OK;
}

int hp3d_pm_destroy(hp3d_pm_ctx* ctx)
{
    if (!ctx) return HP3D_ERR_INVALID_ARG;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ctx->release_all();
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return HP3D_OK;
}

const char* hp3d_pm_last_error(const hp3d_pm_ctx* ctx)
{
    return ctx ? ctx->last_error.c_str() : "NULL context";
}

int hp3d_pm_image_upload(hp3d_pm_ctx* ctx, const void* data, int W, int H, size_t pitch, int fmt,
                         int store, hp3d_img* out)
{
    if (!ctx) return HP3D_ERR_INVALID_ARG;
    if (!data || !out) return ctx->fail(HP3D_ERR_INVALID_ARG, "image upload: NULL pointer");
    if (W < 2 || H < 2) return ctx->fail(HP3D_ERR_INVALID_ARG, "image %dx%d too small", W, H);
    if (fmt != HP3D_FMT_U8 && fmt != HP3D_FMT_F32)
        return ctx->fail(HP3D_ERR_INVALID_ARG, "unknown pixel format %d", fmt);
    if (store != HP3D_STORE_F32 && store != HP3D_STORE_F16)
        return ctx->fail(HP3D_ERR_INVALID_ARG, "unknown storage %d", store);
    const size_t row = (size_t)W * (fmt == HP3D_FMT_U8 ? 1 : 4);
    if (pitch == 0) pitch = row;
    if (pitch < row) return ctx->fail(HP3D_ERR_INVALID_ARG, "pitch %zu < row bytes %zu", pitch, row);
    if (fmt == HP3D_FMT_F32 && pitch % 4 != 0)
        return ctx->fail(HP3D_ERR_INVALID_ARG, "float image pitch %zu is not a multiple of 4", pitch);
    (void)hipSetDevice(ctx->device);

    Image img;
    img.W = W;
    img.H = H;
    img.store = store;
    img.den_scale = 255.0f * 255.0f;
    // taps address a plane with 32-bit buffer offsets and 24-bit row offsets
    if (!image_dims_supported(W, H, store))
        return ctx->fail(HP3D_ERR_UNSUPPORTED,
                         "image %dx%d: plane above 2 GiB - 64 KiB or row pitch of 2^24 elements",
                         W, H);
    image_plane_sizes(W, H, store, &img.pitch_plain, &img.plain_bytes, &img.coef_bytes);
    int rc = ctx->dev_alloc(&img.plain, img.plain_bytes, "image plane");
    if (rc) return rc;
    rc = ctx->dev_alloc(&img.coef, img.coef_bytes, "image row-difference plane");
    if (rc) { ctx->dev_free(img.plain, img.plain_bytes); return rc; }
    auto release = [&] {
        ctx->dev_free(img.plain, img.plain_bytes);
        ctx->dev_free(img.coef, img.coef_bytes);
    };

    // stage the host rows, with their pitch, on the device in one copy (a
    // 2-D copy from pageable memory goes row by row), then convert in one
    // kernel that reads the pitched rows
    const size_t stage_bytes = pitch * (size_t)(H - 1) + row;
    rc = ctx->grow(ctx->staging, stage_bytes, "upload staging");
    if (rc) { release(); return rc; }
    void* stage = ctx->staging.ptr;
    hipError_t e = hipMemcpyAsync(stage, data, stage_bytes, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
        e = launch_img_prep(fmt == HP3D_FMT_U8 ? (const uint8_t*)stage : nullptr,
                            fmt == HP3D_FMT_F32 ? (const float*)stage : nullptr, pitch, W, H,
                            store, img.plain, img.coef, img.pitch_plain, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        release();
        return ctx->hip_fail(e, "image upload");
    }
    const hp3d_img h = ++ctx->next_handle;
    ctx->images[h] = img;
    *out = h;
    return HP3D_OK;
}

int hp3d_pm_image_release(hp3d_pm_ctx* ctx, hp3d_img h)
{
    if (!ctx) return HP3D_ERR_INVALID_ARG;
    auto it = ctx->images.find(h);
    if (it == ctx->images.end())
        return ctx->fail(HP3D_ERR_NOT_FOUND, "unknown image handle %llu", (unsigned long long)h);
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ctx->dev_free(it->second.plain, it->second.plain_bytes);
    ctx->dev_free(it->second.coef, it->second.coef_bytes);
    ctx->images.erase(it);
    return HP3D_OK;
}

int hp3d_pm_image_info(const hp3d_pm_ctx* ctx, hp3d_img h, int* W, int* H, int* store,
                       size_t* bytes)
{
    if (!ctx) return HP3D_ERR_INVALID_ARG;
    const Image* img = ctx->find(h);
    if (!img) return HP3D_ERR_NOT_FOUND;
    if (W) *W = img->W;
    if (H) *H = img->H;
    if (store) *store = img->store;
    if (bytes) *bytes = img->plain_bytes + img->coef_bytes;
    return HP3D_OK;
}

int hp3d_pm_invert3(const double m[9], double inv[9])
{
    if (!m || !inv) return HP3D_ERR_INVALID_ARG;
    const double det = m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
                       m[2] * (m[3] * m[7] - m[4] * m[6]);
    if (det == 0.0 || !std::isfinite(det)) return HP3D_ERR_INVALID_ARG;
    const double id = 1.0 / det;
    inv[0] = (m[4] * m[8] - m[5] * m[7]) * id;
    inv[1] = (m[2] * m[7] - m[1] * m[8]) * id;
    inv[2] = (m[1] * m[5] - m[2] * m[4]) * id;
    inv[3] = (m[5] * m[6] - m[3] * m[8]) * id;
    inv[4] = (m[0] * m[8] - m[2] * m[6]) * id;
    inv[5] = (m[2] * m[3] - m[0] * m[5]) * id;
    inv[6] = (m[3] * m[7] - m[4] * m[6]) * id;
    inv[7] = (m[1] * m[6] - m[0] * m[7]) * id;
    inv[8] = (m[0] * m[4] - m[1] * m[3]) * id;
    return HP3D_OK;
}

int hp3d_pm_compose(const double Kr[9], const double Rr[9], const double tr[3], const double Ks[9],
                    const double Rs[9], const double ts[3], double A[9], double b[3])
{
    if (!Kr || !Rr || !tr || !Ks || !Rs || !ts || !A || !b) return HP3D_ERR_INVALID_ARG;
    double Kri[9];
    if (hp3d_pm_invert3(Kr, Kri) != HP3D_OK) return HP3D_ERR_INVALID_ARG;
    double Rrel[9], M[9];
    for (int r = 0; r < 3; ++r)            // R_rel = R_s R_ref^T
        for (int c = 0; c < 3; ++c)
            Rrel[r * 3 + c] = Rs[r * 3] * Rr[c * 3] + Rs[r * 3 + 1] * Rr[c * 3 + 1] +
                              Rs[r * 3 + 2] * Rr[c * 3 + 2];
    double trel[3];
    for (int r = 0; r < 3; ++r)            // t_rel = t_s - R_rel t_ref
        trel[r] = ts[r] - (Rrel[r * 3] * tr[0] + Rrel[r * 3 + 1] * tr[1] + Rrel[r * 3 + 2] * tr[2]);
    for (int r = 0; r < 3; ++r)            // M = R_rel K_ref^-1
        for (int c = 0; c < 3; ++c)
            M[r * 3 + c] = Rrel[r * 3] * Kri[c] + Rrel[r * 3 + 1] * Kri[3 + c] +
                           Rrel[r * 3 + 2] * Kri[6 + c];
    for (int r = 0; r < 3; ++r)            // A = K_s M, b = K_s t_rel
        for (int c = 0; c < 3; ++c)
            A[r * 3 + c] = Ks[r * 3] * M[c] + Ks[r * 3 + 1] * M[3 + c] + Ks[r * 3 + 2] * M[6 + c];
    for (int r = 0; r < 3; ++r)
        b[r] = Ks[r * 3] * trel[0] + Ks[r * 3 + 1] * trel[1] + Ks[r * 3 + 2] * trel[2];
    return HP3D_OK;
}

int hp3d_pm_run_view(hp3d_pm_ctx* ctx, const hp3d_view_desc* d, const hp3d_view_out* out)
{
    if (!ctx) return HP3D_ERR_INVALID_ARG;
    if (!out || !out->depth) return ctx->fail(HP3D_ERR_INVALID_ARG, "output depth pointer is NULL");
    const Image* ref = nullptr;
    const Image* srcs[kMaxSrc] = {};
    int rc = validate_desc(ctx, d, &ref, srcs);
    if (rc) return rc;
    (void)hipSetDevice(ctx->device);
    return solve_view(ctx, d, ref, srcs, out);
}

int hp3d_pm_run_views(hp3d_pm_ctx* ctx, const hp3d_view_desc* d, int n, const hp3d_view_out* out)
{
    if (!ctx) return HP3D_ERR_INVALID_ARG;
    if (!d || !out || n < 1) return ctx->fail(HP3D_ERR_INVALID_ARG, "run_views: no views");
    std::vector<const Image*> refs(n, nullptr);
    std::vector<std::array<const Image*, kMaxSrc>> srcs(n);
    for (int i = 0; i < n; ++i) {
        if (!out[i].depth)
            return ctx->fail(HP3D_ERR_INVALID_ARG, "run_views: view %d output depth pointer is NULL", i);
        srcs[i].fill(nullptr);
        int rc = validate_desc(ctx, &d[i], &refs[i], srcs[i].data());
        if (rc) return ctx->fail(rc, "run_views view %d: %s", i, ctx->last_error.c_str());
    }
    std::vector<const Image* const*> sp(n);
    for (int i = 0; i < n; ++i) sp[i] = srcs[i].data();
    (void)hipSetDevice(ctx->device);
    return solve_views(ctx, n, d, refs.data(), sp.data(), out);
}

int hp3d_pm_host_alloc(hp3d_pm_ctx* ctx, size_t bytes, void** ptr)
{
    if (!ctx) return HP3D_ERR_INVALID_ARG;
    if (!ptr || bytes == 0) return ctx->fail(HP3D_ERR_INVALID_ARG, "host_alloc: bad arguments");
    *ptr = nullptr;
    (void)hipSetDevice(ctx->device);
    hipError_t e = hipHostMalloc(ptr, bytes, hipHostMallocDefault);
    if (e != hipSuccess) {
        *ptr = nullptr;
        return ctx->hip_fail(e, "host_alloc");
    }
    ctx->host_bufs[*ptr] = bytes;
    return HP3D_OK;
}

int hp3d_pm_host_free(hp3d_pm_ctx* ctx, void* ptr)
{
    if (!ctx) return HP3D_ERR_INVALID_ARG;
    auto it = ctx->host_bufs.find(ptr);
    if (it == ctx->host_bufs.end())
        return ctx->fail(HP3D_ERR_NOT_FOUND, "host_free: %p was not allocated by this context", ptr);
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->copy_stream) (void)hipStreamSynchronize(ctx->copy_stream);
    (void)hipHostFree(ptr);
    ctx->host_bufs.erase(it);
    return HP3D_OK;
}

int hp3d_pm_init_hypotheses(hp3d_pm_ctx* ctx, const hp3d_view_desc* d, float* depth,
                            float* normal)
{
    if (!ctx) return HP3D_ERR_INVALID_ARG;
    if (!depth || !normal) return ctx->fail(HP3D_ERR_INVALID_ARG, "NULL output");
    const Image* ref = nullptr;
    const Image* srcs[kMaxSrc] = {};
    int rc = validate_desc(ctx, d, &ref, srcs);
    if (rc) return rc;
    (void)hipSetDevice(ctx->device);
    return init_only(ctx, d, ref, srcs, depth, normal);
}

int hp3d_pm_cost_eval(hp3d_pm_ctx* ctx, const hp3d_view_desc* d, const float* depth,
                      const float* normal, float* cost)
{
    if (!ctx) return HP3D_ERR_INVALID_ARG;
    if (!depth || !normal || !cost) return ctx->fail(HP3D_ERR_INVALID_ARG, "NULL buffer");
    const Image* ref = nullptr;
    const Image* srcs[kMaxSrc] = {};
    int rc = validate_desc(ctx, d, &ref, srcs);
    if (rc) return rc;
    (void)hipSetDevice(ctx->device);
    return cost_only(ctx, d, ref, srcs, depth, normal, cost);
}

int hp3d_pm_get_stats(const hp3d_pm_ctx* ctx, hp3d_pm_stats* out)
{
    if (!ctx || !out) return HP3D_ERR_INVALID_ARG;
    *out = ctx->stats;
    out->hbm_in_use = ctx->hbm_in_use;
    out->hbm_budget = ctx->hbm_budget;
    return HP3D_OK;
}

int hp3d_pm_device_alloc(hp3d_pm_ctx* ctx, size_t bytes, void** dptr)
{
    if (!ctx) return HP3D_ERR_INVALID_ARG;
    if (!dptr || bytes == 0) return ctx->fail(HP3D_ERR_INVALID_ARG, "device_alloc: bad arguments");
    (void)hipSetDevice(ctx->device);
    int rc = ctx->dev_alloc(dptr, bytes, "device buffer");
    if (rc == HP3D_OK) ctx->user_bufs[*dptr] = bytes;
    return rc;
}

int hp3d_pm_device_free(hp3d_pm_ctx* ctx, void* dptr)
{
    if (!ctx) return HP3D_ERR_INVALID_ARG;
    auto it = ctx->user_bufs.find(dptr);
    if (it == ctx->user_bufs.end())
        return ctx->fail(HP3D_ERR_NOT_FOUND, "device_free: %p was not allocated by this context", dptr);
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ctx->dev_free(it->first, it->second);
    ctx->user_bufs.erase(it);
    return HP3D_OK;
}

int hp3d_pm_device_copy(hp3d_pm_ctx* ctx, void* dst, const void* src, size_t bytes, int kind)
{
    if (!ctx) return HP3D_ERR_INVALID_ARG;
    if (!dst || !src || kind < 0 || kind > 2)
        return ctx->fail(HP3D_ERR_INVALID_ARG, "device_copy: bad arguments");
    static const hipMemcpyKind kinds[3] = {hipMemcpyHostToDevice, hipMemcpyDeviceToHost,
                                           hipMemcpyDeviceToDevice};
    (void)hipSetDevice(ctx->device);
    hipError_t e = hipMemcpyAsync(dst, src, bytes, kinds[kind], ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return e == hipSuccess ? HP3D_OK : ctx->hip_fail(e, "device_copy");
}

}  // extern "C"
