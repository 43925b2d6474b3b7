// pm_noalloc.hip -- hp3d_pm_run_view_noalloc: the per-view solve for callers
// that cannot allocate device memory (code under hipGraph capture, serving
// loops that forbid hipMalloc).
//
// hp3d_pm_run_view keeps a view's hypothesis state (two colour planes of
// float4 hypotheses, two cost planes, two change-flag planes: 21 B per
// pixel) in context buffers grown with hipMalloc.  This entry point keeps
// that state in `scratch`, a file-scope device array of fixed size that
// every GPU the library runs on holds until the process ends, and calls no
// allocating routine: the caller passes the images (uploaded earlier) and
// device output buffers.  Nothing in the array has to start at zero -- the
// init writes every hypothesis, the initial cost pass every cost, and every
// half-pass writes its colour's change flags before any pass reads them --
// so successive calls need no clearing.  Calls that overlap on one GPU
// (several contexts or threads at once) share the array and race.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>

#include "pm_launch.h"
#include "pm_runtime.h"

namespace hp3d {

constexpr size_t kScratchBytes = HP3D_PM_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
// per device: address of that device's copy of `scratch` (the array's name
// in host code is only a stand-in, not a device address), and the status of
// looking it up
void* g_ptr[kMaxDevices] = {};
hipError_t g_status[kMaxDevices] = {};
bool g_done[kMaxDevices] = {};

size_t a256(size_t n) { return (n + 255) & ~(size_t)255; }

}  // namespace

// Called by hp3d_pm_create with the context's device current: resolves the
// device's copy of the array once per device.
void noalloc_setup(int device)
{
    if (device < 0 || device >= kMaxDevices) return;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_done[device]) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_ptr[device] = e == hipSuccess ? p : nullptr;
    g_status[device] = e;
    g_done[device] = true;
}

size_t noalloc_state_bytes(int W, int H)
{
    const size_t plane = (size_t)H * ((W + 1) / 2);
    return a256(2 * plane * sizeof(float4)) + a256(2 * plane * sizeof(float)) + a256(2 * plane);
}

}  // namespace hp3d

using namespace hp3d;

extern "C" int hp3d_pm_run_view_noalloc(hp3d_pm_ctx* ctx, const hp3d_view_desc* d,
                                        const hp3d_view_out* out)
{
    if (!ctx) return HP3D_ERR_INVALID_ARG;
    if (!out || !out->depth || !out->on_device)
        return ctx->fail(HP3D_ERR_INVALID_ARG, "run_view_noalloc: outputs must be device "
                         "buffers (on_device = 1) with a depth pointer");
    const Image* ref = nullptr;
    const Image* srcs[kMaxSrc] = {};
    int rc = validate_desc(ctx, d, &ref, srcs);
    if (rc) return rc;
    if (d->init_depth || d->num_seeds > 0)
        return ctx->fail(HP3D_ERR_INVALID_ARG, "run_view_noalloc: caller hypotheses and SfM "
                         "seeds need staging memory; use hp3d_pm_run_view");
    const size_t need = noalloc_state_bytes(ref->W, ref->H);
    if (need > kScratchBytes)
        return ctx->fail(HP3D_ERR_UNSUPPORTED, "run_view_noalloc: a %dx%d view needs %zu bytes "
                         "of state, above the %zu-byte scratch array (about %zu pixels)",
                         ref->W, ref->H, need, kScratchBytes, kScratchBytes / 21);
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (ctx->device >= 0 && ctx->device < kMaxDevices && g_done[ctx->device]) {
            if (g_status[ctx->device] != hipSuccess)
                return ctx->hip_fail(g_status[ctx->device], "run_view_noalloc: scratch lookup");
            base = g_ptr[ctx->device];
        }
    }
    if (!base)
        return ctx->fail(HP3D_ERR_DEVICE, "run_view_noalloc: no scratch array on device %d",
                         ctx->device);
    (void)hipSetDevice(ctx->device);

    DevView v;
    fill_dev_view(v, d, ref, srcs);
    const size_t plane = (size_t)v.H * v.Wc;
    char* p = static_cast<char*>(base);
    float4* hyp = reinterpret_cast<float4*>(p);
    p += a256(2 * plane * sizeof(float4));
    float* cost = reinterpret_cast<float*>(p);
    p += a256(2 * plane * sizeof(float));
    unsigned char* chg = reinterpret_cast<unsigned char*>(p);
    v.hyp[0] = hyp;
    v.hyp[1] = hyp + plane;
    v.cost[0] = cost;
    v.cost[1] = cost + plane;
    v.chg[0] = chg;
    v.chg[1] = chg + plane;

    // the launch sequence of hp3d_pm_run_view, without timing events
    hipError_t e = launch_init(v, d->seed, ctx->stream);
    for (int parity = 0; parity < 2 && e == hipSuccess; ++parity)
        e = launch_pass(v, ref->store, d->radius, d->stride, kPassCost, 0, parity, 0, ctx->stream);
    for (int it = d->first_iteration; it < d->first_iteration + d->iterations && e == hipSuccess;
         ++it)
        for (int parity = 0; parity < 2 && e == hipSuccess; ++parity)
            e = launch_pass(v, ref->store, d->radius, d->stride, kPassPropagate, it, parity,
                            pass_seed(d->seed, it, parity), ctx->stream);
    if (e == hipSuccess)
        e = launch_finalize(v, out->depth, out->normal, out->cost, out->confidence, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return e == hipSuccess ? HP3D_OK : ctx->hip_fail(e, "run_view_noalloc");
}
