// pm_runtime.h -- host-side context and helpers behind the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/hp3d_pm.h"
#include "pm_device.h"
#include "pm_launch.h"

namespace hp3d {

// One uploaded image: its plane of DN - 128 values and the row-difference gather
// plane of source taps (pm_pixel.h).
struct Image {
    int W = 0, H = 0, store = 0;
    int pitch_plain = 0;                   // elements, both planes
    void* plain = nullptr;
    void* coef = nullptr;
    size_t plain_bytes = 0, coef_bytes = 0;
    float den_scale = 1.0f;                // stored-unit^2 per gray^2
};

// Plane sizes of a W x H image with the given storage.
inline void image_plane_sizes(int W, int H, int store, int* pitch, size_t* plain_bytes,
                              size_t* coef_bytes)
{
    *pitch = (W + 63) & ~63;
    *plain_bytes = (size_t)*pitch * H * (store == HP3D_STORE_F32 ? 4 : 2);
    *coef_bytes = (size_t)*pitch * H * (store == HP3D_STORE_F32 ? 8 : 4);   // row differences
}

// True if a W x H image with this storage can be addressed by the source
// taps: row pitch below 2^24 elements and both planes at most kMaxPlaneBytes.
inline bool image_dims_supported(int W, int H, int store)
{
    if (W < 1 || H < 1 || W > kMaxRowPitch - 64) return false;
    int pitch;
    size_t plain_bytes, coef_bytes;
    image_plane_sizes(W, H, store, &pitch, &plain_bytes, &coef_bytes);
    return pitch < kMaxRowPitch && plain_bytes <= kMaxPlaneBytes && coef_bytes <= kMaxPlaneBytes;
}

struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
};

enum Phase : int { kPhaseStart = 0, kPhaseInit, kPhaseCost, kPhasePropagate, kPhaseFinalize };

struct Mark {
    int phase;
    hipEvent_t ev;
};

}  // namespace hp3d

struct hp3d_pm_ctx {
    int device = 0;
    int flags = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    size_t hbm_budget = 0, hbm_in_use = 0;
    std::string last_error;
    std::unordered_map<hp3d_img, hp3d::Image> images;
    hp3d_img next_handle = 0;
    std::unordered_map<void*, size_t> user_bufs;   // hp3d_pm_device_alloc
    std::unordered_map<void*, size_t> host_bufs;   // hp3d_pm_host_alloc (pinned)

    // multi-view pipeline (hp3d_pm_run_views): a second stream for result
    // downloads, two output staging slots, and per slot the events "finalized
    // into the slot" and "slot downloaded"
    hipStream_t copy_stream = nullptr;
    hp3d::DevBuf out_slot[2];
    hipEvent_t ev_fin[2] = {nullptr, nullptr}, ev_copied[2] = {nullptr, nullptr};

    // per-view working set, grown on demand and reused across views
    hp3d::DevBuf hyp, cost, chg, out_depth, out_normal, out_cost, in_depth, in_normal, staging;
    hp3d::DevBuf seed_pts, seed_owner;

    // batched form: state arena, job table, step items, captured graph
    hp3d::DevBuf batch_state, batch_jobs, batch_items, batch_done;
    hipGraphExec_t batch_graph = nullptr;
    std::vector<unsigned char> batch_graph_key;
    std::map<int, int> tuned;   // (radius, stride, store) key -> variant

    hp3d_pm_stats stats{};
    std::vector<hp3d::Mark> marks;

    int fail(int code, const char* fmt, ...) __attribute__((format(printf, 3, 4)));
    int hip_fail(hipError_t e, const char* what);
    int dev_alloc(void** p, size_t bytes, const char* what);
    void dev_free(void* p, size_t bytes);
    int grow(hp3d::DevBuf& b, size_t bytes, const char* what);
    const hp3d::Image* find(hp3d_img h) const;
    void mark(int phase);
    void collect_marks(int n_prop);

    void release_buf(hp3d::DevBuf& b)
    {
        dev_free(b.ptr, b.bytes);
        b.ptr = nullptr;
        b.bytes = 0;
    }
    void release_all()
    {
        for (auto& kv : images) {
            dev_free(kv.second.plain, kv.second.plain_bytes);
            dev_free(kv.second.coef, kv.second.coef_bytes);
        }
        images.clear();
        for (auto& kv : user_bufs) dev_free(kv.first, kv.second);
        user_bufs.clear();
        for (auto& kv : host_bufs) (void)hipHostFree(kv.first);
        host_bufs.clear();
        if (copy_stream) {
            (void)hipStreamSynchronize(copy_stream);
            (void)hipStreamDestroy(copy_stream);
            copy_stream = nullptr;
        }
        for (int k = 0; k < 2; ++k) {
            release_buf(out_slot[k]);
            if (ev_fin[k]) (void)hipEventDestroy(ev_fin[k]);
            if (ev_copied[k]) (void)hipEventDestroy(ev_copied[k]);
            ev_fin[k] = ev_copied[k] = nullptr;
        }
        if (batch_graph) {
            (void)hipGraphExecDestroy(batch_graph);
            batch_graph = nullptr;
        }
        for (hp3d::DevBuf* b : {&hyp, &cost, &chg, &out_depth, &out_normal, &out_cost, &in_depth,
                                &in_normal, &staging, &seed_pts, &seed_owner, &batch_state,
                                &batch_jobs, &batch_items, &batch_done})
            release_buf(*b);
    }
};

namespace hp3d {

int validate_desc(hp3d_pm_ctx* ctx, const hp3d_view_desc* d, const Image** ref,
                  const Image* srcs[kMaxSrc]);
void fill_dev_view(DevView& v, const hp3d_view_desc* d, const Image* ref, const Image* const* srcs);
double algorithmic_taps(int W, int H, int S, int radius, int stride, int iterations);
uint64_t pass_seed(uint64_t seed, int iteration, int parity);

int solve_view(hp3d_pm_ctx* ctx, const hp3d_view_desc* d, const Image* ref,
               const Image* const* srcs, const hp3d_view_out* out);
// Views solved back to back; host outputs of view i are downloaded on the
// copy stream while view i + 1 computes.  srcs[i] holds view i's sources.
int solve_views(hp3d_pm_ctx* ctx, int n, const hp3d_view_desc* d, const Image* const* refs,
                const Image* const* const* srcs, const hp3d_view_out* out);
// Resolve this device's copy of the no-allocation scratch array (once per device).
void noalloc_setup(int device);

int init_only(hp3d_pm_ctx* ctx, const hp3d_view_desc* d, const Image* ref,
              const Image* const* srcs, float* depth, float* normal);
int cost_only(hp3d_pm_ctx* ctx, const hp3d_view_desc* d, const Image* ref,
              const Image* const* srcs, const float* depth, const float* normal, float* cost);

}  // namespace hp3d
