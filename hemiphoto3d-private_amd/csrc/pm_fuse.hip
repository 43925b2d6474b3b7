// pm_fuse.hip -- depth-map fusion into a point cloud.
//
// Reference: fuseDepthMaps (src/dense_recon/point_cloud.h:7-9) is declared
// as the consumer of the MVS depth and normal maps but is a stub returning an
// empty cloud (point_cloud.cu:4-8).  This is the natural device form of it:
// every pixel with depth > 0 (and confidence >= min_confidence when a
// confidence map is given) becomes one point, in view order then row-major
// pixel order:
//   position = R^T (d K^-1 [x, y, 1] - t)   (world frame, include/types.h:55-74)
//   normal   = R^T n                        (camera -> world)
//   colour   = the pixel's 8-bit value / 255 (gray replicated, or RGB)
// The maps are uploaded once; a count pass, a scan of the per-block counts
// and a write pass place every point at its row-major rank, so the output is
// deterministic.  Single-precision operations are written in one fixed order
// (no contraction), which the oracle restates.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "pm_runtime.h"

namespace hp3d {

namespace {

constexpr int kFuseBlock = 256;

struct FuseViewDev {
    float kinv[9], R[9], t[3];
    int W, H, channels;
    float min_conf;
    const float* depth;
    const float* normal;
    const float* conf;
    const uint8_t* color;
};

__device__ __forceinline__ bool fuse_valid(const FuseViewDev& g, int q)
{
    const float d = g.depth[q];
    return d > 0.0f && (!g.conf || g.conf[q] >= g.min_conf);
}

__global__ __launch_bounds__(kFuseBlock) void fuse_count_kernel(FuseViewDev g, int* counts)
{
    const int q = blockIdx.x * kFuseBlock + threadIdx.x;
    const bool v = q < g.W * g.H && fuse_valid(g, q);
    __shared__ int part[kFuseBlock / 64];
    const unsigned long long m = __ballot(v);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        int s = 0;
        for (int w = 0; w < kFuseBlock / 64; ++w) s += part[w];
        counts[blockIdx.x] = s;
    }
}

// Exclusive scan of n block counts in place (one workgroup; n is at most a
// few hundred thousand), total into *total.
__global__ __launch_bounds__(1024) void fuse_scan_kernel(int* counts, int n, long long base,
                                                         long long* offsets, long long* total)
{
    __shared__ long long carry;
    __shared__ long long part[1024 / 64];
    if (threadIdx.x == 0) carry = base;
    __syncthreads();
    for (int start = 0; start < n; start += 1024) {
        const int i = start + threadIdx.x;
        const long long c = i < n ? counts[i] : 0;
        // inclusive scan within the wave
        long long x = c;
        const int lane = threadIdx.x & 63;
        for (int o = 1; o < 64; o <<= 1) {
            const long long y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (lane == 63) part[threadIdx.x >> 6] = x;
        __syncthreads();
        if (threadIdx.x == 0) {
            long long s = 0;
            for (int w = 0; w < 1024 / 64; ++w) {
                const long long p = part[w];
                part[w] = s;
                s += p;
            }
        }
        __syncthreads();
        if (i < n) offsets[i] = carry + part[threadIdx.x >> 6] + x - c;
        __syncthreads();
        if (threadIdx.x == 1023) carry = carry + part[15] + x;
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ __launch_bounds__(kFuseBlock) void fuse_write_kernel(FuseViewDev g,
                                                                const long long* offsets,
                                                                float* pos, float* nrm, float* col)
{
    const int q = blockIdx.x * kFuseBlock + threadIdx.x;
    const bool v = q < g.W * g.H && fuse_valid(g, q);
    __shared__ int part[kFuseBlock / 64];
    const unsigned long long m = __ballot(v);
    const int lane = threadIdx.x & 63;
    if (lane == 0) part[threadIdx.x >> 6] = __popcll(m);
    __syncthreads();
    if (!v) return;
    int rank = __popcll(m & ((1ull << lane) - 1ull));
    for (int w = 0; w < (int)(threadIdx.x >> 6); ++w) rank += part[w];
    const long long o = offsets[blockIdx.x] + rank;
    const int y = q / g.W, x = q - y * g.W;
    const float d = g.depth[q];
    const float fx = (float)x, fy = (float)y;
    const float* ki = g.kinv;
    const float* R = g.R;
    const float pcx = d * (ki[0] * fx + ki[1] * fy + ki[2]);
    const float pcy = d * (ki[3] * fx + ki[4] * fy + ki[5]);
    const float pcz = d * (ki[6] * fx + ki[7] * fy + ki[8]);
    const float a = pcx - g.t[0], b = pcy - g.t[1], c = pcz - g.t[2];
    pos[3 * o] = R[0] * a + R[3] * b + R[6] * c;
    pos[3 * o + 1] = R[1] * a + R[4] * b + R[7] * c;
    pos[3 * o + 2] = R[2] * a + R[5] * b + R[8] * c;
    if (nrm) {
        float nx = 0.0f, ny = 0.0f, nz = 0.0f;
        if (g.normal) {
            nx = g.normal[3 * (size_t)q];
            ny = g.normal[3 * (size_t)q + 1];
            nz = g.normal[3 * (size_t)q + 2];
        }
        nrm[3 * o] = R[0] * nx + R[3] * ny + R[6] * nz;
        nrm[3 * o + 1] = R[1] * nx + R[4] * ny + R[7] * nz;
        nrm[3 * o + 2] = R[2] * nx + R[5] * ny + R[8] * nz;
    }
    if (col) {
        float r = 0.0f, gg = 0.0f, bb = 0.0f;
        if (g.color && g.channels == 1) {
            r = gg = bb = (float)g.color[q] / 255.0f;
        } else if (g.color && g.channels == 3) {
            r = (float)g.color[3 * (size_t)q] / 255.0f;
            gg = (float)g.color[3 * (size_t)q + 1] / 255.0f;
            bb = (float)g.color[3 * (size_t)q + 2] / 255.0f;
        }
        col[3 * o] = r;
        col[3 * o + 1] = gg;
        col[3 * o + 2] = bb;
    }
}

}  // namespace

}  // namespace hp3d

using namespace hp3d;

extern "C" int hp3d_pm_fuse(hp3d_pm_ctx* ctx, const hp3d_fuse_view* views, int n,
                            float min_confidence, float* positions, float* normals,
                            float* colors, size_t capacity, size_t* count)
{
    if (!ctx) return HP3D_ERR_INVALID_ARG;
    if (!views || n < 1 || !count || !(min_confidence == min_confidence))
        return ctx->fail(HP3D_ERR_INVALID_ARG, "fuse: bad arguments");
    for (int i = 0; i < n; ++i) {
        const hp3d_fuse_view& v = views[i];
        if (v.width < 1 || v.height < 1 || !v.depth ||
            (long long)v.width * v.height > (1ll << 31) - kFuseBlock ||
            (v.color && v.color_channels != 1 && v.color_channels != 3))
            return ctx->fail(HP3D_ERR_INVALID_ARG, "fuse: view %d is malformed", i);
    }
    (void)hipSetDevice(ctx->device);
    // one device block per call: maps of every view, block counts, offsets
    std::vector<size_t> off(n);
    size_t total = 0, max_blocks = 0;
    auto a256 = [](size_t b) { return (b + 255) & ~(size_t)255; };
    for (int i = 0; i < n; ++i) {
        const size_t px = (size_t)views[i].width * views[i].height;
        off[i] = total;
        total += a256(px * 4) + (views[i].normal ? a256(px * 12) : 0) +
                 (views[i].confidence ? a256(px * 4) : 0) +
                 (views[i].color ? a256(px * views[i].color_channels) : 0);
        max_blocks = std::max(max_blocks, (px + kFuseBlock - 1) / kFuseBlock);
    }
    const size_t scan_off = total;
    total += a256(max_blocks * 4) + a256(max_blocks * 8) + 256;
    void* block = nullptr;
    int rc = ctx->dev_alloc(&block, total, "fusion maps");
    if (rc) return rc;
    char* base = static_cast<char*>(block);
    int* d_counts = reinterpret_cast<int*>(base + scan_off);
    long long* d_offsets = reinterpret_cast<long long*>(base + scan_off + a256(max_blocks * 4));
    long long* d_total = reinterpret_cast<long long*>(base + scan_off + a256(max_blocks * 4) +
                                                      a256(max_blocks * 8));
    std::vector<FuseViewDev> g(n);
    hipError_t e = hipSuccess;
    for (int i = 0; i < n && e == hipSuccess; ++i) {
        const hp3d_fuse_view& v = views[i];
        const size_t px = (size_t)v.width * v.height;
        char* p = base + off[i];
        FuseViewDev& d = g[i];
        std::memset(&d, 0, sizeof d);
        double ki[9];
        if (hp3d_pm_invert3(v.K, ki) != HP3D_OK) {
            ctx->dev_free(block, total);
            return ctx->fail(HP3D_ERR_INVALID_ARG, "fuse: view %d K is singular", i);
        }
        for (int k = 0; k < 9; ++k) { d.kinv[k] = (float)ki[k]; d.R[k] = (float)v.R[k]; }
        for (int k = 0; k < 3; ++k) d.t[k] = (float)v.t[k];
        d.W = v.width;
        d.H = v.height;
        d.channels = v.color ? v.color_channels : 0;
        d.min_conf = min_confidence;
        d.depth = reinterpret_cast<float*>(p);
        e = hipMemcpyAsync(p, v.depth, px * 4, hipMemcpyHostToDevice, ctx->stream);
        p += a256(px * 4);
        if (v.normal && e == hipSuccess) {
            d.normal = reinterpret_cast<float*>(p);
            e = hipMemcpyAsync(p, v.normal, px * 12, hipMemcpyHostToDevice, ctx->stream);
            p += a256(px * 12);
        }
        if (v.confidence && e == hipSuccess) {
            d.conf = reinterpret_cast<float*>(p);
            e = hipMemcpyAsync(p, v.confidence, px * 4, hipMemcpyHostToDevice, ctx->stream);
            p += a256(px * 4);
        }
        if (v.color && e == hipSuccess) {
            d.color = reinterpret_cast<uint8_t*>(p);
            e = hipMemcpyAsync(p, v.color, px * v.color_channels, hipMemcpyHostToDevice,
                               ctx->stream);
        }
    }
    // pass 1: count every view, in order, to size the output
    std::vector<long long> view_base(n + 1, 0);
    for (int i = 0; i < n && e == hipSuccess; ++i) {
        const int nb = (int)(((size_t)g[i].W * g[i].H + kFuseBlock - 1) / kFuseBlock);
        hipLaunchKernelGGL(fuse_count_kernel, dim3(nb), dim3(kFuseBlock), 0, ctx->stream, g[i],
                           d_counts);
        hipLaunchKernelGGL(fuse_scan_kernel, dim3(1), dim3(1024), 0, ctx->stream, d_counts, nb,
                           0ll, d_offsets, d_total);
        long long t = 0;
        e = hipGetLastError();
        if (e == hipSuccess)
            e = hipMemcpyAsync(&t, d_total, sizeof t, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        view_base[i + 1] = view_base[i] + t;
    }
    if (e != hipSuccess) {
        ctx->dev_free(block, total);
        return ctx->hip_fail(e, "fusion count");
    }
    const size_t npts = (size_t)view_base[n];
    *count = npts;
    if (!positions || npts == 0) {
        ctx->dev_free(block, total);
        return HP3D_OK;
    }
    if (capacity < npts) {
        ctx->dev_free(block, total);
        return ctx->fail(HP3D_ERR_INVALID_ARG, "fuse: %zu points exceed the capacity %zu", npts,
                         capacity);
    }
    // pass 2: write into one device output block, then copy out
    void* out = nullptr;
    const size_t per = npts * 12;
    const int nbuf = 1 + (normals ? 1 : 0) + (colors ? 1 : 0);
    rc = ctx->dev_alloc(&out, per * nbuf, "fusion output");
    if (rc) {
        ctx->dev_free(block, total);
        return rc;
    }
    float* d_pos = static_cast<float*>(out);
    float* d_nrm = normals ? d_pos + 3 * npts : nullptr;
    float* d_col = colors ? d_pos + 3 * npts * (normals ? 2 : 1) : nullptr;
    for (int i = 0; i < n && e == hipSuccess; ++i) {
        const int nb = (int)(((size_t)g[i].W * g[i].H + kFuseBlock - 1) / kFuseBlock);
        hipLaunchKernelGGL(fuse_count_kernel, dim3(nb), dim3(kFuseBlock), 0, ctx->stream, g[i],
                           d_counts);
        hipLaunchKernelGGL(fuse_scan_kernel, dim3(1), dim3(1024), 0, ctx->stream, d_counts, nb,
                           view_base[i], d_offsets, d_total);
        hipLaunchKernelGGL(fuse_write_kernel, dim3(nb), dim3(kFuseBlock), 0, ctx->stream, g[i],
                           d_offsets, d_pos, d_nrm, d_col);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(positions, d_pos, per, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess && normals)
        e = hipMemcpyAsync(normals, d_nrm, per, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess && colors)
        e = hipMemcpyAsync(colors, d_col, per, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    ctx->dev_free(out, per * nbuf);
    ctx->dev_free(block, total);
    if (e != hipSuccess) return ctx->hip_fail(e, "fusion");
    return HP3D_OK;
}
