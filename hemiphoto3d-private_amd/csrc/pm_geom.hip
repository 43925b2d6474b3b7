// pm_geom.hip -- geometric consistency filter over a set of depth maps.
//
// Reference: geometric_consistency_filter_kernel (src/dense_recon/
// patchmatch_mvs.cu:668-781) and its driver (:1267-1384).  The driver walks
// the views in order; view i is filtered against its sources' depth maps as
// they are at that moment, so a source earlier in the order contributes its
// already-filtered map.  Here every map is uploaded once and stays on the
// device; one launch per view in the same order reproduces that dependence,
// and the single-precision operations are written in the reference's order
// (this file is compiled without contraction), so the result is bit-exact
// with an FP32 restatement of the same loop.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "pm_runtime.h"
#include "pm_xcd.h"

namespace hp3d {

struct GeomSrcDev {
    float K[9], R[9], t[3];
    const float* depth;
    const float* normal;
};

struct GeomViewDev {
    float kinv[9], R[9], t[3];
    int W, H, S, min_consistent;
    float tol, cos_tol;
    float* depth;
    const float* normal;
    float* conf;
    GeomSrcDev src[kMaxSrc];
};

// One 64x4-pixel tile per workgroup, tiles in XCD order: the source maps'
// rows that a band of the view projects to stay in one XCD's L2.
__global__ __launch_bounds__(256) void geom_filter_kernel(GeomViewDev g, int tiles_x, int tiles)
{
    int bx, by;
    if (!xcd_tile(tiles_x, tiles, bx, by)) return;
    const int x = bx * 64 + (threadIdx.x & 63);
    const int y = by * 4 + (threadIdx.x >> 6);
    if (x >= g.W || y >= g.H) return;
    const size_t idx = (size_t)y * g.W + x;
    const float d = g.depth[idx];
    if (d <= 0.0f) {
        if (g.conf) g.conf[idx] = 0.0f;
        return;
    }
    const float* ki = g.kinv;
    const float* R = g.R;
    const float fx = (float)x, fy = (float)y;
    // back-project, then to world: P_w = R^T (P_c - t)
    const float pcx = d * (ki[0] * fx + ki[1] * fy + ki[2]);
    const float pcy = d * (ki[3] * fx + ki[4] * fy + ki[5]);
    const float pcz = d * (ki[6] * fx + ki[7] * fy + ki[8]);
    const float a = pcx - g.t[0], b = pcy - g.t[1], c = pcz - g.t[2];
    const float pwx = R[0] * a + R[3] * b + R[6] * c;
    const float pwy = R[1] * a + R[4] * b + R[7] * c;
    const float pwz = R[2] * a + R[5] * b + R[8] * c;
    const float ncx = g.normal[3 * idx], ncy = g.normal[3 * idx + 1], ncz = g.normal[3 * idx + 2];
    const float nwx = R[0] * ncx + R[3] * ncy + R[6] * ncz;
    const float nwy = R[1] * ncx + R[4] * ncy + R[7] * ncz;
    const float nwz = R[2] * ncx + R[5] * ncy + R[8] * ncz;
    int consistent = 0;
    for (int s = 0; s < g.S; ++s) {
        const GeomSrcDev& sv = g.src[s];
        const float psx = sv.R[0] * pwx + sv.R[1] * pwy + sv.R[2] * pwz + sv.t[0];
        const float psy = sv.R[3] * pwx + sv.R[4] * pwy + sv.R[5] * pwz + sv.t[1];
        const float psz = sv.R[6] * pwx + sv.R[7] * pwy + sv.R[8] * pwz + sv.t[2];
        if (psz <= 0.0f) continue;
        const float prx = sv.K[0] * psx + sv.K[1] * psy + sv.K[2] * psz;
        const float pry = sv.K[3] * psx + sv.K[4] * psy + sv.K[5] * psz;
        const float prz = sv.K[6] * psx + sv.K[7] * psy + sv.K[8] * psz;
        const float u = prx / prz, w = pry / prz;
        const float ru = rintf(u), rw = rintf(w);   // round to nearest even (__float2int_rn)
        // bounds tested in float first: values beyond int range are out anyway
        if (!(ru >= 0.0f && ru < (float)g.W && rw >= 0.0f && rw < (float)g.H)) continue;
        const size_t sp = (size_t)(int)rw * g.W + (int)ru;
        const float ds = sv.depth[sp];
        if (ds <= 0.0f) continue;
        if (fabsf(ds - psz) / psz > g.tol) continue;
        const float nsx = sv.normal[3 * sp], nsy = sv.normal[3 * sp + 1], nsz = sv.normal[3 * sp + 2];
        const float swx = sv.R[0] * nsx + sv.R[3] * nsy + sv.R[6] * nsz;
        const float swy = sv.R[1] * nsx + sv.R[4] * nsy + sv.R[7] * nsz;
        const float swz = sv.R[2] * nsx + sv.R[5] * nsy + sv.R[8] * nsz;
        if (nwx * swx + nwy * swy + nwz * swz < g.cos_tol) continue;
        consistent++;
    }
    if (consistent < g.min_consistent) {
        g.depth[idx] = 0.0f;
        if (g.conf) g.conf[idx] = 0.0f;
    }
}

}  // namespace hp3d

using namespace hp3d;

extern "C" int hp3d_pm_geom_filter(hp3d_pm_ctx* ctx, hp3d_geom_view* views, int n,
                                   float depth_tol, float normal_tol_deg, int min_consistent,
                                   int* retained)
{
    if (!ctx) return HP3D_ERR_INVALID_ARG;
    if (!views || n < 1 || min_consistent < 0 || !(depth_tol >= 0.0f))
        return ctx->fail(HP3D_ERR_INVALID_ARG, "geom_filter: bad arguments");
    for (int i = 0; i < n; ++i) {
        const hp3d_geom_view& v = views[i];
        if (v.width < 1 || v.height < 1 || !v.depth || !v.normal || v.num_src < 0 ||
            v.num_src > kMaxSrc)
            return ctx->fail(HP3D_ERR_INVALID_ARG, "geom_filter: view %d is malformed", i);
        // a view's own map is filtered in place while its sources are read,
        // so it must not be one of them; a repeated source would count twice
        for (int s = 0; s < v.num_src; ++s) {
            if (v.src[s] < 0 || v.src[s] >= n)
                return ctx->fail(HP3D_ERR_INVALID_ARG, "geom_filter: view %d source %d out of range", i, s);
            if (v.src[s] == i)
                return ctx->fail(HP3D_ERR_INVALID_ARG, "geom_filter: view %d lists itself as a source", i);
            for (int r = 0; r < s; ++r)
                if (v.src[r] == v.src[s])
                    return ctx->fail(HP3D_ERR_INVALID_ARG, "geom_filter: view %d lists source %d twice",
                                     i, v.src[s]);
        }
    }
    (void)hipSetDevice(ctx->device);
    // one device block: per view depth, normal, confidence
    std::vector<size_t> off(n);
    size_t total = 0;
    for (int i = 0; i < n; ++i) {
        off[i] = total;
        const size_t px = (size_t)views[i].width * views[i].height;
        total += ((px * 5 * sizeof(float)) + 255) & ~(size_t)255;
    }
    void* block = nullptr;
    int rc = ctx->dev_alloc(&block, total, "geometric filter maps");
    if (rc) return rc;
    auto dptr = [&](int i, int k) {   // k: 0 depth, 1 normal, 4 confidence
        const size_t px = (size_t)views[i].width * views[i].height;
        return reinterpret_cast<float*>(static_cast<char*>(block) + off[i]) + (k == 4 ? 4 * px : k * px);
    };
    hipError_t e = hipSuccess;
    for (int i = 0; i < n && e == hipSuccess; ++i) {
        const size_t px = (size_t)views[i].width * views[i].height;
        e = hipMemcpyAsync(dptr(i, 0), views[i].depth, px * 4, hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(dptr(i, 1), views[i].normal, 3 * px * 4, hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess && views[i].confidence)
            e = hipMemcpyAsync(dptr(i, 4), views[i].confidence, px * 4, hipMemcpyHostToDevice, ctx->stream);
    }
    const float cos_tol = std::cos(normal_tol_deg * 3.14159265358979323846f / 180.0f);
    std::vector<char> filtered(n, 0);
    double ki[9];
    for (int i = 0; i < n && e == hipSuccess; ++i) {
        const hp3d_geom_view& v = views[i];
        GeomViewDev g;
        std::memset(&g, 0, sizeof g);
        // sources with a different size are skipped (:1310)
        for (int s = 0; s < v.num_src; ++s) {
            const hp3d_geom_view& sv = views[v.src[s]];
            if (sv.width != v.width || sv.height != v.height) continue;
            GeomSrcDev& d = g.src[g.S++];
            for (int k = 0; k < 9; ++k) { d.K[k] = (float)sv.K[k]; d.R[k] = (float)sv.R[k]; }
            for (int k = 0; k < 3; ++k) d.t[k] = (float)sv.t[k];
            d.depth = dptr(v.src[s], 0);
            d.normal = dptr(v.src[s], 1);
        }
        if (g.S < min_consistent) continue;   // view left unfiltered (:1330)
        if (hp3d_pm_invert3(v.K, ki) != HP3D_OK) {
            (void)ctx->dev_free(block, total);
            return ctx->fail(HP3D_ERR_INVALID_ARG, "geom_filter: view %d K is singular", i);
        }
        for (int k = 0; k < 9; ++k) { g.kinv[k] = (float)ki[k]; g.R[k] = (float)v.R[k]; }
        for (int k = 0; k < 3; ++k) g.t[k] = (float)v.t[k];
        g.W = v.width;
        g.H = v.height;
        g.min_consistent = min_consistent;
        g.tol = depth_tol;
        g.cos_tol = cos_tol;
        g.depth = dptr(i, 0);
        g.normal = dptr(i, 1);
        g.conf = v.confidence ? dptr(i, 4) : nullptr;
        const int tiles_x = (v.width + 63) / 64, tiles = tiles_x * ((v.height + 3) / 4);
        hipLaunchKernelGGL(geom_filter_kernel, dim3(xcd_grid(tiles)), dim3(256), 0, ctx->stream, g,
                           tiles_x, tiles);
        e = hipGetLastError();
        filtered[i] = 1;
    }
    for (int i = 0; i < n && e == hipSuccess; ++i) {
        const size_t px = (size_t)views[i].width * views[i].height;
        e = hipMemcpyAsync(views[i].depth, dptr(i, 0), px * 4, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess && views[i].confidence)
            e = hipMemcpyAsync(views[i].confidence, dptr(i, 4), px * 4, hipMemcpyDeviceToHost, ctx->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    ctx->dev_free(block, total);
    if (e != hipSuccess) return ctx->hip_fail(e, "geometric filter");
    if (retained)
        for (int i = 0; i < n; ++i) {
            const size_t px = (size_t)views[i].width * views[i].height;
            int r = 0;
            for (size_t p = 0; p < px; ++p) r += views[i].depth[p] > 0.0f;
            retained[i] = filtered[i] ? r : -1;
        }
    return HP3D_OK;
}
