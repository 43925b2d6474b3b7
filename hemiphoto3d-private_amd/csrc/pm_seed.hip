// pm_seed.hip -- SfM seeding of the hypothesis planes on the device.
//
// Reference: the seeding block of runPatchMatchMVS (src/dense_recon/
// patchmatch_mvs.cu:1111-1177) copies the random init to the host, paints a
// (2r+1)^2 square of (z, n) around every seed in point order, so a pixel ends
// with the LAST seed whose square covers it, and copies both maps back.
// Here the maps never leave the device: one workgroup per seed records, with
// an atomic max, the highest seed index covering each pixel of its square;
// one lane per pixel then copies that seed's (z, n) into its colour plane.
// The winner per pixel is the same as in the sequential loop, so the result
// is identical.
#include <hip/hip_runtime.h>

#include "../../include/hp3d_pm.h"
#include "pm_device.h"
#include "pm_launch.h"

namespace hp3d {

__global__ __launch_bounds__(256) void seed_clear_kernel(int* owner, int n)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) owner[i] = 0;
}

// owner[q] = 1 + max seed index whose square covers pixel q (0: none).
__global__ __launch_bounds__(256) void seed_owner_kernel(const hp3d_seed_point* seeds, int W,
                                                         int H, int radius, int* owner)
{
    const int k = blockIdx.x;
    const int cx = seeds[k].x, cy = seeds[k].y;
    const int x0 = max(cx - radius, 0), x1 = min(cx + radius, W - 1);
    const int y0 = max(cy - radius, 0), y1 = min(cy + radius, H - 1);
    if (x0 > x1 || y0 > y1) return;
    const int w = x1 - x0 + 1, n = w * (y1 - y0 + 1);
    for (int i = threadIdx.x; i < n; i += 256) {
        const int y = y0 + i / w, x = x0 + i % w;
        atomicMax(&owner[y * W + x], k + 1);
    }
}

__global__ __launch_bounds__(256) void seed_apply_kernel(DevView v, const hp3d_seed_point* seeds,
                                                         const int* owner)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= v.W || y >= v.H) return;
    const int o = owner[y * v.W + x];
    if (o == 0) return;
    const hp3d_seed_point s = seeds[o - 1];
    v.hyp[(x + y) & 1][y * v.Wc + (x >> 1)] =
        make_float4(s.depth, s.normal[0], s.normal[1], s.normal[2]);
}

hipError_t launch_seed(const DevView& v, const hp3d_seed_point* seeds, int n, int radius,
                       int* owner, hipStream_t st)
{
    const int np = v.W * v.H;
    hipLaunchKernelGGL(seed_clear_kernel, dim3((np + 255) / 256), dim3(256), 0, st, owner, np);
    hipLaunchKernelGGL(seed_owner_kernel, dim3(n), dim3(256), 0, st, seeds, v.W, v.H, radius,
                       owner);
    hipLaunchKernelGGL(seed_apply_kernel, dim3((v.W + 63) / 64, (v.H + 3) / 4), dim3(256), 0, st,
                       v, seeds, owner);
    return hipGetLastError();
}

}  // namespace hp3d
