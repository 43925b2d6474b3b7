// pm_kernels.hip -- per-view (plain form) gfx950 kernels of the PatchMatch path.
//
// Reference behaviour (src/dense_recon/patchmatch_mvs.cu):
//   pm_init_kernel      <- init_depth_normal_kernel      (:164-216)
//   pm_cost_kernel      <- compute_ncc_cost_kernel        (:349-513)
//   pm_halfpass_kernel  <- propagate_and_refine_kernel    (:519-655)
//
// Design (gfx950, wave64):
//   * one lane per active pixel of one colour: lanes map to the compacted
//     colour lattice x = 2i + ((y + parity) & 1), so no lane idles on the
//     checkerboard;
//   * a workgroup of TX x TY lanes stages its reference window tile (with a
//     radius-wide halo, clamped at the image edge) in LDS once;
//   * per candidate and source the plane homography H = A + b (K^-T n)^T / (n.P)
//     is formed once (window centre in FP64), so a tap costs 3 FMAs for the
//     warp, one v_rcp_f32, one buffer load (the quad from the
//     row-difference plane) and 4 instructions for the bilinear sample;
//   * neighbour candidates that provably cannot win are skipped and the rest
//     shared across the wave's lanes (halfpass_pixel in pm_pixel.h);
//   * the five NCC moments are accumulated in FP32 on values shifted by
//     -128 DN; the variance cancellation is done in FP64.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pm_device.h"
#include "pm_launch.h"
#include "pm_pixel.h"
#include "pm_xcd.h"

namespace hp3d {

__global__ __launch_bounds__(256) void pm_init_kernel(DevView v, uint64_t seed)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= v.W || y >= v.H) return;
    v.hyp[(x + y) & 1][y * v.Wc + (x >> 1)] = init_hyp(v, x, y, seed);
}

// Cost of the current hypotheses of one colour, run once per colour before
// the first iteration.
template <int NT, int ST, int STORE, int TX, int TY>
__global__ __launch_bounds__(TX * TY) void pm_cost_kernel(DevView v, int parity, int tiles_x,
                                                          int tiles)
{
    constexpr int R = (NT - 1) * ST / 2;
    using G = TileGeom<R, STORE, TX, TY>;
    __shared__ float tile[G::LH * G::LP];
    __shared__ WaveXchg xw[TX * TY / 64];
    int bx, by;
    if (!xcd_tile(tiles_x, tiles, bx, by)) return;
    stage_ref_tile<R, STORE, TX, TY>(v, tile, bx, by);
    cost_tile_sorted<NT, ST, STORE, TX, TY>(v, tile, bx, by, threadIdx.x % TX, threadIdx.x / TX,
                                            parity, xw[threadIdx.x >> 6]);
}

// Build-time tuning knob of the per-view kernels: minimum waves per SIMD the
// register allocation must allow (1: no constraint).
#ifndef HP3D_PLAIN_WPE
#define HP3D_PLAIN_WPE 1
#endif

template <int NT, int ST, int STORE, int TX, int TY>
__global__ __launch_bounds__(TX * TY) __attribute__((amdgpu_waves_per_eu(HP3D_PLAIN_WPE, 8)))
void pm_halfpass_kernel(DevView v, int iteration, int parity, uint64_t pass_seed, int tiles_x,
                        int tiles)
{
    constexpr int R = (NT - 1) * ST / 2;
    using G = TileGeom<R, STORE, TX, TY>;
    __shared__ float tile[G::LH * G::LP];
    __shared__ WaveXchg xw[TX * TY / 64];
    int bx, by;
    if (!xcd_tile(tiles_x, tiles, bx, by)) return;
    stage_ref_tile<R, STORE, TX, TY>(v, tile, bx, by);
    halfpass_pixel<NT, ST, STORE, TX, TY>(v, tile, bx, by, threadIdx.x % TX, threadIdx.x / TX,
                                          iteration, parity, pass_seed, xw[threadIdx.x >> 6]);
}

// Colour planes -> row-major depth (H*W), AoS normal (H*W*3), cost and
// confidence (H*W).  Null outputs are skipped.
__global__ __launch_bounds__(256) void pm_finalize_kernel(DevView v, float* depth, float* normal,
                                                          float* cost, float* conf)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= v.W || y >= v.H) return;
    emit_pixel(v, x, y, depth, normal, cost, conf);
}

// Row-major depth + AoS normal -> colour planes (caller-given hypotheses).
__global__ __launch_bounds__(256) void pm_import_kernel(DevView v, const float* depth,
                                                        const float* normal)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= v.W || y >= v.H) return;
    const size_t q = (size_t)y * v.W + x;
    v.hyp[(x + y) & 1][y * v.Wc + (x >> 1)] =
        make_float4(depth[q], normal[3 * q], normal[3 * q + 1], normal[3 * q + 2]);
}

// 8-bit or [0,1] float gray -> the image planes (DN - 128, see pm_device.h):
// the plain plane and the row-difference gather plane
// (I[y][x], I[y+1][x] - I[y][x]), in FP32 or FP16.  in_u8 / in_f32: exactly one is non-null; in_pitch in bytes.
template <int STORE>
__global__ __launch_bounds__(256) void img_prep_kernel(const uint8_t* in_u8, const float* in_f32,
                                                       size_t in_pitch, int W, int H, void* plain,
                                                       void* coef, int pitch)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= W || y >= H) return;
    auto val = [&](int xx, int yy) -> float {
        if (in_u8) return (float)in_u8[(size_t)yy * in_pitch + xx] - kShiftDN;
        const float* row = (const float*)((const char*)in_f32 + (size_t)yy * in_pitch);
        return row[xx] * 255.0f - kShiftDN;
    };
    const float a = val(x, y);
    const size_t q = (size_t)y * pitch + x;
    // the last row's difference is never read: taps need v < H - 1
    const float dn = y < H - 1 ? val(x, y + 1) - a : 0.0f;
    if (STORE == kStoreF32) {
        static_cast<float*>(plain)[q] = a;
        static_cast<float2*>(coef)[q] = make_float2(a, dn);
    } else {
        static_cast<_Float16*>(plain)[q] = (_Float16)a;
        half2_t h;
        h.x = (_Float16)a;
        h.y = (_Float16)dn;
        static_cast<half2_t*>(coef)[q] = h;
    }
}

// One pyramid level from the previous one: element (x, y) is the mean of
// the 2 x 2 block at (2x, 2y) of the previous level's plain plane (stored
// DN - 128 values), written as the level's plain and row-difference planes.
// The sums are exact in FP32 (values carry 2 more fraction bits per level);
// FP16 storage holds the values of the first two levels and the row
// differences of the first level exactly.
template <int STORE>
__global__ __launch_bounds__(256) void img_down_kernel(const void* prev, int prev_pitch, int W,
                                                       int H, void* plain, void* coef, int pitch)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= W || y >= H) return;
    auto val = [&](int xx, int yy) -> float {
        const int q = 2 * yy * prev_pitch + 2 * xx;
        const float a = load_plain<STORE>(prev, q), b = load_plain<STORE>(prev, q + 1);
        const float c = load_plain<STORE>(prev, q + prev_pitch);
        const float d = load_plain<STORE>(prev, q + prev_pitch + 1);
        return ((a + b) + (c + d)) * 0.25f;
    };
    const size_t q = (size_t)y * pitch + x;
    if (STORE == kStoreF32) {
        const float a = val(x, y);
        const float dn = y < H - 1 ? val(x, y + 1) - a : 0.0f;
        static_cast<float*>(plain)[q] = a;
        static_cast<float2*>(coef)[q] = make_float2(a, dn);
    } else {
        // the level's stored (FP16) values, and their difference as stored
        const _Float16 a = (_Float16)val(x, y);
        const _Float16 b = (_Float16)(y < H - 1 ? val(x, y + 1) : 0.0f);
        static_cast<_Float16*>(plain)[q] = a;
        half2_t h;
        h.x = a;
        h.y = y < H - 1 ? (_Float16)((float)b - (float)a) : (_Float16)0.0f;
        static_cast<half2_t*>(coef)[q] = h;
    }
}

// ---------------------------------------------------------------------------
// Host-side launchers
// ---------------------------------------------------------------------------
static dim3 grid_2d(int W, int H) { return dim3((W + 63) / 64, (H + 3) / 4); }

hipError_t launch_init(const DevView& v, uint64_t seed, hipStream_t st)
{
    hipLaunchKernelGGL(pm_init_kernel, grid_2d(v.W, v.H), dim3(256), 0, st, v, seed);
    return hipGetLastError();
}

hipError_t launch_finalize(const DevView& v, float* depth, float* normal, float* cost, float* conf,
                           hipStream_t st)
{
    hipLaunchKernelGGL(pm_finalize_kernel, grid_2d(v.W, v.H), dim3(256), 0, st, v, depth, normal,
                       cost, conf);
    return hipGetLastError();
}

hipError_t launch_import(const DevView& v, const float* depth, const float* normal,
                         hipStream_t st)
{
    hipLaunchKernelGGL(pm_import_kernel, grid_2d(v.W, v.H), dim3(256), 0, st, v, depth, normal);
    return hipGetLastError();
}

hipError_t launch_img_prep(const uint8_t* in_u8, const float* in_f32, size_t in_pitch, int W,
                           int H, int store, void* plain, void* coef, int pitch, hipStream_t st)
{
    if (store == kStoreF32)
        hipLaunchKernelGGL(img_prep_kernel<kStoreF32>, grid_2d(W, H), dim3(256), 0, st, in_u8,
                           in_f32, in_pitch, W, H, plain, coef, pitch);
    else
        hipLaunchKernelGGL(img_prep_kernel<kStoreF16>, grid_2d(W, H), dim3(256), 0, st, in_u8,
                           in_f32, in_pitch, W, H, plain, coef, pitch);
    return hipGetLastError();
}

hipError_t launch_img_down(const void* prev, int prev_pitch, int W, int H, int store, void* plain,
                           void* coef, int pitch, hipStream_t st)
{
    if (store == kStoreF32)
        hipLaunchKernelGGL(img_down_kernel<kStoreF32>, grid_2d(W, H), dim3(256), 0, st, prev,
                           prev_pitch, W, H, plain, coef, pitch);
    else
        hipLaunchKernelGGL(img_down_kernel<kStoreF16>, grid_2d(W, H), dim3(256), 0, st, prev,
                           prev_pitch, W, H, plain, coef, pitch);
    return hipGetLastError();
}

// Workgroup tile (compacted columns x rows), from sweeps on 1600x1200 views:
// 32 x 8 (64 x 8 pixels) for windows of up to 8 taps a side (6 x 6 at
// stride 2: 32x8, 32x4 and 32x2 within 1 %, 16x16 and 64x4 slower); 32 x 4
// for the 11 x 11 stride-1 window, whose kernel holds 3 waves per SIMD:
// half-size workgroups leave a shorter tail per launch, +8.5 % with FP32
// storage (HP3D_PLAIN_TX / _TY override both at build time).
template <int NT>
struct PlainTile {
#ifdef HP3D_PLAIN_TX
    static constexpr int TX = HP3D_PLAIN_TX, TY = HP3D_PLAIN_TY;
#else
    static constexpr int TX = 32, TY = NT >= 9 ? 4 : 8;
#endif
};

template <int NT, int ST, int STORE>
static void launch_pass_t(const DevView& v, int mode, int iteration, int parity,
                          uint64_t pass_seed, hipStream_t st)
{
    constexpr int kTX = PlainTile<NT>::TX, kTY = PlainTile<NT>::TY;
    const int tiles_x = (v.Wc + kTX - 1) / kTX;
    const int tiles = tiles_x * ((v.H + kTY - 1) / kTY);
    const dim3 grid(((tiles + 7) / 8) * 8);
    if (mode == kPassCost)
        hipLaunchKernelGGL((pm_cost_kernel<NT, ST, STORE, kTX, kTY>), grid, dim3(kTX * kTY), 0, st,
                           v, parity, tiles_x, tiles);
    else
        hipLaunchKernelGGL((pm_halfpass_kernel<NT, ST, STORE, kTX, kTY>), grid, dim3(kTX * kTY), 0,
                           st, v, iteration, parity, pass_seed, tiles_x, tiles);
}

template <int STORE>
static bool dispatch_window(const DevView& v, int radius, int stride, int mode, int iteration,
                            int parity, uint64_t pass_seed, hipStream_t st)
{
#define HP3D_CASE(R_, S_)                                                                   \
    if (radius == R_ && stride == S_) {                                                     \
        launch_pass_t<(2 * R_) / S_ + 1, S_, STORE>(v, mode, iteration, parity, pass_seed, st); \
        return true;                                                                        \
    }
    HP3D_CASE(1, 1) HP3D_CASE(2, 1) HP3D_CASE(3, 1) HP3D_CASE(4, 1) HP3D_CASE(5, 1)
    HP3D_CASE(1, 2) HP3D_CASE(2, 2) HP3D_CASE(3, 2) HP3D_CASE(4, 2) HP3D_CASE(5, 2)
#undef HP3D_CASE
    return false;
}

bool window_supported(int radius, int stride)
{
    return radius >= 1 && radius <= 5 && (stride == 1 || stride == 2);
}

hipError_t launch_pass(const DevView& v, int store, int radius, int stride, int mode,
                       int iteration, int parity, uint64_t pass_seed, hipStream_t st)
{
    const bool ok = store == kStoreF32
                        ? dispatch_window<kStoreF32>(v, radius, stride, mode, iteration, parity,
                                                     pass_seed, st)
                        : dispatch_window<kStoreF16>(v, radius, stride, mode, iteration, parity,
                                                     pass_seed, st);
    if (!ok) return hipErrorInvalidValue;
    return hipGetLastError();
}

}  // namespace hp3d
