// pm_pixel.h -- per-pixel device functions of the PatchMatch path, shared by
// the per-view kernels (pm_kernels.hip) and the multi-view step kernel
// (pm_step.hip).  Every function here is a pure function of its arguments and
// the view's state, so a pixel computes the same bits in either kernel.
//
// Reference behaviour (src/dense_recon/patchmatch_mvs.cu):
//   init_hyp        <- init_depth_normal_kernel      (:164-216)
//   cand_cost       <- computePixelCost               (:222-343)
//   cost_pixel      <- compute_ncc_cost_kernel body   (:349-513)
//   halfpass_pixel  <- propagate_and_refine_kernel    (:519-655)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pm_device.h"

// Build-time choice of the half-pass's depth-ordered candidate list: the
// number of leading iterations of a solve that order a wave's candidates by
// depth (1, the default: the first; 0: owner order in every iteration).
#ifndef HP3D_BUCKET
#define HP3D_BUCKET 1
#endif
// Build-time choice of the windows (by tap count) whose tap coordinates are
// taken relative to an integer base near the window centre's source position
// (default: up to 5x5 taps, the windows whose NCC a sub-pixel coordinate
// error moves the most; 0: none; taken to every window it measured 1.5-4 %
// slower at 36 and 121 taps, for no gain in accuracy there).
#ifndef HP3D_LOCAL_TAPS
#define HP3D_LOCAL_TAPS 25
#endif
// Build-time choice of the windows (by tap count) whose taps run from the
// centre on (centre row, the rows below it, then the rows above; the same for
// the columns of a row), so that the window's first tap is its centre; the
// others run top-left first.  Default: from 7x7 taps up, where the moment
// sums' rounding needed it; below, the centre-first order measured slower
// (its gathers hit the L1 cache less) for no gain in accuracy.
#ifndef HP3D_CENTRE_FIRST_TAPS
#define HP3D_CENTRE_FIRST_TAPS 49
#endif

namespace hp3d {

typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
typedef _Float16 half4_t __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// Philox4x32-10 and the u01 mapping ((x >> 8) + 1) * 2^-24 in (0, 1].
// ---------------------------------------------------------------------------
__device__ __forceinline__ void philox(uint32_t& c0, uint32_t& c1, uint32_t& c2, uint32_t& c3,
                                       uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += kPhiloxW0; k1 += kPhiloxW1; }
        const uint32_t hi0 = __umulhi(kPhiloxM0, c0), lo0 = kPhiloxM0 * c0;
        const uint32_t hi1 = __umulhi(kPhiloxM1, c2), lo1 = kPhiloxM1 * c2;
        const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
}

__device__ __forceinline__ float u01(uint32_t x)
{
    return (float)((x >> 8) + 1u) * (1.0f / 16777216.0f);
}

// Unit view ray K^-1 [x, y, 1] / |.|, written so it rounds like the host code.
__device__ __forceinline__ void unit_ray(const float* ki, int x, int y, float& rx, float& ry,
                                         float& rz)
{
    const float fx = (float)x, fy = (float)y;
    float a = ki[0] * fx + ki[1] * fy + ki[2];
    float b = ki[3] * fx + ki[4] * fy + ki[5];
    float c = ki[6] * fx + ki[7] * fy + ki[8];
    const float len = sqrtf(a * a + b * b + c * c);
    if (len > 1e-12f) { a = a / len; b = b / len; c = c / len; }
    rx = a; ry = b; rz = c;
}

// Unnormalised ray K^-1 [x, y, 1] in FP64 (the cost's back-projection, :241-243).
__device__ __forceinline__ void ray_fp64(const double* ki, int x, int y, double& rx, double& ry,
                                         double& rz)
{
    const double X = x, Y = y;
    rx = ki[0] * X + ki[1] * Y + ki[2];
    ry = ki[3] * X + ki[4] * Y + ki[5];
    rz = ki[6] * X + ki[7] * Y + ki[8];
}

// ---------------------------------------------------------------------------
// Image element access
// ---------------------------------------------------------------------------
template <int STORE> __device__ __forceinline__ float load_plain(const void* p, int off);
template <> __device__ __forceinline__ float load_plain<kStoreF32>(const void* p, int off)
{
    return static_cast<const float*>(p)[off];
}
template <> __device__ __forceinline__ float load_plain<kStoreF16>(const void* p, int off)
{
    return (float)static_cast<const _Float16*>(p)[off];
}

// Planes are read through buffer resources: the hardware range-checks every
// access and returns 0 outside the plane, so a tap whose computed address is
// wild (a lane that will be masked) can never fault.  Dword 3 = 0x00020000 is
// the raw 32-bit data format of gfx9-family buffer descriptors.
//   stride 0 (raw): num_records counts bytes and the load takes a byte
//     offset, checked against it;
//   stride s (structured, the source gather planes): num_records counts
//     records of s bytes and the load takes a record index, which the buffer
//     unit checks against num_records and scales by s.  The check is on the
//     index alone: a 16-byte load at an 8-byte stride returns the record and
//     its successor in full (measured on gfx950).
constexpr int kBufferRsrcWord3 = 0x00020000;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* p, int num_records, short stride = 0)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), stride, num_records, kBufferRsrcWord3);
}

typedef int v4i32_t __attribute__((ext_vector_type(4)));
typedef int v2i32_t __attribute__((ext_vector_type(2)));
// Structured buffer loads (the LLVM intrinsics; hipcc has no builtin for them).
__device__ v4i32_t struct_load_b128(__amdgpu_buffer_rsrc_t rsrc, int vindex, int voffset, int soffset,
                                    int aux) __asm("llvm.amdgcn.struct.ptr.buffer.load.v4i32");
__device__ v2i32_t struct_load_b64(__amdgpu_buffer_rsrc_t rsrc, int vindex, int voffset, int soffset,
                                   int aux) __asm("llvm.amdgcn.struct.ptr.buffer.load.v2i32");
// A record index past every gather plane (a plane holds at most
// kMaxPlaneBytes, so fewer records): the load is range-checked away and
// returns zeros.
constexpr unsigned kIdxNowhere = 0x7ffffff0u;
static_assert(kIdxNowhere >= kMaxPlaneBytes, "index inside a plane");

// Access to the hypothesis state planes (hyp, cost, change flags).
// COH = false: ordinary loads and stores (one kernel launch per pass: the
// launch boundary orders passes).  COH = true: device-coherent accesses
// (the sc1 cache policy, bypassing the per-CU L1 and the per-XCD L2 caches
// that are not coherent across XCDs), for the persistent batch kernel, where
// a tile's pass reads what workgroups on other XCDs wrote in the same launch.
// The image planes are read-only within a launch and keep ordinary cached
// loads either way.  Planes addressed this way hold less than 2 GiB.
// How the persistent kernel keeps the state planes coherent across XCDs
// (build-time choice, measured): bit 0 set = its state stores are ordinary
// stores followed by an agent-scope release fence (an L2 write-back) before a
// tile's counter moves, else device-coherent (sc1) stores; bit 1 set = its
// state loads are ordinary loads after an agent-scope acquire fence (an L2
// invalidate) once the tile's dependencies are met, else sc1 loads.
#ifndef HP3D_PERSIST_MEM
#define HP3D_PERSIST_MEM 0
#endif
constexpr bool kPersistPlainStores = (HP3D_PERSIST_MEM & 1) != 0;
constexpr bool kPersistPlainLoads = (HP3D_PERSIST_MEM & 2) != 0;

template <bool COH> struct StateIO;
template <> struct StateIO<false> {
    static __device__ __forceinline__ float4 ld4(const float4* p, int i) { return p[i]; }
    static __device__ __forceinline__ float ld1(const float* p, int i) { return p[i]; }
    static __device__ __forceinline__ unsigned ldb(const unsigned char* p, int i) { return p[i]; }
    static __device__ __forceinline__ void st4(float4* p, int i, float4 x) { p[i] = x; }
    static __device__ __forceinline__ void st1(float* p, int i, float x) { p[i] = x; }
    static __device__ __forceinline__ void stb(unsigned char* p, int i, unsigned x)
    {
        p[i] = (unsigned char)x;
    }
};
template <> struct StateIO<true> {
    static constexpr int kSC1 = 16;   // cache-policy operand: sc1 (device scope)
    static __device__ __forceinline__ __amdgpu_buffer_rsrc_t rs(const void* p)
    {
        return make_rsrc(p, 0x7fffffff);
    }
    static constexpr int kLd = kPersistPlainLoads ? 0 : kSC1;
    static constexpr int kSt = kPersistPlainStores ? 0 : kSC1;
    static __device__ __forceinline__ float4 ld4(const float4* p, int i)
    {
        return __builtin_bit_cast(float4,
                                  __builtin_amdgcn_raw_buffer_load_b128(rs(p), i * 16, 0, kLd));
    }
    static __device__ __forceinline__ float ld1(const float* p, int i)
    {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs(p), i * 4, 0, kLd));
    }
    static __device__ __forceinline__ unsigned ldb(const unsigned char* p, int i)
    {
        return __builtin_amdgcn_raw_buffer_load_b8(rs(p), i, 0, kLd);
    }
    static __device__ __forceinline__ void st4(float4* p, int i, float4 x)
    {
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(__attribute__((ext_vector_type(4))) unsigned, x),
                                               rs(p), i * 16, 0, kSt);
    }
    static __device__ __forceinline__ void st1(float* p, int i, float x)
    {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, x), rs(p), i * 4, 0, kSt);
    }
    static __device__ __forceinline__ void stb(unsigned char* p, int i, unsigned x)
    {
        __builtin_amdgcn_raw_buffer_store_b8((unsigned char)x, rs(p), i, 0, kSt);
    }
};

// Bilinear sample at (ix + au, iy + aw) of the texel quad a = I[iy][ix],
// b = I[iy][ix+1], c = I[iy+1][ix], d = I[iy+1][ix+1].  Both storages keep a
// row-difference plane for source gathers, element (x, y) =
// (I[y][x], I[y+1][x] - I[y][x]), so that one load per tap at (ix, iy)
// returns the whole quad as (a, c - a, b, d - b): 16 bytes from 8 bytes per
// pixel (F32) or 8 bytes from 4 (F16).  The sample is then two FMAs down the
// columns and a subtract and an FMA across -- 4 VALU instructions, against 6
// for three lerps of a row-pair plane (a, c, b, d) -- in FP32 either way.
// For 8-bit images the values (DN - 128) and the row differences (at most
// 255 in magnitude) are integers, exact in FP16, so both storages give
// identical results.  (Measured on 1600x1200 views: one load per tap instead
// of one per texel row, 1.6x at stride 2; the F16 plane's halved footprint
// 1.13x more at stride 1; differences instead of row pairs, +2 % (F32) /
// +5 % (F16) at stride 2 and +16 % at stride 1.)
// A source's gather plane as records of one texel (FP32: 8 bytes, FP16: 4):
// the load takes the record index iy * pitch + ix (one multiply-add per tap)
// and reads that texel's record and its right neighbour's, (a, c-a, b, d-b).
// A valid tap has ix <= W-2 and iy <= H-2, so both records lie inside the
// plane; an invalid tap's index goes nowhere and reads zeros.
template <int STORE> struct SrcImage;
template <> struct SrcImage<kStoreF32> {
    __amdgpu_buffer_rsrc_t rsrc;
    unsigned pitch;
    __device__ SrcImage(const DevView& v, int s)
        : rsrc(make_rsrc(v.src_coef[s], v.coef_bytes / 8, 8)), pitch((unsigned)v.pitch_plain) {}
    __device__ __forceinline__ float sample(int ix, int iy, float au, float aw, bool ok = true) const
    {
        unsigned idx = __umul24((unsigned)iy, pitch) + (unsigned)ix;
        idx = ok ? idx : kIdxNowhere;
        const float4 q = __builtin_bit_cast(float4, struct_load_b128(rsrc, (int)idx, 0, 0, 0));
        const float c0 = fmaf(aw, q.y, q.x), c1 = fmaf(aw, q.w, q.z);
        return fmaf(au, c1 - c0, c0);
    }
};
template <> struct SrcImage<kStoreF16> {
    __amdgpu_buffer_rsrc_t rsrc;
    unsigned pitch;
    __device__ SrcImage(const DevView& v, int s)
        : rsrc(make_rsrc(v.src_coef[s], v.coef_bytes / 4, 4)), pitch((unsigned)v.pitch_plain) {}
    __device__ __forceinline__ float sample(int ix, int iy, float au, float aw, bool ok = true) const
    {
        unsigned idx = __umul24((unsigned)iy, pitch) + (unsigned)ix;
        idx = ok ? idx : kIdxNowhere;
        const half4_t q = __builtin_bit_cast(half4_t, struct_load_b64(rsrc, (int)idx, 0, 0, 0));
        const float c0 = fmaf(aw, (float)q.y, (float)q.x), c1 = fmaf(aw, (float)q.w, (float)q.z);
        return fmaf(au, c1 - c0, c0);
    }
};

// floor(x) converted to int in one instruction (v_cvt_flr_i32_f32; out of
// range saturates, which the unsigned bounds test then rejects).
__device__ __forceinline__ int cvt_flr(float x)
{
    int r;
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

struct TapSums {
    float sr, ss, srr, sss, srs;
    int cnt;
};

// Sums of the reference window over every tap, in tap order: they do not
// depend on the candidate or the source, so a window whose taps are all
// valid takes them from here instead of accumulating them per source.
struct WinRef {
    float sr, srr;
};
template <int NT, int ST>
__device__ __forceinline__ WinRef window_ref_sums(const float* __restrict__ lwin, int lp)
{
    constexpr int F = NT * NT >= HP3D_CENTRE_FIRST_TAPS ? NT / 2 : 0;   // first row and column
    float sr = 0.f, srr = 0.f;
#pragma unroll 1
    for (int jr = 0; jr < NT; ++jr) {
        const int j = jr + F < NT ? jr + F : jr + F - NT;
        const float* lrow = lwin + j * ST * lp;
#pragma unroll
        for (int ir = 0; ir < NT; ++ir) {
            const int i = ir + F < NT ? ir + F : ir + F - NT;
            const float r = lrow[i * ST];
            sr += r;
            srr = fmaf(r, r, srr);
        }
    }
    return WinRef{sr, srr};
}

// The window's taps for one candidate and source (patchmatch_mvs.cu:273-307).
// MASKED = false is used only when every tap is known to be inside the image;
// it takes the reference sums from wr, which window_ref_sums forms with the
// same additions in the same order as the masked loop does when every tap is
// valid, so a window's sums do not depend on which variant its wave ran.
//
// The source moments are taken of s - s0, with s0 the window's first valid
// sample (the NCC is invariant to a constant shift of the samples).  Unshifted,
// a low-contrast window's samples sit far from zero (up to 128 DN against a
// spread of a few DN), and the FP32 rounding of s^2 and of the running sums
// carries over into n*sum(s^2) - sum(s)^2 amplified by mean^2 / variance: up
// to 1.2e-3 of cost error on small views.  Shifted, the summands are of the
// size of the spread.  When every tap is valid both loops take the first
// tap's sample, so the variants still agree bit for bit.  In windows of 7x7
// taps and more the taps run from the window centre on (window_ref_sums in
// the same order), so s0 is the centre's sample: near the window mean, the
// running sums of s - s0 wander instead of growing with the tap count, and
// sum(r (s - s0)), whose reference values are not shifted, rounds 3-4x less
// over 121 taps.
//
// (cx, cy) and the columns (h00, h10), (h01, h11) may be relative to an
// integer base (bu, bv): the taps' coordinates are then u - bu and v - bv,
// and the base is added back to the texel index.
template <int NT, int ST, int STORE, bool MASKED, bool ZCHK = true, int RU = (NT >= 9 ? 2 : 1)>
__device__ __forceinline__ void tap_window(const SrcImage<STORE>& img,
                                           const float* __restrict__ lwin, int lp, float cx,
                                           float cy, float cz, float h00, float h01, float h10,
                                           float h11, float h20, float h21, unsigned umax,
                                           unsigned vmax, const WinRef& wr, TapSums& m,
                                           int bu = 0, int bv = 0)
{
    constexpr int R = (NT - 1) * ST / 2;
    constexpr int F = NT * NT >= HP3D_CENTRE_FIRST_TAPS ? NT / 2 : 0;   // first row and column
    float sr = 0.f, ss = 0.f, srr = 0.f, sss = 0.f, srs = 0.f;
    float s0 = 0.f;     // the shift: first valid sample of the window
    bool have = false;  // (masked loop) a valid tap has been seen
    int cnt = 0;
    // window rows (RU at a time): by default two for the 11-tap (stride 1)
    // rows of the half-pass, where the extra loads in flight pay, one for the
    // 6-tap rows, where unrolling measured slower; the initial-cost pass takes
    // its 11-tap rows one at a time too (its random hypotheses scatter the
    // gathers: 0.751 -> 0.442 ms per colour with FP32 storage at 1600x1200)
#pragma unroll RU
    for (int jr = 0; jr < NT; ++jr) {
        const int j = jr + F < NT ? jr + F : jr + F - NT;
        const float oy = (float)(j * ST - R);
        const float bx = fmaf(h01, oy, cx);
        const float by = fmaf(h11, oy, cy);
        const float bz = fmaf(h21, oy, cz);
        const float* lrow = lwin + j * ST * lp;
#pragma unroll
        for (int ir = 0; ir < NT; ++ir) {
            const int i = ir + F < NT ? ir + F : ir + F - NT;
            const float ox = (float)(i * ST - R);
            const float wx = fmaf(h00, ox, bx);
            const float wy = fmaf(h10, ox, by);
            const float wz = fmaf(h20, ox, bz);
            const float rw = __builtin_amdgcn_rcpf(wz);
            const float u = wx * rw, w = wy * rw;
            // (a wrapped sum lands far outside [0, umax], as any invalid tap)
            const int ix = (int)((unsigned)bu + (unsigned)cvt_flr(u));
            const int iy = (int)((unsigned)bv + (unsigned)cvt_flr(w));
            const float au = __builtin_amdgcn_fractf(u), aw = __builtin_amdgcn_fractf(w);
            // |w_z| guard and 0 <= u < W-1, 0 <= v < H-1 (:289-295); an
            // invalid tap's load goes nowhere and reads zeros (so sval = 0)
            const bool ok = !MASKED || ((!ZCHK || fabsf(wz) >= 1e-8f) & ((unsigned)ix <= umax) &
                                        ((unsigned)iy <= vmax));
            const float sraw = img.sample(ix, iy, au, aw, ok);
            float r = lrow[i * ST];
            float sval;
            if (MASKED) {
                // an invalid tap read zeros; before the first valid tap s0
                // stays 0, at it s0 becomes that tap's sample
                s0 = have ? s0 : sraw;
                have = have || ok;
                sval = ok ? sraw - s0 : 0.0f;
                r = ok ? r : 0.0f;
                cnt += ok ? 1 : 0;
            } else {
                if (ir == 0) s0 = jr == 0 ? sraw : s0;
                sval = sraw - s0;
            }
            if (MASKED) {
                sr += r;
                srr = fmaf(r, r, srr);
            }
            ss += sval;
            sss = fmaf(sval, sval, sss);
            srs = fmaf(r, sval, srs);
        }
    }
    m.sr = MASKED ? sr : wr.sr;
    m.srr = MASKED ? srr : wr.srr;
    m.ss = ss; m.sss = sss; m.srs = srs;
    m.cnt = MASKED ? cnt : NT * NT;
}

// ---------------------------------------------------------------------------
// True if the tap loop, with its FP32 arithmetic, will find every tap of the
// window valid (|w_z| >= 1e-8, 0 <= u < W-1, 0 <= v < H-1), so the unmasked
// loop gives the masked loop's result.  A tap's w = c + ox h_0 + oy h_1 is
// affine in the offset, so over the square |ox|, |oy| <= R its extremes are
// c -/+ R (|a| + |b|): no corners to enumerate.  The conditions used are
//   min w_x > 0, min w_y > 0, min w_z > 1e-8      (then u, v >= 0), and
//   max (w_x - Wq w_z) < 0, max (w_y - Hq w_z) < 0,
// with Wq = (W-1)(1 - 2^-20), so that u = w_x * rcp(w_z) < W-1 even with
// v_rcp_f32's 1-ulp error and the rounding of the product.  Each bound is
// tightened by 2^-21 (8 FP32 rounding units) times the magnitude of the terms
// it is built from (T = |c| + R (|a| + |b|)): a tap's coordinate carries at
// most 2 units of that (two FMAs), the bound's own evaluation at most 4.  So
// the test is exact in what it promises, also for grazing windows where the
// terms nearly cancel.  NaN or infinite inputs fail every comparison.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool window_inside(float cx, float cy, float cz, float h00, float h01,
                                              float h10, float h11, float h20, float h21, float R,
                                              float Wq, float Hq)
{
    constexpr float kE = 4.76837158203125e-07f;   // 2^-21
    const float ax = fabsf(h00) + fabsf(h01), ay = fabsf(h10) + fabsf(h11);
    const float az = fabsf(h20) + fabsf(h21);
    const float Tx = fmaf(R, ax, fabsf(cx)), Ty = fmaf(R, ay, fabsf(cy));
    const float Tz = fmaf(R, az, fabsf(cz));
    const float lox = fmaf(-R, ax, cx), loy = fmaf(-R, ay, cy), loz = fmaf(-R, az, cz);
    const float gx = fmaf(R, fabsf(fmaf(-Wq, h20, h00)) + fabsf(fmaf(-Wq, h21, h01)),
                          fmaf(-Wq, cz, cx));
    const float gy = fmaf(R, fabsf(fmaf(-Hq, h20, h10)) + fabsf(fmaf(-Hq, h21, h11)),
                          fmaf(-Hq, cz, cy));
    return (fmaf(-kE, Tx, lox) >= 0.0f) && (fmaf(-kE, Ty, loy) >= 0.0f) &&
           (fmaf(-kE, Tz, loz) >= 1e-8f) && (fmaf(kE, fmaf(Wq, Tz, Tx), gx) < 0.0f) &&
           (fmaf(kE, fmaf(Hq, Tz, Ty), gy) < 0.0f);
}

// True if every tap of the window has |w_z| >= 1e-8 in its FP32 evaluation:
// w_z is affine in the offset, so its extremes over |ox|, |oy| <= R are
// c_z -/+ R (|h20| + |h21|), each bound tightened by 2^-21 of the terms'
// magnitude as in window_inside (the taps' two FMAs, the bound's own
// rounding).  Then the masked loop's per-tap |w_z| test always passes.
__device__ __forceinline__ bool z_clear(float cz, float h20, float h21, float R)
{
    constexpr float kE = 4.76837158203125e-07f;   // 2^-21
    const float az = fabsf(h20) + fabsf(h21);
    const float Tz = fmaf(R, az, fabsf(cz));
    return (fmaf(-kE, Tz, fmaf(-R, az, cz)) >= 1e-8f) || (fmaf(kE, Tz, fmaf(R, az, cz)) <= -1e-8f);
}

// ---------------------------------------------------------------------------
// Cost of one candidate plane (d, n) at pixel (x, y): patchmatch_mvs.cu:222-343.
// lwin points at the pixel's window top-left tap in the LDS reference tile.
// (rx, ry, rz) is the unnormalised ray K^-1 [x, y, 1] in FP64.
//
// Warp: w(p) = H p with H = A + b g^T, g = K^-T n / (n.P) (the reference's
// q + b * factor, :280-292, in closed form).  The window centre w(x, y) and
// the columns of H are formed in FP64; the taps are w(x, y) + dx h0 + dy h1
// in FP32 with |dx|, |dy| <= 5, so the coordinates are accurate to the FP32
// representation of w even for grazing planes.
// ---------------------------------------------------------------------------
template <int NT, int ST, int STORE, int RU = (NT >= 9 ? 2 : 1)>
__device__ __forceinline__ float cand_cost(const DevView& v, const float* __restrict__ lwin,
                                           int lp, const WinRef& wr, int x, int y, double rx,
                                           double ry, double rz, float d, float nx, float ny,
                                           float nz)
{
    constexpr int R = (NT - 1) * ST / 2;
    constexpr int kMinCount = (NT * NT) / 2;
    constexpr bool kLocalCoords = NT * NT <= HP3D_LOCAL_TAPS;
    const double dd = d, dnx = nx, dny = ny, dnz = nz;
    const double nP = dd * fma(dnx, rx, fma(dny, ry, dnz * rz));
    if (fabs(nP) < 1e-8) return 2.0f;
    const double inv = 1.0 / nP;
    const double* ki = v.kinv_d;
    // g = K^-T n / (n.P): the reference's factor(p) = g.p
    const double g0 = fma(ki[0], dnx, fma(ki[3], dny, ki[6] * dnz)) * inv;
    const double g1 = fma(ki[1], dnx, fma(ki[4], dny, ki[7] * dnz)) * inv;
    const double g2 = fma(ki[2], dnx, fma(ki[5], dny, ki[8] * dnz)) * inv;
    const double X = x, Y = y;
    const double gc = fma(g0, X, fma(g1, Y, g2));   // factor at the window centre

    float top[kMaxTopK];
#pragma unroll
    for (int k = 0; k < kMaxTopK; ++k) top[k] = 2.0f;

    const unsigned umax = (unsigned)(v.W - 2), vmax = (unsigned)(v.H - 2);
    constexpr float kShrink = 0.99999904632568359375f;   // 1 - 2^-20
    const float Wq = (float)(v.W - 1) * kShrink, Hq = (float)(v.H - 1) * kShrink;
    const bool small_k = v.top_k <= 3;   // wave-uniform

    for (int s = 0; s < v.S; ++s) {
        const double* A = v.src[s].A;
        const double* b = v.src[s].b;
        // centre of the window and the two tap-step columns of H
        const float cx = (float)fma(b[0], gc, fma(A[0], X, fma(A[1], Y, A[2])));
        const float cy = (float)fma(b[1], gc, fma(A[3], X, fma(A[4], Y, A[5])));
        const float cz = (float)fma(b[2], gc, fma(A[6], X, fma(A[7], Y, A[8])));
        const float h00 = (float)fma(b[0], g0, A[0]), h01 = (float)fma(b[0], g1, A[1]);
        const float h10 = (float)fma(b[1], g0, A[3]), h11 = (float)fma(b[1], g1, A[4]);
        const float h20 = (float)fma(b[2], g0, A[6]), h21 = (float)fma(b[2], g1, A[7]);
        const SrcImage<STORE> img(v, s);

        // Small windows' taps relative to an integer base (bu, bv) near the
        // centre's source position: the base comes out of the centre and the
        // columns in FP64 (u - bu = (w_x - bu w_z) / w_z, still affine in the
        // tap offset over w_z), so the taps' FP32 arithmetic works on u - bu,
        // a few pixels, and rounds to ~ulp(R) instead of ~ulp(u) (1e-5 px at
        // u ~ 100), an error a 3x3 window's NCC amplifies past 1e-4 of cost.
        // A centre that projects beyond 2^22 pixels keeps the base 0.
        float tx = cx, ty = cy, t00 = h00, t01 = h01, t10 = h10, t11 = h11;
        int bu = 0, bv = 0;
        if constexpr (kLocalCoords) {
            const float rc = __builtin_amdgcn_rcpf(cz);
            const float pu = cx * rc, pv = cy * rc;
            bu = fabsf(pu) < 4194304.0f ? (int)rintf(pu) : 0;
            bv = fabsf(pv) < 4194304.0f ? (int)rintf(pv) : 0;
            const double nu = -(double)bu, nv = -(double)bv;
            const double CZ = fma(b[2], gc, fma(A[6], X, fma(A[7], Y, A[8])));
            const double H20 = fma(b[2], g0, A[6]), H21 = fma(b[2], g1, A[7]);
            tx = (float)fma(nu, CZ, fma(b[0], gc, fma(A[0], X, fma(A[1], Y, A[2]))));
            t00 = (float)fma(nu, H20, fma(b[0], g0, A[0]));
            t01 = (float)fma(nu, H21, fma(b[0], g1, A[1]));
            ty = (float)fma(nv, CZ, fma(b[1], gc, fma(A[3], X, fma(A[4], Y, A[5]))));
            t10 = (float)fma(nv, H20, fma(b[1], g0, A[3]));
            t11 = (float)fma(nv, H21, fma(b[1], g1, A[4]));
        }

        // The unmasked tap loop when every lane's window is inside the
        // source image (windows with w_z < 0 throughout take the masked
        // loop, which gives the same result).
        const bool inside = window_inside(cx, cy, cz, h00, h01, h10, h11, h20, h21, (float)R,
                                          Wq, Hq);
        TapSums m;
        if (__all(inside)) {
            tap_window<NT, ST, STORE, false, true, RU>(img, lwin, lp, tx, ty, cz, t00, t01, t10,
                                                       t11, h20, h21, umax, vmax, wr, m, bu, bv);
        } else if (STORE == kStoreF16 && __all(z_clear(cz, h20, h21, (float)R))) {
            // FP16 storage only: a third loop without the per-tap |w_z| test
            // measured +2 % there and -12 % with FP32 storage at stride 1
            tap_window<NT, ST, STORE, true, false, RU>(img, lwin, lp, tx, ty, cz, t00, t01, t10,
                                                       t11, h20, h21, umax, vmax, wr, m, bu, bv);
        } else {
            tap_window<NT, ST, STORE, true, true, RU>(img, lwin, lp, tx, ty, cz, t00, t01, t10,
                                                      t11, h20, h21, umax, vmax, wr, m, bu, bv);
        }
        const int cnt = m.cnt;
        const float sr = m.sr, ss = m.ss, srr = m.srr, sss = m.sss, srs = m.srs;
        float c = 2.0f;
        if (cnt >= kMinCount) {
            const double fc = (double)cnt;
            const double num = fc * (double)srs - (double)sr * (double)ss;
            const double dr = fc * (double)srr - (double)sr * (double)sr;
            const double ds = fc * (double)sss - (double)ss * (double)ss;
            if (!(dr < (double)v.den_eps) && !(ds < (double)v.den_eps)) {
                // dr, ds >= den_eps (~0.065): far from the denormal range,
                // where rsqrtf would first rescale, so the bare v_rsq_f32
                float ncc = (float)num * __builtin_amdgcn_rsqf((float)dr * (float)ds);
                ncc = fmaxf(-1.0f, fminf(1.0f, ncc));
                c = 1.0f - ncc;
            }
        }
        // insert into the ascending list of the smallest costs: 3 entries
        // suffice when at most 3 are averaged (the reference's top-3), else 8
        if (small_k) {
#pragma unroll
            for (int k = 2; k > 0; --k)
                top[k] = (c < top[k - 1]) ? top[k - 1] : (c < top[k] ? c : top[k]);
        } else {
#pragma unroll
            for (int k = kMaxTopK - 1; k > 0; --k)
                top[k] = (c < top[k - 1]) ? top[k - 1] : (c < top[k] ? c : top[k]);
        }
        top[0] = c < top[0] ? c : top[0];
    }
    const int k = v.top_k < v.S ? v.top_k : v.S;
    float sum = 0.0f;
#pragma unroll
    for (int i = 0; i < kMaxTopK; ++i)
        if (i < k) sum += top[i];
    return sum / (float)k;
}

// ---------------------------------------------------------------------------
// Reference tile staging: rows y0-R .. y0+TY-1+R, columns 2*bx*TX-R ..
// 2*bx*TX+2*TX-1+R, clamped to the image (texture clamp, :807-808).
// ---------------------------------------------------------------------------
template <int R, int STORE, int TX, int TY>
struct TileGeom {
    static constexpr int LW = 2 * TX + 2 * R;   // LDS tile width (pixels)
    static constexpr int LH = TY + 2 * R;       // LDS tile height
    static constexpr int LP = LW + 1;           // LDS row pitch
};

// Tile (bx, by) covers compacted columns bx*TX .. bx*TX+TX-1 (pixel columns
// 2*bx*TX .. 2*bx*TX+2*TX-1) of rows by*TY .. by*TY+TY-1.
template <int R, int STORE, int TX, int TY>
__device__ __forceinline__ void stage_ref_tile(const DevView& v, float* tile, int bx, int by)
{
    using G = TileGeom<R, STORE, TX, TY>;
    const int X0 = 2 * bx * TX - R;
    const int Y0 = by * TY - R;
    for (int k = threadIdx.x; k < G::LW * G::LH; k += TX * TY) {
        const int ly = k / G::LW, lx = k - ly * G::LW;
        const int gx = min(max(X0 + lx, 0), v.W - 1);
        const int gy = min(max(Y0 + ly, 0), v.H - 1);
        tile[ly * G::LP + lx] = load_plain<STORE>(v.ref_plain, gy * v.pitch_plain + gx);
    }
    __syncthreads();
}

// Random initial hypothesis of pixel (x, y): init_depth_normal_kernel (:164-216).
__device__ __forceinline__ float4 init_hyp(const DevView& v, int x, int y, uint64_t seed)
{
    uint32_t c0 = (uint32_t)(y * v.W + x), c1 = 0u, c2 = 0u, c3 = kDomainInit;
    philox(c0, c1, c2, c3, (uint32_t)seed, (uint32_t)(seed >> 32));
    const float d = v.dmin + u01(c0) * (v.dmax - v.dmin);
    float rx, ry, rz;
    unit_ray(v.kinv, x, y, rx, ry, rz);
    const float theta = u01(c1) * 6.28318530f;
    const float cos_phi = u01(c2);
    const float sin_phi = sqrtf(1.0f - cos_phi * cos_phi);
    float nx = sin_phi * (float)cos((double)theta);
    float ny = sin_phi * (float)sin((double)theta);
    float nz = cos_phi;
    if (nx * rx + ny * ry + nz * rz > 0.0f) { nx = -nx; ny = -ny; nz = -nz; }
    const float nlen = sqrtf(nx * nx + ny * ny + nz * nz);
    if (nlen > 1e-12f) { nx = nx / nlen; ny = ny / nlen; nz = nz / nlen; }
    return make_float4(d, nx, ny, nz);
}

// Cost of the current hypothesis of the colour-`parity` pixel of lane (tx, ty)
// in tile (bx, by); returns false if the lane has no pixel.
template <int NT, int ST, int STORE, int TX, int TY, bool COH = false>
__device__ __forceinline__ bool cost_pixel(const DevView& v, const float* tile, int bx, int by,
                                           int tx, int ty, int parity)
{
    constexpr int R = (NT - 1) * ST / 2;
    using G = TileGeom<R, STORE, TX, TY>;
    const int i = bx * TX + tx;
    const int y = by * TY + ty;
    const int x = 2 * i + ((y + parity) & 1);
    if (y >= v.H || x >= v.W) return false;
    const float* lwin = tile + ty * G::LP + (x - 2 * bx * TX);
    double rx, ry, rz;
    ray_fp64(v.kinv_d, x, y, rx, ry, rz);
    const int p = y * v.Wc + i;
    const float4 h = StateIO<COH>::ld4(v.hyp[parity], p);
    const WinRef wr = window_ref_sums<NT, ST>(lwin, G::LP);
    StateIO<COH>::st1(v.cost[parity], p,
                      cand_cost<NT, ST, STORE, 1>(v, lwin, G::LP, wr, x, y, rx, ry, rz, h.x, h.y,
                                                  h.z, h.w));
    return true;
}

// Per-wave LDS scratch of a half-pass: the wave's list of neighbour
// candidates that must be evaluated (owner lane | candidate << 6), their
// hypotheses (indexed the same way: each is loaded once, by its owner, and
// read from here by the lane that evaluates it and by the owner when it
// accepts it) and their costs, at most 4 per lane; and the depth-bucket
// counts of the first iteration's ordering.
struct WaveXchg {
    unsigned char item[256];
    float4 hyp[256];
    float cost[256];
    int hist[64];
};

// Cost of the current hypotheses of colour `parity` at this workgroup's
// pixels (compute_ncc_cost_kernel, patchmatch_mvs.cu:349-513), evaluated in
// inverse-depth order within each wave: the hypotheses are random, so each
// lane's windows would gather from an unrelated part of every source; a
// counting sort of the wave's hypotheses by inverse depth (disparity) hands
// lane k the k-th, so a wave's gathers fall in a narrower band of each
// source.  Any lane can evaluate any pixel's candidate (the whole window
// tile is in LDS) and a cost is a pure function of the pixel and the
// hypothesis, so the result is the same as one lane per pixel.  Every
// thread of the workgroup must call this (it synchronises the group).
template <int NT, int ST, int STORE, int TX, int TY, bool COH = false>
__device__ __forceinline__ void cost_tile_sorted(const DevView& v, const float* tile, int bx,
                                                 int by, int tx, int ty, int parity, WaveXchg& xw)
{
    constexpr int R = (NT - 1) * ST / 2;
    using G = TileGeom<R, STORE, TX, TY>;
    using IO = StateIO<COH>;
    constexpr int kBuckets = 64;
    const int lane = threadIdx.x & 63;
    const int wave0 = threadIdx.x & ~63;
    const int i = bx * TX + tx, y = by * TY + ty;
    const int x = 2 * i + ((y + parity) & 1);
    const bool active = y < v.H && x < v.W;
    float4 h = make_float4(0.f, 0.f, 0.f, 0.f);
    if (active) h = IO::ld4(v.hyp[parity], y * v.Wc + i);
    xw.hist[lane] = 0;
    __builtin_amdgcn_wave_barrier();
    int bk = 0, rk = 0;
    if (active) {
        const float ilo = 1.0f / v.dmax, iscale = (float)kBuckets / (1.0f / v.dmin - ilo);
        const int b = (int)((1.0f / h.x - ilo) * iscale);
        bk = b < 0 ? 0 : (b > kBuckets - 1 ? kBuckets - 1 : b);
        rk = atomicAdd(&xw.hist[bk], 1);
    }
    __builtin_amdgcn_wave_barrier();
    const int cnt = xw.hist[lane];
    int incl = cnt;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int t = __shfl_up(incl, d, 64);
        if (lane >= d) incl += t;
    }
    __builtin_amdgcn_wave_barrier();
    xw.hist[lane] = incl - cnt;
    __builtin_amdgcn_wave_barrier();
    if (active) {
        const int slot = xw.hist[bk] + rk;
        xw.item[slot] = (unsigned char)lane;
        xw.hyp[slot] = h;
    }
    const unsigned total = (unsigned)__popcll(__ballot(active));
    __syncthreads();
    if ((unsigned)lane < total) {
        const int ol = xw.item[lane];
        const float4 hh = xw.hyp[lane];
        const int t_o = wave0 + ol;
        const int tx_o = t_o % TX, ty_o = t_o / TX;
        const int i_o = bx * TX + tx_o, y_o = by * TY + ty_o;
        const int x_o = 2 * i_o + ((y_o + parity) & 1);
        const float* lwin = tile + ty_o * G::LP + (x_o - 2 * bx * TX);
        const WinRef wr = window_ref_sums<NT, ST>(lwin, G::LP);
        double rx, ry, rz;
        ray_fp64(v.kinv_d, x_o, y_o, rx, ry, rz);
        IO::st1(v.cost[parity], y_o * v.Wc + i_o,
                cand_cost<NT, ST, STORE, 1>(v, lwin, G::LP, wr, x_o, y_o, rx, ry, rz, hh.x, hh.y,
                                         hh.z, hh.w));
    }
}

// One pixel of a red (parity 0) or black (parity 1) half-pass:
// propagate_and_refine_kernel (:519-655).  Candidates L, R, U, D from the
// other colour plane, then two random refinements with ranges
// (dmax-dmin)/2 * 0.6^iter and 0.6^iter, halved per step.  Strict '<'
// acceptance throughout.
//
// A neighbour candidate that is bit-identical to the pixel's own hypothesis,
// or to an earlier valid candidate, cannot be accepted: its cost is the same
// deterministic function of (pixel, hypothesis), and bc is already at most
// that cost.  Once propagation has spread planes, 40-60 % of the neighbour
// candidates are such repeats, but rarely all 64 lanes of a wave at once.  So
// the wave lists the candidates its lanes still need (also leaving out
// neighbours unchanged since this pixel's previous pass, see below), any
// lane evaluates any
// pixel's candidate (all windows of the workgroup are in LDS), and each
// owner then applies the costs in L, R, U, D order.  The result is the one
// the plain per-lane loop gives, with fewer evaluation rounds per wave.
// Every thread of the workgroup must call this (it synchronises the group).
template <int NT, int ST, int STORE, int TX, int TY, bool COH = false>
__device__ __forceinline__ void halfpass_pixel(const DevView& v, const float* tile, int bx, int by,
                                               int tx, int ty, int iteration, int parity,
                                               uint64_t pass_seed, WaveXchg& xw)
{
    constexpr int R = (NT - 1) * ST / 2;
    using G = TileGeom<R, STORE, TX, TY>;
    using IO = StateIO<COH>;
    const int lane = threadIdx.x & 63;
    const int wave0 = threadIdx.x & ~63;
    const int i = bx * TX + tx;
    const int y = by * TY + ty;
    const int o = (y + parity) & 1;
    const int x = 2 * i + o;
    const bool active = y < v.H && x < v.W;
    const float* lwin = tile + ty * G::LP + (x - 2 * bx * TX);
    const float dmin = v.dmin, dmax = v.dmax;
    const int Wc = v.Wc;
    float4* own = v.hyp[parity];
    const float4* oth = v.hyp[parity ^ 1];
    const int p = y * Wc + i;
    // neighbour positions in the other plane: L = i+o-1, R = i+o (row y),
    // U, D = i (rows y-1, y+1)
    auto nb = [&](int k, int pp, int oo) {
        return k == 0 ? pp + oo - 1 : k == 1 ? pp + oo : k == 2 ? pp - Wc : pp + Wc;
    };
    auto same = [](const float4& a, const float4& b) {
        return __float_as_uint(a.x) == __float_as_uint(b.x) &&
               __float_as_uint(a.y) == __float_as_uint(b.y) &&
               __float_as_uint(a.z) == __float_as_uint(b.z) &&
               __float_as_uint(a.w) == __float_as_uint(b.w);
    };

    float4 best = make_float4(0.f, 0.f, 0.f, 0.f);
    float bc = 2.0f;
    bool changed = false;
    unsigned need = 0;   // bit k: neighbour k has to be evaluated
    float hd[4] = {0.f, 0.f, 0.f, 0.f};   // the candidates' depths
    if (active) {
        best = IO::ld4(own, p);
        bc = IO::ld1(v.cost[parity], p);
        // From the second iteration of a solve on, this pixel evaluated its
        // neighbours' hypotheses one iteration ago; a neighbour that did not
        // change since (its flag from the other colour's last half-pass)
        // still holds that hypothesis, whose cost was then >= bc, and bc has
        // only decreased: it cannot be accepted now.
        const bool seen = iteration > v.first_iteration;
        const unsigned char* chg_oth = v.chg[parity ^ 1];
        float4 h[4];
        bool ok[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const bool has = k == 0 ? x >= 1 : k == 1 ? x + 1 < v.W : k == 2 ? y >= 1 : y + 1 < v.H;
            ok[k] = false;
            h[k] = best;
            bool stale = false;
            if (has) {
                const int q = nb(k, p, o);
                h[k] = IO::ld4(oth, q);
                ok[k] = !(h[k].x < dmin || h[k].x > dmax);
                stale = seen && IO::ldb(chg_oth, q) == 0;
            }
            bool n = ok[k] && !stale && !same(h[k], best);
#pragma unroll
            for (int j = 0; j < k; ++j) n = n && !(ok[j] && same(h[j], h[k]));
            need |= (unsigned)n << k;
            hd[k] = h[k].x;
            if (n) xw.hyp[lane | (k << 6)] = h[k];
        }
    }
    const WinRef wr = window_ref_sums<NT, ST>(lwin, G::LP);

    // Wave-wide list ordered by owner lane, then candidate: a round of 64
    // items then covers ~25 neighbouring pixels whose candidates project to
    // nearly the same source rows and columns, so its gathers touch far
    // fewer 64-B segments than 64 pixels' worth would (the L1 access rate,
    // not bandwidth, bounds this kernel).
    const unsigned long long below = (1ull << lane) - 1ull;
    int slot[4];
    unsigned total = 0;
    int base = 0;   // items of the lanes below this one
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const unsigned long long m = __ballot((need >> k) & 1u);
        base += __popcll(m & below);
        total += (unsigned)__popcll(m);
    }
#if HP3D_BUCKET
    if (NT < 9 && iteration < v.first_iteration + HP3D_BUCKET) {
        // First iteration: the neighbours hold random, unrelated depths, so
        // owner order does not cluster the gathers.  Order the items by
        // inverse depth (disparity) instead, in kBuckets buckets (a counting
        // sort in LDS; the order within a bucket does not matter), so each
        // round's windows lie in a narrower band of every source.  (+1.4 %
        // on 6x6-tap windows; with 11x11 windows it measured slower.)
        constexpr int kBuckets = 64;
        const float ilo = 1.0f / dmax, iscale = (float)kBuckets / (1.0f / dmin - ilo);
        xw.hist[lane] = 0;
        __builtin_amdgcn_wave_barrier();
        int bk[4], rk[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            bk[k] = 0;
            rk[k] = 0;
            if ((need >> k) & 1u) {
                const int b = (int)((1.0f / hd[k] - ilo) * iscale);
                bk[k] = b < 0 ? 0 : (b > kBuckets - 1 ? kBuckets - 1 : b);
                rk[k] = atomicAdd(&xw.hist[bk[k]], 1);
            }
        }
        __builtin_amdgcn_wave_barrier();
        const int cnt = xw.hist[lane];
        int incl = cnt;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int t = __shfl_up(incl, d, 64);
            if (lane >= d) incl += t;
        }
        __builtin_amdgcn_wave_barrier();
        xw.hist[lane] = incl - cnt;
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            slot[k] = 0;
            if ((need >> k) & 1u) {
                slot[k] = xw.hist[bk[k]] + rk[k];
                xw.item[slot[k]] = (unsigned char)(lane | (k << 6));
            }
        }
    } else
#endif
    {
        int off = base;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            slot[k] = off;
            if ((need >> k) & 1u) {
                xw.item[off] = (unsigned char)(lane | (k << 6));
                ++off;
            }
        }
    }
    __syncthreads();
    for (unsigned r = 0; r < total; r += 64) {
        const unsigned g = r + lane;
        const unsigned it = xw.item[g < total ? g : total - 1];
        const int ol = it & 63;
        const WinRef wro{__shfl(wr.sr, ol, 64), __shfl(wr.srr, ol, 64)};
        if (g < total) {
            const int t_o = wave0 + ol;
            const int tx_o = t_o % TX, ty_o = t_o / TX;
            const int i_o = bx * TX + tx_o, y_o = by * TY + ty_o;
            const int o_o = (y_o + parity) & 1, x_o = 2 * i_o + o_o;
            const float4 h = xw.hyp[it];
            double rx, ry, rz;
            ray_fp64(v.kinv_d, x_o, y_o, rx, ry, rz);
            xw.cost[g] = cand_cost<NT, ST, STORE>(v, tile + ty_o * G::LP + (x_o - 2 * bx * TX),
                                                  G::LP, wro, x_o, y_o, rx, ry, rz, h.x, h.y, h.z,
                                                  h.w);
        }
    }
    __syncthreads();
    if (!active) return;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (!((need >> k) & 1u)) continue;
        const float c = xw.cost[slot[k]];
        if (c < bc) { bc = c; best = xw.hyp[lane | (k << 6)]; changed = true; }
    }

    double rx, ry, rz;
    ray_fp64(v.kinv_d, x, y, rx, ry, rz);
    float ux, uy, uz;
    unit_ray(v.kinv, x, y, ux, uy, uz);
    float depth_range = (dmax - dmin) * 0.5f;
    float normal_range = 1.0f;
    for (int it = 0; it < iteration; ++it) { depth_range *= 0.6f; normal_range *= 0.6f; }
    const uint32_t idx = (uint32_t)(y * v.W + x);
#pragma unroll 1
    for (int step = 0; step < 2; ++step) {
        uint32_t c0 = idx, c1 = (uint32_t)iteration, c2 = (uint32_t)step, c3 = kDomainRefine;
        philox(c0, c1, c2, c3, (uint32_t)pass_seed, (uint32_t)(pass_seed >> 32));
        const float rand_d = u01(c0) * 2.0f - 1.0f;
        float cd = best.x + rand_d * depth_range;
        cd = fmaxf(dmin, fminf(dmax, cd));
        float cnx = best.y + (u01(c1) * 2.0f - 1.0f) * normal_range;
        float cny = best.z + (u01(c2) * 2.0f - 1.0f) * normal_range;
        float cnz = best.w + (u01(c3) * 2.0f - 1.0f) * normal_range;
        const float nlen = sqrtf(cnx * cnx + cny * cny + cnz * cnz);
        if (nlen > 1e-12f) { cnx = cnx / nlen; cny = cny / nlen; cnz = cnz / nlen; }
        if (cnx * ux + cny * uy + cnz * uz > 0.0f) { cnx = -cnx; cny = -cny; cnz = -cnz; }
        const float c = cand_cost<NT, ST, STORE>(v, lwin, G::LP, wr, x, y, rx, ry, rz, cd, cnx,
                                                 cny, cnz);
        if (c < bc) { bc = c; best = make_float4(cd, cnx, cny, cnz); changed = true; }
        depth_range *= 0.5f;
        normal_range *= 0.5f;
    }
    IO::st4(own, p, best);
    IO::st1(v.cost[parity], p, bc);
    IO::stb(v.chg[parity], p, changed ? 1u : 0u);
}

// Row-major outputs of pixel (x, y) from the colour planes.
template <bool COH = false>
__device__ __forceinline__ void emit_pixel(const DevView& v, int x, int y, float* depth,
                                           float* normal, float* cost, float* conf)
{
    const int c = (x + y) & 1, p = y * v.Wc + (x >> 1);
    const size_t q = (size_t)y * v.W + x;
    const float4 h = StateIO<COH>::ld4(v.hyp[c], p);
    if (depth) depth[q] = h.x;
    if (normal) { normal[3 * q] = h.y; normal[3 * q + 1] = h.z; normal[3 * q + 2] = h.w; }
    if (cost || conf) {
        const float k = StateIO<COH>::ld1(v.cost[c], p);
        if (cost) cost[q] = k;
        if (conf) conf[q] = 1.0f - k * 0.5f;   // patchmatch_mvs.cu:1249-1252
    }
}

}  // namespace hp3d
