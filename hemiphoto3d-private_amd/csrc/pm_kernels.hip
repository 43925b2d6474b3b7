// pm_kernels.hip -- gfx950 kernels of the PatchMatch-MVS hot path.
//
// Reference behaviour (src/dense_recon/patchmatch_mvs.cu):
//   pm_init_kernel      <- init_depth_normal_kernel      (:164-216)
//   cand_cost           <- computePixelCost               (:222-343)
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
//     is formed once, so a tap costs 3 FMAs for the warp, one v_rcp_f32, one
//     load of the 4 bilinear coefficients and 3 FMAs for the sample;
//   * the five NCC moments are accumulated in FP32 on values shifted by
//     -128 DN; the variance cancellation is done in FP64.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pm_device.h"
#include "pm_launch.h"

namespace hp3d {

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

template <int STORE> __device__ __forceinline__ float4 load_coef(const void* p, int off);
template <> __device__ __forceinline__ float4 load_coef<kStoreF32>(const void* p, int off)
{
    return static_cast<const float4*>(p)[off];
}
template <> __device__ __forceinline__ float4 load_coef<kStoreF16>(const void* p, int off)
{
    const half4_t h = static_cast<const half4_t*>(p)[off];
    return make_float4((float)h.x, (float)h.y, (float)h.z, (float)h.w);
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
template <int NT, int ST, int STORE>
__device__ __forceinline__ float cand_cost(const DevView& v, const float* __restrict__ lwin,
                                           int lp, int x, int y, double rx, double ry, double rz,
                                           float d, float nx, float ny, float nz)
{
    constexpr int R = (NT - 1) * ST / 2;
    constexpr int kMinCount = (NT * NT) / 2;
    const double dd = d, dnx = nx, dny = ny, dnz = nz;
    const double nP = dnx * (dd * rx) + dny * (dd * ry) + dnz * (dd * rz);
    if (fabs(nP) < 1e-8) return 2.0f;
    const double inv = 1.0 / nP;
    const double* ki = v.kinv_d;
    // g = K^-T n / (n.P): the reference's factor(p) = g.p
    const double g0 = (ki[0] * dnx + ki[3] * dny + ki[6] * dnz) * inv;
    const double g1 = (ki[1] * dnx + ki[4] * dny + ki[7] * dnz) * inv;
    const double g2 = (ki[2] * dnx + ki[5] * dny + ki[8] * dnz) * inv;
    const double X = x, Y = y;
    const double gc = g0 * X + g1 * Y + g2;   // factor at the window centre

    float top[kMaxTopK];
#pragma unroll
    for (int k = 0; k < kMaxTopK; ++k) top[k] = 2.0f;

    const float Wm1 = (float)(v.W - 1), Hm1 = (float)(v.H - 1);
    const int pc = v.pitch_coef;

    for (int s = 0; s < v.S; ++s) {
        const double* A = v.src[s].A;
        const double* b = v.src[s].b;
        // centre of the window and the two tap-step columns of H
        const float cx = (float)(A[0] * X + A[1] * Y + A[2] + b[0] * gc);
        const float cy = (float)(A[3] * X + A[4] * Y + A[5] + b[1] * gc);
        const float cz = (float)(A[6] * X + A[7] * Y + A[8] + b[2] * gc);
        const float h00 = (float)(A[0] + b[0] * g0), h01 = (float)(A[1] + b[0] * g1);
        const float h10 = (float)(A[3] + b[1] * g0), h11 = (float)(A[4] + b[1] * g1);
        const float h20 = (float)(A[6] + b[2] * g0), h21 = (float)(A[7] + b[2] * g1);
        const void* img = v.src_coef[s];

        float sr = 0.f, ss = 0.f, srr = 0.f, sss = 0.f, srs = 0.f;
        int cnt = 0;
#pragma unroll 1
        for (int j = 0; j < NT; ++j) {
            const float oy = (float)(j * ST - R);
            const float bx = fmaf(h01, oy, cx);
            const float by = fmaf(h11, oy, cy);
            const float bz = fmaf(h21, oy, cz);
            const float* lrow = lwin + j * ST * lp;
#pragma unroll
            for (int i = 0; i < NT; ++i) {
                const float ox = (float)(i * ST - R);
                const float wx = fmaf(h00, ox, bx);
                const float wy = fmaf(h10, ox, by);
                const float wz = fmaf(h20, ox, bz);
                const float rw = __builtin_amdgcn_rcpf(wz);
                const float u = wx * rw, w = wy * rw;
                const bool ok = (fabsf(wz) >= 1e-8f) & (u >= 0.0f) & (u < Wm1) & (w >= 0.0f) &
                                (w < Hm1);
                const float fu = floorf(u), fw = floorf(w);
                const int off = ok ? ((int)fw * pc + (int)fu) : 0;
                const float4 c = load_coef<STORE>(img, off);
                const float au = u - fu, aw = w - fw;
                const float sval = fmaf(aw, fmaf(au, c.w, c.z), fmaf(au, c.y, c.x));
                const float r = lrow[i * ST];
                const float sm = ok ? sval : 0.0f;
                const float rm = ok ? r : 0.0f;
                sr += rm;
                ss += sm;
                srr = fmaf(rm, rm, srr);
                sss = fmaf(sm, sm, sss);
                srs = fmaf(rm, sm, srs);
                cnt += ok ? 1 : 0;
            }
        }
        float c = 2.0f;
        if (cnt >= kMinCount) {
            const double fc = (double)cnt;
            const double num = fc * (double)srs - (double)sr * (double)ss;
            const double dr = fc * (double)srr - (double)sr * (double)sr;
            const double ds = fc * (double)sss - (double)ss * (double)ss;
            if (!(dr < (double)v.den_eps) && !(ds < (double)v.den_eps)) {
                float ncc = (float)num * rsqrtf((float)dr * (float)ds);
                ncc = fmaxf(-1.0f, fminf(1.0f, ncc));
                c = 1.0f - ncc;
            }
        }
        // insert into the ascending list of the kMaxTopK smallest costs
#pragma unroll
        for (int k = kMaxTopK - 1; k > 0; --k)
            top[k] = (c < top[k - 1]) ? top[k - 1] : (c < top[k] ? c : top[k]);
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
template <int R, int STORE, int TX, int TY, int LW, int LH, int LP>
__device__ __forceinline__ void stage_ref_tile(const DevView& v, float* tile)
{
    const int X0 = 2 * (int)blockIdx.x * TX - R;
    const int Y0 = (int)blockIdx.y * TY - R;
    for (int k = threadIdx.x; k < LW * LH; k += TX * TY) {
        const int ly = k / LW, lx = k - ly * LW;
        const int gx = min(max(X0 + lx, 0), v.W - 1);
        const int gy = min(max(Y0 + ly, 0), v.H - 1);
        tile[ly * LP + lx] = load_plain<STORE>(v.ref_plain, gy * v.pitch_plain + gx);
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// Initial hypotheses: init_depth_normal_kernel (:164-216)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void pm_init_kernel(DevView v, uint64_t seed)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= v.W || y >= v.H) return;
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
    v.hyp[(x + y) & 1][y * v.Wc + (x >> 1)] = make_float4(d, nx, ny, nz);
}

// ---------------------------------------------------------------------------
// Cost of the current hypotheses of one colour: compute_ncc_cost_kernel
// (:349-513), run once per colour before the first iteration.
// ---------------------------------------------------------------------------
template <int NT, int ST, int STORE, int TX, int TY>
__global__ __launch_bounds__(TX * TY) void pm_cost_kernel(DevView v, int parity)
{
    constexpr int R = (NT - 1) * ST / 2;
    constexpr int LW = 2 * TX + 2 * R, LH = TY + 2 * R, LP = LW + 1;
    __shared__ float tile[LH * LP];
    stage_ref_tile<R, STORE, TX, TY, LW, LH, LP>(v, tile);
    const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    const int i = blockIdx.x * TX + tx;
    const int y = blockIdx.y * TY + ty;
    const int x = 2 * i + ((y + parity) & 1);
    if (y >= v.H || x >= v.W) return;
    const float* lwin = tile + ty * LP + (x - 2 * (int)blockIdx.x * TX);
    double rx, ry, rz;
    ray_fp64(v.kinv_d, x, y, rx, ry, rz);
    const int p = y * v.Wc + i;
    const float4 h = v.hyp[parity][p];
    v.cost[parity][p] = cand_cost<NT, ST, STORE>(v, lwin, LP, x, y, rx, ry, rz, h.x, h.y, h.z, h.w);
}

// ---------------------------------------------------------------------------
// One red (parity 0) or black (parity 1) half-pass: propagate_and_refine_kernel
// (:519-655).  Candidates L, R, U, D from the other colour plane, then two
// random refinements with ranges (dmax-dmin)/2 * 0.6^iter and 0.6^iter,
// halved per step.  Strict '<' acceptance throughout.
// ---------------------------------------------------------------------------
template <int NT, int ST, int STORE, int TX, int TY>
__global__ __launch_bounds__(TX * TY) void pm_halfpass_kernel(DevView v, int iteration, int parity,
                                                              uint64_t pass_seed)
{
    constexpr int R = (NT - 1) * ST / 2;
    constexpr int LW = 2 * TX + 2 * R, LH = TY + 2 * R, LP = LW + 1;
    __shared__ float tile[LH * LP];
    stage_ref_tile<R, STORE, TX, TY, LW, LH, LP>(v, tile);
    const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    const int i = blockIdx.x * TX + tx;
    const int y = blockIdx.y * TY + ty;
    const int o = (y + parity) & 1;
    const int x = 2 * i + o;
    if (y >= v.H || x >= v.W) return;
    const float* lwin = tile + ty * LP + (x - 2 * (int)blockIdx.x * TX);
    double rx, ry, rz;
    ray_fp64(v.kinv_d, x, y, rx, ry, rz);
    const float dmin = v.dmin, dmax = v.dmax;
    const int Wc = v.Wc;

    float4* own = v.hyp[parity];
    const float4* oth = v.hyp[parity ^ 1];
    const int p = y * Wc + i;
    float4 best = own[p];
    float bc = v.cost[parity][p];

    // neighbour positions in the other plane: L = i+o-1, R = i+o (row y),
    // U, D = i (rows y-1, y+1)
#pragma unroll 1
    for (int k = 0; k < 4; ++k) {
        const bool has = k == 0 ? x >= 1 : k == 1 ? x + 1 < v.W : k == 2 ? y >= 1 : y + 1 < v.H;
        if (!has) continue;
        const int q = k == 0 ? p + o - 1 : k == 1 ? p + o : k == 2 ? p - Wc : p + Wc;
        const float4 h = oth[q];
        if (h.x < dmin || h.x > dmax) continue;
        const float c = cand_cost<NT, ST, STORE>(v, lwin, LP, x, y, rx, ry, rz, h.x, h.y, h.z, h.w);
        if (c < bc) { bc = c; best = h; }
    }

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
        const float c = cand_cost<NT, ST, STORE>(v, lwin, LP, x, y, rx, ry, rz, cd, cnx, cny, cnz);
        if (c < bc) { bc = c; best = make_float4(cd, cnx, cny, cnz); }
        depth_range *= 0.5f;
        normal_range *= 0.5f;
    }
    own[p] = best;
    v.cost[parity][p] = bc;
}

// ---------------------------------------------------------------------------
// Layout conversion kernels
// ---------------------------------------------------------------------------

// Colour planes -> row-major depth (H*W), AoS normal (H*W*3), cost and
// confidence (H*W).  Null outputs are skipped.
__global__ __launch_bounds__(256) void pm_finalize_kernel(DevView v, float* depth, float* normal,
                                                          float* cost, float* conf)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= v.W || y >= v.H) return;
    const int c = (x + y) & 1, p = y * v.Wc + (x >> 1);
    const size_t q = (size_t)y * v.W + x;
    const float4 h = v.hyp[c][p];
    if (depth) depth[q] = h.x;
    if (normal) { normal[3 * q] = h.y; normal[3 * q + 1] = h.z; normal[3 * q + 2] = h.w; }
    if (cost || conf) {
        const float k = v.cost[c][p];
        if (cost) cost[q] = k;
        if (conf) conf[q] = 1.0f - k * 0.5f;   // patchmatch_mvs.cu:1249-1252
    }
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

// 8-bit or [0,1] float gray -> plain and bilinear-coefficient planes (see
// pm_device.h).  in_u8 / in_f32: exactly one is non-null; in_pitch in bytes.
template <int STORE>
__global__ __launch_bounds__(256) void img_prep_kernel(const uint8_t* in_u8, const float* in_f32,
                                                       size_t in_pitch, int W, int H, void* plain,
                                                       int pitch_plain, void* coef, int pitch_coef)
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
    if (STORE == kStoreF32) static_cast<float*>(plain)[(size_t)y * pitch_plain + x] = a;
    else static_cast<_Float16*>(plain)[(size_t)y * pitch_plain + x] = (_Float16)a;
    if (x < W - 1 && y < H - 1) {
        const float b = val(x + 1, y), c = val(x, y + 1), d = val(x + 1, y + 1);
        const float cb = b - a, cc = c - a, cd = (a - b) - (c - d);
        const size_t q = (size_t)y * pitch_coef + x;
        if (STORE == kStoreF32) {
            static_cast<float4*>(coef)[q] = make_float4(a, cb, cc, cd);
        } else {
            half4_t h;
            h.x = (_Float16)a; h.y = (_Float16)cb; h.z = (_Float16)cc; h.w = (_Float16)cd;
            static_cast<half4_t*>(coef)[q] = h;
        }
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
                           int H, int store, void* plain, int pitch_plain, void* coef,
                           int pitch_coef, hipStream_t st)
{
    if (store == kStoreF32)
        hipLaunchKernelGGL(img_prep_kernel<kStoreF32>, grid_2d(W, H), dim3(256), 0, st, in_u8,
                           in_f32, in_pitch, W, H, plain, pitch_plain, coef, pitch_coef);
    else
        hipLaunchKernelGGL(img_prep_kernel<kStoreF16>, grid_2d(W, H), dim3(256), 0, st, in_u8,
                           in_f32, in_pitch, W, H, plain, pitch_plain, coef, pitch_coef);
    return hipGetLastError();
}

constexpr int kTX = 16, kTY = 16;

template <int NT, int ST, int STORE>
static void launch_pass_t(const DevView& v, int mode, int iteration, int parity,
                          uint64_t pass_seed, hipStream_t st)
{
    const dim3 grid((v.Wc + kTX - 1) / kTX, (v.H + kTY - 1) / kTY);
    if (mode == kPassCost)
        hipLaunchKernelGGL((pm_cost_kernel<NT, ST, STORE, kTX, kTY>), grid, dim3(kTX * kTY), 0, st,
                           v, parity);
    else
        hipLaunchKernelGGL((pm_halfpass_kernel<NT, ST, STORE, kTX, kTY>), grid, dim3(kTX * kTY), 0,
                           st, v, iteration, parity, pass_seed);
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
