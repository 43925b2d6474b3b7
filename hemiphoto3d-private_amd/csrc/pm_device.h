// pm_device.h -- types shared by the host runtime and the gfx950 kernels of
// the PatchMatch-MVS path.
//
// Data layout in HBM (one reference view being solved):
//   * images: plain[y][x] = DN - 128, row pitch a multiple of 64 elements.
//     The shift keeps window sums small (the NCC is invariant to it).
//     Beside the plain plane (reference windows, staged in LDS) every image
//     has a row-difference plane diff[y][x] = (plain[y][x], plain[y+1][x] -
//     plain[y][x]) for source gathers: one load at (x, y) returns the whole
//     bilinear quad.
//     F32 storage: 4 + 8 B/px; F16 storage: 2 + 4 B/px.  For 8-bit input
//     every value is an integer in [-128, 127], exact in FP16, so F16
//     storage is lossless and gives the same results as F32.
//   * hypotheses: two compacted colour planes (red: (x+y) even, black: odd),
//     each H x Wc with Wc = (W+1)/2.  Pixel (x, y) of colour c sits at
//     column i = x >> 1 of plane c.  hyp = (depth, nx, ny, nz) as float4,
//     cost as float, and a byte per pixel: changed in its last half-pass.
//     A half-pass reads the other plane only.
#pragma once
#include <stdint.h>

namespace hp3d {

constexpr int kMaxSrc = 8;
constexpr int kMaxTopK = 8;
// SfM seeds: half side of a seeded square, and |pixel coordinate| of a seed
// (bounds that keep every square's corner arithmetic inside int).
constexpr int kMaxSeedRadius = 4096;
constexpr int kMaxSeedCoord = 1 << 24;
constexpr float kShiftDN = 128.0f;
// Bound on iterations and first_iteration of a solve (keeps every step and
// iteration index, 2 * iterations + 1 included, inside int).
constexpr int kMaxIterations = 1 << 20;
// Largest image plane: taps address a plane with 32-bit buffer offsets, and
// offsets at or above this bound are a guaranteed miss of every plane.
constexpr unsigned long long kMaxPlaneBytes = 0x7fff0000ull;
// Source taps form row offsets with a 24-bit multiply (iy * pitch), so a
// plane's row pitch (in elements) must stay below 2^24.
constexpr int kMaxRowPitch = 1 << 24;

enum Store : int { kStoreF32 = 0, kStoreF16 = 1 };

// Homography terms of one source, kept in FP64: the warp of a window centre
// is formed in FP64 and only the small tap offsets in FP32, so the warped
// coordinates carry no cancellation error from large pixel coordinates.
struct DevSrc {
    double A[9];   // A = K_s R_rel K_ref^-1, row-major
    double b[3];   // b = K_s t_rel
};

struct DevView {
    const void* ref_plain;              // reference plane
    const void* src_plain[kMaxSrc];     // source planes
    const void* src_coef[kMaxSrc];      // source gather planes (row differences)
    int pitch_plain;                    // elements per row of a plane (both kinds)
    int plain_bytes;                    // bytes of one plane (buffer range check)
    int coef_bytes;                     // bytes of one gather plane
    float kinv[9];                      // K_ref^-1 (row-major), hypothesis arithmetic
    double kinv_d[9];                   // K_ref^-1 for the cost
    DevSrc src[kMaxSrc];
    int W, H, S, Wc;
    float dmin, dmax;
    float den_eps;                      // NCC variance guard in stored units^2
    int top_k;
    float4* hyp[2];                     // colour planes (depth, nx, ny, nz)
    float* cost[2];
    // 1 if the pixel's hypothesis changed in its colour's last half-pass;
    // written by every half-pass, read from iteration first_iteration + 1 on
    unsigned char* chg[2];
    int first_iteration;
};

// Philox4x32-10 (Salmon et al., SC'11).  Used on host and device.
constexpr uint32_t kPhiloxM0 = 0xD2511F53u, kPhiloxM1 = 0xCD9E8D57u;
constexpr uint32_t kPhiloxW0 = 0x9E3779B9u, kPhiloxW1 = 0xBB67AE85u;
constexpr uint32_t kDomainInit = 0x494E4954u;    // "INIT"
constexpr uint32_t kDomainRefine = 0x52464E45u;  // "RFNE"

}  // namespace hp3d
