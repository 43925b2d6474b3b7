// pm_launch.h -- host-callable launchers of the kernels in pm_kernels.hip.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/hp3d_pm.h"
#include "pm_device.h"

namespace hp3d {

enum PassMode : int { kPassCost = 0, kPassPropagate = 1 };

bool window_supported(int radius, int stride);
hipError_t launch_init(const DevView& v, uint64_t seed, hipStream_t st);
hipError_t launch_finalize(const DevView& v, float* depth, float* normal, float* cost, float* conf,
                           hipStream_t st);
hipError_t launch_import(const DevView& v, const float* depth, const float* normal,
                         hipStream_t st);
// coef: the row-difference gather plane; both planes use `pitch` elements per row.
hipError_t launch_img_prep(const uint8_t* in_u8, const float* in_f32, size_t in_pitch, int W,
                           int H, int store, void* plain, void* coef, int pitch, hipStream_t st);
// Pyramid level (W x H) from the previous level's plain plane (2W x 2H or
// larger, prev_pitch elements per row): 2 x 2 box means, both planes.
hipError_t launch_img_down(const void* prev, int prev_pitch, int W, int H, int store, void* plain,
                           void* coef, int pitch, hipStream_t st);
// SfM seeds (device array of n) painted over the hypothesis planes; owner is
// W*H ints of scratch.
hipError_t launch_seed(const DevView& v, const hp3d_seed_point* seeds, int n, int radius,
                       int* owner, hipStream_t st);
// mode kPassCost: cost of the current hypotheses of colour `parity`;
// mode kPassPropagate: one half-pass of colour `parity` at `iteration`.
hipError_t launch_pass(const DevView& v, int store, int radius, int stride, int mode,
                       int iteration, int parity, uint64_t pass_seed, hipStream_t st);

}  // namespace hp3d
