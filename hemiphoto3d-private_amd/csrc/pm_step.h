// pm_step.h -- types and launcher of the multi-view step kernel (pm_step.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pm_device.h"

namespace hp3d {

// One view resident on the device: its images, cameras and state planes, and
// where its outputs go.
struct DevJob {
    DevView v;
    uint64_t seed;          // view seed of the random init
    float* out_depth;       // H*W, device
    float* out_normal;      // H*W*3, device (nullable)
    float* out_cost;        // H*W, device (nullable)
    float* out_conf;        // H*W, device (nullable)
};

enum StepMode : int { kStepStart = 0, kStepPropagate = 1 };

// One view's share of one step.
struct StepItem {
    int job;           // index into the job table
    int mode;          // StepMode
    int iteration;     // PatchMatch iteration of a PROPAGATE item
    int parity;        // 0 red, 1 black
    int emit;          // write the view's outputs after this item
    int tile_begin;    // first logical tile of this item in the step
    int tiles_x;       // tiles per row of the view
    int tiles;         // tiles of the view
    uint64_t pass_seed;
};

struct StepVariant {
    int tx, ty;        // compacted columns x rows per workgroup
    int wpe;           // minimum waves per SIMD requested from the compiler
};
extern const StepVariant kStepVariants[];
extern const int kNumStepVariants;

// Only radius-5 windows (11x11 at stride 1 or 2) have every variant; other
// windows run variant 0 (16 x 16 tiles).
bool step_variant_tunable(int radius, int stride);
// Variant a radius-5 batch or server runs when none is requested or tuned:
// 32 x 8 tiles, within 1 % of the fastest shape on mixed-resolution batches
// at both strides, where 16 x 16 is 8-10 % slower.
constexpr int kDefaultStepVariant = 1;
int step_tile_tx(int variant);
int step_tile_ty(int variant);

hipError_t launch_step(int store, int radius, int stride, int variant, const DevJob* jobs,
                       const StepItem* items, int n_items, int total_tiles, hipStream_t st);

}  // namespace hp3d
