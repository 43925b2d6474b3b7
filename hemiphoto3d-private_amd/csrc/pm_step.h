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
    int step;          // the view's step index: 0 start, 1 + half-pass index
    int done_base;     // persistent kernel: first of the view's tile counters
};

// Persistent batch kernel: control words after the per-tile counters -- the
// next chunk of each of the 8 per-XCD work queues, and the abort flag.
enum PersistCtl : int { kCtlQueue0 = 0, kCtlAbort = 8, kCtlWords = 9 };

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

// One launch for a whole batch: items of every step in step order, their
// tile_begin counting units (tiles) across all steps; done holds one counter
// per (view, tile), followed by kCtlWords control words, all zero at launch
// (zero_ints).  chunk: consecutive units a workgroup takes at a time.
// wait_ticks: bound, in wall-clock ticks, on any one dependency wait; a wait
// that exceeds it sets done[n_counters + kCtlAbort] and every workgroup exits.
hipError_t launch_persistent(int store, int radius, int stride, int variant, const DevJob* jobs,
                             const StepItem* items, int n_items, int total_units, int* done,
                             int n_counters, int chunk, unsigned long long wait_ticks,
                             hipStream_t st);
hipError_t zero_ints(int* p, int n, hipStream_t st);

}  // namespace hp3d
