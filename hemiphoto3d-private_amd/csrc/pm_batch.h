// pm_batch.h -- helpers shared by the batched form and the serving loop.
#pragma once
#include "pm_runtime.h"
#include "pm_step.h"

namespace hp3d {

// Persistent batch kernel: units a workgroup takes at a time, and the bound
// on one dependency wait.
constexpr int kDefaultPersistChunk = 16;
constexpr int kPersistWaitSeconds = 10;

// Bytes of one view's state planes plus the staged outputs named by mask
// (1 depth, 2 normal, 4 cost, 8 confidence).
size_t job_state_bytes(int W, int H, int want_outputs_mask);
// Fill a device job for view d with its state (and staged outputs) laid out at base.
void bind_job(DevJob& j, const hp3d_view_desc* d, const Image* ref, const Image* const* srcs,
              char* base, int outputs_mask);
// Step `step` (0 = start) of a view with `iterations` iterations.
StepItem make_item(int job, int step, int iterations, uint64_t seed, int W, int H, int variant);
int tune_key(int radius, int stride, int store, double pixels, double mean_src);
int batch_tune_key(const hp3d_view_desc* descs, const std::vector<const Image*>& refs, int n);
int chosen_variant(const hp3d_pm_ctx* ctx, int radius, int stride, int key, int requested);

}  // namespace hp3d
