// pm_step.hip -- multi-view step kernel: one launch advances every active
// view of a batch or of the serving loop by one pass.
//
// A step is a list of items, one per active view (job): START (random init
// and initial cost of both colours, patchmatch_mvs.cu:1107-1196) or PROPAGATE
// (one red/black half-pass, :1208-1226), optionally followed by writing the
// view's outputs (:1243-1252).  Each view advances at its own pass index, so
// views of any size, source count and iteration count share a launch and new
// views can join between steps.
//
// Work decomposition: every view is cut into TX x TY compacted tiles; all
// items' tiles form one logical range, cut into chunks of kStepChunk
// consecutive tiles.  Workgroups are dispatched to the 8 XCDs round-robin
// (block b -> XCD b % 8); XCD x takes chunks x, x + 8, x + 16, ... in order.
// Tiles of a chunk are neighbours in one view and read the same source rows
// through that XCD's L2, while every XCD sees an even sample of the whole
// step -- views differ in source count, START items cost a third of a
// half-pass, and early passes cost more than late ones, so contiguous
// eighths of the range would leave XCDs idle at the end of the step.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "pm_device.h"
#include "pm_pixel.h"
#include "pm_step.h"

namespace hp3d {

constexpr int kStepChunk = 16;   // tiles per chunk of the XCD interleave
constexpr int kPersistMaxGrid = 4096;

template <int NT, int ST, int STORE, int TX, int TY, int WPE>
__global__ __launch_bounds__(TX * TY) __attribute__((amdgpu_waves_per_eu(WPE, 8)))
void pm_step_kernel(const DevJob* __restrict__ jobs, const StepItem* __restrict__ items,
                    int n_items, int total_tiles)
{
    constexpr int R = (NT - 1) * ST / 2;
    using G = TileGeom<R, STORE, TX, TY>;
    __shared__ float tile[G::LH * G::LP];
    __shared__ WaveXchg xw[TX * TY / 64];

    const int b = blockIdx.x;                       // gridDim.x: a multiple of 8 * kStepChunk
    const int sidx = b >> 3;                        // position in this XCD's sequence
    const int t = ((sidx / kStepChunk) * 8 + (b & 7)) * kStepChunk + sidx % kStepChunk;
    if (t >= total_tiles) return;
    int lo = 0, hi = n_items - 1;                   // last item with tile_begin <= t
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (items[mid].tile_begin <= t) lo = mid;
        else hi = mid - 1;
    }
    const StepItem& it = items[lo];
    const DevJob& job = jobs[it.job];
    const DevView& v = job.v;
    const int lt = t - it.tile_begin;
    const int by = lt / it.tiles_x, bx = lt - by * it.tiles_x;
    const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;

    stage_ref_tile<R, STORE, TX, TY>(v, tile, bx, by);
    const int i = bx * TX + tx, y = by * TY + ty;
    if (it.mode == kStepStart) {
        if (y < v.H) {
            for (int c = 0; c < 2; ++c) {
                const int x = 2 * i + ((y + c) & 1);
                if (x < v.W) v.hyp[c][y * v.Wc + i] = init_hyp(v, x, y, job.seed);
            }
        }
        cost_pixel<NT, ST, STORE, TX, TY>(v, tile, bx, by, tx, ty, 0);
        cost_pixel<NT, ST, STORE, TX, TY>(v, tile, bx, by, tx, ty, 1);
    } else {
        halfpass_pixel<NT, ST, STORE, TX, TY>(v, tile, bx, by, tx, ty, it.iteration, it.parity,
                                              it.pass_seed, xw[threadIdx.x >> 6]);
    }
    // outputs: this lane's pixel of each colour; the other colour was
    // completed by the previous step
    if (it.emit && y < v.H) {
        for (int c = 0; c < 2; ++c) {
            const int x = 2 * i + ((y + c) & 1);
            if (x < v.W) emit_pixel(v, x, y, job.out_depth, job.out_normal, job.out_cost,
                                    job.out_conf);
        }
    }
}

// ---------------------------------------------------------------------------
// Persistent batch kernel: one launch runs every step of every view.
//
// The units (view, step, tile) are listed step by step, so every unit a
// unit depends on comes earlier in the list.  The list is cut into chunks of
// `chunk` consecutive units, dealt to 8 work queues as the step kernel deals
// them to the XCDs (chunk c to queue c % 8, in order), and workgroup b serves
// queue b % 8 -- the XCD the dispatcher places it on -- so neighbouring
// tiles, which gather the same source rows, share that XCD's L2 (one shared
// queue spread them over every XCD and ran the batch 1.85x slower).  A
// workgroup takes its queue's next chunk and runs its units in order; when
// its queue is empty it takes chunks from the other queues.
// A half-pass of tile T at step s reads the other colour of T and of its
// left, right, upper and lower neighbour tiles (the L, R, U, D candidates of
// its edge pixels) as step s - 1 left them, and overwrites its own colour,
// which step s - 1 of those same tiles read: so it waits until the five
// tiles' counters (steps completed) reach s, and afterwards sets T's counter
// to s + 1.  A start unit (random init and initial cost of both colours)
// depends on nothing.
//
// Progress: take the first unit of the list that is not finished.  Every
// unit before it is finished.  Its queue hands out chunks in list order, so
// if its chunk was not taken yet, every chunk its queue handed out so far is
// finished and the queue's own workgroups (at least one of them is running:
// the grid has at least 8 workgroups and workgroups start in index order)
// take it next; once taken, its holder has reached it (earlier units of the
// chunk are finished), and its dependencies come earlier in the list, so its
// wait ends.  No co-residency of the whole grid is needed.  Each wait is still
// bounded in time; a wait that exceeds the bound sets the abort word, every
// workgroup then leaves, and the host reports the failure.
//
// Memory: the state planes are read and written with device-coherent (sc1)
// accesses (StateIO<true>), which bypass the per-CU L1 and per-XCD L2 caches
// that are not coherent across XCDs; before a tile's counter is set, every
// wave waits for its state stores to complete (vmcnt 0), and the counter
// itself is a device-scope store that waiting workgroups poll with
// device-scope loads.  A workgroup reads the planes only after its poll has
// seen the counters, so it reads what the producer stored.  The image planes
// are read-only during the launch and keep cached loads.
// ---------------------------------------------------------------------------
template <int NT, int ST, int STORE, int TX, int TY, int WPE>
__global__ __launch_bounds__(TX * TY) __attribute__((amdgpu_waves_per_eu(WPE, 8)))
void pm_persist_kernel(const DevJob* __restrict__ jobs, const StepItem* __restrict__ items,
                       int n_items, int total_units, int* done, int n_counters, int chunk,
                       unsigned long long wait_ticks)
{
    constexpr int R = (NT - 1) * ST / 2;
    using G = TileGeom<R, STORE, TX, TY>;
    __shared__ float tile[G::LH * G::LP];
    __shared__ WaveXchg xw[TX * TY / 64];
    __shared__ int s_word;
    int* const ctl = done + n_counters;
    const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;

    const int home = blockIdx.x & 7;
    const int chunks = (total_units + chunk - 1) / chunk;
    for (;;) {
        if (threadIdx.x == 0) {
            // the home queue first, then the others; a queue whose counter is
            // past its last chunk is empty
            int u = total_units;
            if (!__hip_atomic_load(&ctl[kCtlAbort], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                for (int k = 0; k < 8; ++k) {
                    const int q = (home + k) & 7;
                    if (q >= chunks) continue;
                    const int per_q = (chunks - q + 7) / 8;   // chunks q, q + 8, ... below chunks
                    if (__hip_atomic_load(&ctl[kCtlQueue0 + q], __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT) >= per_q)
                        continue;
                    const int pos = __hip_atomic_fetch_add(&ctl[kCtlQueue0 + q], 1,
                                                           __ATOMIC_RELAXED,
                                                           __HIP_MEMORY_SCOPE_AGENT);
                    if (pos < per_q) {
                        u = (8 * pos + q) * chunk;
                        break;
                    }
                }
            }
            s_word = u;
        }
        __syncthreads();
        const int u0 = __builtin_amdgcn_readfirstlane(s_word);
        __syncthreads();
        if (u0 >= total_units) return;
        const int u1 = min(u0 + chunk, total_units);
        for (int u = u0; u < u1; ++u) {
            int lo = 0, hi = n_items - 1;     // last item with tile_begin <= u
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (items[mid].tile_begin <= u) lo = mid;
                else hi = mid - 1;
            }
            const StepItem& it = items[lo];
            const DevJob& job = jobs[it.job];
            const DevView& v = job.v;
            const int lt = u - it.tile_begin;
            const int by = lt / it.tiles_x, bx = lt - by * it.tiles_x;
            // the reference tile depends only on the images: staged before the wait
            stage_ref_tile<R, STORE, TX, TY>(v, tile, bx, by);
            if (it.mode != kStepStart) {
                if (threadIdx.x == 0) {
                    const int* cnt = done + it.done_base;
                    const int ty_n = it.tiles / it.tiles_x;
                    int dep[5] = {lt, -1, -1, -1, -1};
                    if (bx > 0) dep[1] = lt - 1;
                    if (bx + 1 < it.tiles_x) dep[2] = lt + 1;
                    if (by > 0) dep[3] = lt - it.tiles_x;
                    if (by + 1 < ty_n) dep[4] = lt + it.tiles_x;
                    int abort = 0;
                    const unsigned long long t0 = wall_clock64();
                    for (int k = 0; k < 5 && !abort; ++k) {
                        if (dep[k] < 0) continue;
                        while (__hip_atomic_load(&cnt[dep[k]], __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT) < it.step) {
                            __builtin_amdgcn_s_sleep(2);
                            if (__hip_atomic_load(&ctl[kCtlAbort], __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_AGENT)) {
                                abort = 1;
                                break;
                            }
                            if (wall_clock64() - t0 > wait_ticks) {
                                __hip_atomic_store(&ctl[kCtlAbort], 1, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT);
                                abort = 1;
                                break;
                            }
                        }
                    }
                    s_word = abort;
                }
                __syncthreads();
                const int abort = __builtin_amdgcn_readfirstlane(s_word);
                __syncthreads();
                if (abort) return;
                if (kPersistPlainLoads) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            }
            const int i = bx * TX + tx, y = by * TY + ty;
            if (it.mode == kStepStart) {
                if (y < v.H) {
                    for (int c = 0; c < 2; ++c) {
                        const int x = 2 * i + ((y + c) & 1);
                        if (x < v.W) StateIO<true>::st4(v.hyp[c], y * v.Wc + i, init_hyp(v, x, y, job.seed));
                    }
                }
                cost_pixel<NT, ST, STORE, TX, TY, true>(v, tile, bx, by, tx, ty, 0);
                cost_pixel<NT, ST, STORE, TX, TY, true>(v, tile, bx, by, tx, ty, 1);
            } else {
                halfpass_pixel<NT, ST, STORE, TX, TY, true>(v, tile, bx, by, tx, ty, it.iteration,
                                                            it.parity, it.pass_seed,
                                                            xw[threadIdx.x >> 6]);
            }
            if (it.emit && y < v.H) {
                for (int c = 0; c < 2; ++c) {
                    const int x = 2 * i + ((y + c) & 1);
                    if (x < v.W)
                        emit_pixel<true>(v, x, y, job.out_depth, job.out_normal, job.out_cost,
                                         job.out_conf);
                }
            }
            // every wave's state stores complete before the tile's counter moves
            if (kPersistPlainStores) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            else __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            if (threadIdx.x == 0)
                __hip_atomic_store(&done[it.done_base + lt], it.step + 1, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

__global__ void zero_ints_kernel(int* p, int n)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) p[k] = 0;
}

hipError_t zero_ints(int* p, int n, hipStream_t st)
{
    const int blocks = std::min((n + 255) / 256, 1024);
    hipLaunchKernelGGL(zero_ints_kernel, dim3(std::max(blocks, 1)), dim3(256), 0, st, p, n);
    return hipGetLastError();
}

template <int NT, int ST, int STORE, int TX, int TY, int WPE>
static hipError_t launch_persist_t(const DevJob* jobs, const StepItem* items, int n_items,
                                   int total_units, int* done, int n_counters, int chunk,
                                   unsigned long long wait_ticks, hipStream_t st)
{
    // More workgroups than the GPU holds at once (256 CUs x at most 8 of
    // these per CU), a multiple of 8 so every queue has its workgroups: the
    // dispatcher keeps every CU full, and workgroups started after the work
    // is gone exit at once.  No co-residency is assumed (see the progress
    // argument above).
    const int chunks = (total_units + chunk - 1) / chunk;
    const int grid = std::max(8, std::min((chunks + 7) / 8 * 8, kPersistMaxGrid));
    hipLaunchKernelGGL((pm_persist_kernel<NT, ST, STORE, TX, TY, WPE>), dim3(grid), dim3(TX * TY),
                       0, st, jobs, items, n_items, total_units, done, n_counters, chunk,
                       wait_ticks);
    return hipGetLastError();
}

template <int NT, int ST, int STORE>
static hipError_t launch_persist_tuned(int variant, const DevJob* jobs, const StepItem* items,
                                       int n_items, int total_units, int* done, int n_counters,
                                       int chunk, unsigned long long wait_ticks, hipStream_t st)
{
    constexpr int kW = 1;
#define HP3D_PV(V_, TX_, TY_)                                                                     \
    case V_:                                                                                    \
        return launch_persist_t<NT, ST, STORE, TX_, TY_, kW>(jobs, items, n_items, total_units, \
                                                             done, n_counters, chunk, wait_ticks, st);
    switch (variant) {
        HP3D_PV(0, 16, 16) HP3D_PV(1, 32, 8) HP3D_PV(2, 8, 32) HP3D_PV(3, 64, 4)
        HP3D_PV(4, 32, 4) HP3D_PV(5, 16, 8)
        default: return hipErrorInvalidValue;
    }
#undef HP3D_PV
}

hipError_t launch_persistent(int store, int radius, int stride, int variant, const DevJob* jobs,
                             const StepItem* items, int n_items, int total_units, int* done,
                             int n_counters, int chunk, unsigned long long wait_ticks,
                             hipStream_t st)
{
    if (variant < 0 || variant >= kNumStepVariants || chunk < 1) return hipErrorInvalidValue;
    if (!step_variant_tunable(radius, stride) && variant != 0) return hipErrorInvalidValue;
#define HP3D_PT(R_, S_, ST_)                                                                   \
    if (radius == R_ && stride == S_ && store == ST_)                                         \
        return launch_persist_tuned<(2 * R_) / S_ + 1, S_, ST_>(variant, jobs, items, n_items,  \
                                                               total_units, done, n_counters,  \
                                                               chunk, wait_ticks, st);
#define HP3D_PP(R_, S_, ST_)                                                                   \
    if (radius == R_ && stride == S_ && store == ST_)                                         \
        return launch_persist_t<(2 * R_) / S_ + 1, S_, ST_, 16, 16, 1>(                        \
            jobs, items, n_items, total_units, done, n_counters, chunk, wait_ticks, st);
    HP3D_PT(5, 1, kStoreF32) HP3D_PT(5, 2, kStoreF32)
    HP3D_PT(5, 1, kStoreF16) HP3D_PT(5, 2, kStoreF16)
    HP3D_PP(1, 1, kStoreF32) HP3D_PP(2, 1, kStoreF32) HP3D_PP(3, 1, kStoreF32)
    HP3D_PP(4, 1, kStoreF32) HP3D_PP(1, 2, kStoreF32) HP3D_PP(2, 2, kStoreF32)
    HP3D_PP(3, 2, kStoreF32) HP3D_PP(4, 2, kStoreF32)
    HP3D_PP(1, 1, kStoreF16) HP3D_PP(2, 1, kStoreF16) HP3D_PP(3, 1, kStoreF16)
    HP3D_PP(4, 1, kStoreF16) HP3D_PP(1, 2, kStoreF16) HP3D_PP(2, 2, kStoreF16)
    HP3D_PP(3, 2, kStoreF16) HP3D_PP(4, 2, kStoreF16)
#undef HP3D_PT
#undef HP3D_PP
    return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------
// Kernel variants.  Every window has the default variant; the windows of the
// batched and serving workloads (radius 5, stride 1 or 2) also have the tile
// shapes and occupancies the autotuner chooses from.
// ---------------------------------------------------------------------------
template <int NT, int ST, int STORE, int TX, int TY, int WPE>
static hipError_t launch_t(const DevJob* jobs, const StepItem* items, int n_items, int total_tiles,
                           hipStream_t st)
{
    constexpr int kGroup = 8 * kStepChunk;
    const int grid = ((total_tiles + kGroup - 1) / kGroup) * kGroup;
    hipLaunchKernelGGL((pm_step_kernel<NT, ST, STORE, TX, TY, WPE>), dim3(grid), dim3(TX * TY), 0,
                       st, jobs, items, n_items, total_tiles);
    return hipGetLastError();
}

// 256-lane tiles of four shapes and two 128-lane tiles (finer granularity
// for small views).  Variants that force more waves per SIMD than the
// register budget allows spill to scratch and are not offered.
const StepVariant kStepVariants[] = {
    {16, 16, 1}, {32, 8, 1}, {8, 32, 1}, {64, 4, 1}, {32, 4, 1}, {16, 8, 1},
};
const int kNumStepVariants = (int)(sizeof(kStepVariants) / sizeof(kStepVariants[0]));

template <int NT, int ST, int STORE>
static hipError_t launch_tuned(int variant, const DevJob* jobs, const StepItem* items, int n_items,
                               int total_tiles, hipStream_t st)
{
    switch (variant) {
        case 0: return launch_t<NT, ST, STORE, 16, 16, 1>(jobs, items, n_items, total_tiles, st);
        case 1: return launch_t<NT, ST, STORE, 32, 8, 1>(jobs, items, n_items, total_tiles, st);
        case 2: return launch_t<NT, ST, STORE, 8, 32, 1>(jobs, items, n_items, total_tiles, st);
        case 3: return launch_t<NT, ST, STORE, 64, 4, 1>(jobs, items, n_items, total_tiles, st);
        case 4: return launch_t<NT, ST, STORE, 32, 4, 1>(jobs, items, n_items, total_tiles, st);
        case 5: return launch_t<NT, ST, STORE, 16, 8, 1>(jobs, items, n_items, total_tiles, st);
        default: return hipErrorInvalidValue;
    }
}

bool step_variant_tunable(int radius, int stride) { return radius == 5; }

int step_tile_tx(int variant) { return kStepVariants[variant].tx; }
int step_tile_ty(int variant) { return kStepVariants[variant].ty; }

hipError_t launch_step(int store, int radius, int stride, int variant, const DevJob* jobs,
                       const StepItem* items, int n_items, int total_tiles, hipStream_t st)
{
    if (variant < 0 || variant >= kNumStepVariants) return hipErrorInvalidValue;
    if (!step_variant_tunable(radius, stride) && variant != 0) return hipErrorInvalidValue;
#define HP3D_TUNED(R_, S_, ST_)                                                                  \
    if (radius == R_ && stride == S_ && store == ST_)                                           \
        return launch_tuned<(2 * R_) / S_ + 1, S_, ST_>(variant, jobs, items, n_items, total_tiles, \
                                                        st);
#define HP3D_PLAIN(R_, S_, ST_)                                                                 \
    if (radius == R_ && stride == S_ && store == ST_)                                          \
        return launch_t<(2 * R_) / S_ + 1, S_, ST_, 16, 16, 1>(jobs, items, n_items, total_tiles, \
                                                                st);
    HP3D_TUNED(5, 1, kStoreF32) HP3D_TUNED(5, 2, kStoreF32)
    HP3D_TUNED(5, 1, kStoreF16) HP3D_TUNED(5, 2, kStoreF16)
    HP3D_PLAIN(1, 1, kStoreF32) HP3D_PLAIN(2, 1, kStoreF32) HP3D_PLAIN(3, 1, kStoreF32)
    HP3D_PLAIN(4, 1, kStoreF32) HP3D_PLAIN(1, 2, kStoreF32) HP3D_PLAIN(2, 2, kStoreF32)
    HP3D_PLAIN(3, 2, kStoreF32) HP3D_PLAIN(4, 2, kStoreF32)
    HP3D_PLAIN(1, 1, kStoreF16) HP3D_PLAIN(2, 1, kStoreF16) HP3D_PLAIN(3, 1, kStoreF16)
    HP3D_PLAIN(4, 1, kStoreF16) HP3D_PLAIN(1, 2, kStoreF16) HP3D_PLAIN(2, 2, kStoreF16)
    HP3D_PLAIN(3, 2, kStoreF16) HP3D_PLAIN(4, 2, kStoreF16)
#undef HP3D_TUNED
#undef HP3D_PLAIN
    return hipErrorInvalidValue;
}

}  // namespace hp3d
