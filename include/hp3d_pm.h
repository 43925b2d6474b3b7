/*
 * hp3d_pm.h -- C ABI of the MI355X (gfx950) PatchMatch multi-view-stereo path.
 *
 * This is the boundary the reference's C++ API sits on.  The reference's
 * entry points and what replaces each (paths relative to the reference):
 *
 *   runPatchMatchMVS(images, cameras, sfm, config)
 *       src/dense_recon/patchmatch_mvs.h:39-44, .cu:1028-1388
 *       -> hp3d_pm_image_upload (textures, .cu:1042-1055)
 *        + hp3d_pm_compose      (precomputeSourceViews, .cu:847-884, in FP64)
 *        + hp3d_pm_run_view     (per-view body, .cu:1076-1258)
 *        + hp3d_pm_geom_filter  (filter driver, .cu:1267-1384)
 *   init_depth_normal_kernel launch     .cu:890-902   -> hp3d_pm_init_hypotheses
 *   compute_ncc_cost_kernel launch      .cu:904-934   -> hp3d_pm_cost_eval
 *   selectSourceViews(sfm, n)           patchmatch_mvs.h:32-35 -> hp3d_select_source_views
 *   SfM seeding in the driver           .cu:1111-1177  -> hp3d_project_seed_points
 *                                                         + hp3d_view_desc.seeds
 *   saveDepthMapPFM / loadDepthMapPFM   patchmatch_mvs.h:59,62 -> hp3d_pfm_save / hp3d_pfm_load
 *
 * Conventions
 *   - Every function returns an int status: HP3D_OK (0) or a negative
 *     hp3d_status.  Nothing here calls exit(); the reference's CUDA_CHECK
 *     exits the process (src/utils/cuda_utils.cuh:7-15).
 *   - A context is used by one host thread at a time; contexts are
 *     independent.  All device work of a context runs on its stream.
 *   - Matrices are row-major doubles.  Cameras map world to camera:
 *     P_cam = R P_world + t (include/types.h:55-74).
 *   - Outputs: depth H*W float (z-depth), normal H*W*3 float (camera frame),
 *     cost H*W float in [0, 2]; confidence = 1 - cost/2 (.cu:1246-1252).
 */
#ifndef HP3D_PM_H
#define HP3D_PM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HP3D_PM_MAX_SRC 8   /* MAX_SRC_VIEWS, patchmatch_mvs.cu:18 */
#define HP3D_PM_MAX_TOPK 8

typedef enum {
    HP3D_OK = 0,
    HP3D_ERR_INVALID_ARG = -1,
    HP3D_ERR_OOM_BUDGET = -2,   /* request exceeds the context's HBM budget */
    HP3D_ERR_DEVICE = -3,       /* a HIP call failed */
    HP3D_ERR_CANCELLED = -4,
    HP3D_ERR_NOT_FOUND = -5,    /* unknown handle or job */
    HP3D_ERR_UNSUPPORTED = -6,  /* e.g. window radius/stride not built */
    HP3D_ERR_IO = -7,
    HP3D_ERR_AGAIN = -8         /* job not finished yet */
} hp3d_status;

typedef enum { HP3D_FMT_U8 = 0, HP3D_FMT_F32 = 1 } hp3d_pixel_format;
typedef enum { HP3D_STORE_F32 = 0, HP3D_STORE_F16 = 1 } hp3d_storage;

#define HP3D_CTX_TIMING 1   /* record HIP events around every launch */

typedef struct hp3d_pm_ctx hp3d_pm_ctx;
typedef uint64_t hp3d_img;

typedef struct {
    int device;          /* HIP device ordinal */
    size_t hbm_budget;   /* bytes of device memory the context may hold; 0 = 90% of free */
    void* stream;        /* hipStream_t to run on; NULL = the context creates one */
    int flags;           /* HP3D_CTX_* */
} hp3d_pm_opts;

/* One reference view to solve (the body of the per-view loop,
 * patchmatch_mvs.cu:1076-1258). */
typedef struct {
    hp3d_img ref;
    hp3d_img src[HP3D_PM_MAX_SRC];
    int num_src;                         /* 1..8 */
    double K_ref_inv[9];                 /* K_ref^-1 */
    double A[HP3D_PM_MAX_SRC][9];        /* from hp3d_pm_compose */
    double b[HP3D_PM_MAX_SRC][3];
    float depth_min, depth_max;          /* MVSViewParams, patchmatch_mvs.h:23-28 */
    int radius;                          /* window half size 1..5 (.cu:1065-1070) */
    int stride;                          /* window stride 1 or 2 (1 = reference) */
    int iterations;                      /* PipelineConfig::mvs_iterations, 0..2^20 */
    int top_k;                           /* NCC_TOP_K = 3 in the reference */
    uint64_t seed;                       /* view seed; reference: 42 + 1e6*ref_id (.cu:1106) */
    const float* init_depth;             /* optional host hypotheses replacing the random */
    const float* init_normal;            /*   init (H*W and H*W*3); both or neither */
    const struct hp3d_seed_point* seeds; /* optional SfM seeds (hp3d_project_seed_points), */
    int num_seeds;                       /*   applied after the init, before the first cost */
    int seed_radius;                     /* half side of each seeded square; reference 50 */
    int first_iteration;                 /* schedule index of the first iteration run: 0, or
                                            k <= 2^20 to resume a run saved after k iterations from
                                            its depth and normals (init_depth/init_normal);
                                            the result is then identical to an uninterrupted
                                            run.  hp3d_pm_run_view only. */
} hp3d_view_desc;

typedef struct {
    float* depth;        /* H*W, required */
    float* normal;       /* H*W*3, nullable */
    float* cost;         /* H*W, nullable */
    float* confidence;   /* H*W, nullable: 1 - cost/2 */
    int on_device;       /* 1: the pointers are device memory of the context's device */
} hp3d_view_out;

typedef struct {
    double ms_total;       /* last run: first launch to last completion (device events) */
    double ms_init;
    double ms_cost;        /* initial-cost launches */
    double ms_propagate;   /* sum over half-pass launches */
    int n_propagate;       /* half-pass launches in the last run */
    double ms_finalize;
    double taps;           /* algorithmic taps of the last run (upper bound) */
    size_t hbm_in_use;     /* bytes currently held by the context */
    size_t hbm_budget;
} hp3d_pm_stats;

const char* hp3d_pm_version(void);

int hp3d_pm_create(const hp3d_pm_opts* opts, hp3d_pm_ctx** out);
int hp3d_pm_destroy(hp3d_pm_ctx* ctx);
/* Message of the last failed call on ctx ("" if none).  Owned by ctx. */
const char* hp3d_pm_last_error(const hp3d_pm_ctx* ctx);

/* Upload one gray image.  fmt U8: 8-bit DN; fmt F32: gray in [0,1] as in
 * ImageData::gray (include/types.h:17).  pitch in bytes (0: packed rows; a
 * multiple of 4 for F32); the pitch * (height - 1) + row bytes from data are
 * copied.  HP3D_ERR_UNSUPPORTED if a device plane of the image would exceed
 * 2 GiB - 64 KiB (about 16000 x 16000 px with FP32 storage).  Replaces
 * createGrayscaleTexture (patchmatch_mvs.cu:789-814). */
int hp3d_pm_image_upload(hp3d_pm_ctx* ctx, const void* data, int width, int height,
                         size_t pitch, int fmt, int store, hp3d_img* out);
int hp3d_pm_image_release(hp3d_pm_ctx* ctx, hp3d_img img);
int hp3d_pm_image_info(const hp3d_pm_ctx* ctx, hp3d_img img, int* width, int* height,
                       int* store, size_t* bytes);

/* FP64 composition of the plane-homography terms for one source:
 * R_rel = R_s R_ref^T, t_rel = t_s - R_rel t_ref, A = K_s R_rel K_ref^-1,
 * b = K_s t_rel (precomputeSourceViews, patchmatch_mvs.cu:847-884). */
int hp3d_pm_compose(const double K_ref[9], const double R_ref[9], const double t_ref[3],
                    const double K_src[9], const double R_src[9], const double t_src[3],
                    double A[9], double b[3]);
/* FP64 3x3 inverse (Eigen's K.inverse(), patchmatch_mvs.cu:834). */
int hp3d_pm_invert3(const double M[9], double Minv[9]);

/* Solve one view: random init (or desc->init_*), initial cost, then
 * desc->iterations red/black half-passes.  Blocks until out is written. */
int hp3d_pm_run_view(hp3d_pm_ctx* ctx, const hp3d_view_desc* desc, const hp3d_view_out* out);

/* Solve n views back to back, as the reference's per-view loop in
 * runPatchMatchMVS does (patchmatch_mvs.cu:1076-1258, one view after the
 * other, each result downloaded before the next view starts, :1243-1252).
 * Here the host outputs of view i are downloaded on a second stream while
 * view i + 1 computes, so the downloads overlap the solves; outputs are
 * identical to n hp3d_pm_run_view calls.  out[i] as for hp3d_pm_run_view
 * (host or device pointers per view); host outputs in memory from
 * hp3d_pm_host_alloc download at full link speed.  Blocks until every
 * output is written. */
int hp3d_pm_run_views(hp3d_pm_ctx* ctx, const hp3d_view_desc* desc, int n,
                      const hp3d_view_out* out);

/* The solve of hp3d_pm_run_view for callers that cannot allocate device
 * memory (code under hipGraph capture, serving loops that forbid hipMalloc).
 * No counterpart in the reference.  The view's hypothesis state lives in a
 * fixed device array of HP3D_PM_NOALLOC_SCRATCH_BYTES that the library holds
 * on every GPU it runs on until the process ends; the call allocates nothing.
 * out must be device buffers (on_device = 1); caller hypotheses and SfM seeds
 * are refused (they need staging memory), as is a view whose state (21 bytes
 * per pixel) exceeds the array: about 3.2 Mpixel.  Results equal
 * hp3d_pm_run_view's.  Calls that overlap on one GPU share the array and
 * race: one call at a time per GPU. */
#define HP3D_PM_NOALLOC_SCRATCH_BYTES (64u << 20)
int hp3d_pm_run_view_noalloc(hp3d_pm_ctx* ctx, const hp3d_view_desc* desc,
                             const hp3d_view_out* out);

/* Page-locked host memory owned by the context, for output buffers that
 * downloads write at full link speed (the reference returns cv::Mat in
 * pageable memory, patchmatch_mvs.cu:1243-1252).  Freed by hp3d_pm_host_free
 * or when the context is destroyed. */
int hp3d_pm_host_alloc(hp3d_pm_ctx* ctx, size_t bytes, void** ptr);
int hp3d_pm_host_free(hp3d_pm_ctx* ctx, void* ptr);

/* Random initial hypotheses only (the reference's mvs_iterations = 0 maps,
 * minus the cost).  depth H*W, normal H*W*3, host memory. */
int hp3d_pm_init_hypotheses(hp3d_pm_ctx* ctx, const hp3d_view_desc* desc, float* depth,
                            float* normal);

/* Cost of caller-given hypotheses (host depth H*W, normal H*W*3) -> host cost H*W. */
int hp3d_pm_cost_eval(hp3d_pm_ctx* ctx, const hp3d_view_desc* desc, const float* depth,
                      const float* normal, float* cost);

int hp3d_pm_get_stats(const hp3d_pm_ctx* ctx, hp3d_pm_stats* out);

/* ---------------------------------------------------------------------------
 * Geometric consistency filter over a set of depth maps (replaces
 * geometric_consistency_filter_kernel and its driver, patchmatch_mvs.cu:668-781
 * and :1267-1384).  Views are filtered in array order, in place: a pixel keeps
 * its depth if at least min_consistent sources see it within depth_tol
 * relative depth (and normal_tol_deg, 180 = off); later views see earlier
 * views already filtered, as in the reference's host loop.  Sources of a
 * different size are skipped; a view with fewer usable sources than
 * min_consistent is left as is.  The reference runs it with 0.12, 180, 2
 * (:1273-1275).  Host memory in and out; the maps stay on the device for
 * the whole set.  retained (nullable): per view, pixels left (-1: unfiltered).
 * ------------------------------------------------------------------------- */
typedef struct {
    int width, height;
    float* depth;           /* H*W, filtered in place (0 = invalid) */
    const float* normal;    /* H*W*3, camera frame */
    float* confidence;      /* H*W, zeroed with depth (nullable) */
    double K[9], R[9], t[3];
    int num_src;
    int src[HP3D_PM_MAX_SRC];   /* indices into the view array */
} hp3d_geom_view;

int hp3d_pm_geom_filter(hp3d_pm_ctx* ctx, hp3d_geom_view* views, int n, float depth_tol,
                        float normal_tol_deg, int min_consistent, int* retained);

/* ---------------------------------------------------------------------------
 * Depth-map fusion into a point cloud.  The reference declares fuseDepthMaps
 * (src/dense_recon/point_cloud.h:7-9) as the consumer of the MVS maps but
 * ships a stub that returns an empty cloud (point_cloud.cu:4-8); this is its
 * device form.  Every pixel with depth > 0 (and, where a confidence map is
 * given, confidence >= min_confidence) becomes a point, in view order then
 * row-major pixel order: position R^T (d K^-1 [x y 1] - t), normal R^T n,
 * colour = 8-bit value / 255 (gray replicated to three channels, or RGB).
 * Host memory in and out.  With positions == NULL only *count is reported;
 * otherwise the arrays hold capacity points (3 floats each); normals and
 * colors are nullable.
 * ------------------------------------------------------------------------- */
typedef struct {
    int width, height;
    const float* depth;         /* H*W, 0 = invalid */
    const float* normal;        /* H*W*3 camera frame, nullable (zero normals) */
    const float* confidence;    /* H*W, nullable */
    const uint8_t* color;       /* H*W*color_channels, nullable */
    int color_channels;         /* 1 (gray) or 3 (RGB) when color is given */
    double K[9], R[9], t[3];
} hp3d_fuse_view;

int hp3d_pm_fuse(hp3d_pm_ctx* ctx, const hp3d_fuse_view* views, int n, float min_confidence,
                 float* positions, float* normals, float* colors, size_t capacity, size_t* count);

/* ---------------------------------------------------------------------------
 * Batched form: many reference views of any size, source count and
 * iteration count solved together.  The reference solves views one by one
 * in a host loop (patchmatch_mvs.cu:1076-1258) with a device synchronisation
 * after every launch; here one step kernel advances every view by one pass
 * and the whole step sequence is captured in one hipGraph and launched once.
 * All views of a batch share radius, stride and image storage.  Results are
 * bit-identical to hp3d_pm_run_view on the same view.
 * ------------------------------------------------------------------------- */
typedef struct {
    int variant;     /* step-kernel variant; -1 = the autotuned choice, else variant 1
                        (32 x 8 tiles); windows other than 11x11 always run variant 0 */
    int use_graph;   /* 1 = capture the launches in one hipGraph (default) */
    int persistent;  /* 1 = one persistent kernel for the whole batch: a device work list
                        of (view, step, tile) units, each half-pass of a tile starting as
                        soon as that tile and its 4 neighbours finished the previous pass
                        (per-tile counters), instead of one launch per step */
    int chunk;       /* persistent: units a workgroup takes at a time (0 = 16) */
} hp3d_batch_opts;

int hp3d_pm_batch_run(hp3d_pm_ctx* ctx, const hp3d_view_desc* descs, int n,
                      const hp3d_view_out* outs, const hp3d_batch_opts* opts);

/* Autotuner: times every step-kernel variant (tile shape and workgroup size)
 * on the given views for one start step and two half-passes, on scratch
 * state, and records the fastest for their (radius, stride, storage, total
 * pixels in power-of-two buckets from 0.5 Mpixel, mean source count); a later
 * batch with the same key runs that variant.
 * times_ms (nullable) receives HP3D_PM_NUM_VARIANTS entries, -1 where a
 * variant does not apply. */
#define HP3D_PM_NUM_VARIANTS 6
int hp3d_pm_autotune(hp3d_pm_ctx* ctx, const hp3d_view_desc* descs, int n, int* best_variant,
                     float* times_ms);
/* Tile shape and waves-per-SIMD hint of a variant. */
int hp3d_pm_variant_info(int variant, int* tile_cols, int* tile_rows, int* waves_per_simd);
/* Persist / restore the autotuner's choices (text, one line per key). */
int hp3d_pm_tune_save(const hp3d_pm_ctx* ctx, const char* path);
int hp3d_pm_tune_load(hp3d_pm_ctx* ctx, const char* path);

/* ---------------------------------------------------------------------------
 * Serving form: a request loop with continuous batching.  Jobs (one
 * reference view each, with its own images) are submitted at any time; a
 * scheduler thread admits them into free slots when their device footprint
 * (images + state + outputs) fits the server's HBM budget, and launches one
 * step at a time over every admitted job, each at its own pass, with up to
 * two steps queued on the device so the scheduler prepares the next step
 * while one runs.  Jobs join and leave between steps; images, hypothesis
 * state and outputs stay on the device until the job completes, and a
 * completion thread downloads finished jobs' outputs off the scheduler.  A job's result is bit-identical to
 * hp3d_pm_run_view on the same view, whatever else runs beside it.
 * There is no counterpart in the reference, whose MVS runs only inside the
 * blocking runPatchMatchMVS call (patchmatch_mvs.cu:1028).
 * ------------------------------------------------------------------------- */
typedef struct hp3d_pm_server hp3d_pm_server;

typedef struct {
    int device;
    size_t hbm_budget;    /* device arena reserved for admitted jobs (bytes) */
    int max_jobs;         /* concurrent job slots */
    int radius, stride;   /* window shared by every job */
    int store;            /* hp3d_storage of every job's images */
    int variant;          /* step-kernel variant, -1 = variant 1 (32 x 8 tiles) */
} hp3d_server_opts;

typedef struct {
    int width, height, num_src;              /* width, height: of the job's images (at
                                                `level` for resident images) */
    const uint8_t* ref;                      /* host 8-bit images, copied at submit; ignored
                                                where a resident image is named below */
    const uint8_t* src[HP3D_PM_MAX_SRC];
    double K_ref_inv[9];
    double A[HP3D_PM_MAX_SRC][9];
    double b[HP3D_PM_MAX_SRC][3];
    float depth_min, depth_max;
    int iterations, top_k;                   /* iterations 0..2^20 */
    uint64_t seed;
    float* out_depth;                        /* host, H*W: written when the job completes */
    float* out_normal;                       /* host, H*W*3, nullable */
    float* out_cost;                         /* host, H*W, nullable */
    int fail_at_step;                        /* test hook: > 0 fails the job at that step */
    uint64_t ref_image;                      /* resident image (hp3d_server_image_add) used as
                                                the reference instead of `ref`; 0 = none */
    uint64_t src_image[HP3D_PM_MAX_SRC];     /* resident images used as sources; 0 = none */
    int level;                               /* pyramid level of the resident images (0 = full
                                                size); K_ref_inv, A, b are the caller's for
                                                that level's pixel grid */
} hp3d_job_desc;

typedef enum {
    HP3D_JOB_QUEUED = 0, HP3D_JOB_RUNNING = 1, HP3D_JOB_DONE = 2,
    HP3D_JOB_FAILED = 3, HP3D_JOB_CANCELLED = 4
} hp3d_job_state;

typedef struct {
    int state;             /* hp3d_job_state */
    int error;             /* hp3d_status of a failed job */
    double t_submit;       /* seconds on the server's steady clock */
    double t_admit;
    double t_done;
    int steps;             /* steps the job took part in */
} hp3d_job_status;

typedef struct {
    long long steps;            /* step launches */
    double busy_ms;             /* device time of all steps */
    double wall_s;              /* since the server started */
    long long submitted, done, failed, cancelled, rejected;
    double taps;                /* algorithmic taps of completed jobs */
    double mpix_iters;          /* W*H*iterations of completed jobs / 1e6 */
    size_t arena_bytes, arena_peak;
    int max_active;             /* most jobs in one step */
    long long image_uploads;    /* images uploaded: per-job host images and resident images */
    long long image_reuses;     /* job images served by a resident image (no upload) */
    long long resident_images;  /* resident images currently held */
    double wait_ms;             /* scheduler time spent waiting for a step to finish */
} hp3d_server_stats;

int hp3d_server_create(const hp3d_server_opts* opts, hp3d_pm_server** out);
/* Queue a job; HP3D_ERR_OOM_BUDGET if it could never fit the budget. */
int hp3d_server_submit(hp3d_pm_server* srv, const hp3d_job_desc* job, uint64_t* id);
int hp3d_server_poll(hp3d_pm_server* srv, uint64_t id, hp3d_job_status* st);
/* Block until the job leaves the queue and the device (timeout_s < 0: forever);
 * HP3D_ERR_AGAIN on timeout. */
int hp3d_server_wait(hp3d_pm_server* srv, uint64_t id, double timeout_s, hp3d_job_status* st);
int hp3d_server_cancel(hp3d_pm_server* srv, uint64_t id);
/* Drop the record of a finished (done, failed or cancelled) job: later polls
 * return HP3D_ERR_NOT_FOUND.  HP3D_ERR_AGAIN while the job is queued or
 * running -- its output buffers must stay valid until then.  A long-running
 * server keeps one small record per job until it is forgotten.  (No
 * counterpart in the reference, which has no serving loop.) */
int hp3d_server_forget(hp3d_pm_server* srv, uint64_t id);
int hp3d_server_get_stats(hp3d_pm_server* srv, hp3d_server_stats* st);
/* Resident images: one 8-bit image uploaded once, with `levels` (1..4)
 * pyramid levels -- level l is the 2x2 box average of level l - 1, of size
 * (W >> l) x (H >> l), in the server's storage -- kept in the server's HBM
 * budget until removed, for any number of jobs to use by handle (the
 * reference keeps every registered image resident for its whole run,
 * patchmatch_mvs.cu:1042-1055).  Blocks until the upload is done;
 * HP3D_ERR_OOM_BUDGET if the planes do not fit the budget now.  A removed
 * image's memory is freed once no queued or running job uses it. */
int hp3d_server_image_add(hp3d_pm_server* srv, const uint8_t* gray, int width, int height,
                          size_t pitch, int levels, uint64_t* image);
int hp3d_server_image_remove(hp3d_pm_server* srv, uint64_t image);
/* A resident image's level as DN values (gray * 255, H_l x W_l floats). */
int hp3d_server_image_read(hp3d_pm_server* srv, uint64_t image, int level, int* width,
                           int* height, float* out, size_t capacity);
const char* hp3d_server_last_error(hp3d_pm_server* srv);
int hp3d_server_destroy(hp3d_pm_server* srv);

/* Device buffers owned by the context and counted against its HBM budget,
 * for callers without a HIP runtime of their own (the reference keeps its
 * device state in DeviceBuffer<T>, src/utils/cuda_utils.cuh:26-57).
 * hp3d_pm_device_copy kind: 0 host->device, 1 device->host, 2 device->device;
 * it completes before returning. */
int hp3d_pm_device_alloc(hp3d_pm_ctx* ctx, size_t bytes, void** dptr);
int hp3d_pm_device_free(hp3d_pm_ctx* ctx, void* dptr);
int hp3d_pm_device_copy(hp3d_pm_ctx* ctx, void* dst, const void* src, size_t bytes, int kind);

/* ---------------------------------------------------------------------------
 * Host view set-up: the CPU side of runPatchMatchMVS (patchmatch_mvs.cu:
 * 1036-1039, 1111-1177) on plain arrays, for callers that hold an SfM result
 * (SfMResult, src/sfm/camera_pose.h:8-14).
 * ------------------------------------------------------------------------- */
typedef struct {
    int image_id;            /* must equal the camera's index in the array: the
                                reference indexes sfm.cameras by image id (.cu:104) */
    int is_registered;
    double K[9], R[9], t[3]; /* world -> camera, row-major (include/types.h:55-74) */
} hp3d_sfm_camera;

typedef struct {
    double position[3];      /* world point (SparsePoint, include/types.h:76-83) */
    const int* image_ids;    /* views observing it */
    int num_image_ids;
} hp3d_sfm_point;

typedef struct {             /* SourceViewInfo, patchmatch_mvs.h:16-20 */
    int image_id;
    float score;             /* shared points x sin(clamp(baseline, 5, 30 deg)) */
    float baseline_deg;
} hp3d_source_view;

typedef struct {             /* MVSViewParams, patchmatch_mvs.h:23-28 */
    int ref_image_id;
    int num_src;
    hp3d_source_view src[HP3D_PM_MAX_SRC];
    float depth_min, depth_max;
} hp3d_view_params;

/* selectSourceViews (patchmatch_mvs.h:32-35, .cu:74-157): one entry per
 * registered camera, in camera order, written to out (capacity num_cams);
 * *num_out receives the count.  Sources: every other registered camera whose
 * baseline angle at the SfM centroid is in [5, 90] deg and that shares at
 * least one point, by descending score (ties: lower image id first; the
 * reference's std::sort leaves tie order unspecified), at most num_sources
 * (1..8).  Depth range: [0.5 P5, 1.5 P95] of the visible points' z, or
 * [0.3 z_c, 3 z_c] at the centroid when fewer than 2 are visible. */
int hp3d_select_source_views(const hp3d_sfm_camera* cams, int num_cams, const hp3d_sfm_point* pts,
                             int num_pts, int num_sources, hp3d_view_params* out, int* num_out);

/* One SfM seed: the pixel a sparse point projects to, its z-depth and the
 * unit vector from the point to the camera centre in the camera frame. */
typedef struct hp3d_seed_point {
    int x, y;
    float depth;
    float normal[3];
} hp3d_seed_point;

/* Seeds of reference camera cam (.cu:1126-1155): every point observed by
 * cam->image_id, in front of it, with depth in [depth_min, depth_max], in
 * point order; points whose seeded square (half side seed_radius) misses the
 * width x height image are left out, as they seed nothing.  out holds up to
 * num_pts entries.  hp3d_view_desc.seeds applies them on the device with the
 * reference's rule: a pixel takes the last seed whose square covers it. */
int hp3d_project_seed_points(const hp3d_sfm_camera* cam, const hp3d_sfm_point* pts, int num_pts,
                             float depth_min, float depth_max, int width, int height,
                             int seed_radius, hp3d_seed_point* out, int* num_out);

/* Portable FloatMap depth I/O, byte-compatible with saveDepthMapPFM /
 * loadDepthMapPFM (patchmatch_mvs.cu:990-1022): "Pf\n<W> <H>\n-1.0\n", rows
 * bottom to top.  hp3d_pfm_load with data == NULL only reports the size. */
int hp3d_pfm_save(const char* path, int width, int height, const float* depth);
int hp3d_pfm_load(const char* path, int* width, int* height, float* data, size_t capacity);
/* Any PFM: channels 1 ("Pf", e.g. depth or confidence) or 3 ("PF", e.g. the
 * H*W*3 normal map), same row order.  With these a run's depth and normals
 * can be saved and resumed (hp3d_view_desc.first_iteration); the reference
 * saves depth only (.cu:1254-1257).  Load with data == NULL reports the size
 * and channel count only; capacity is in floats. */
int hp3d_pfm_save_ex(const char* path, int width, int height, int channels, const float* data);
int hp3d_pfm_load_ex(const char* path, int* width, int* height, int* channels, float* data,
                     size_t capacity);

#ifdef __cplusplus
}
#endif
#endif /* HP3D_PM_H */
