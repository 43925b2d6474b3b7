/*
 * pm_oracle.c -- CPU restatement of Hemiphoto3D's PatchMatch-MVS hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline); see
 * pm_oracle.h for what may load it and for the list of deviations.
 *
 * Build: oracle/Makefile, with -ffp-contract=off so that single-precision
 * expressions round exactly in the order they are written.
 */
#include "pm_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon et al., "Parallel Random Numbers: As Easy as 1, 2,  */
/* 3", SC'11, section 3.3): multipliers M0, M1, Weyl key increments W0, W1.  */
/* ------------------------------------------------------------------------ */
#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; ++round) {
        if (round > 0) { k0 += PHILOX_W0; k1 += PHILOX_W1; }
        uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
        uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

float orc_u01(uint32_t x)
{
    return (float)((x >> 8) + 1u) * (1.0f / 16777216.0f);
}

static void philox_block(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2,
                         uint32_t c3, uint32_t out[4])
{
    uint32_t ctr[4] = {c0, c1, c2, c3};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    orc_philox4x32_10(ctr, key, out);
}

/* ------------------------------------------------------------------------ */
/* FP64 host composition (patchmatch_mvs.cu:847-884, math_utils.h:88-164)   */
/* ------------------------------------------------------------------------ */
static void mat3_mul(const double A[9], const double B[9], double C[9])
{
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += A[r * 3 + k] * B[k * 3 + c];
            C[r * 3 + c] = s;
        }
}

int orc_inv3(const double m[9], double inv[9])
{
    double det = m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
                 m[2] * (m[3] * m[7] - m[4] * m[6]);
    if (det == 0.0 || !isfinite(det)) return -1;
    double id = 1.0 / det;
    inv[0] = (m[4] * m[8] - m[5] * m[7]) * id;
    inv[1] = (m[2] * m[7] - m[1] * m[8]) * id;
    inv[2] = (m[1] * m[5] - m[2] * m[4]) * id;
    inv[3] = (m[5] * m[6] - m[3] * m[8]) * id;
    inv[4] = (m[0] * m[8] - m[2] * m[6]) * id;
    inv[5] = (m[2] * m[3] - m[0] * m[5]) * id;
    inv[6] = (m[3] * m[7] - m[4] * m[6]) * id;
    inv[7] = (m[1] * m[6] - m[0] * m[7]) * id;
    inv[8] = (m[0] * m[4] - m[1] * m[3]) * id;
    return 0;
}

int orc_compose(const double Kr[9], const double Rr[9], const double tr[3],
                const double Ks[9], const double Rs[9], const double ts[3],
                double A[9], double b[3])
{
    double Kri[9], RrT[9], Rrel[9], tmp[9];
    if (orc_inv3(Kr, Kri) != 0) return -1;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) RrT[r * 3 + c] = Rr[c * 3 + r];
    mat3_mul(Rs, RrT, Rrel);                       /* R_rel = R_s R_ref^T */
    double trel[3];
    for (int r = 0; r < 3; ++r)                    /* t_rel = t_s - R_rel t_ref */
        trel[r] = ts[r] - (Rrel[r * 3] * tr[0] + Rrel[r * 3 + 1] * tr[1] + Rrel[r * 3 + 2] * tr[2]);
    mat3_mul(Rrel, Kri, tmp);                      /* A = K_s (R_rel K_ref^-1) */
    mat3_mul(Ks, tmp, A);
    for (int r = 0; r < 3; ++r)                    /* b = K_s t_rel */
        b[r] = Ks[r * 3] * trel[0] + Ks[r * 3 + 1] * trel[1] + Ks[r * 3 + 2] * trel[2];
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Window geometry: taps at -r + stride*j, j < n, n = 2r/stride + 1.         */
/* stride 1 reproduces the reference loops (patchmatch_mvs.cu:273-274).      */
/* ------------------------------------------------------------------------ */
static int win_taps(const orc_view* v) { return (2 * v->radius) / v->stride + 1; }

static inline int clampi(int a, int lo, int hi) { return a < lo ? lo : (a > hi ? hi : a); }

static int near_rel(double a, double thr, double rel)
{
    return fabs(a - thr) <= rel * thr;
}

/* ------------------------------------------------------------------------ */
/* FP64 cost (patchmatch_mvs.cu:222-343 evaluated in double)                 */
/* ------------------------------------------------------------------------ */
#define EDGE_EPS 2e-3

double orc_cost_fp64(const orc_view* v, int x, int y, float d, float nxf, float nyf,
                     float nzf, uint32_t* flags_out, double* per_src)
{
    const double* ki = v->Kinv;
    const double fx = x, fy = y, dd = d;
    const double nx = nxf, ny = nyf, nz = nzf;
    uint32_t flags = 0;

    const double Px = dd * (ki[0] * fx + ki[1] * fy + ki[2]);
    const double Py = dd * (ki[3] * fx + ki[4] * fy + ki[5]);
    const double Pz = dd * (ki[6] * fx + ki[7] * fy + ki[8]);
    const double nP = nx * Px + ny * Py + nz * Pz;
    if (near_rel(fabs(nP), 1e-8, 1e-3)) flags |= ORC_FLAG_NP;
    if (fabs(nP) < 1e-8) {
        if (flags_out) *flags_out = flags;
        if (per_src) for (int s = 0; s < v->S; ++s) per_src[s] = 2.0;
        return 2.0;
    }
    const double nK0 = ki[0] * nx + ki[3] * ny + ki[6] * nz;
    const double nK1 = ki[1] * nx + ki[4] * ny + ki[7] * nz;
    const double nK2 = ki[2] * nx + ki[5] * ny + ki[8] * nz;

    const int n = win_taps(v);
    const int min_count = (n * n) / 2;
    const int W = v->W, H = v->H;
    const double umax = W - 1.0, vmax = H - 1.0;

    double top[ORC_MAX_SRC];
    const int K = v->top_k;
    for (int i = 0; i < K; ++i) top[i] = 2.0;

    for (int s = 0; s < v->S; ++s) {
        const double* A = v->A[s];
        const double* b = v->b[s];
        const uint8_t* img = v->src[s];
        const float* imgf = v->src_gray[s];
        double sr = 0, ss = 0, srr = 0, sss = 0, srs = 0;
        int count = 0;
        for (int j = 0; j < n; ++j) {
            const int dy = -v->radius + v->stride * j;
            for (int i = 0; i < n; ++i) {
                const int dx = -v->radius + v->stride * i;
                const size_t ri = (size_t)clampi(y + dy, 0, H - 1) * W + clampi(x + dx, 0, W - 1);
                const double rv = v->ref_gray ? (double)v->ref_gray[ri] : v->ref[ri] / 255.0;
                const double px = fx + dx, py = fy + dy;
                const double qx = A[0] * px + A[1] * py + A[2];
                const double qy = A[3] * px + A[4] * py + A[5];
                const double qz = A[6] * px + A[7] * py + A[8];
                const double factor = (nK0 * px + nK1 * py + nK2) / nP;
                const double wz = qz + b[2] * factor;
                if (near_rel(fabs(wz), 1e-8, 1e-3)) flags |= ORC_FLAG_WZ;
                if (fabs(wz) < 1e-8) continue;
                const double u = (qx + b[0] * factor) / wz;
                const double w = (qy + b[1] * factor) / wz;
                const int inside_loose = u >= -EDGE_EPS && u <= umax + EDGE_EPS &&
                                         w >= -EDGE_EPS && w <= vmax + EDGE_EPS;
                const int inside_tight = u >= EDGE_EPS && u <= umax - EDGE_EPS &&
                                         w >= EDGE_EPS && w <= vmax - EDGE_EPS;
                if (inside_loose && !inside_tight) flags |= ORC_FLAG_EDGE;
                if (u < 0.0 || u >= umax || w < 0.0 || w >= vmax) continue;
                const int x0 = (int)floor(u), y0 = (int)floor(w);
                const double ax = u - x0, ay = w - y0;
                const size_t q0 = (size_t)y0 * W + x0, q1 = q0 + W;
                double sv;
                if (imgf) {
                    sv = (1 - ax) * (1 - ay) * (double)imgf[q0] + ax * (1 - ay) * (double)imgf[q0 + 1] +
                         (1 - ax) * ay * (double)imgf[q1] + ax * ay * (double)imgf[q1 + 1];
                } else {
                    sv = ((1 - ax) * (1 - ay) * img[q0] + ax * (1 - ay) * img[q0 + 1] +
                          (1 - ax) * ay * img[q1] + ax * ay * img[q1 + 1]) / 255.0;
                }
                sr += rv; ss += sv; srr += rv * rv; sss += sv * sv; srs += rv * sv;
                ++count;
            }
        }
        double c;
        if (count < min_count) {
            c = 2.0;
        } else {
            const double fc = count;
            const double num = fc * srs - sr * ss;
            const double den_r = fc * srr - sr * sr;
            const double den_s = fc * sss - ss * ss;
            if (near_rel(den_r, 1e-6, 1e-3) || near_rel(den_s, 1e-6, 1e-3)) flags |= ORC_FLAG_DEN;
            if (den_r < 1e-6 || den_s < 1e-6) {
                c = 2.0;
            } else {
                double ncc = num / (sqrt(den_r) * sqrt(den_s));
                if (ncc < -1.0) ncc = -1.0;
                if (ncc > 1.0) ncc = 1.0;
                c = 1.0 - ncc;
            }
        }
        if (per_src) per_src[s] = c;
        if (c < top[K - 1]) {                      /* strict <, :326-336 */
            top[K - 1] = c;
            for (int i = K - 2; i >= 0; --i) {
                if (top[i + 1] < top[i]) { double t = top[i]; top[i] = top[i + 1]; top[i + 1] = t; }
                else break;
            }
        }
    }
    const int k = K < v->S ? K : v->S;
    double sum = 0.0;
    for (int i = 0; i < k; ++i) sum += top[i];
    if (flags_out) *flags_out = flags;
    return sum / k;
}

/* ------------------------------------------------------------------------ */
/* Thread helpers: row-interleaved work split, joined per half-pass (the     */
/* join is the barrier between the red and the black pass).                  */
/* ------------------------------------------------------------------------ */
typedef void (*row_fn)(void* ctx, int y);
typedef struct { row_fn fn; void* ctx; int H, t, T; } row_job;

static void* row_worker(void* p)
{
    row_job* j = (row_job*)p;
    for (int y = j->t; y < j->H; y += j->T) j->fn(j->ctx, y);
    return NULL;
}

static void run_rows(row_fn fn, void* ctx, int H, int threads)
{
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    if (threads == 1) {
        for (int y = 0; y < H; ++y) fn(ctx, y);
        return;
    }
    pthread_t tid[256];
    row_job jobs[256];
    int started = 0;
    for (int t = 0; t < threads; ++t) {
        jobs[t] = (row_job){fn, ctx, H, t, threads};
        if (pthread_create(&tid[t], NULL, row_worker, &jobs[t]) != 0) break;
        ++started;
    }
    /* rows a failed thread would have taken are done here */
    for (int t = started; t < threads; ++t) row_worker(&jobs[t]);
    for (int t = 0; t < started; ++t) pthread_join(tid[t], NULL);
}

typedef struct {
    const orc_view* v;
    const float* depth;
    const float* normal;
    double* cost;
    uint32_t* flags;
} costmap_ctx;

static void costmap_row(void* p, int y)
{
    costmap_ctx* c = (costmap_ctx*)p;
    const int W = c->v->W;
    for (int x = 0; x < W; ++x) {
        const size_t i = (size_t)y * W + x;
        uint32_t f = 0;
        c->cost[i] = orc_cost_fp64(c->v, x, y, c->depth[i], c->normal[3 * i],
                                   c->normal[3 * i + 1], c->normal[3 * i + 2], &f, NULL);
        if (c->flags) c->flags[i] = f;
    }
}

void orc_cost_map_fp64(const orc_view* v, const float* depth, const float* normal,
                       double* cost, uint32_t* flags, int threads)
{
    costmap_ctx c = {v, depth, normal, cost, flags};
    run_rows(costmap_row, &c, v->H, threads);
}

typedef struct {
    const orc_view* v;
    int n;
    const int* xs;
    const int* ys;
    const float* depth;
    const float* normal;
    double* cost;
    uint32_t* flags;
} costlist_ctx;

/* "row" r of the work split handles list entries r, r + rows, ... */
#define LIST_ROWS 64
static void costlist_row(void* p, int r)
{
    costlist_ctx* c = (costlist_ctx*)p;
    for (int i = r; i < c->n; i += LIST_ROWS) {
        uint32_t f = 0;
        c->cost[i] = orc_cost_fp64(c->v, c->xs[i], c->ys[i], c->depth[i], c->normal[3 * i],
                                   c->normal[3 * i + 1], c->normal[3 * i + 2], &f, NULL);
        if (c->flags) c->flags[i] = f;
    }
}

void orc_cost_list_fp64(const orc_view* v, int n, const int* xs, const int* ys,
                        const float* depth, const float* normal, double* cost, uint32_t* flags,
                        int threads)
{
    costlist_ctx c = {v, n, xs, ys, depth, normal, cost, flags};
    run_rows(costlist_row, &c, LIST_ROWS, threads);
}

/* ------------------------------------------------------------------------ */
/* FP32 cost, reference operation order (patchmatch_mvs.cu:222-343)          */
/* ------------------------------------------------------------------------ */
typedef struct {
    const orc_view* v;
    const float* ref;            /* gray = DN / 255, H*W */
    const float* src[ORC_MAX_SRC];
    float ki[9];
    float A[ORC_MAX_SRC][9], b[ORC_MAX_SRC][3];
    int cost_mode;               /* ORC_COST_FP32 or ORC_COST_FP64 */
} f32_view;

static float gray_of(uint8_t dn) { return (float)dn / 255.0f; }

static float cost_fp32_ref(const f32_view* fv, int x, int y, float cand_depth, float cnx,
                           float cny, float cnz);

/* Cost used by the full run: the reference's FP32 arithmetic, or the FP64
 * evaluation rounded once to float. */
static float cost_fp32(const f32_view* fv, int x, int y, float d, float nx, float ny, float nz)
{
    if (fv->cost_mode == ORC_COST_FP64)
        return (float)orc_cost_fp64(fv->v, x, y, d, nx, ny, nz, NULL, NULL);
    return cost_fp32_ref(fv, x, y, d, nx, ny, nz);
}

static float cost_fp32_ref(const f32_view* fv, int x, int y, float cand_depth, float cnx,
                           float cny, float cnz)
{
    const orc_view* v = fv->v;
    const float* ki = fv->ki;
    const float fx = (float)x, fy = (float)y;
    const float Px = cand_depth * (ki[0] * fx + ki[1] * fy + ki[2]);
    const float Py = cand_depth * (ki[3] * fx + ki[4] * fy + ki[5]);
    const float Pz = cand_depth * (ki[6] * fx + ki[7] * fy + ki[8]);
    const float n_dot_P = cnx * Px + cny * Py + cnz * Pz;
    if (fabsf(n_dot_P) < 1e-8f) return 2.0f;
    const float inv_n_dot_P = 1.0f / n_dot_P;
    const float nK0 = ki[0] * cnx + ki[3] * cny + ki[6] * cnz;
    const float nK1 = ki[1] * cnx + ki[4] * cny + ki[7] * cnz;
    const float nK2 = ki[2] * cnx + ki[5] * cny + ki[8] * cnz;

    const int n = win_taps(v);
    const int min_count = (n * n) / 2;
    const int W = v->W, H = v->H;
    const float fwidth = (float)W, fheight = (float)H;
    const int K = v->top_k;
    float top[ORC_MAX_SRC];
    for (int i = 0; i < K; ++i) top[i] = 2.0f;

    for (int s = 0; s < v->S; ++s) {
        const float* A = fv->A[s];
        const float* b = fv->b[s];
        const float* img = fv->src[s];
        float sum_ref = 0.0f, sum_src = 0.0f, sum_ref2 = 0.0f, sum_src2 = 0.0f, sum_rs = 0.0f;
        int count = 0;
        for (int j = 0; j < n; ++j) {
            const int dy = -v->radius + v->stride * j;
            for (int i = 0; i < n; ++i) {
                const int dx = -v->radius + v->stride * i;
                const float ref_val =
                    fv->ref[clampi(y + dy, 0, H - 1) * W + clampi(x + dx, 0, W - 1)];
                const float wpx = fx + (float)dx;
                const float wpy = fy + (float)dy;
                const float qx = A[0] * wpx + A[1] * wpy + A[2];
                const float qy = A[3] * wpx + A[4] * wpy + A[5];
                const float qz = A[6] * wpx + A[7] * wpy + A[8];
                const float sval = nK0 * wpx + nK1 * wpy + nK2;
                const float factor = sval * inv_n_dot_P;
                const float wz = qz + b[2] * factor;
                if (fabsf(wz) < 1e-8f) continue;
                const float inv_wz = 1.0f / wz;
                const float src_u = (qx + b[0] * factor) * inv_wz;
                const float src_v = (qy + b[1] * factor) * inv_wz;
                if (src_u < 0.0f || src_u >= fwidth - 1.0f || src_v < 0.0f ||
                    src_v >= fheight - 1.0f)
                    continue;
                /* exact bilinear, math_utils.h:227-249 */
                const int x0 = (int)floorf(src_u), y0 = (int)floorf(src_v);
                const float ax = src_u - floorf(src_u), ay = src_v - floorf(src_v);
                const float* r0 = img + (size_t)y0 * W + x0;
                const float* r1 = r0 + W;
                const float src_val = (1 - ax) * (1 - ay) * r0[0] + ax * (1 - ay) * r0[1] +
                                      (1 - ax) * ay * r1[0] + ax * ay * r1[1];
                sum_ref += ref_val;
                sum_src += src_val;
                sum_ref2 += ref_val * ref_val;
                sum_src2 += src_val * src_val;
                sum_rs += ref_val * src_val;
                count++;
            }
        }
        float c;
        if (count < min_count) {
            c = 2.0f;
        } else {
            const float fc = (float)count;
            const float num = fc * sum_rs - sum_ref * sum_src;
            const float den_r = fc * sum_ref2 - sum_ref * sum_ref;
            const float den_s = fc * sum_src2 - sum_src * sum_src;
            if (den_r < 1e-6f || den_s < 1e-6f) {
                c = 2.0f;
            } else {
                float ncc = num / (sqrtf(den_r) * sqrtf(den_s));
                ncc = fmaxf(-1.0f, fminf(1.0f, ncc));
                c = 1.0f - ncc;
            }
        }
        if (c < top[K - 1]) {
            top[K - 1] = c;
            for (int i = K - 2; i >= 0; --i) {
                if (top[i + 1] < top[i]) { float t = top[i]; top[i] = top[i + 1]; top[i + 1] = t; }
                else break;
            }
        }
    }
    const int k = K < v->S ? K : v->S;
    float sum = 0.0f;
    for (int i = 0; i < k; ++i) sum += top[i];
    return sum / (float)k;
}

static float* gray_plane(const uint8_t* dn, const float* gray, size_t n)
{
    float* g = (float*)malloc(n * sizeof(float));
    if (gray) memcpy(g, gray, n * sizeof(float));
    else for (size_t i = 0; i < n; ++i) g[i] = gray_of(dn[i]);
    return g;
}

static void f32_view_open(f32_view* fv, const orc_view* v)
{
    const size_t n = (size_t)v->W * v->H;
    fv->v = v;
    fv->cost_mode = ORC_COST_FP32;
    fv->ref = gray_plane(v->ref, v->ref_gray, n);
    for (int s = 0; s < v->S; ++s) {
        fv->src[s] = gray_plane(v->src[s], v->src_gray[s], n);
        for (int i = 0; i < 9; ++i) fv->A[s][i] = (float)v->A[s][i];
        for (int i = 0; i < 3; ++i) fv->b[s][i] = (float)v->b[s][i];
    }
    for (int i = 0; i < 9; ++i) fv->ki[i] = (float)v->Kinv[i];
}

static void f32_view_close(f32_view* fv)
{
    free((void*)fv->ref);
    for (int s = 0; s < fv->v->S; ++s) free((void*)fv->src[s]);
}

float orc_cost_fp32(const orc_view* v, int x, int y, float d, float nx, float ny, float nz)
{
    f32_view fv;
    f32_view_open(&fv, v);
    float c = cost_fp32(&fv, x, y, d, nx, ny, nz);
    f32_view_close(&fv);
    return c;
}

/* ------------------------------------------------------------------------ */
/* Hypothesis arithmetic shared by init and refinement.  Every operation is  */
/* written out so that it rounds in a fixed order.                           */
/* ------------------------------------------------------------------------ */

/* Unit view ray K^-1 [x, y, 1] / |.| (patchmatch_mvs.cu:190-194, :590-594). */
static void view_ray(const float ki[9], int x, int y, float* rx, float* ry, float* rz)
{
    const float fx = (float)x, fy = (float)y;
    float a = ki[0] * fx + ki[1] * fy + ki[2];
    float b = ki[3] * fx + ki[4] * fy + ki[5];
    float c = ki[6] * fx + ki[7] * fy + ki[8];
    const float len = sqrtf(a * a + b * b + c * c);
    if (len > 1e-12f) { a = a / len; b = b / len; c = c / len; }
    *rx = a; *ry = b; *rz = c;
}

static void init_pixel(const float ki[9], float dmin, float dmax, uint64_t seed, int W,
                       int x, int y, float* d_out, float* n_out)
{
    const uint32_t idx = (uint32_t)(y * W + x);
    uint32_t r[4];
    philox_block(seed, idx, 0u, 0u, ORC_DOMAIN_INIT, r);
    /* depth U(dmin, dmax] (:184-186) */
    *d_out = dmin + orc_u01(r[0]) * (dmax - dmin);
    float rx, ry, rz;
    view_ray(ki, x, y, &rx, &ry, &rz);
    /* hemisphere normal (:196-215); sin/cos evaluated in double, rounded once */
    const float theta = orc_u01(r[1]) * 6.28318530f;
    const float cos_phi = orc_u01(r[2]);
    const float sin_phi = sqrtf(1.0f - cos_phi * cos_phi);
    float nx = sin_phi * (float)cos((double)theta);
    float ny = sin_phi * (float)sin((double)theta);
    float nz = cos_phi;
    if (nx * rx + ny * ry + nz * rz > 0.0f) { nx = -nx; ny = -ny; nz = -nz; }
    const float nlen = sqrtf(nx * nx + ny * ny + nz * nz);
    if (nlen > 1e-12f) { nx = nx / nlen; ny = ny / nlen; nz = nz / nlen; }
    n_out[0] = nx; n_out[1] = ny; n_out[2] = nz;
}

void orc_init_fp32(const orc_view* v, uint64_t seed, float* depth, float* normal)
{
    float ki[9];
    for (int i = 0; i < 9; ++i) ki[i] = (float)v->Kinv[i];
    for (int y = 0; y < v->H; ++y)
        for (int x = 0; x < v->W; ++x) {
            const size_t i = (size_t)y * v->W + x;
            init_pixel(ki, v->dmin, v->dmax, seed, v->W, x, y, &depth[i], &normal[3 * i]);
        }
}

/* ------------------------------------------------------------------------ */
/* Half-pass (patchmatch_mvs.cu:519-655)                                     */
/* ------------------------------------------------------------------------ */
typedef struct {
    const f32_view* fv;
    int iteration, parity;
    uint64_t pass_seed;
    float* depth;
    float* normal;
    float* cost;
} pass_ctx;

static void halfpass_row(void* p, int y)
{
    pass_ctx* pc = (pass_ctx*)p;
    const f32_view* fv = pc->fv;
    const orc_view* v = fv->v;
    const int W = v->W, H = v->H;
    const float dmin = v->dmin, dmax = v->dmax;
    static const int dx4[4] = {-1, 1, 0, 0};   /* L, R, U, D (:551-552) */
    static const int dy4[4] = {0, 0, -1, 1};
    for (int x = (y + pc->parity) & 1; x < W; x += 2) {
        const size_t idx = (size_t)y * W + x;
        float best_d = pc->depth[idx];
        float best_nx = pc->normal[3 * idx], best_ny = pc->normal[3 * idx + 1],
              best_nz = pc->normal[3 * idx + 2];
        float best_c = pc->cost[idx];
        for (int k = 0; k < 4; ++k) {
            const int qx = x + dx4[k], qy = y + dy4[k];
            if (qx < 0 || qx >= W || qy < 0 || qy >= H) continue;
            const size_t q = (size_t)qy * W + qx;
            const float nd = pc->depth[q];
            if (nd < dmin || nd > dmax) continue;
            const float nnx = pc->normal[3 * q], nny = pc->normal[3 * q + 1],
                        nnz = pc->normal[3 * q + 2];
            const float c = cost_fp32(fv, x, y, nd, nnx, nny, nnz);
            if (c < best_c) { best_c = c; best_d = nd; best_nx = nnx; best_ny = nny; best_nz = nnz; }
        }
        float rx, ry, rz;
        view_ray(fv->ki, x, y, &rx, &ry, &rz);
        float depth_range = (dmax - dmin) * 0.5f;
        float normal_range = 1.0f;
        for (int i = 0; i < pc->iteration; ++i) { depth_range *= 0.6f; normal_range *= 0.6f; }
        for (int step = 0; step < 2; ++step) {          /* NUM_REFINE_STEPS = 2 (:600) */
            uint32_t r[4];
            philox_block(pc->pass_seed, (uint32_t)idx, (uint32_t)pc->iteration, (uint32_t)step,
                         ORC_DOMAIN_REFINE, r);
            const float rand_d = orc_u01(r[0]) * 2.0f - 1.0f;
            float cd = best_d + rand_d * depth_range;
            cd = fmaxf(dmin, fminf(dmax, cd));
            float cnx = best_nx + (orc_u01(r[1]) * 2.0f - 1.0f) * normal_range;
            float cny = best_ny + (orc_u01(r[2]) * 2.0f - 1.0f) * normal_range;
            float cnz = best_nz + (orc_u01(r[3]) * 2.0f - 1.0f) * normal_range;
            const float nlen = sqrtf(cnx * cnx + cny * cny + cnz * cnz);
            if (nlen > 1e-12f) { cnx = cnx / nlen; cny = cny / nlen; cnz = cnz / nlen; }
            if (cnx * rx + cny * ry + cnz * rz > 0.0f) { cnx = -cnx; cny = -cny; cnz = -cnz; }
            const float c = cost_fp32(fv, x, y, cd, cnx, cny, cnz);
            if (c < best_c) { best_c = c; best_d = cd; best_nx = cnx; best_ny = cny; best_nz = cnz; }
            depth_range *= 0.5f;
            normal_range *= 0.5f;
        }
        pc->depth[idx] = best_d;
        pc->normal[3 * idx] = best_nx;
        pc->normal[3 * idx + 1] = best_ny;
        pc->normal[3 * idx + 2] = best_nz;
        pc->cost[idx] = best_c;
    }
}

/* Per-pass seed: seed_view + 100*iter (red), + 50 more for black (:1216,1225). */
static uint64_t pass_seed_of(uint64_t seed, int iteration, int parity)
{
    return seed + 100ull * (uint64_t)iteration + (parity ? 50ull : 0ull);
}

void orc_halfpass_fp32(const orc_view* v, int iteration, int parity, uint64_t seed,
                       float* depth, float* normal, float* cost, int threads)
{
    f32_view fv;
    f32_view_open(&fv, v);
    pass_ctx pc = {&fv, iteration, parity, pass_seed_of(seed, iteration, parity), depth, normal, cost};
    run_rows(halfpass_row, &pc, v->H, threads);
    f32_view_close(&fv);
}

typedef struct {
    const f32_view* fv;
    const float* depth;
    const float* normal;
    float* cost;
} cost32_ctx;

static void cost32_row(void* p, int y)
{
    cost32_ctx* c = (cost32_ctx*)p;
    const int W = c->fv->v->W;
    for (int x = 0; x < W; ++x) {
        const size_t i = (size_t)y * W + x;
        c->cost[i] = cost_fp32(c->fv, x, y, c->depth[i], c->normal[3 * i], c->normal[3 * i + 1],
                               c->normal[3 * i + 2]);
    }
}

void orc_patchmatch_fp32(const orc_view* v, int iterations, uint64_t seed, int init_given,
                         float* depth, float* normal, float* cost, int threads)
{
    orc_patchmatch_ex(v, iterations, seed, init_given, ORC_COST_FP32, depth, normal, cost, threads);
}

void orc_patchmatch_ex(const orc_view* v, int iterations, uint64_t seed, int init_given,
                       int cost_mode, float* depth, float* normal, float* cost, int threads)
{
    f32_view fv;
    f32_view_open(&fv, v);
    fv.cost_mode = cost_mode;
    if (!init_given) orc_init_fp32(v, seed, depth, normal);
    cost32_ctx cc = {&fv, depth, normal, cost};
    run_rows(cost32_row, &cc, v->H, threads);           /* initial cost (:1192-1196) */
    for (int it = 0; it < iterations; ++it)
        for (int parity = 0; parity < 2; ++parity) {     /* red, then black */
            pass_ctx pc = {&fv, it, parity, pass_seed_of(seed, it, parity), depth, normal, cost};
            run_rows(halfpass_row, &pc, v->H, threads);
        }
    f32_view_close(&fv);
}

/* ------------------------------------------------------------------------ */
/* Geometric consistency filter (patchmatch_mvs.cu:668-781)                  */
/* ------------------------------------------------------------------------ */
void orc_geom_filter_fp32(int W, int H, float* depth, const float* normal, float* confidence,
                          const float ki[9], const float R[9], const float t[3],
                          const orc_geom_src* srcs, int num_src, float depth_tol,
                          float normal_tol_cos, int min_consistent)
{
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const size_t idx = (size_t)y * W + x;
            const float d = depth[idx];
            if (d <= 0.0f) { confidence[idx] = 0.0f; continue; }
            const float fx = (float)x, fy = (float)y;
            const float pcx = d * (ki[0] * fx + ki[1] * fy + ki[2]);
            const float pcy = d * (ki[3] * fx + ki[4] * fy + ki[5]);
            const float pcz = d * (ki[6] * fx + ki[7] * fy + ki[8]);
            const float a = pcx - t[0], b = pcy - t[1], c = pcz - t[2];
            const float pwx = R[0] * a + R[3] * b + R[6] * c;
            const float pwy = R[1] * a + R[4] * b + R[7] * c;
            const float pwz = R[2] * a + R[5] * b + R[8] * c;
            const float ncx = normal[3 * idx], ncy = normal[3 * idx + 1], ncz = normal[3 * idx + 2];
            const float nwx = R[0] * ncx + R[3] * ncy + R[6] * ncz;
            const float nwy = R[1] * ncx + R[4] * ncy + R[7] * ncz;
            const float nwz = R[2] * ncx + R[5] * ncy + R[8] * ncz;
            int consistent = 0;
            for (int s = 0; s < num_src; ++s) {
                const orc_geom_src* sv = &srcs[s];
                const float psx = sv->R[0] * pwx + sv->R[1] * pwy + sv->R[2] * pwz + sv->t[0];
                const float psy = sv->R[3] * pwx + sv->R[4] * pwy + sv->R[5] * pwz + sv->t[1];
                const float psz = sv->R[6] * pwx + sv->R[7] * pwy + sv->R[8] * pwz + sv->t[2];
                if (psz <= 0.0f) continue;
                const float prx = sv->K[0] * psx + sv->K[1] * psy + sv->K[2] * psz;
                const float pry = sv->K[3] * psx + sv->K[4] * psy + sv->K[5] * psz;
                const float prz = sv->K[6] * psx + sv->K[7] * psy + sv->K[8] * psz;
                const float u = prx / prz, w = pry / prz;
                const int ui = (int)rintf(u), vi = (int)rintf(w);   /* round to nearest even */
                if (ui < 0 || ui >= W || vi < 0 || vi >= H) continue;
                const size_t sp = (size_t)vi * W + ui;
                const float ds = sv->depth[sp];
                if (ds <= 0.0f) continue;
                if (fabsf(ds - psz) / psz > depth_tol) continue;
                const float nsx = sv->normal[3 * sp], nsy = sv->normal[3 * sp + 1],
                            nsz = sv->normal[3 * sp + 2];
                const float swx = sv->R[0] * nsx + sv->R[3] * nsy + sv->R[6] * nsz;
                const float swy = sv->R[1] * nsx + sv->R[4] * nsy + sv->R[7] * nsz;
                const float swz = sv->R[2] * nsx + sv->R[5] * nsy + sv->R[8] * nsz;
                if (nwx * swx + nwy * swy + nwz * swz < normal_tol_cos) continue;
                consistent++;
            }
            if (consistent < min_consistent) { depth[idx] = 0.0f; confidence[idx] = 0.0f; }
        }
}

/* ==========================================================================
 * Host view set-up.  selectSourceViews and its helpers
 * (src/dense_recon/patchmatch_mvs.cu:41-157) and the seeding block of the
 * driver (:1111-1177), restated literally: shared points by scanning every
 * point for every camera pair, the depth range from sorted visible depths.
 * ======================================================================== */
static int obs(const orc_point* p, int id)
{
    for (int k = 0; k < p->num_image_ids; ++k)
        if (p->image_ids[k] == id) return 1;
    return 0;
}

/* countSharedPoints (:41-52) */
static int count_shared(const orc_point* pts, int n, int a, int b)
{
    int c = 0;
    for (int k = 0; k < n; ++k)
        if (obs(&pts[k], a) && obs(&pts[k], b)) c++;
    return c;
}

static void cam_centre(const orc_camera* c, double C[3])
{
    for (int i = 0; i < 3; ++i)
        C[i] = -(c->R[i] * c->t[0] + c->R[3 + i] * c->t[1] + c->R[6 + i] * c->t[2]);
}

static void cam_point(const orc_camera* c, const double X[3], double p[3])
{
    for (int i = 0; i < 3; ++i)
        p[i] = (c->R[3 * i] * X[0] + c->R[3 * i + 1] * X[1] + c->R[3 * i + 2] * X[2]) + c->t[i];
}

static void unit3(double v[3])
{
    const double n = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (n > 0.0)
        for (int i = 0; i < 3; ++i) v[i] /= n;
}

/* computeBaselineAngle (:55-63) */
static float baseline_angle(const orc_camera* a, const orc_camera* b, const double cen[3])
{
    double ca[3], cb[3], va[3], vb[3];
    cam_centre(a, ca);
    cam_centre(b, cb);
    for (int i = 0; i < 3; ++i) { va[i] = cen[i] - ca[i]; vb[i] = cen[i] - cb[i]; }
    unit3(va);
    unit3(vb);
    double c = va[0] * vb[0] + va[1] * vb[1] + va[2] * vb[2];
    c = c < -1.0 ? -1.0 : (c > 1.0 ? 1.0 : c);
    return (float)(acos(c) * 180.0 / 3.14159265358979323846);
}

/* baselineScore (:66-72): BASELINE_MIN_DEG 5, BASELINE_MAX_DEG 30 (:19-20) */
static float baseline_score(float deg)
{
    if (deg < 5.0f || deg > 30.0f * 3.0f) return 0.0f;
    const float c = deg < 5.0f ? 5.0f : (deg > 30.0f ? 30.0f : deg);
    return sinf(c * (float)3.14159265358979323846 / 180.0f);
}

static int cmp_float(const void* a, const void* b)
{
    const float x = *(const float*)a, y = *(const float*)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

int orc_select_source_views(const orc_camera* cams, int num_cams, const orc_point* pts,
                            int num_pts, int num_sources, orc_view_params* out)
{
    double cen[3] = {0.0, 0.0, 0.0};
    if (num_pts > 0) {
        for (int k = 0; k < num_pts; ++k)
            for (int i = 0; i < 3; ++i) cen[i] += pts[k].position[i];
        for (int i = 0; i < 3; ++i) cen[i] /= (double)num_pts;
    }
    int n_out = 0;
    int* cid = (int*)malloc(sizeof(int) * (num_cams > 0 ? num_cams : 1));
    float* csc = (float*)malloc(sizeof(float) * (num_cams > 0 ? num_cams : 1));
    float* cdeg = (float*)malloc(sizeof(float) * (num_cams > 0 ? num_cams : 1));
    float* dep = (float*)malloc(sizeof(float) * (num_pts > 0 ? num_pts : 1));
    for (int r = 0; r < num_cams; ++r) {
        if (!cams[r].is_registered) continue;
        const int ref = cams[r].image_id;
        orc_view_params* vp = &out[n_out++];
        memset(vp, 0, sizeof(*vp));
        vp->ref_image_id = ref;
        int nc = 0;
        for (int s = 0; s < num_cams; ++s) {
            if (!cams[s].is_registered) continue;
            const int src = cams[s].image_id;
            if (src == ref) continue;
            const float ang = baseline_angle(&cams[ref], &cams[src], cen);
            const float bs = baseline_score(ang);
            if (bs <= 0.0f) continue;
            const int sh = count_shared(pts, num_pts, ref, src);
            if (sh == 0) continue;
            /* insertion keeping descending score, equal scores in arrival order */
            const float sc = (float)sh * bs;
            int pos = nc;
            while (pos > 0 && csc[pos - 1] < sc) {
                cid[pos] = cid[pos - 1];
                csc[pos] = csc[pos - 1];
                cdeg[pos] = cdeg[pos - 1];
                --pos;
            }
            cid[pos] = src;
            csc[pos] = sc;
            cdeg[pos] = ang;
            ++nc;
        }
        vp->num_src = nc < num_sources ? nc : num_sources;
        for (int k = 0; k < vp->num_src; ++k) {
            vp->src_id[k] = cid[k];
            vp->score[k] = csc[k];
            vp->baseline_deg[k] = cdeg[k];
        }
        int nd = 0;
        for (int k = 0; k < num_pts; ++k) {
            if (!obs(&pts[k], ref)) continue;
            double p[3];
            cam_point(&cams[ref], pts[k].position, p);
            if (p[2] > 0.0) dep[nd++] = (float)p[2];
        }
        if (nd >= 2) {
            qsort(dep, (size_t)nd, sizeof(float), cmp_float);
            const size_t i5 = (size_t)(nd * 0.05);
            size_t i95 = (size_t)(nd * 0.95);
            if (i95 > (size_t)nd - 1) i95 = (size_t)nd - 1;
            vp->depth_min = dep[i5] * 0.5f;
            vp->depth_max = dep[i95] * 1.5f;
        } else {
            double p[3];
            cam_point(&cams[ref], cen, p);
            const float d = (float)p[2];
            vp->depth_min = d * 0.3f;
            vp->depth_max = d * 3.0f;
        }
    }
    free(cid);
    free(csc);
    free(cdeg);
    free(dep);
    return n_out;
}

int orc_project_seeds(const orc_camera* cam, const orc_point* pts, int num_pts, float dmin,
                      float dmax, orc_seed* out)
{
    double C[3];
    cam_centre(cam, C);
    int n = 0;
    for (int k = 0; k < num_pts; ++k) {
        if (!obs(&pts[k], cam->image_id)) continue;
        double p[3], q[3];
        cam_point(cam, pts[k].position, p);
        if (p[2] <= 0.0) continue;
        for (int i = 0; i < 3; ++i)
            q[i] = cam->K[3 * i] * p[0] + cam->K[3 * i + 1] * p[1] + cam->K[3 * i + 2] * p[2];
        const int px = (int)round(q[0] / q[2]), py = (int)round(q[1] / q[2]);
        const float z = (float)p[2];
        if (z < dmin || z > dmax) continue;
        double tc[3] = {C[0] - pts[k].position[0], C[1] - pts[k].position[1],
                        C[2] - pts[k].position[2]};
        unit3(tc);
        orc_seed* s = &out[n++];
        s->x = px;
        s->y = py;
        s->depth = z;
        for (int i = 0; i < 3; ++i)
            s->normal[i] = (float)(cam->R[3 * i] * tc[0] + cam->R[3 * i + 1] * tc[1] +
                                   cam->R[3 * i + 2] * tc[2]);
    }
    return n;
}

void orc_apply_seeds(int W, int H, const orc_seed* seeds, int n, int radius, float* depth,
                     float* normal)
{
    for (int k = 0; k < n; ++k)
        for (int dy = -radius; dy <= radius; ++dy)
            for (int dx = -radius; dx <= radius; ++dx) {
                const int sx = seeds[k].x + dx, sy = seeds[k].y + dy;
                if (sx < 0 || sx >= W || sy < 0 || sy >= H) continue;
                const size_t q = (size_t)sy * W + sx;
                depth[q] = seeds[k].depth;
                for (int i = 0; i < 3; ++i) normal[3 * q + i] = seeds[k].normal[i];
            }
}

/* ==========================================================================
 * Depth-map fusion: pixel -> world point, camera normal -> world normal,
 * 8-bit colour -> [0, 1], in view order then row-major order.
 * ======================================================================== */
long long orc_fuse(const orc_fuse_view* views, int n, float min_conf, float* pos, float* nrm,
                   float* col)
{
    long long o = 0;
    for (int i = 0; i < n; ++i) {
        const orc_fuse_view* g = &views[i];
        const float* ki = g->kinv;
        const float* R = g->R;
        for (int y = 0; y < g->H; ++y)
            for (int x = 0; x < g->W; ++x) {
                const size_t q = (size_t)y * g->W + x;
                const float d = g->depth[q];
                if (!(d > 0.0f)) continue;
                if (g->confidence && !(g->confidence[q] >= min_conf)) continue;
                if (pos) {
                    const float fx = (float)x, fy = (float)y;
                    const float pcx = d * (ki[0] * fx + ki[1] * fy + ki[2]);
                    const float pcy = d * (ki[3] * fx + ki[4] * fy + ki[5]);
                    const float pcz = d * (ki[6] * fx + ki[7] * fy + ki[8]);
                    const float a = pcx - g->t[0], b = pcy - g->t[1], c = pcz - g->t[2];
                    pos[3 * o] = R[0] * a + R[3] * b + R[6] * c;
                    pos[3 * o + 1] = R[1] * a + R[4] * b + R[7] * c;
                    pos[3 * o + 2] = R[2] * a + R[5] * b + R[8] * c;
                }
                if (nrm) {
                    float nx = 0.0f, ny = 0.0f, nz = 0.0f;
                    if (g->normal) {
                        nx = g->normal[3 * q];
                        ny = g->normal[3 * q + 1];
                        nz = g->normal[3 * q + 2];
                    }
                    nrm[3 * o] = R[0] * nx + R[3] * ny + R[6] * nz;
                    nrm[3 * o + 1] = R[1] * nx + R[4] * ny + R[7] * nz;
                    nrm[3 * o + 2] = R[2] * nx + R[5] * ny + R[8] * nz;
                }
                if (col) {
                    float r = 0.0f, gg = 0.0f, bb = 0.0f;
                    if (g->color && g->channels == 1) {
                        r = gg = bb = (float)g->color[q] / 255.0f;
                    } else if (g->color && g->channels == 3) {
                        r = (float)g->color[3 * q] / 255.0f;
                        gg = (float)g->color[3 * q + 1] / 255.0f;
                        bb = (float)g->color[3 * q + 2] / 255.0f;
                    }
                    col[3 * o] = r;
                    col[3 * o + 1] = gg;
                    col[3 * o + 2] = bb;
                }
                ++o;
            }
    }
    return o;
}
