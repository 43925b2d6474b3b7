// test_patchmatch.cpp -- the reference's PatchMatch test suite
// (tests/test_patchmatch.cpp:72-539) run against the C++ host API of this
// implementation (include/hp3d_mvs.hpp).
//
// Input: the reference's own layout (src/utils/synthetic_data.cpp:17-28):
// 36 views 10 deg apart on a circle of radius 0.25 around a textured sphere
// of radius 0.05, K = [2000 0 640; 0 2000 480; 0 0 1], 1280x960, dark
// background; rendered by hp3d_scene_render with ground-truth cameras.  The
// reference runs SIFT + SfM to get cameras and sparse points; here the SfM
// result is exact: ground-truth cameras and Fibonacci-lattice surface points
// observed by the views that face them and see them in frame.  Ground-truth
// depth is the exact sphere (the reference fits a sphere to its SfM points,
// :459-468).
//
// Usage: test_patchmatch [--cpu]   (--cpu: only the tests without a GPU)
// Exit status: number of failed tests.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "hp3d_mvs.hpp"
#include "hp3d_scene.h"

using namespace hp3d;

// ---------------------------------------------------------------------------
// minimal harness
// ---------------------------------------------------------------------------
static int g_fail_in_test = 0;
#define EXPECT(cond, ...)                                                   \
    do {                                                                    \
        if (!(cond)) {                                                      \
            std::printf("    EXPECT failed (%s:%d): %s  ", __FILE__, __LINE__, #cond); \
            std::printf(__VA_ARGS__);                                       \
            std::printf("\n");                                              \
            ++g_fail_in_test;                                               \
        }                                                                   \
    } while (0)
#define ASSERT(cond, ...)                                                   \
    do {                                                                    \
        if (!(cond)) {                                                      \
            std::printf("    ASSERT failed (%s:%d): %s  ", __FILE__, __LINE__, #cond); \
            std::printf(__VA_ARGS__);                                       \
            std::printf("\n");                                              \
            ++g_fail_in_test;                                               \
            return;                                                         \
        }                                                                   \
    } while (0)

struct TestCase {
    const char* name;
    bool gpu;
    std::function<void()> fn;
};

// ---------------------------------------------------------------------------
// turntable fixture
// ---------------------------------------------------------------------------
static constexpr int kViews = 36, kW = 1280, kH = 960;
static constexpr double kF = 2000.0, kRadius = 0.05, kDist = 0.25;
static const double kCentre[3] = {0.0, 0.0, 0.25};

static std::vector<ImageData> images_;
static SfMResult sfm_;

static void arc_pose(double deg, Mat3& R, Vec3& t)
{
    const double th = deg * M_PI / 180.0;
    const double C[3] = {kDist * std::sin(th), 0.0, kCentre[2] - kDist * std::cos(th)};
    const double rows[9] = {std::cos(th), 0.0, std::sin(th), 0.0, 1.0, 0.0,
                            -std::sin(th), 0.0, std::cos(th)};
    for (int i = 0; i < 9; ++i) R[i] = rows[i];
    for (int i = 0; i < 3; ++i) t[i] = -(R[3 * i] * C[0] + R[3 * i + 1] * C[1] + R[3 * i + 2] * C[2]);
}

static void setup_fixture()
{
    hp3d_scene s{};
    s.kind = HP3D_SCENE_SPHERE;
    s.seed = 11;
    s.octaves = 4;
    s.tex_freq = 400.0;
    for (int i = 0; i < 3; ++i) s.sphere_center[i] = kCentre[i];
    s.sphere_radius = kRadius;
    const Mat3 K = {kF, 0.0, kW / 2.0, 0.0, kF, kH / 2.0, 0.0, 0.0, 1.0};
    for (int v = 0; v < kViews; ++v) {
        CameraPose c;
        c.image_id = v;
        c.K = K;
        c.is_registered = true;
        arc_pose(10.0 * v, c.R, c.t);
        sfm_.cameras.push_back(c);
        hp3d_camera hc{};
        hc.width = kW;
        hc.height = kH;
        for (int i = 0; i < 9; ++i) {
            hc.K[i] = K[i];
            hc.R[i] = c.R[i];
        }
        for (int i = 0; i < 3; ++i) hc.t[i] = c.t[i];
        std::vector<uint8_t> u8((size_t)kW * kH);
        if (hp3d_scene_render(&s, &hc, 2.0, 5000 + v, u8.data(), nullptr, nullptr, 0) != 0) {
            std::printf("render failed\n");
            std::exit(100);
        }
        ImageData im;
        im.width = kW;
        im.height = kH;
        im.id = v;
        im.K = K;
        im.gray.resize(u8.size());
        for (size_t p = 0; p < u8.size(); ++p) im.gray[p] = (float)u8[p] / 255.0f;
        images_.push_back(std::move(im));
    }
    sfm_.num_registered = kViews;
    // Fibonacci lattice on the sphere; a view observes a point that faces it
    // within 60 degrees (cos > 0.5: features much closer to grazing are not
    // matched by SIFT) and projects into its frame
    const int n = 3000;
    for (int k = 0; k < n; ++k) {
        const double i = k + 0.5, phi = std::acos(1.0 - 2.0 * i / n);
        const double th = M_PI * (1.0 + std::sqrt(5.0)) * i;
        const double nrm[3] = {std::cos(th) * std::sin(phi), std::cos(phi), std::sin(th) * std::sin(phi)};
        SparsePoint p;
        for (int a = 0; a < 3; ++a) p.position[a] = kCentre[a] + kRadius * nrm[a];
        for (int v = 0; v < kViews; ++v) {
            const CameraPose& c = sfm_.cameras[v];
            const Vec3 C = c.cam_center();
            double tc[3], len = 0.0;
            for (int a = 0; a < 3; ++a) {
                tc[a] = C[a] - p.position[a];
                len += tc[a] * tc[a];
            }
            len = std::sqrt(len);
            if ((nrm[0] * tc[0] + nrm[1] * tc[1] + nrm[2] * tc[2]) / len <= 0.5) continue;
            double pc[3];
            for (int a = 0; a < 3; ++a)
                pc[a] = c.R[3 * a] * p.position[0] + c.R[3 * a + 1] * p.position[1] +
                        c.R[3 * a + 2] * p.position[2] + c.t[a];
            const double u = kF * pc[0] / pc[2] + kW / 2.0, w = kF * pc[1] / pc[2] + kH / 2.0;
            if (pc[2] > 0 && u >= 0 && u < kW && w >= 0 && w < kH) p.image_ids.push_back(v);
        }
        if (p.image_ids.size() >= 2) sfm_.points.push_back(p);
    }
}

// exact depth of the sphere along the ray of pixel (x, y) of camera c (0: miss)
static float gt_depth(const CameraPose& c, int x, int y)
{
    const Vec3 C = c.cam_center();
    const double rc[3] = {(x - kW / 2.0) / kF, (y - kH / 2.0) / kF, 1.0};
    double d[3];
    for (int a = 0; a < 3; ++a) d[a] = c.R[a] * rc[0] + c.R[3 + a] * rc[1] + c.R[6 + a] * rc[2];
    const double n = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    for (double& v : d) v /= n;
    double oc[3];
    for (int a = 0; a < 3; ++a) oc[a] = C[a] - kCentre[a];
    const double b = 2.0 * (oc[0] * d[0] + oc[1] * d[1] + oc[2] * d[2]);
    const double cc = oc[0] * oc[0] + oc[1] * oc[1] + oc[2] * oc[2] - kRadius * kRadius;
    const double disc = b * b - 4.0 * cc;
    if (disc < 0) return 0.0f;
    const double th = (-b - std::sqrt(disc)) / 2.0;
    if (th < 0) return 0.0f;
    double hit[3];
    for (int a = 0; a < 3; ++a) hit[a] = C[a] + th * d[a];
    return (float)(c.R[6] * hit[0] + c.R[7] * hit[1] + c.R[8] * hit[2] + c.t[2]);
}

static PipelineConfig cfg_iters(int iters)
{
    PipelineConfig c;
    c.mvs_iterations = iters;
    c.verbose = false;
    return c;
}

// ---------------------------------------------------------------------------
// tests (names as in the reference suite)
// ---------------------------------------------------------------------------
static void SourceViewSelection()
{
    auto vps = selectSourceViews(sfm_, 5);
    ASSERT((int)vps.size() == sfm_.num_registered, "%zu views", vps.size());
    for (const auto& vp : vps) {
        EXPECT(vp.source_views.size() > 0, "image %d has no source views", vp.ref_image_id);
        EXPECT(vp.source_views.size() <= 5, "image %d", vp.ref_image_id);
    }
}

static void SourceViewBaselines()
{
    for (const auto& vp : selectSourceViews(sfm_, 5))
        for (const auto& sv : vp.source_views) {
            EXPECT(sv.baseline_deg >= 3.0f, "src %d ref %d: %f", sv.image_id, vp.ref_image_id,
                   sv.baseline_deg);
            EXPECT(sv.baseline_deg <= 120.0f, "src %d ref %d: %f", sv.image_id, vp.ref_image_id,
                   sv.baseline_deg);
            EXPECT(sv.score > 0.0f, "score %f", sv.score);
        }
}

static void DepthRangeFromSfM()
{
    for (const auto& vp : selectSourceViews(sfm_, 5)) {
        EXPECT(vp.depth_min > 0.0f, "image %d", vp.ref_image_id);
        EXPECT(vp.depth_max > vp.depth_min, "image %d", vp.ref_image_id);
        const float r = vp.depth_max / vp.depth_min;
        EXPECT(r > 1.1f && r < 10.0f, "image %d ratio %f", vp.ref_image_id, r);
    }
}

static void InitKernelDepthRange()
{
    auto dms = runPatchMatchMVS(images_, sfm_.cameras, sfm_, cfg_iters(0));
    ASSERT(!dms.empty(), "no maps");
    const DepthMap& dm = dms[0];
    EXPECT(dm.width == images_[dm.image_id].width && dm.height == images_[dm.image_id].height,
           "size %dx%d", dm.width, dm.height);
    float dmin = 0.0f, dmax = 0.0f;
    for (const auto& vp : selectSourceViews(sfm_, 5))
        if (vp.ref_image_id == dm.image_id) {
            dmin = vp.depth_min;
            dmax = vp.depth_max;
        }
    ASSERT(dmax > 0.0f, "no view params for image %d", dm.image_id);
    int out = 0;
    for (float d : dm.depth) out += !(d >= dmin && d <= dmax);
    EXPECT(out == 0, "%d/%zu pixels out of range", out, dm.depth.size());
}

static void InitKernelNormals()
{
    auto dms = runPatchMatchMVS(images_, sfm_.cameras, sfm_, cfg_iters(0));
    ASSERT(!dms.empty(), "no maps");
    const DepthMap& dm = dms[0];
    int unit = 0;
    const int total = dm.width * dm.height;
    for (int p = 0; p < total; ++p) {
        const float* n = &dm.normal[3 * p];
        unit += std::abs(std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]) - 1.0f) < 0.01f;
    }
    EXPECT((float)unit / total > 0.99f, "%d/%d unit normals", unit, total);
}

static void PFMSaveLoad()
{
    auto dms = runPatchMatchMVS(images_, sfm_.cameras, sfm_, cfg_iters(0));
    ASSERT(!dms.empty(), "no maps");
    const DepthMap& dm = dms[0];
    const std::string path = "/tmp/hp3d_test_depth.pfm";
    ASSERT(saveDepthMapPFM(dm, path), "save");
    DepthMap loaded;
    ASSERT(loadDepthMapPFM(loaded, path), "load");
    EXPECT(loaded.width == dm.width && loaded.height == dm.height, "size");
    EXPECT(std::memcmp(loaded.depth.data(), dm.depth.data(), dm.depth.size() * 4) == 0,
           "values differ");
    std::remove(path.c_str());
}

static void PFMSaveLoadHost()   // the same round trip without a GPU
{
    DepthMap dm;
    dm.width = 37;
    dm.height = 11;
    for (int i = 0; i < 37 * 11; ++i) dm.depth.push_back(0.5f + 0.001f * i);
    const std::string path = "/tmp/hp3d_test_depth_host.pfm";
    ASSERT(saveDepthMapPFM(dm, path), "save");
    DepthMap loaded;
    ASSERT(loadDepthMapPFM(loaded, path), "load");
    EXPECT(loaded.width == 37 && loaded.height == 11 && loaded.depth == dm.depth, "round trip");
    EXPECT(!loadDepthMapPFM(loaded, "/tmp/hp3d_no_such_file.pfm"), "missing file must fail");
    std::remove(path.c_str());
}

static void NCCCostValuesInRange()
{
    auto dms = runPatchMatchMVS(images_, sfm_.cameras, sfm_, cfg_iters(0));
    ASSERT(!dms.empty(), "no maps");
    const DepthMap& dm = dms[0];
    int in = 0;
    double sum = 0.0;
    for (float c : dm.confidence) {
        in += c >= -0.01f && c <= 1.01f;
        sum += c;
    }
    EXPECT((float)in / dm.confidence.size() > 0.99f, "%d in range", in);
    EXPECT(std::isfinite(sum / dm.confidence.size()), "mean not finite");
}

static void NCCCostNotAllSame()
{
    auto dms = runPatchMatchMVS(images_, sfm_.cameras, sfm_, cfg_iters(0));
    ASSERT(!dms.empty(), "no maps");
    float lo = 1e10f, hi = -1e10f;
    for (float c : dms[0].confidence) {
        lo = std::min(lo, c);
        hi = std::max(hi, c);
    }
    EXPECT(hi - lo > 0.01f, "min %f max %f", lo, hi);
}

static void NCCCostCorrectDepthBetter()
{
    auto dms = runPatchMatchMVS(images_, sfm_.cameras, sfm_, cfg_iters(0));
    ASSERT(!dms.empty(), "no maps");
    const DepthMap& dm = dms[0];
    const CameraPose& cam = sfm_.cameras[dm.image_id];
    double sum = 0.0;
    int n = 0;
    for (const auto& pt : sfm_.points) {
        bool vis = false;
        for (int id : pt.image_ids) vis |= id == dm.image_id;
        if (!vis) continue;
        double pc[3];
        for (int a = 0; a < 3; ++a)
            pc[a] = cam.R[3 * a] * pt.position[0] + cam.R[3 * a + 1] * pt.position[1] +
                    cam.R[3 * a + 2] * pt.position[2] + cam.t[a];
        if (pc[2] <= 0) continue;
        const int px = (int)(kF * pc[0] / pc[2] + kW / 2.0), py = (int)(kF * pc[1] / pc[2] + kH / 2.0);
        if (px < 0 || px >= dm.width || py < 0 || py >= dm.height) continue;
        sum += 2.0 * (1.0 - dm.confidence[(size_t)py * dm.width + px]);
        ++n;
    }
    ASSERT(n > 10, "%d sparse points visible", n);
    const double mean = sum / n;
    EXPECT(std::isfinite(mean) && mean < 2.0, "mean cost %f", mean);
    // with exact SfM seeds the sparse points sit on seeded hypotheses
    std::printf("    [Test] NCC cost at %d SfM points: mean=%.3f\n", n, mean);
}

static void PropagationReducesCost()
{
    auto d0 = runPatchMatchMVS(images_, sfm_.cameras, sfm_, cfg_iters(0));
    ASSERT(!d0.empty(), "no maps");
    double s0 = 0.0;
    for (float c : d0[0].confidence) s0 += 2.0 * (1.0 - c);
    const double m0 = s0 / d0[0].confidence.size();
    auto d3 = runPatchMatchMVS(images_, sfm_.cameras, sfm_, cfg_iters(3));
    ASSERT(!d3.empty(), "no maps");
    double s3 = 0.0;
    int n3 = 0;
    for (size_t p = 0; p < d3[0].depth.size(); ++p) {
        if (d3[0].depth[p] <= 0.0f) continue;
        s3 += 2.0 * (1.0 - d3[0].confidence[p]);
        ++n3;
    }
    ASSERT(n3 > 0, "no pixels survived the geometric filter");
    const double m3 = s3 / n3;
    std::printf("    [Test] Mean cost: 0 iters=%.4f (all px), 3 iters=%.4f (%d retained px)\n", m0,
                m3, n3);
    EXPECT(m3 < m0, "%f !< %f", m3, m0);
}

static void EndToEndFilterAndAccuracy()
{
    PipelineConfig cfg = cfg_iters(12);
    const auto t0 = std::chrono::steady_clock::now();
    auto dms = runPatchMatchMVS(images_, sfm_.cameras, sfm_, cfg);
    const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    ASSERT((int)dms.size() == sfm_.num_registered, "%zu maps", dms.size());
    long long tot_s = 0, ret_s = 0, tot = 0, ret = 0, acc = 0, chk = 0;
    double err_sum = 0.0;
    for (const auto& dm : dms) {
        const CameraPose& cam = sfm_.cameras[dm.image_id];
        for (int y = 0; y < dm.height; ++y)
            for (int x = 0; x < dm.width; ++x) {
                ++tot;
                const float d = dm.depth[(size_t)y * dm.width + x];
                ret += d > 0.0f;
                const float gt = gt_depth(cam, x, y);
                if (gt <= 0.0f) continue;
                ++tot_s;
                if (d <= 0.0f) continue;
                ++ret_s;
                const double e = std::abs(d - gt) / gt;
                err_sum += e;
                ++chk;
                acc += e < 0.05;
            }
    }
    const double rs = (double)ret_s / std::max(1LL, tot_s), ra = (double)ret / std::max(1LL, tot);
    std::printf("    [Test] Retention: sphere=%.1f%% (%lld/%lld), total=%.1f%% (%lld/%lld)\n",
                100 * rs, ret_s, tot_s, 100 * ra, ret, tot);
    if (chk > 0) {
        const double me = err_sum / chk;
        std::printf("    [Test] Depth accuracy: %.1f%% within 5%% (%lld/%lld), mean error=%.4f\n",
                    100.0 * acc / chk, acc, chk, me);
        EXPECT(me < 0.15, "mean error %f", me);
    }
    std::printf("    [Test] Total MVS time: %.2f sec for %d views at %dx%d\n", sec,
                sfm_.num_registered, dms[0].width, dms[0].height);
    EXPECT(rs >= 0.50, "sphere retention %f", rs);
    EXPECT(ra < 0.80, "total retention %f", ra);
    EXPECT(sec < 300.0, "%f s", sec);

    // fusion of the filtered maps: every point near the sphere surface
    DensePointCloud pc = fuseDepthMaps(dms, sfm_.cameras, images_);
    EXPECT(pc.size() == (int)ret, "%d points for %lld retained pixels", pc.size(), ret);
    long long near = 0;
    for (const auto& p : pc.positions) {
        const double r = std::sqrt(p[0] * p[0] + p[1] * p[1] + (p[2] - kCentre[2]) * (p[2] - kCentre[2]));
        near += std::abs(r - kRadius) < 0.1 * kRadius;
    }
    const double fn = (double)near / std::max(1, pc.size());
    std::printf("    [Test] Fusion: %d points, %.1f%% within 10%% of the sphere radius\n", pc.size(),
                100.0 * fn);
    EXPECT(fn > 0.5, "fraction near the surface %f", fn);
}

static void LegacyStubs()
{
    DepthMap dm = patchMatchMVS(images_[0], {}, sfm_.cameras[0], {});
    EXPECT(dm.image_id == -1 && dm.width == 0 && dm.depth.empty(), "legacy patchMatchMVS");
    EXPECT(runDenseReconstruction(images_, sfm_.cameras, PipelineConfig{}).empty(),
           "legacy runDenseReconstruction");
}

int main(int argc, char** argv)
{
    const bool cpu_only = argc > 1 && std::strcmp(argv[1], "--cpu") == 0;
    setup_fixture();
    std::printf("fixture: %d views %dx%d, %zu SfM points\n", kViews, kW, kH, sfm_.points.size());
    const std::vector<TestCase> tests = {
        {"SourceViewSelection", false, SourceViewSelection},
        {"SourceViewBaselines", false, SourceViewBaselines},
        {"DepthRangeFromSfM", false, DepthRangeFromSfM},
        {"PFMSaveLoadHost", false, PFMSaveLoadHost},
        {"LegacyStubs", false, LegacyStubs},
        {"InitKernelDepthRange", true, InitKernelDepthRange},
        {"InitKernelNormals", true, InitKernelNormals},
        {"PFMSaveLoad", true, PFMSaveLoad},
        {"NCCCostValuesInRange", true, NCCCostValuesInRange},
        {"NCCCostNotAllSame", true, NCCCostNotAllSame},
        {"NCCCostCorrectDepthBetter", true, NCCCostCorrectDepthBetter},
        {"PropagationReducesCost", true, PropagationReducesCost},
        {"EndToEndFilterAndAccuracy", true, EndToEndFilterAndAccuracy},
    };
    int failed = 0;
    for (const auto& t : tests) {
        if (cpu_only && t.gpu) continue;
        std::printf("[ RUN      ] %s\n", t.name);
        std::fflush(stdout);
        g_fail_in_test = 0;
        try {
            t.fn();
        } catch (const MVSError& e) {
            std::printf("    exception: %s (code %d)\n", e.what(), e.code());
            ++g_fail_in_test;
        }
        std::printf("[ %s ] %s\n", g_fail_in_test ? "  FAILED" : "      OK", t.name);
        std::fflush(stdout);
        failed += g_fail_in_test != 0;
    }
    std::printf("%d test(s) failed\n", failed);
    return failed;
}
