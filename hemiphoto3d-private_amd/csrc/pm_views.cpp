// pm_views.cpp -- host view set-up: source-view selection and SfM seed
// projection (the CPU side of runPatchMatchMVS, src/dense_recon/patchmatch_mvs.cu).
//
//   hp3d_select_source_views  <- selectSourceViews + countSharedPoints,
//                                computeBaselineAngle, baselineScore (:41-157)
//   hp3d_project_seed_points  <- the per-point part of the seeding block
//                                (:1126-1155)
//
// Arithmetic follows the reference's types: geometry in double, angles and
// scores in float.  Shared-point counts come from one pass over the points'
// observation lists (each point adds one to every pair of its distinct
// views), which gives the reference's per-pair scan counts in O(P t^2)
// instead of O(V^2 P t).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "hp3d_pm.h"

namespace {

constexpr float kBaselineMinDeg = 5.0f;    // BASELINE_MIN_DEG, .cu:19
constexpr float kBaselineMaxDeg = 30.0f;   // BASELINE_MAX_DEG, .cu:20
constexpr double kPi = 3.14159265358979323846;

// P_cam = R X + t
void to_camera(const hp3d_sfm_camera& c, const double X[3], double p[3])
{
    for (int i = 0; i < 3; ++i)
        p[i] = (c.R[3 * i] * X[0] + c.R[3 * i + 1] * X[1] + c.R[3 * i + 2] * X[2]) + c.t[i];
}

// C = -R^T t
void centre_of(const hp3d_sfm_camera& c, double C[3])
{
    for (int i = 0; i < 3; ++i)
        C[i] = -(c.R[i] * c.t[0] + c.R[3 + i] * c.t[1] + c.R[6 + i] * c.t[2]);
}

void normalize3(double v[3])
{
    const double n = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (n > 0.0)
        for (int i = 0; i < 3; ++i) v[i] /= n;
}

// Angle (deg) between the rays from two camera centres to the centroid.
float baseline_deg(const hp3d_sfm_camera& a, const hp3d_sfm_camera& b, const double cen[3])
{
    double ca[3], cb[3], va[3], vb[3];
    centre_of(a, ca);
    centre_of(b, cb);
    for (int i = 0; i < 3; ++i) {
        va[i] = cen[i] - ca[i];
        vb[i] = cen[i] - cb[i];
    }
    normalize3(va);
    normalize3(vb);
    const double c = std::min(1.0, std::max(-1.0, va[0] * vb[0] + va[1] * vb[1] + va[2] * vb[2]));
    return (float)(std::acos(c) * 180.0 / kPi);
}

// sin(clamp(angle, 5, 30) deg), 0 outside [5, 90] deg.
float baseline_score(float deg)
{
    if (deg < kBaselineMinDeg || deg > kBaselineMaxDeg * 3.0f) return 0.0f;
    const float c = std::min(kBaselineMaxDeg, std::max(kBaselineMinDeg, deg));
    return std::sin(c * (float)kPi / 180.0f);
}

bool observes(const hp3d_sfm_point& p, int id)
{
    for (int k = 0; k < p.num_image_ids; ++k)
        if (p.image_ids[k] == id) return true;
    return false;
}

}  // namespace

extern "C" int hp3d_select_source_views(const hp3d_sfm_camera* cams, int num_cams,
                                        const hp3d_sfm_point* pts, int num_pts, int num_sources,
                                        hp3d_view_params* out, int* num_out)
{
    if (!cams || num_cams < 0 || num_pts < 0 || (num_pts > 0 && !pts) || !out || !num_out ||
        num_sources < 1 || num_sources > HP3D_PM_MAX_SRC)
        return HP3D_ERR_INVALID_ARG;
    for (int i = 0; i < num_cams; ++i)
        if (cams[i].is_registered && cams[i].image_id != i) return HP3D_ERR_INVALID_ARG;
    for (int k = 0; k < num_pts; ++k)
        if (pts[k].num_image_ids < 0 || (pts[k].num_image_ids > 0 && !pts[k].image_ids))
            return HP3D_ERR_INVALID_ARG;

    double cen[3] = {0.0, 0.0, 0.0};
    if (num_pts > 0) {
        for (int k = 0; k < num_pts; ++k)
            for (int i = 0; i < 3; ++i) cen[i] += pts[k].position[i];
        for (int i = 0; i < 3; ++i) cen[i] /= (double)num_pts;
    }

    // shared[i * V + j]: points observed by both camera i and camera j
    const int V = num_cams;
    std::vector<int> shared((size_t)V * V, 0);
    std::vector<int> ids;
    for (int k = 0; k < num_pts; ++k) {
        ids.assign(pts[k].image_ids, pts[k].image_ids + pts[k].num_image_ids);
        std::sort(ids.begin(), ids.end());
        ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
        for (size_t a = 0; a < ids.size(); ++a) {
            if (ids[a] < 0 || ids[a] >= V) continue;
            for (size_t b = a + 1; b < ids.size(); ++b) {
                if (ids[b] < 0 || ids[b] >= V) continue;
                ++shared[(size_t)ids[a] * V + ids[b]];
                ++shared[(size_t)ids[b] * V + ids[a]];
            }
        }
    }

    std::vector<int> reg;
    for (int i = 0; i < V; ++i)
        if (cams[i].is_registered) reg.push_back(i);

    int n_out = 0;
    std::vector<hp3d_source_view> cand;
    std::vector<float> depths;
    for (int ref : reg) {
        hp3d_view_params vp;
        std::memset(&vp, 0, sizeof(vp));
        vp.ref_image_id = ref;
        cand.clear();
        for (int src : reg) {
            if (src == ref) continue;
            const float ang = baseline_deg(cams[ref], cams[src], cen);
            const float bs = baseline_score(ang);
            if (bs <= 0.0f) continue;
            const int sh = shared[(size_t)ref * V + src];
            if (sh == 0) continue;
            cand.push_back(hp3d_source_view{src, (float)sh * bs, ang});
        }
        std::stable_sort(cand.begin(), cand.end(),
                         [](const hp3d_source_view& a, const hp3d_source_view& b) {
                             return a.score > b.score;
                         });
        vp.num_src = std::min(num_sources, (int)cand.size());
        for (int s = 0; s < vp.num_src; ++s) vp.src[s] = cand[s];

        depths.clear();
        for (int k = 0; k < num_pts; ++k) {
            if (!observes(pts[k], ref)) continue;
            double p[3];
            to_camera(cams[ref], pts[k].position, p);
            if (p[2] > 0.0) depths.push_back((float)p[2]);
        }
        if (depths.size() >= 2) {
            std::sort(depths.begin(), depths.end());
            const size_t i5 = (size_t)(depths.size() * 0.05);
            const size_t i95 = std::min((size_t)(depths.size() * 0.95), depths.size() - 1);
            vp.depth_min = depths[i5] * 0.5f;
            vp.depth_max = depths[i95] * 1.5f;
        } else {
            double p[3];
            to_camera(cams[ref], cen, p);
            const float d = (float)p[2];
            vp.depth_min = d * 0.3f;
            vp.depth_max = d * 3.0f;
        }
        out[n_out++] = vp;
    }
    *num_out = n_out;
    return HP3D_OK;
}

extern "C" int hp3d_project_seed_points(const hp3d_sfm_camera* cam, const hp3d_sfm_point* pts,
                                        int num_pts, float depth_min, float depth_max, int width,
                                        int height, int seed_radius, hp3d_seed_point* out,
                                        int* num_out)
{
    if (!cam || num_pts < 0 || (num_pts > 0 && (!pts || !out)) || !num_out || width < 1 ||
        height < 1 || seed_radius < 0)
        return HP3D_ERR_INVALID_ARG;
    double C[3];
    centre_of(*cam, C);
    int n = 0;
    for (int k = 0; k < num_pts; ++k) {
        const hp3d_sfm_point& pt = pts[k];
        if (!observes(pt, cam->image_id)) continue;
        double p[3];
        to_camera(*cam, pt.position, p);
        if (p[2] <= 0.0) continue;
        double q[3];
        for (int i = 0; i < 3; ++i)
            q[i] = cam->K[3 * i] * p[0] + cam->K[3 * i + 1] * p[1] + cam->K[3 * i + 2] * p[2];
        const double fx = std::round(q[0] / q[2]), fy = std::round(q[1] / q[2]);
        const float z = (float)p[2];
        if (z < depth_min || z > depth_max) continue;
        // a square that misses the image seeds nothing (also keeps the casts in range)
        if (!(fx + seed_radius >= 0.0 && fx - seed_radius <= width - 1.0 &&
              fy + seed_radius >= 0.0 && fy - seed_radius <= height - 1.0))
            continue;
        double to_cam[3] = {C[0] - pt.position[0], C[1] - pt.position[1], C[2] - pt.position[2]};
        normalize3(to_cam);
        hp3d_seed_point s;
        s.x = (int)fx;
        s.y = (int)fy;
        s.depth = z;
        for (int i = 0; i < 3; ++i)
            s.normal[i] = (float)(cam->R[3 * i] * to_cam[0] + cam->R[3 * i + 1] * to_cam[1] +
                                  cam->R[3 * i + 2] * to_cam[2]);
        out[n++] = s;
    }
    *num_out = n;
    return HP3D_OK;
}
