// mvs.cpp -- runPatchMatchMVS and friends (include/hp3d_mvs.hpp) over the C ABI.
//
// The driver follows the reference's runPatchMatchMVS
// (src/dense_recon/patchmatch_mvs.cu:1028-1388) step by step:
//   1. source selection with num_sources = min(8, max(3, registered - 1))  (:1036-1039)
//   2. every registered, non-empty image resident on the device           (:1042-1055)
//   3. window radius min(5, max(1, (patch_size - 1) / 2))                  (:1065-1070)
//   4. per reference view with sources: seed 42 + 1e6 id, random init,
//      SfM seeds (radius 50), initial cost, red/black iterations,
//      confidence = 1 - cost / 2, optional PFM dump                       (:1076-1258)
//   5. geometric filter (0.12 relative depth, normals off, >= 2 views)
//      over all maps in order, if iterations > 0 and >= 2 maps             (:1267-1384)
// Views without sources are left out of the result, as in the reference.
#include "../../include/hp3d_mvs.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <memory>

namespace hp3d {

namespace {

constexpr int kMaxSources = HP3D_PM_MAX_SRC;   // MAX_SRC_VIEWS, .cu:18
constexpr int kMaxRadius = 5;                  // NCC_PATCH_RADIUS, .cu:23
constexpr int kSeedRadius = 50;                // SEED_RADIUS, .cu:1124
constexpr float kGeomDepthTol = 0.12f;         // .cu:1273-1275
constexpr float kGeomNormalTolDeg = 180.0f;
constexpr int kGeomMinConsistent = 2;

void check(int rc, hp3d_pm_ctx* ctx, const char* what)
{
    if (rc == HP3D_OK) return;
    std::string msg = std::string(what) + ": ";
    msg += ctx ? hp3d_pm_last_error(ctx) : "invalid argument";
    throw MVSError(rc, msg);
}

struct CtxDeleter {
    void operator()(hp3d_pm_ctx* c) const { hp3d_pm_destroy(c); }
};

hp3d_sfm_camera to_c(const CameraPose& c, int index)
{
    hp3d_sfm_camera o{};
    o.image_id = index;
    o.is_registered = c.is_registered ? 1 : 0;
    for (int i = 0; i < 9; ++i) {
        o.K[i] = c.K[i];
        o.R[i] = c.R[i];
    }
    for (int i = 0; i < 3; ++i) o.t[i] = c.t[i];
    return o;
}

struct CPoints {
    std::vector<hp3d_sfm_point> pts;
    explicit CPoints(const std::vector<SparsePoint>& in)
    {
        pts.resize(in.size());
        for (size_t k = 0; k < in.size(); ++k) {
            for (int i = 0; i < 3; ++i) pts[k].position[i] = in[k].position[i];
            pts[k].image_ids = in[k].image_ids.data();
            pts[k].num_image_ids = (int)in[k].image_ids.size();
        }
    }
};

std::string dir_of(const std::string& path)
{
    const size_t p = path.find_last_of('/');
    return p == std::string::npos ? std::string(".") : path.substr(0, p);
}

}  // namespace

Vec3 CameraPose::cam_center() const
{
    return {-(R[0] * t[0] + R[3] * t[1] + R[6] * t[2]), -(R[1] * t[0] + R[4] * t[1] + R[7] * t[2]),
            -(R[2] * t[0] + R[5] * t[1] + R[8] * t[2])};
}

std::vector<MVSViewParams> selectSourceViews(const SfMResult& sfm, int num_sources)
{
    std::vector<hp3d_sfm_camera> cams(sfm.cameras.size());
    for (size_t i = 0; i < cams.size(); ++i) {
        if (sfm.cameras[i].is_registered && sfm.cameras[i].image_id != (int)i)
            throw MVSError(HP3D_ERR_INVALID_ARG, "selectSourceViews: camera " + std::to_string(i) +
                                                     " has image_id " +
                                                     std::to_string(sfm.cameras[i].image_id));
        cams[i] = to_c(sfm.cameras[i], (int)i);
    }
    CPoints pts(sfm.points);
    std::vector<hp3d_view_params> out(std::max<size_t>(1, cams.size()));
    int n = 0;
    check(hp3d_select_source_views(cams.data(), (int)cams.size(), pts.pts.data(),
                                   (int)pts.pts.size(), std::min(num_sources, kMaxSources),
                                   out.data(), &n),
          nullptr, "selectSourceViews");
    std::vector<MVSViewParams> res(n);
    for (int i = 0; i < n; ++i) {
        res[i].ref_image_id = out[i].ref_image_id;
        res[i].depth_min = out[i].depth_min;
        res[i].depth_max = out[i].depth_max;
        for (int s = 0; s < out[i].num_src; ++s)
            res[i].source_views.push_back(
                {out[i].src[s].image_id, out[i].src[s].score, out[i].src[s].baseline_deg});
    }
    return res;
}

std::vector<DepthMap> runPatchMatchMVS(const std::vector<ImageData>& images,
                                       const std::vector<CameraPose>& cameras,
                                       const SfMResult& sfm, const PipelineConfig& config)
{
    if (config.verbose)
        std::printf("[MVS] Starting PatchMatch MVS on %d cameras, %d iterations\n",
                    sfm.num_registered, config.mvs_iterations);
    if (cameras.size() < images.size())
        throw MVSError(HP3D_ERR_INVALID_ARG, "runPatchMatchMVS: fewer cameras than images");
    const int num_sources = std::min(kMaxSources, std::max(3, sfm.num_registered - 1));
    const std::vector<MVSViewParams> view_params = selectSourceViews(sfm, num_sources);

    hp3d_pm_opts o{};
    o.device = config.device;
    o.hbm_budget = config.hbm_budget;
    hp3d_pm_ctx* raw = nullptr;
    check(hp3d_pm_create(&o, &raw), nullptr, "hp3d_pm_create");
    std::unique_ptr<hp3d_pm_ctx, CtxDeleter> ctx(raw);

    const int n_img = (int)images.size();
    std::vector<hp3d_img> handle(n_img, 0);
    for (int i = 0; i < n_img; ++i) {
        if (!cameras[i].is_registered) continue;
        const ImageData& im = images[i];
        const size_t px = (size_t)im.width * im.height;
        if (!im.gray.empty() && im.gray.size() == px)
            check(hp3d_pm_image_upload(ctx.get(), im.gray.data(), im.width, im.height,
                                       (size_t)im.width * 4, HP3D_FMT_F32, config.mvs_storage,
                                       &handle[i]),
                  ctx.get(), "image upload");
        else if (!im.gray_u8.empty() && im.gray_u8.size() == px)
            check(hp3d_pm_image_upload(ctx.get(), im.gray_u8.data(), im.width, im.height,
                                       (size_t)im.width, HP3D_FMT_U8, config.mvs_storage,
                                       &handle[i]),
                  ctx.get(), "image upload");
    }

    int radius = std::max(1, (config.mvs_patch_size - 1) / 2);
    if (radius > kMaxRadius) {
        std::printf("[MVS] Warning: requested patch_radius %d exceeds max %d, clamping\n", radius,
                    kMaxRadius);
        radius = kMaxRadius;
    }

    CPoints pts(sfm.points);
    std::vector<DepthMap> maps;
    std::vector<const MVSViewParams*> map_params;
    maps.reserve(view_params.size());
    // every view's descriptor first, then one call that solves them back to
    // back with each view's download overlapping the next view's solve
    std::vector<hp3d_view_desc> descs;
    std::vector<std::vector<hp3d_seed_point>> seed_sets;
    descs.reserve(view_params.size());
    seed_sets.reserve(view_params.size());
    for (size_t vi = 0; vi < view_params.size(); ++vi) {
        const MVSViewParams& vp = view_params[vi];
        const int ref_id = vp.ref_image_id;
        if (vp.source_views.empty()) {
            std::printf("[MVS] Warning: no source views for image %d, skipping\n", ref_id);
            continue;
        }
        if (ref_id >= n_img || !handle[ref_id])
            throw MVSError(HP3D_ERR_INVALID_ARG,
                           "runPatchMatchMVS: no image data for view " + std::to_string(ref_id));
        const ImageData& ref_img = images[ref_id];
        const int w = ref_img.width, h = ref_img.height;
        if (config.verbose)
            std::printf("[MVS] Processing view %zu/%zu (image %d): %dx%d, depth=[%.4f, %.4f], "
                        "%zu sources\n",
                        vi + 1, view_params.size(), ref_id, w, h, vp.depth_min, vp.depth_max,
                        vp.source_views.size());

        hp3d_view_desc d{};
        d.ref = handle[ref_id];
        d.num_src = (int)vp.source_views.size();
        const CameraPose& rc = cameras[ref_id];
        check(hp3d_pm_invert3(rc.K.data(), d.K_ref_inv), nullptr, "K_ref inverse");
        for (int s = 0; s < d.num_src; ++s) {
            const int sid = vp.source_views[s].image_id;
            if (sid >= n_img || !handle[sid])
                throw MVSError(HP3D_ERR_INVALID_ARG,
                               "runPatchMatchMVS: no image data for source " + std::to_string(sid));
            d.src[s] = handle[sid];
            const CameraPose& sc = cameras[sid];
            check(hp3d_pm_compose(rc.K.data(), rc.R.data(), rc.t.data(), sc.K.data(), sc.R.data(),
                                  sc.t.data(), d.A[s], d.b[s]),
                  nullptr, "homography composition");
        }
        d.depth_min = vp.depth_min;
        d.depth_max = vp.depth_max;
        d.radius = radius;
        d.stride = config.mvs_window_stride;
        d.iterations = std::max(0, config.mvs_iterations);
        d.top_k = 3;   // NCC_TOP_K, .cu:25
        d.seed = 42ull + (unsigned long long)ref_id * 1000000ull;

        // SfM seeds of the reference camera (cameras[], as the reference's driver uses)
        const hp3d_sfm_camera cref = to_c(rc, ref_id);
        seed_sets.emplace_back(std::max<size_t>(1, pts.pts.size()));
        std::vector<hp3d_seed_point>& seeds = seed_sets.back();
        int n_seeds = 0;
        check(hp3d_project_seed_points(&cref, pts.pts.data(), (int)pts.pts.size(), vp.depth_min,
                                       vp.depth_max, w, h, kSeedRadius, seeds.data(), &n_seeds),
              nullptr, "seed projection");
        d.seeds = seeds.data();
        d.num_seeds = n_seeds;
        d.seed_radius = kSeedRadius;
        descs.push_back(d);

        DepthMap dm;
        dm.image_id = ref_id;
        dm.width = w;
        dm.height = h;
        dm.depth.resize((size_t)w * h);
        dm.normal.resize((size_t)w * h * 3);
        dm.confidence.resize((size_t)w * h);
        maps.push_back(std::move(dm));
        map_params.push_back(&vp);
    }
    if (!maps.empty()) {
        std::vector<hp3d_view_out> outs(maps.size());
        for (size_t i = 0; i < maps.size(); ++i) {
            outs[i].depth = maps[i].depth.data();
            outs[i].normal = maps[i].normal.data();
            outs[i].confidence = maps[i].confidence.data();
        }
        check(hp3d_pm_run_views(ctx.get(), descs.data(), (int)descs.size(), outs.data()),
              ctx.get(), "hp3d_pm_run_views");
    }
    if (config.save_intermediate && !config.output_path.empty())
        for (const DepthMap& dm : maps)
            saveDepthMapPFM(dm, dir_of(config.output_path) + "/depth_" +
                                    std::to_string(dm.image_id) + ".pfm");
    for (hp3d_img hnd : handle)
        if (hnd) hp3d_pm_image_release(ctx.get(), hnd);

    if (config.mvs_iterations > 0 && maps.size() >= 2) {
        std::vector<int> id_to_map(n_img, -1);
        for (size_t i = 0; i < maps.size(); ++i) id_to_map[maps[i].image_id] = (int)i;
        std::vector<hp3d_geom_view> gv(maps.size());
        for (size_t i = 0; i < maps.size(); ++i) {
            hp3d_geom_view& g = gv[i];
            g.width = maps[i].width;
            g.height = maps[i].height;
            g.depth = maps[i].depth.data();
            g.normal = maps[i].normal.data();
            g.confidence = maps[i].confidence.data();
            const CameraPose& c = cameras[maps[i].image_id];
            for (int k = 0; k < 9; ++k) {
                g.K[k] = c.K[k];
                g.R[k] = c.R[k];
            }
            for (int k = 0; k < 3; ++k) g.t[k] = c.t[k];
            for (const SourceViewInfo& sv : map_params[i]->source_views) {
                if (sv.image_id < 0 || sv.image_id >= n_img) continue;
                const int m = id_to_map[sv.image_id];
                if (m >= 0 && g.num_src < HP3D_PM_MAX_SRC) g.src[g.num_src++] = m;
            }
        }
        std::vector<int> retained(maps.size(), -1);
        check(hp3d_pm_geom_filter(ctx.get(), gv.data(), (int)gv.size(), kGeomDepthTol,
                                  kGeomNormalTolDeg, kGeomMinConsistent, retained.data()),
              ctx.get(), "geometric filter");
        if (config.verbose) {
            long long kept = 0, total = 0;
            for (size_t i = 0; i < maps.size(); ++i)
                if (retained[i] >= 0) {
                    kept += retained[i];
                    total += (long long)maps[i].width * maps[i].height;
                }
            if (total > 0)
                std::printf("[MVS] Geometric filter: %lld/%lld total pixels retained (%.1f%%)\n",
                            kept, total, 100.0 * kept / total);
        }
    }
    if (config.verbose)
        std::printf("[MVS] PatchMatch MVS complete: %zu depth maps generated\n", maps.size());
    return maps;
}

DepthMap patchMatchMVS(const ImageData&, const std::vector<ImageData>&, const CameraPose&,
                       const std::vector<CameraPose>&, int, int)
{
    DepthMap dm;   // the reference's legacy entry point returns an empty map
    dm.image_id = -1;
    return dm;
}

std::vector<DepthMap> runDenseReconstruction(const std::vector<ImageData>&,
                                             const std::vector<CameraPose>&,
                                             const PipelineConfig&)
{
    return {};   // legacy stub in the reference as well
}

DensePointCloud fuseDepthMaps(const std::vector<DepthMap>& depth_maps,
                              const std::vector<CameraPose>& poses,
                              const std::vector<ImageData>& images, float min_confidence,
                              int device)
{
    DensePointCloud pc;
    std::vector<hp3d_fuse_view> fv;
    std::vector<std::vector<uint8_t>> gray;
    for (const DepthMap& dm : depth_maps) {
        const size_t px = (size_t)dm.width * dm.height;
        if (dm.image_id < 0 || dm.image_id >= (int)poses.size() || dm.depth.size() != px)
            throw MVSError(HP3D_ERR_INVALID_ARG, "fuseDepthMaps: map without a pose or depth");
        hp3d_fuse_view v{};
        v.width = dm.width;
        v.height = dm.height;
        v.depth = dm.depth.data();
        v.normal = dm.normal.size() == 3 * px ? dm.normal.data() : nullptr;
        v.confidence = dm.confidence.size() == px ? dm.confidence.data() : nullptr;
        if (dm.image_id < (int)images.size()) {
            const ImageData& im = images[dm.image_id];
            if (im.gray_u8.size() == px) {
                v.color = im.gray_u8.data();
            } else if (im.gray.size() == px) {
                gray.emplace_back(px);
                for (size_t p = 0; p < px; ++p)
                    gray.back()[p] = (uint8_t)std::lround(
                        std::min(1.0f, std::max(0.0f, im.gray[p])) * 255.0f);
                v.color = gray.back().data();
            }
            v.color_channels = v.color ? 1 : 0;
        }
        const CameraPose& c = poses[dm.image_id];
        for (int k = 0; k < 9; ++k) {
            v.K[k] = c.K[k];
            v.R[k] = c.R[k];
        }
        for (int k = 0; k < 3; ++k) v.t[k] = c.t[k];
        fv.push_back(v);
    }
    if (fv.empty()) return pc;
    hp3d_pm_opts o{};
    o.device = device;
    hp3d_pm_ctx* raw = nullptr;
    check(hp3d_pm_create(&o, &raw), nullptr, "hp3d_pm_create");
    std::unique_ptr<hp3d_pm_ctx, CtxDeleter> ctx(raw);
    size_t n = 0;
    check(hp3d_pm_fuse(ctx.get(), fv.data(), (int)fv.size(), min_confidence, nullptr, nullptr,
                       nullptr, 0, &n),
          ctx.get(), "fusion count");
    pc.positions.resize(n);
    pc.normals.resize(n);
    pc.colors.resize(n);
    if (n)
        check(hp3d_pm_fuse(ctx.get(), fv.data(), (int)fv.size(), min_confidence,
                           pc.positions[0].data(), pc.normals[0].data(), pc.colors[0].data(), n,
                           &n),
              ctx.get(), "fusion");
    return pc;
}

bool saveDepthMapPFM(const DepthMap& dm, const std::string& path)
{
    if (dm.width < 1 || dm.height < 1 || dm.depth.size() != (size_t)dm.width * dm.height)
        return false;
    return hp3d_pfm_save(path.c_str(), dm.width, dm.height, dm.depth.data()) == HP3D_OK;
}

bool loadDepthMapPFM(DepthMap& dm, const std::string& path)
{
    int w = 0, h = 0;
    if (hp3d_pfm_load(path.c_str(), &w, &h, nullptr, 0) != HP3D_OK) return false;
    std::vector<float> data((size_t)w * h);
    if (hp3d_pfm_load(path.c_str(), &w, &h, data.data(), data.size()) != HP3D_OK) return false;
    dm.width = w;
    dm.height = h;
    dm.depth = std::move(data);   // normal and confidence are left as they were
    return true;
}

}  // namespace hp3d
