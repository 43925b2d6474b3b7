// hp3d_mvs.hpp -- C++ host API of the PatchMatch-MVS path, over the C ABI.
//
// Mirrors the reference's dense-reconstruction entry points
// (src/dense_recon/patchmatch_mvs.h:16-62) with the same names, argument
// meaning and results, on plain containers instead of cv::Mat / Eigen:
//
//   selectSourceViews(sfm, n)        patchmatch_mvs.h:32-35
//   runPatchMatchMVS(images, cameras, sfm, config)
//                                    patchmatch_mvs.h:39-44, .cu:1028-1388
//   patchMatchMVS(...)               patchmatch_mvs.h:47-51 (legacy stub, .cu:1394-1404)
//   runDenseReconstruction(...)      patchmatch_mvs.h:54-56 (legacy stub, .cu:1406-1410)
//   saveDepthMapPFM / loadDepthMapPFM  patchmatch_mvs.h:59,62
//   fuseDepthMaps(maps, poses, images) src/dense_recon/point_cloud.h:7-9 (a
//                                    stub returning an empty cloud in the
//                                    reference; here every valid pixel becomes
//                                    a world point with normal and colour)
//
// Every device step goes through hp3d_pm.h; this layer holds no HIP code.
// Differences in error behaviour: a device or argument failure throws
// hp3d::MVSError (status code + message) where the reference prints and
// exits (src/utils/cuda_utils.cuh:7-15); the PFM functions return bool like
// the reference's.
#pragma once

#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "hp3d_pm.h"

namespace hp3d {

using Mat3 = std::array<double, 9>;   // row-major
using Vec3 = std::array<double, 3>;

// ImageData (include/types.h:15-24): the fields this path reads.
struct ImageData {
    std::vector<float> gray;       // H*W gray in [0, 1], row-major (the CV_32F gray)
    std::vector<uint8_t> gray_u8;  // or 8-bit DN; used when gray is empty
    Mat3 K{};                      // intrinsics
    int width = 0, height = 0;
    int id = 0;
    std::string filename;
};

// CameraPose (include/types.h:55-74): P_cam = R P_world + t.
struct CameraPose {
    int image_id = 0;
    Mat3 R{}, K{};
    Vec3 t{};
    bool is_registered = false;
    Vec3 cam_center() const;   // -R^T t
};

// SparsePoint (include/types.h:76-83): the fields this path reads.
struct SparsePoint {
    Vec3 position{};
    std::vector<int> image_ids;
};

// SfMResult (src/sfm/camera_pose.h:8-14).
struct SfMResult {
    std::vector<CameraPose> cameras;
    std::vector<SparsePoint> points;
    int num_registered = 0;
};

// PipelineConfig (include/types.h:94-151): the dense-reconstruction fields,
// plus the device options of this implementation.
struct PipelineConfig {
    bool save_intermediate = false;
    bool verbose = false;
    std::string output_path;
    int mvs_iterations = 12;
    int mvs_patch_size = 11;
    // MI355X options (defaults reproduce the reference's arithmetic)
    int mvs_window_stride = 1;            // 1 = every tap of the window
    int mvs_storage = HP3D_STORE_F32;     // HP3D_STORE_F16: FP16 image planes
    int device = 0;
    size_t hbm_budget = 0;                // 0 = 90 % of free device memory
};

// DepthMap (include/types.h:85-91).  normal is in the reference camera's
// frame (the reference stores camera-frame normals too, .cu:1236-1244).
struct DepthMap {
    std::vector<float> depth;        // H*W, 0 = invalid
    std::vector<float> normal;       // H*W*3
    std::vector<float> confidence;   // H*W, 1 - cost/2
    int image_id = -1;
    int width = 0, height = 0;
};

// SourceViewInfo / MVSViewParams (patchmatch_mvs.h:16-28).
struct SourceViewInfo {
    int image_id;
    float score;
    float baseline_deg;
};
struct MVSViewParams {
    int ref_image_id = -1;
    std::vector<SourceViewInfo> source_views;
    float depth_min = 0.0f, depth_max = 0.0f;
};

// DensePointCloud (include/types.h:93-105): host arrays only.
struct DensePointCloud {
    std::vector<std::array<float, 3>> positions, normals, colors;
    int size() const { return (int)positions.size(); }
};

class MVSError : public std::runtime_error {
public:
    MVSError(int code, const std::string& msg) : std::runtime_error(msg), code_(code) {}
    int code() const { return code_; }

private:
    int code_;
};

std::vector<MVSViewParams> selectSourceViews(const SfMResult& sfm, int num_sources = 5);

std::vector<DepthMap> runPatchMatchMVS(const std::vector<ImageData>& images,
                                       const std::vector<CameraPose>& cameras,
                                       const SfMResult& sfm, const PipelineConfig& config);

DepthMap patchMatchMVS(const ImageData& ref_image, const std::vector<ImageData>& src_images,
                       const CameraPose& ref_pose, const std::vector<CameraPose>& src_poses,
                       int iterations = 8, int patch_size = 11);

std::vector<DepthMap> runDenseReconstruction(const std::vector<ImageData>& images,
                                             const std::vector<CameraPose>& poses,
                                             const PipelineConfig& config);

// Pixels with depth > 0 and confidence >= min_confidence (when the map has
// a confidence channel) in map order then row-major order; colours from the
// image's 8-bit gray (or gray * 255) of the map's image_id.
DensePointCloud fuseDepthMaps(const std::vector<DepthMap>& depth_maps,
                              const std::vector<CameraPose>& poses,
                              const std::vector<ImageData>& images, float min_confidence = 0.0f,
                              int device = 0);

bool saveDepthMapPFM(const DepthMap& dm, const std::string& path);
bool loadDepthMapPFM(DepthMap& dm, const std::string& path);

}  // namespace hp3d
