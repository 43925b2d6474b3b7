/*
 * hp3d_scene.h -- seeded synthetic scenes with exact ground truth.
 *
 * Replaces, for this path's inputs, the reference's sphere-turntable
 * generator generateSyntheticData (src/utils/synthetic_data.cpp:222), whose
 * noise is unseeded (cv::randn, :75-161).  Here every image is a pure
 * function of (scene, camera, noise seed):
 *   - texture: 3-D value noise (smoothstep-interpolated hashed lattice),
 *     `octaves` octaves, frequency x2 and amplitude x1/2 per octave,
 *     evaluated at the ray's first surface hit in world coordinates;
 *   - DN = clamp(round(32 + 192 v + sigma N(0,1)), 0, 255);
 *   - ground truth: camera-frame z-depth and camera-frame unit normal of the
 *     first hit (0 where the ray hits nothing).
 */
#ifndef HP3D_SCENE_H
#define HP3D_SCENE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    HP3D_SCENE_PLANE = 0,         /* one infinite textured plane */
    HP3D_SCENE_SPHERE_GROUND = 1, /* sphere + ground plane y = ground_y + backdrop z = backdrop_z */
    HP3D_SCENE_SPHERE = 2         /* sphere alone on a flat DN 30 background (depth 0 and
                                     no added noise there), the layout of the reference's
                                     turntable data (src/utils/synthetic_data.cpp:17-28,80-84) */
} hp3d_scene_kind;

typedef struct {
    int kind;
    uint64_t seed;             /* texture lattice seed */
    int octaves;               /* >= 1 */
    double tex_freq;           /* lattice cells per world unit at octave 0 */
    double plane_point[3];     /* HP3D_SCENE_PLANE */
    double plane_normal[3];
    double sphere_center[3];   /* HP3D_SCENE_SPHERE_GROUND */
    double sphere_radius;
    double ground_y;
    double backdrop_z;
} hp3d_scene;

typedef struct {
    int width, height;
    double K[9];   /* row-major intrinsics */
    double R[9];   /* world -> camera rotation */
    double t[3];   /* world -> camera translation: P_cam = R P_world + t */
} hp3d_camera;

/* Render one 8-bit gray view.  gt_depth (H*W) and gt_normal (H*W*3) are
 * nullable.  threads <= 0 uses the hardware concurrency.  Returns 0 or -1. */
int hp3d_scene_render(const hp3d_scene* scene, const hp3d_camera* cam, double noise_sigma,
                      uint64_t noise_seed, uint8_t* image, float* gt_depth, float* gt_normal,
                      int threads);

#ifdef __cplusplus
}
#endif
#endif
