// scene.cpp -- seeded ray-cast scenes for the PatchMatch path (include/hp3d_scene.h).
#include <algorithm>
#include <cmath>
#include <thread>
#include <vector>

#include "../../include/hp3d_scene.h"

namespace {

inline uint64_t mix64(uint64_t z)
{
    // splitmix64 finaliser (Steele, Lea, Flood, "Fast splittable
    // pseudorandom number generators", OOPSLA 2014)
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

inline double lattice(int64_t ix, int64_t iy, int64_t iz, int oct, uint64_t seed)
{
    uint64_t h = seed;
    h = mix64(h ^ (uint64_t)ix);
    h = mix64(h ^ (uint64_t)iy);
    h = mix64(h ^ (uint64_t)iz);
    h = mix64(h ^ (uint64_t)oct);
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

inline double smooth(double f) { return f * f * (3.0 - 2.0 * f); }

double value_noise(double x, double y, double z, int oct, uint64_t seed)
{
    const double fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
    const int64_t ix = (int64_t)fx, iy = (int64_t)fy, iz = (int64_t)fz;
    const double sx = smooth(x - fx), sy = smooth(y - fy), sz = smooth(z - fz);
    double c[2][2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int e = 0; e < 2; ++e) c[a][b][e] = lattice(ix + a, iy + b, iz + e, oct, seed);
    auto lerp = [](double p, double q, double t) { return p + (q - p) * t; };
    const double x00 = lerp(c[0][0][0], c[1][0][0], sx), x10 = lerp(c[0][1][0], c[1][1][0], sx);
    const double x01 = lerp(c[0][0][1], c[1][0][1], sx), x11 = lerp(c[0][1][1], c[1][1][1], sx);
    return lerp(lerp(x00, x10, sy), lerp(x01, x11, sy), sz);
}

double fbm(const hp3d_scene& s, const double p[3])
{
    double sum = 0.0, norm = 0.0, amp = 1.0, f = s.tex_freq;
    for (int o = 0; o < s.octaves; ++o) {
        sum += amp * value_noise(p[0] * f, p[1] * f, p[2] * f, o, s.seed);
        norm += amp;
        amp *= 0.5;
        f *= 2.0;
    }
    return sum / norm;
}

// Nearest positive hit of the ray C + t*dir; returns t (<= 0: miss) and the
// world normal at the hit.
double intersect(const hp3d_scene& s, const double C[3], const double dir[3], double n[3])
{
    double best = -1.0;
    auto plane = [&](const double p0[3], const double pn[3]) {
        const double den = pn[0] * dir[0] + pn[1] * dir[1] + pn[2] * dir[2];
        if (std::fabs(den) < 1e-15) return;
        const double t = (pn[0] * (p0[0] - C[0]) + pn[1] * (p0[1] - C[1]) + pn[2] * (p0[2] - C[2])) / den;
        if (t > 1e-9 && (best < 0.0 || t < best)) {
            best = t;
            n[0] = pn[0]; n[1] = pn[1]; n[2] = pn[2];
        }
    };
    if (s.kind == HP3D_SCENE_PLANE) {
        plane(s.plane_point, s.plane_normal);
    } else {
        const double* c = s.sphere_center;
        const double oc[3] = {C[0] - c[0], C[1] - c[1], C[2] - c[2]};
        const double a = dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2];
        const double b = 2.0 * (oc[0] * dir[0] + oc[1] * dir[1] + oc[2] * dir[2]);
        const double cc = oc[0] * oc[0] + oc[1] * oc[1] + oc[2] * oc[2] - s.sphere_radius * s.sphere_radius;
        const double disc = b * b - 4.0 * a * cc;
        if (disc >= 0.0) {
            const double sq = std::sqrt(disc);
            double t = (-b - sq) / (2.0 * a);
            if (t <= 1e-9) t = (-b + sq) / (2.0 * a);
            if (t > 1e-9) {
                best = t;
                for (int k = 0; k < 3; ++k) n[k] = (C[k] + t * dir[k] - c[k]) / s.sphere_radius;
            }
        }
        if (s.kind == HP3D_SCENE_SPHERE) return best;
        const double gp[3] = {0.0, s.ground_y, 0.0}, gn[3] = {0.0, -1.0, 0.0};
        plane(gp, gn);
        const double bp[3] = {0.0, 0.0, s.backdrop_z}, bn[3] = {0.0, 0.0, -1.0};
        plane(bp, bn);
    }
    return best;
}

bool invert3(const double m[9], double inv[9])
{
    const double det = m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
                       m[2] * (m[3] * m[7] - m[4] * m[6]);
    if (det == 0.0 || !std::isfinite(det)) return false;
    const double id = 1.0 / det;
    inv[0] = (m[4] * m[8] - m[5] * m[7]) * id;
    inv[1] = (m[2] * m[7] - m[1] * m[8]) * id;
    inv[2] = (m[1] * m[5] - m[2] * m[4]) * id;
    inv[3] = (m[5] * m[6] - m[3] * m[8]) * id;
    inv[4] = (m[0] * m[8] - m[2] * m[6]) * id;
    inv[5] = (m[2] * m[3] - m[0] * m[5]) * id;
    inv[6] = (m[3] * m[7] - m[4] * m[6]) * id;
    inv[7] = (m[1] * m[6] - m[0] * m[7]) * id;
    inv[8] = (m[0] * m[4] - m[1] * m[3]) * id;
    return true;
}

}  // namespace

extern "C" int hp3d_scene_render(const hp3d_scene* scene, const hp3d_camera* cam,
                                 double noise_sigma, uint64_t noise_seed, uint8_t* image,
                                 float* gt_depth, float* gt_normal, int threads)
{
    if (!scene || !cam || !image || cam->width <= 0 || cam->height <= 0 || scene->octaves < 1)
        return -1;
    double Ki[9];
    if (!invert3(cam->K, Ki)) return -1;
    const hp3d_scene s = *scene;
    const double* R = cam->R;
    const double* t = cam->t;
    // camera centre C = -R^T t
    const double C[3] = {-(R[0] * t[0] + R[3] * t[1] + R[6] * t[2]),
                         -(R[1] * t[0] + R[4] * t[1] + R[7] * t[2]),
                         -(R[2] * t[0] + R[5] * t[1] + R[8] * t[2])};
    const int W = cam->width, H = cam->height;
    auto row = [&](int y) {
        for (int x = 0; x < W; ++x) {
            const double pc[3] = {Ki[0] * x + Ki[1] * y + Ki[2], Ki[3] * x + Ki[4] * y + Ki[5],
                                  Ki[6] * x + Ki[7] * y + Ki[8]};
            const double dir[3] = {R[0] * pc[0] + R[3] * pc[1] + R[6] * pc[2],
                                   R[1] * pc[0] + R[4] * pc[1] + R[7] * pc[2],
                                   R[2] * pc[0] + R[5] * pc[1] + R[8] * pc[2]};
            double nw[3] = {0, 0, 0};
            const double th = intersect(s, C, dir, nw);
            const size_t q = (size_t)y * W + x;
            double v = s.kind == HP3D_SCENE_SPHERE ? 30.0 : 0.0;
            double depth = 0.0, nc[3] = {0, 0, 0};
            if (th > 0.0) {
                const double X[3] = {C[0] + th * dir[0], C[1] + th * dir[1], C[2] + th * dir[2]};
                v = 32.0 + 192.0 * fbm(s, X);
                depth = R[6] * X[0] + R[7] * X[1] + R[8] * X[2] + t[2];
                for (int k = 0; k < 3; ++k) nc[k] = R[3 * k] * nw[0] + R[3 * k + 1] * nw[1] + R[3 * k + 2] * nw[2];
                if (nc[0] * pc[0] + nc[1] * pc[1] + nc[2] * pc[2] > 0.0)
                    for (double& c : nc) c = -c;
            }
            // the sphere-only scene keeps its background flat (no texture to match)
            if (noise_sigma > 0.0 && !(s.kind == HP3D_SCENE_SPHERE && th <= 0.0)) {
                // Box-Muller on two hashed uniforms
                const uint64_t h1 = mix64(noise_seed ^ mix64(2 * q)), h2 = mix64(noise_seed ^ mix64(2 * q + 1));
                const double u1 = ((h1 >> 11) + 1) * (1.0 / 9007199254740992.0);
                const double u2 = (h2 >> 11) * (1.0 / 9007199254740992.0);
                v += noise_sigma * std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
            }
            image[q] = (uint8_t)std::min(255.0, std::max(0.0, std::nearbyint(v)));
            if (gt_depth) gt_depth[q] = (float)depth;
            if (gt_normal)
                for (int k = 0; k < 3; ++k) gt_normal[3 * q + k] = (float)nc[k];
        }
    };
    int T = threads > 0 ? threads : (int)std::thread::hardware_concurrency();
    T = std::max(1, std::min(T, 64));
    if (T == 1) {
        for (int y = 0; y < H; ++y) row(y);
        return 0;
    }
    std::vector<std::thread> pool;
    for (int k = 0; k < T; ++k)
        pool.emplace_back([&, k] {
            for (int y = k; y < H; y += T) row(y);
        });
    for (auto& th : pool) th.join();
    return 0;
}
