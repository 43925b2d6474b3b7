// pm_pfm.cpp -- PFM depth-map I/O byte-compatible with saveDepthMapPFM /
// loadDepthMapPFM (src/dense_recon/patchmatch_mvs.cu:990-1022): header
// "Pf\n<W> <H>\n-1.0\n" (little-endian float), rows stored bottom to top.
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../../include/hp3d_pm.h"

extern "C" {

int hp3d_pfm_save(const char* path, int W, int H, const float* depth)
{
    if (!path || !depth || W <= 0 || H <= 0) return HP3D_ERR_INVALID_ARG;
    FILE* f = std::fopen(path, "wb");
    if (!f) return HP3D_ERR_IO;
    bool ok = std::fprintf(f, "Pf\n%d %d\n-1.0\n", W, H) > 0;
    for (int y = H - 1; ok && y >= 0; --y)
        ok = std::fwrite(depth + (size_t)y * W, sizeof(float), (size_t)W, f) == (size_t)W;
    ok = (std::fclose(f) == 0) && ok;
    return ok ? HP3D_OK : HP3D_ERR_IO;
}

int hp3d_pfm_load(const char* path, int* W, int* H, float* data, size_t capacity)
{
    if (!path || !W || !H) return HP3D_ERR_INVALID_ARG;
    FILE* f = std::fopen(path, "rb");
    if (!f) return HP3D_ERR_IO;
    char magic[3] = {0, 0, 0};
    int w = 0, h = 0;
    float scale = 0.f;
    // the reference reads: magic token, width, height, scale, then one
    // whitespace byte (patchmatch_mvs.cu:1006-1014)
    if (std::fscanf(f, "%2s %d %d %f", magic, &w, &h, &scale) != 4 || std::strcmp(magic, "Pf") != 0 ||
        w <= 0 || h <= 0) {
        std::fclose(f);
        return HP3D_ERR_IO;
    }
    std::fgetc(f);
    *W = w;
    *H = h;
    if (!data) {
        std::fclose(f);
        return HP3D_OK;
    }
    if (capacity < (size_t)w * h) {
        std::fclose(f);
        return HP3D_ERR_INVALID_ARG;
    }
    bool ok = true;
    for (int y = h - 1; ok && y >= 0; --y)
        ok = std::fread(data + (size_t)y * w, sizeof(float), (size_t)w, f) == (size_t)w;
    std::fclose(f);
    if (!ok) return HP3D_ERR_IO;
    if (scale > 0.f) {   // big-endian file: swap bytes (PFM convention)
        for (size_t i = 0; i < (size_t)w * h; ++i) {
            uint32_t u;
            std::memcpy(&u, &data[i], 4);
            u = __builtin_bswap32(u);
            std::memcpy(&data[i], &u, 4);
        }
    }
    return HP3D_OK;
}

}  // extern "C"
