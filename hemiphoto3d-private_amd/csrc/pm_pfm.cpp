// pm_pfm.cpp -- PFM depth-map I/O byte-compatible with saveDepthMapPFM /
// loadDepthMapPFM (src/dense_recon/patchmatch_mvs.cu:990-1022): header
// "Pf\n<W> <H>\n-1.0\n" (little-endian float), rows stored bottom to top;
// plus the three-channel "PF" form for normal maps.
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../../include/hp3d_pm.h"

extern "C" {

int hp3d_pfm_save_ex(const char* path, int W, int H, int C, const float* data)
{
    if (!path || !data || W <= 0 || H <= 0 || (C != 1 && C != 3)) return HP3D_ERR_INVALID_ARG;
    FILE* f = std::fopen(path, "wb");
    if (!f) return HP3D_ERR_IO;
    bool ok = std::fprintf(f, "%s\n%d %d\n-1.0\n", C == 1 ? "Pf" : "PF", W, H) > 0;
    const size_t row = (size_t)W * C;
    for (int y = H - 1; ok && y >= 0; --y)
        ok = std::fwrite(data + (size_t)y * row, sizeof(float), row, f) == row;
    ok = (std::fclose(f) == 0) && ok;
    return ok ? HP3D_OK : HP3D_ERR_IO;
}

int hp3d_pfm_save(const char* path, int W, int H, const float* depth)
{
    return hp3d_pfm_save_ex(path, W, H, 1, depth);
}

int hp3d_pfm_load_ex(const char* path, int* W, int* H, int* C, float* data, size_t capacity)
{
    if (!path || !W || !H || !C) return HP3D_ERR_INVALID_ARG;
    FILE* f = std::fopen(path, "rb");
    if (!f) return HP3D_ERR_IO;
    char magic[3] = {0, 0, 0};
    int w = 0, h = 0;
    float scale = 0.f;
    // the reference reads: magic token, width, height, scale, then one
    // whitespace byte (patchmatch_mvs.cu:1006-1014)
    if (std::fscanf(f, "%2s %d %d %f", magic, &w, &h, &scale) != 4 || w <= 0 || h <= 0 ||
        (std::strcmp(magic, "Pf") != 0 && std::strcmp(magic, "PF") != 0)) {
        std::fclose(f);
        return HP3D_ERR_IO;
    }
    std::fgetc(f);
    const int c = magic[1] == 'f' ? 1 : 3;
    *W = w;
    *H = h;
    *C = c;
    if (!data) {
        std::fclose(f);
        return HP3D_OK;
    }
    const size_t row = (size_t)w * c, n = row * h;
    if (capacity < n) {
        std::fclose(f);
        return HP3D_ERR_INVALID_ARG;
    }
    bool ok = true;
    for (int y = h - 1; ok && y >= 0; --y)
        ok = std::fread(data + (size_t)y * row, sizeof(float), row, f) == row;
    std::fclose(f);
    if (!ok) return HP3D_ERR_IO;
    if (scale > 0.f) {   // big-endian file: swap bytes (PFM convention)
        for (size_t i = 0; i < n; ++i) {
            uint32_t u;
            std::memcpy(&u, &data[i], 4);
            u = __builtin_bswap32(u);
            std::memcpy(&data[i], &u, 4);
        }
    }
    return HP3D_OK;
}

int hp3d_pfm_load(const char* path, int* W, int* H, float* data, size_t capacity)
{
    int w = 0, h = 0, c = 0;
    int rc = hp3d_pfm_load_ex(path, &w, &h, &c, nullptr, 0);
    if (rc) return rc;
    if (c != 1) return HP3D_ERR_IO;   // the depth format is "Pf" only
    rc = hp3d_pfm_load_ex(path, &w, &h, &c, data, capacity);
    if (rc == HP3D_OK) {
        *W = w;
        *H = h;
    }
    return rc;
}

}  // extern "C"
