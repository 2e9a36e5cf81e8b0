/*
 * Blackframe XPU/HIP — texture lookups and RGB-to-spectrum reflectance, shared by host and
 * device code (the same arithmetic in the same order on both, built with -ffp-contract=off).
 *
 * The reference has no albedo or roughness textures (its device textures only perturb normals);
 * these are this backend's extension. Conventions:
 *  - texel (x, y) channel c lives at data[(y * width + x) * channels + c], row 0 at v = 0;
 *  - lookups repeat in u and v and filter bilinearly between texel centres ((x + 0.5) / width);
 *  - an RGB texel t maps to the reflectance spectrum rho(lambda) = clamp01(sum_c t_c B_c(lambda))
 *    with the three basis spectra B_c of rgb_basis.inc: the smoothest partition of unity whose
 *    film colours are the scaled primaries (see oracle/gen_tables.py for the construction,
 *    after Mallett & Yuksel, EGSR 2019). It is linear in t, so its texel gradient is
 *    w_k * B_c(lambda) for bilinear weight w_k.
 */
#ifndef BLACKFRAME_TEXTURE_H
#define BLACKFRAME_TEXTURE_H

#include "blackframe_math.h"

#define BF_RGB_BASIS_SAMPLES 48u

/* The four texels a bilinear lookup reads, and their weights (summing to 1). */
typedef struct bf_texel_footprint {
    uint32_t index[4]; /* texel numbers y * width + x */
    float weight[4];
} bf_texel_footprint;

BF_HD uint32_t bf_wrap_texel(float coordinate, uint32_t size) {
    const int64_t i = (int64_t)coordinate % (int64_t)size;
    return (uint32_t)(i < 0 ? i + (int64_t)size : i);
}

/* Returns 0 for coordinates that are not finite or beyond 2^24 (no wrap-around is representable). */
BF_HD int bf_texture_footprint(uint32_t width, uint32_t height, float u, float v, bf_texel_footprint* f) {
    if (!(fabsf(u) < 16777216.0f) || !(fabsf(v) < 16777216.0f) || width == 0u || height == 0u) return 0;
    const float x = u * (float)width - 0.5f;
    const float y = v * (float)height - 0.5f;
    const float x0 = floorf(x), y0 = floorf(y);
    const float fx = x - x0, fy = y - y0;
    const uint32_t ix0 = bf_wrap_texel(x0, width), iy0 = bf_wrap_texel(y0, height);
    const uint32_t ix1 = ix0 + 1u == width ? 0u : ix0 + 1u;
    const uint32_t iy1 = iy0 + 1u == height ? 0u : iy0 + 1u;
    f->index[0] = iy0 * width + ix0;
    f->index[1] = iy0 * width + ix1;
    f->index[2] = iy1 * width + ix0;
    f->index[3] = iy1 * width + ix1;
    f->weight[0] = (1.0f - fx) * (1.0f - fy);
    f->weight[1] = fx * (1.0f - fy);
    f->weight[2] = (1.0f - fx) * fy;
    f->weight[3] = fx * fy;
    return 1;
}

/* Filtered value of `channels` channels (<= 3) at a footprint. */
BF_HD void bf_texture_filter(const float* texels, uint32_t channels, const bf_texel_footprint* f, float* out) {
    for (uint32_t c = 0; c < channels; ++c) {
        float sum = f->weight[0] * texels[(size_t)f->index[0] * channels + c];
        sum = sum + f->weight[1] * texels[(size_t)f->index[1] * channels + c];
        sum = sum + f->weight[2] * texels[(size_t)f->index[2] * channels + c];
        sum = sum + f->weight[3] * texels[(size_t)f->index[3] * channels + c];
        out[c] = sum;
    }
}

/* B_c(lambda) for c = 0, 1, 2: linear interpolation on the basis grid (the interpolation the
 * scene's tabulated spectra use). basis is [3][BF_RGB_BASIS_SAMPLES]. */
BF_HD void bf_rgb_basis_at(const float* basis, float wavelength, float out[3]) {
    const float step = 470.0f / (float)(BF_RGB_BASIS_SAMPLES - 1u);
    const float t = (wavelength - 360.0f) / step;
    uint32_t i = t <= 0.0f ? 0u : (uint32_t)t;
    if (i > BF_RGB_BASIS_SAMPLES - 2u) i = BF_RGB_BASIS_SAMPLES - 2u;
    const float f = t - (float)i;
    for (int c = 0; c < 3; ++c) {
        const float* v = basis + c * BF_RGB_BASIS_SAMPLES;
        out[c] = fmaf(v[i + 1] - v[i], f, v[i]);
    }
}

/* Reflectance of an RGB (channels 3) or grey (channels 1) texel value at one wavelength. */
BF_HD float bf_texel_reflectance(const float* basis, const float* value, uint32_t channels, float wavelength) {
    float rho;
    if (channels == 1u) {
        rho = value[0];
    } else {
        float b[3];
        bf_rgb_basis_at(basis, wavelength, b);
        rho = value[0] * b[0];
        rho = rho + value[1] * b[1];
        rho = rho + value[2] * b[2];
    }
    return rho < 0.0f ? 0.0f : (rho > 1.0f ? 1.0f : rho);
}

#endif /* BLACKFRAME_TEXTURE_H */
