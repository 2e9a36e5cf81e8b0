/*
 * Device scalar library of the HIP wavefront backend: vectors, sampling, checked products, MIS,
 * Russian roulette and the closure lobes. Each function states the reference function whose
 * arithmetic it reproduces operation for operation (file:line relative to the reference root);
 * compiled with -ffp-contract=off so that only the explicit fmaf calls fuse.
 */
#ifndef BF_DEVICE_MATH_H
#define BF_DEVICE_MATH_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/blackframe_hip_abi.h"
#include "../../include/blackframe_math.h"

#define BF_D __device__ __forceinline__

#define BF_EPS 0x1p-23f
#define BF_DENORM_MIN 0x1p-149f
#define BF_PI 3.14159265358979323846f
#define BF_INV_PI 0.31830988618379067154f
#define BF_QUARTER_PI 0.78539816339744830962f

#define BF_DIM_RASTER_X 0xA24BAED4963EE407ull
#define BF_DIM_RASTER_Y 0x9FB21C651E98DF25ull

enum { BF_L_OK = 0, BF_L_INVALID = 1, BF_L_NOT_REPRESENTABLE = 2, BF_L_OUTSIDE = 3 };

struct v3f {
    float x, y, z;
};

BF_D v3f mk(float x, float y, float z) { return v3f{x, y, z}; }
BF_D bool fin3(v3f v) { return bf_isfinite(v.x) && bf_isfinite(v.y) && bf_isfinite(v.z); }
BF_D v3f add3(v3f a, v3f b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
BF_D v3f sub3(v3f a, v3f b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
BF_D v3f mul3(v3f a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
/* CudaWavefrontStages.cu:314-324 */
BF_D float dot3(v3f a, v3f b) { return fmaf(a.x, b.x, fmaf(a.y, b.y, a.z * b.z)); }
BF_D v3f cross3(v3f a, v3f b) {
    return mk(fmaf(a.y, b.z, -a.z * b.y), fmaf(a.z, b.x, -a.x * b.z), fmaf(a.x, b.y, -a.y * b.x));
}
BF_D float comp3(v3f v, uint32_t a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }
BF_D void set3(v3f& v, uint32_t a, float value) {
    if (a == 0) v.x = value;
    else if (a == 1) v.y = value;
    else v.z = value;
}
/* CudaSceneClosestHit.cu:514-536 */
BF_D v3f xform_point(const float* m, v3f p) {
    return mk(fmaf(m[0], p.x, fmaf(m[1], p.y, fmaf(m[2], p.z, m[3]))),
              fmaf(m[4], p.x, fmaf(m[5], p.y, fmaf(m[6], p.z, m[7]))),
              fmaf(m[8], p.x, fmaf(m[9], p.y, fmaf(m[10], p.z, m[11]))));
}
BF_D v3f xform_vector(const float* m, v3f v) {
    return mk(fmaf(m[0], v.x, fmaf(m[1], v.y, m[2] * v.z)), fmaf(m[4], v.x, fmaf(m[5], v.y, m[6] * v.z)),
              fmaf(m[8], v.x, fmaf(m[9], v.y, m[10] * v.z)));
}

/* ---- sampling: SampleStreamDevice.cuh:12-33, OwenScrambling.cpp:8-43 ----------------------- */
BF_D uint64_t mix_bits(uint64_t v) {
    v ^= v >> 30;
    v *= 0xBF58476D1CE4E5B9ull;
    v ^= v >> 27;
    v *= 0x94D049BB133111EBull;
    return v ^ (v >> 31);
}

struct sample_key {
    uint32_t px, py;
    uint64_t s, seed;
};

BF_D uint64_t indexed_bits(const sample_key& k, uint64_t dim) {
    const uint64_t packed = ((uint64_t)k.px << 32) | (uint64_t)k.py;
    uint64_t state = mix_bits(k.seed ^ 0x9E3779B97F4A7C15ull);
    state = mix_bits(state ^ packed ^ 0xD1B54A32D192ED03ull);
    state = mix_bits(state ^ k.s ^ 0x8CB92BA72F3D8DD7ull);
    return mix_bits(state ^ dim);
}

/* Owen-scrambled Sobol, evaluated only to the 24 leading bits the binary32 sample keeps: the
 * nested scramble of OwenScrambling.cpp:22-36 decides output bit 63-d from input bits above it,
 * so depths 24..63 never reach (bits >> 40). */
BF_D uint32_t owen_sobol_top24(const sample_key& k, const uint64_t* directions, uint32_t dims, uint64_t sobol_dim,
                               bool* ok) {
    if (sobol_dim >= dims) {
        *ok = false;
        return 0u;
    }
    uint64_t gray = k.s ^ (k.s >> 1);
    const uint64_t* d = directions + sobol_dim * 64u;
    uint64_t sample = 0;
    while (gray != 0) {
        sample ^= d[__builtin_ctzll(gray)];
        gray &= gray - 1u;
    }
    const uint64_t packed = ((uint64_t)k.px << 32) | (uint64_t)k.py;
    const uint64_t key = mix_bits(k.seed ^ mix_bits(sobol_dim + 1u)) ^ mix_bits(packed);
    const uint32_t input = (uint32_t)(sample >> 40);
    uint32_t prefix = 0u, output = 0u;
    for (uint32_t depth = 0; depth < 24u; ++depth) {
        const uint32_t bit = (input >> (23u - depth)) & 1u;
        const uint64_t node = (1ull << depth) | (uint64_t)prefix;
        const uint32_t flip = (uint32_t)(mix_bits(key ^ node) >> 63);
        output |= (bit ^ flip) << (23u - depth);
        prefix = (prefix << 1) | bit;
    }
    *ok = true;
    return output;
}

struct sampler_ctx {
    uint32_t kind;
    uint32_t sobol_dims;
    const uint64_t* directions;
};

/* One transport dimension -> binary32 in [0, 1): (bits >> 40) * 2^-24 (SampleStreamDevice.cuh:29-32). */
BF_D float sample_1d(const sampler_ctx& sc, const sample_key& k, uint64_t dim) {
    if (sc.kind == BF_SAMPLER_SOBOL_OWEN) {
        const uint64_t sd = dim == BF_DIM_RASTER_X ? 0ull : (dim == BF_DIM_RASTER_Y ? 1ull : dim + 2u);
        bool ok;
        const uint32_t top = owen_sobol_top24(k, sc.directions, sc.sobol_dims, sd, &ok);
        return ok ? (float)top * 0x1p-24f : -1.0f;
    }
    return (float)(indexed_bits(k, dim) >> 40) * 0x1p-24f;
}

/* ---- checked arithmetic: TransportDevice.cuh:196-259 --------------------------------------- */
BF_D bool checked_product(float l, float r, float& out) {
    if (!bf_isfinite(l) || l < 0.0f || !bf_isfinite(r) || r < 0.0f) return false;
    const float p = l * r;
    if (!bf_isfinite(p) || p < 0.0f || (l != 0.0f && r != 0.0f && p == 0.0f)) return false;
    out = p;
    return true;
}

template <int NN, int ND>
BF_D int checked_product_quotient(const float (&num)[NN], const float (&den)[ND], float& out) {
    bool has_zero = false;
#pragma unroll
    for (int i = 0; i < ND; ++i)
        if (!bf_isfinite(den[i]) || !(den[i] > 0.0f)) return BF_L_INVALID;
#pragma unroll
    for (int i = 0; i < NN; ++i) {
        if (!bf_isfinite(num[i]) || num[i] < 0.0f) return BF_L_INVALID;
        has_zero = has_zero || num[i] == 0.0f;
    }
    if (has_zero) {
        out = 0.0f;
        return BF_L_OK;
    }
    float significand = 1.0f;
    int exponent = 0;
#pragma unroll
    for (int i = 0; i < NN; ++i) {
        int e;
        significand *= bf_frexpf(num[i], &e);
        exponent += e;
    }
#pragma unroll
    for (int i = 0; i < ND; ++i) {
        int e;
        significand /= bf_frexpf(den[i], &e);
        exponent -= e;
    }
    if (!bf_isfinite(significand) || !(significand > 0.0f)) return BF_L_NOT_REPRESENTABLE;
    int ne;
    significand = bf_frexpf(significand, &ne);
    exponent += ne;
    const float r = bf_ldexpf(significand, exponent);
    if (!bf_isfinite(r) || !(r > 0.0f)) return BF_L_NOT_REPRESENTABLE;
    out = r;
    return BF_L_OK;
}

/* TransportDevice.cuh:552-590 */
BF_D bool joint_light_pdf(float selection, float conditional, float& out) {
    if (!bf_isfinite(selection) || !(selection > 0.0f) || selection > 1.0f || !bf_isfinite(conditional) ||
        conditional < 0.0f)
        return false;
    if (conditional == 0.0f) {
        out = 0.0f;
        return true;
    }
    const float num[2] = {selection, conditional};
    const float den[1] = {1.0f};
    return checked_product_quotient(num, den, out) == BF_L_OK;
}

/* TransportDevice.cuh:592-628 */
BF_D bool ratio_at_most_one(float n, float d, bool squared, float& out) {
    if (!bf_isfinite(n) || !(n > 0.0f) || !bf_isfinite(d) || !(d > 0.0f) || n > d) return false;
    int ne, de;
    const float ns = bf_frexpf(n, &ne);
    const float ds = bf_frexpf(d, &de);
    const float ratio = ns / ds;
    const float r = squared ? bf_ldexpf(ratio * ratio, 2 * (ne - de)) : bf_ldexpf(ratio, ne - de);
    if (!bf_isfinite(r) || r < 0.0f || r > 1.0f) return false;
    out = r;
    return true;
}

/* TransportDevice.cuh:630-689 */
BF_D bool mis_weight(uint32_t heuristic, float sampled, float competing, float& out) {
    if (!bf_isfinite(sampled) || !(sampled > 0.0f) || !bf_isfinite(competing) || competing < 0.0f || heuristic > 1u)
        return false;
    if (competing == 0.0f) {
        out = 1.0f;
        return true;
    }
    float ratio;
    if (sampled < competing) {
        if (!ratio_at_most_one(sampled, competing, heuristic == 1u, ratio) || !(ratio > 0.0f)) return false;
        const float w = ratio / (1.0f + ratio);
        if (!bf_isfinite(w) || !(w > 0.0f) || w > 1.0f) return false;
        out = w;
        return true;
    }
    if (!ratio_at_most_one(competing, sampled, heuristic == 1u, ratio)) return false;
    const float w = 1.0f / (1.0f + ratio);
    if (!bf_isfinite(w) || !(w > 0.0f) || w > 1.0f) return false;
    out = w;
    return true;
}

/* TransportDevice.cuh:711-782 */
/* The roulette decision with survival probability p: u < p survives and divides the throughput
 * by p. */
BF_D bool russian_roulette_at(float4& t, float p, float u, bool& survived) {
    if (!bf_isfinite(u) || u < 0.0f || !(u < 1.0f)) return false;
    const float v[4] = {t.x, t.y, t.z, t.w};
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (!bf_isfinite(v[k]) || v[k] < 0.0f) return false;
    if (!bf_isfinite(p) || !(p > 0.0f) || p > 1.0f) return false;
    if (!(u < p)) {
        survived = false;
        return true;
    }
    float o[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        o[k] = v[k] / p;
        if (!bf_isfinite(o[k]) || (v[k] != 0.0f && o[k] == 0.0f)) return false;
    }
    t = make_float4(o[0], o[1], o[2], o[3]);
    survived = true;
    return true;
}

/* The survival probability clamp(max(throughput) * eta, minimum, maximum), then the decision. */
BF_D bool russian_roulette(float4& t, float eta, float u, float minimum, float maximum, bool& survived, float& p) {
    if (!bf_isfinite(eta) || !(eta > 0.0f) || !bf_isfinite(u) || u < 0.0f || !(u < 1.0f)) return false;
    const float v[4] = {t.x, t.y, t.z, t.w};
    float m = 0.0f;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (!bf_isfinite(v[k]) || v[k] < 0.0f) return false;
        m = fmaxf(m, v[k]);
    }
    if (m == 0.0f) return false;
    if (m >= maximum / eta) p = maximum;
    else if (m <= minimum / eta) p = minimum;
    else {
        const float c = m * eta;
        p = c < minimum ? minimum : (maximum < c ? maximum : c);
    }
    return russian_roulette_at(t, p, u, survived);
}

/* TransportDevice.cuh:327-333 */
BF_D bool unit_vector(v3f d) {
    if (!fin3(d)) return false;
    const float sq = fmaf(d.x, d.x, fmaf(d.y, d.y, d.z * d.z));
    return bf_isfinite(sq) && fabsf(sq - 1.0f) <= 128.0f * BF_EPS;
}

/* CudaWavefrontStages.cu:326-349 */
BF_D bool normalize3(v3f v, v3f& out) {
    if (!fin3(v)) return false;
    const float m = fmaxf(fabsf(v.x), fmaxf(fabsf(v.y), fabsf(v.z)));
    if (!bf_isfinite(m) || m == 0.0f) return false;
    const v3f s = mk(v.x / m, v.y / m, v.z / m);
    const float mag = sqrtf(dot3(s, s));
    if (!bf_isfinite(mag) || !(mag > 0.0f)) return false;
    out = mk(s.x / mag, s.y / mag, s.z / mag);
    return fin3(out);
}

BF_D bool valid_surface_event(uint32_t bits) {
    const uint32_t family = bits & 0x7u;
    const uint32_t direction = bits & 0x18u;
    return (bits & ~0x3Fu) == 0u && (bits & BF_LOBE_VOLUME) == 0u && family != 0u && (family & (family - 1u)) == 0u &&
           direction != 0u && (direction & (direction - 1u)) == 0u;
}

#endif
