/*
 * ORACLE — test infrastructure only (see bf_oracle.h).
 * Scalar transport helpers and the closure library restated from
 *   Blackframe/XPU/CUDA/Include/Blackframe/XPU/CUDA/TransportDevice.cuh
 *   Blackframe/XPU/CUDA/Include/Blackframe/XPU/CUDA/TransportLobesDevice.cuh
 */
#ifndef BF_ORACLE_LOBES_H
#define BF_ORACLE_LOBES_H

#include "bf_oracle_internal.h"

enum { LOBE_OK = 0, LOBE_INVALID = 1, LOBE_NOT_REPRESENTABLE = 2, LOBE_OUTSIDE_SUPPORT = 3 };
enum { LOBE_MEASURE_DISCRETE = 0, LOBE_MEASURE_SOLID_ANGLE = 1 };

typedef struct {
    uint32_t kind, lobes;
    float weight[4];
    float parameters[10];
} closure_rec;

typedef struct {
    uint32_t active_count;
    closure_rec closures[BF_MAX_CLOSURES];
    float probabilities[BF_MAX_CLOSURES];
    float cdf[BF_MAX_CLOSURES + 1];
} lobe_mixture;

typedef struct {
    uint32_t source_indices[BF_MAX_CLOSURES];
    uint32_t allowed_directions[BF_MAX_CLOSURES];
    uint32_t source_count;
    uint32_t blocked_lobes;
    uint32_t status;
} lobe_filter_state;

typedef struct {
    float value[4];
    v3 incoming;
    float pdf;
    uint32_t measure;
    uint32_t lobes;
    float eta_scale;
    uint32_t status;
} lobe_sample;

typedef struct {
    float value[4];
    v3 incoming;
    float pdf;
    uint32_t measure;
    uint32_t selected;
    uint32_t lobes;
    float eta_scale;
    uint32_t status;
} lobe_mixture_sample;

/* TransportDevice.cuh:327-333 */
static inline int lobe_unit_vector(v3 d) {
    if (!v3_finite(d)) return 0;
    const float sq = fmaf(d.x, d.x, fmaf(d.y, d.y, d.z * d.z));
    return bf_isfinite(sq) && fabsf(sq - 1.0f) <= 128.0f * BFO_FLOAT_EPSILON;
}

/* TransportDevice.cuh:196-207 */
static inline int lobe_checked_product(float l, float r, float* out) {
    if (!bf_isfinite(l) || l < 0.0f || !bf_isfinite(r) || r < 0.0f) return 0;
    const float p = l * r;
    if (!bf_isfinite(p) || p < 0.0f || (l != 0.0f && r != 0.0f && p == 0.0f)) return 0;
    *out = p;
    return 1;
}

/* TransportDevice.cuh:209-259 */
static inline int lobe_checked_product_quotient(const float* num, int nn, const float* den, int nd, float* out) {
    int has_zero = 0;
    for (int i = 0; i < nd; ++i)
        if (!bf_isfinite(den[i]) || !(den[i] > 0.0f)) return LOBE_INVALID;
    for (int i = 0; i < nn; ++i) {
        if (!bf_isfinite(num[i]) || num[i] < 0.0f) return LOBE_INVALID;
        has_zero = has_zero || num[i] == 0.0f;
    }
    if (has_zero) {
        *out = 0.0f;
        return LOBE_OK;
    }
    float significand = 1.0f;
    int exponent = 0;
    for (int i = 0; i < nn; ++i) {
        int e;
        significand *= bf_frexpf(num[i], &e);
        exponent += e;
    }
    for (int i = 0; i < nd; ++i) {
        int e;
        significand /= bf_frexpf(den[i], &e);
        exponent -= e;
    }
    if (!bf_isfinite(significand) || !(significand > 0.0f)) return LOBE_NOT_REPRESENTABLE;
    int ne;
    significand = bf_frexpf(significand, &ne);
    exponent += ne;
    const float r = bf_ldexpf(significand, exponent);
    if (!bf_isfinite(r) || !(r > 0.0f)) return LOBE_NOT_REPRESENTABLE;
    *out = r;
    return LOBE_OK;
}

/* TransportDevice.cuh:552-590 */
static inline int lobe_joint_light_pdf(float selection, float conditional, float* out) {
    if (!bf_isfinite(selection) || !(selection > 0.0f) || selection > 1.0f || !bf_isfinite(conditional) ||
        conditional < 0.0f)
        return 0;
    if (conditional == 0.0f) {
        *out = 0.0f;
        return 1;
    }
    const float num[2] = {selection, conditional};
    const float den[1] = {1.0f};
    return lobe_checked_product_quotient(num, 2, den, 1, out) == LOBE_OK;
}

/* TransportDevice.cuh:592-628 */
static inline int lobe_ratio_at_most_one(float n, float d, int squared, float* out) {
    if (!bf_isfinite(n) || !(n > 0.0f) || !bf_isfinite(d) || !(d > 0.0f) || n > d) return 0;
    int ne, de;
    const float ns = bf_frexpf(n, &ne);
    const float ds = bf_frexpf(d, &de);
    const float ratio = ns / ds;
    const float r = squared ? bf_ldexpf(ratio * ratio, 2 * (ne - de)) : bf_ldexpf(ratio, ne - de);
    if (!bf_isfinite(r) || r < 0.0f || r > 1.0f) return 0;
    *out = r;
    return 1;
}

/* TransportDevice.cuh:630-689 (same measure on both densities) */
static inline int lobe_mis_weight(uint32_t heuristic, float sampled, float competing, float* out) {
    if (!bf_isfinite(sampled) || !(sampled > 0.0f) || !bf_isfinite(competing) || competing < 0.0f) return 0;
    if (heuristic > 1u) return 0;
    if (competing == 0.0f) {
        *out = 1.0f;
        return 1;
    }
    float ratio;
    if (sampled < competing) {
        if (!lobe_ratio_at_most_one(sampled, competing, heuristic == 1u, &ratio) || !(ratio > 0.0f)) return 0;
        const float w = ratio / (1.0f + ratio);
        if (!bf_isfinite(w) || !(w > 0.0f) || w > 1.0f) return 0;
        *out = w;
        return 1;
    }
    if (!lobe_ratio_at_most_one(competing, sampled, heuristic == 1u, &ratio)) return 0;
    const float w = 1.0f / (1.0f + ratio);
    if (!bf_isfinite(w) || !(w > 0.0f) || w > 1.0f) return 0;
    *out = w;
    return 1;
}

/* The decision with a given survival probability p: u < p survives, the throughput is divided by p
 * (RussianRoulette.hpp; a recorded p replays the decision with the texels moved). */
static inline int lobe_russian_roulette_at(float* throughput, float p, float u, int* survived) {
    if (!bf_isfinite(u) || u < 0.0f || !(u < 1.0f)) return 0;
    for (int k = 0; k < 4; ++k)
        if (!bf_isfinite(throughput[k]) || throughput[k] < 0.0f) return 0;
    if (!bf_isfinite(p) || !(p > 0.0f) || p > 1.0f) return 0;
    if (!(u < p)) {
        *survived = 0;
        return 1;
    }
    float out[4];
    for (int k = 0; k < 4; ++k) {
        out[k] = throughput[k] / p;
        if (!bf_isfinite(out[k]) || (throughput[k] != 0.0f && out[k] == 0.0f)) return 0;
    }
    memcpy(throughput, out, sizeof out);
    *survived = 1;
    return 1;
}

/* TransportDevice.cuh:711-782 (policy validated by the host): p = clamp(max(throughput) * eta, minimum, maximum), then the decision. */
static inline int lobe_russian_roulette(float* throughput, float eta_scale, float u, float minimum, float maximum,
                                        int* survived, float* probability) {
    if (!bf_isfinite(eta_scale) || !(eta_scale > 0.0f) || !bf_isfinite(u) || u < 0.0f || !(u < 1.0f)) return 0;
    float m = 0.0f;
    for (int k = 0; k < 4; ++k) {
        if (!bf_isfinite(throughput[k]) || throughput[k] < 0.0f) return 0;
        m = fmaxf(m, throughput[k]);
    }
    if (m == 0.0f) return 0;
    float p;
    if (m >= maximum / eta_scale) p = maximum;
    else if (m <= minimum / eta_scale) p = minimum;
    else {
        const float candidate = m * eta_scale;
        p = candidate < minimum ? minimum : (maximum < candidate ? maximum : candidate);
    }
    *probability = p;
    return lobe_russian_roulette_at(throughput, p, u, survived);
}

/* ---- Lambert (TransportDevice.cuh:335-478) ------------------------------------------------- */
static inline int lobe_reflectance(const float* w) {
    for (int k = 0; k < 4; ++k)
        if (!bf_isfinite(w[k]) || w[k] < 0.0f || w[k] > 1.0f) return 0;
    return 1;
}

static inline int lobe_lambert_eval(const float* reflectance, v3 wo, v3 wi, float* out) {
    if (!lobe_reflectance(reflectance) || !lobe_unit_vector(wo) || !lobe_unit_vector(wi)) return LOBE_INVALID;
    if (!(wo.z > 0.0f) || !(wi.z > 0.0f)) {
        out[0] = out[1] = out[2] = out[3] = 0.0f;
        return LOBE_OK;
    }
    for (int k = 0; k < 4; ++k) out[k] = reflectance[k] * BFO_INV_PI;
    return LOBE_OK;
}

static inline int lobe_lambert_pdf(v3 wo, v3 wi, float* pdf) {
    if (!lobe_unit_vector(wo) || !lobe_unit_vector(wi)) return LOBE_INVALID;
    if (!(wo.z > 0.0f) || !(wi.z > 0.0f)) {
        *pdf = 0.0f;
        return LOBE_OK;
    }
    float d;
    if (!lobe_checked_product(wi.z, BFO_INV_PI, &d) || !(d > 0.0f)) return LOBE_NOT_REPRESENTABLE;
    *pdf = d;
    return LOBE_OK;
}

/* TransportDevice.cuh:388-427 */
static inline int lobe_cosine_hemisphere(float u, float v, v3* out) {
    if (!bf_isfinite(u) || u < 0.0f || !(u < 1.0f) || !bf_isfinite(v) || v < 0.0f || !(v < 1.0f)) return LOBE_INVALID;
    const float ox = 2.0f * u - 1.0f;
    const float oy = 2.0f * v - 1.0f;
    if (ox == 0.0f && oy == 0.0f) {
        *out = v3_make(0.0f, 0.0f, 1.0f);
        return LOBE_OK;
    }
    float r, phi;
    if (fabsf(ox) > fabsf(oy)) {
        r = ox;
        phi = BFO_QUARTER_PI * (oy / ox);
    } else {
        r = oy;
        phi = 2.0f * BFO_QUARTER_PI - BFO_QUARTER_PI * (ox / oy);
    }
    const float radius = fabsf(r);
    float s, c;
    bf_sincosf(phi, &s, &c);
    *out = v3_make(r * c, r * s, sqrtf((1.0f - radius) * (1.0f + radius)));
    if (!lobe_unit_vector(*out)) return LOBE_NOT_REPRESENTABLE;
    return LOBE_OK;
}

static inline int lobe_valid_surface_event(uint32_t bits) {
    const uint32_t family = bits & 0x7u;
    const uint32_t direction = bits & 0x18u;
    return (bits & ~0x3Fu) == 0u && (bits & BF_LOBE_VOLUME) == 0u && family != 0u && (family & (family - 1u)) == 0u &&
           direction != 0u && (direction & (direction - 1u)) == 0u;
}

static inline void lobe_empty_sample(lobe_sample* s, uint32_t status) {
    memset(s, 0, sizeof *s);
    s->measure = LOBE_MEASURE_SOLID_ANGLE;
    s->eta_scale = 1.0f;
    s->status = status;
}

#include "bf_oracle_microfacet.h"

/* ---- specular delta closures (TransportLobesDevice.cuh:1689-1960) --------------------------- */
static inline int lobe_specular_zero(const closure_rec* c, v3 wo, v3 wi) {
    if (c->kind == BF_CLOSURE_SPECULAR_TRANSMISSION &&
        (!bf_isfinite(c->parameters[0]) || !(c->parameters[0] > 0.0f) || !bf_isfinite(c->parameters[1]) ||
         !(c->parameters[1] > 0.0f)))
        return LOBE_INVALID;
    return lobe_reflectance(c->weight) && lobe_unit_vector(wo) && lobe_unit_vector(wi) ? LOBE_OK : LOBE_INVALID;
}

/* checked_positive_product_quotient (:1756-1782) */
static inline int lobe_positive_product_quotient(float left, float right, float denominator, float* out) {
    *out = 0.0f;
    if (left == 0.0f) return LOBE_OK;
    if (!bf_isfinite(left) || !(left > 0.0f) || !bf_isfinite(right) || !(right > 0.0f) || !bf_isfinite(denominator) ||
        !(denominator > 0.0f))
        return LOBE_INVALID;
    int le, re, de, ne;
    const float ls = bf_frexpf(left, &le), rs = bf_frexpf(right, &re), ds = bf_frexpf(denominator, &de);
    const float ns = bf_frexpf((ls * rs) / ds, &ne);
    const float r = bf_ldexpf(ns, le + re - de + ne);
    if (!bf_isfinite(r) || !(r > 0.0f)) return LOBE_NOT_REPRESENTABLE;
    *out = r;
    return LOBE_OK;
}

/* scaled_delta_value (:1784-1797) */
static inline int lobe_delta_value(const float* coefficient, float factor, float cosine, float* out) {
    for (int k = 0; k < 4; ++k) {
        const int st = lobe_positive_product_quotient(coefficient[k], factor, cosine, &out[k]);
        if (st != LOBE_OK) return st;
    }
    return LOBE_OK;
}

/* checked_squared_ratio (:1799-1819) */
static inline int lobe_squared_ratio(float numerator, float denominator, float* out) {
    *out = 0.0f;
    if (!bf_isfinite(numerator) || !(numerator > 0.0f) || !bf_isfinite(denominator) || !(denominator > 0.0f))
        return LOBE_INVALID;
    int ne, de, re;
    const float ns = bf_frexpf(numerator, &ne), ds = bf_frexpf(denominator, &de);
    const float q = ns / ds;
    const float rs = bf_frexpf(q * q, &re);
    const float r = bf_ldexpf(rs, 2 * (ne - de) + re);
    if (!bf_isfinite(r) || !(r > 0.0f)) return LOBE_NOT_REPRESENTABLE;
    *out = r;
    return LOBE_OK;
}

static inline void lobe_delta_sample(lobe_sample* out, v3 wi, const float* value, uint32_t lobes, float eta_scale) {
    memcpy(out->value, value, 4 * sizeof(float));
    out->incoming = wi;
    out->pdf = 1.0f;
    out->measure = LOBE_MEASURE_DISCRETE;
    out->lobes = lobes;
    out->eta_scale = eta_scale;
    out->status = LOBE_OK;
}

/* sample_specular_reflection (:1821-1845) */
static inline void lobe_specular_reflection_sample(const float* w, v3 wo, lobe_sample* out) {
    lobe_empty_sample(out, LOBE_INVALID);
    if (!lobe_reflectance(w) || !lobe_unit_vector(wo)) return;
    if (wo.z == 0.0f) {
        out->status = LOBE_OUTSIDE_SUPPORT;
        return;
    }
    const v3 wi = v3_make(-wo.x, -wo.y, wo.z);
    float value[4];
    const int st = lobe_delta_value(w, 1.0f, fabsf(wi.z), value);
    if (st != LOBE_OK) {
        out->status = st;
        return;
    }
    lobe_delta_sample(out, wi, value, BF_LOBE_SPECULAR | BF_LOBE_REFLECTION, 1.0f);
}

/* sample_specular_transmission (:1847-1958), radiance transport mode */
static inline void lobe_specular_transmission_sample(const float* w, float exterior, float interior, v3 wo, lobe_sample* out) {
    lobe_empty_sample(out, LOBE_INVALID);
    if (!lobe_reflectance(w) || !bf_isfinite(exterior) || !(exterior > 0.0f) || !bf_isfinite(interior) ||
        !(interior > 0.0f) || !lobe_unit_vector(wo))
        return;
    if (wo.z == 0.0f) {
        out->status = LOBE_OUTSIDE_SUPPORT;
        return;
    }
    const int entering = wo.z > 0.0f;
    const float eta_i = entering ? exterior : interior, eta_t = entering ? interior : exterior;
    float value[4];
    if (eta_i == eta_t) {
        const v3 wi = v3_make(-wo.x, -wo.y, -wo.z);
        const int st = lobe_delta_value(w, 1.0f, fabsf(wi.z), value);
        if (st != LOBE_OK) {
            out->status = st;
            return;
        }
        lobe_delta_sample(out, wi, value, BF_LOBE_SPECULAR | BF_LOBE_TRANSMISSION, 1.0f);
        return;
    }
    out->status = LOBE_NOT_REPRESENTABLE;
    const float raw_sine = bf_hypotf(wo.x, wo.y);
    if (!bf_isfinite(raw_sine)) return;
    float sine = raw_sine, tangent_normalization = 1.0f;
    if (raw_sine > 1.0f) {
        const float cosine = fabsf(wo.z);
        if (cosine > 1.0f) return;
        sine = sqrtf((1.0f - cosine) * (1.0f + cosine));
        tangent_normalization = sine / raw_sine;
        if (!bf_isfinite(sine) || !bf_isfinite(tangent_normalization) || !(tangent_normalization > 0.0f)) return;
    }
    if (sine > 0.0f) {
        const float critical = eta_t / eta_i;
        if (critical <= 1.0f && sine >= critical) {
            out->status = LOBE_OUTSIDE_SUPPORT;
            return;
        }
    }
    const float ratio = eta_i / eta_t;
    if (!bf_isfinite(ratio) || !(ratio > 0.0f)) return;
    const float ix = -ratio * tangent_normalization * wo.x;
    const float iy = -ratio * tangent_normalization * wo.y;
    if (!bf_isfinite(ix) || !bf_isfinite(iy) || (wo.x != 0.0f && ix == 0.0f) || (wo.y != 0.0f && iy == 0.0f)) return;
    const float t_sine = bf_hypotf(ix, iy);
    if (!bf_isfinite(t_sine) || !(t_sine < 1.0f)) return;
    const float t_cosine = sqrtf((1.0f - t_sine) * (1.0f + t_sine));
    if (!bf_isfinite(t_cosine) || !(t_cosine > 0.0f)) return;
    const v3 wi = v3_make(ix, iy, entering ? -t_cosine : t_cosine);
    if (!lobe_unit_vector(wi)) return;
    float factor, eta_scale;
    int st = lobe_squared_ratio(eta_i, eta_t, &factor);
    const int st2 = lobe_squared_ratio(eta_t, eta_i, &eta_scale);
    if (st != LOBE_OK || st2 != LOBE_OK) {
        out->status = st != LOBE_OK ? st : st2;
        return;
    }
    st = lobe_delta_value(w, factor, t_cosine, value);
    if (st != LOBE_OK) {
        out->status = st;
        return;
    }
    lobe_delta_sample(out, wi, value, BF_LOBE_SPECULAR | BF_LOBE_TRANSMISSION, eta_scale);
}

/* ---- closure record dispatch (TransportLobesDevice.cuh:2092-2260); transport mode radiance --- */
static inline int lobe_record_eval(const closure_rec* c, uint32_t allowed, v3 wo, v3 wi, float* out) {
    if (!mf_valid_closure(c)) return LOBE_INVALID;
    const float* p = c->parameters;
    switch (c->kind) {
    case BF_CLOSURE_LAMBERTIAN:
        return lobe_lambert_eval(c->weight, wo, wi, out);
    case BF_CLOSURE_ROUGH_DIFFUSE:
        return mf_rough_diffuse_eval(c->weight, p[0], wo, wi, out);
    case BF_CLOSURE_ROUGH_CONDUCTOR:
        return mf_conductor_eval(c->weight, p, p + 4, p[8], p[9], wo, wi, out);
    case BF_CLOSURE_ROUGH_DIELECTRIC:
        return mf_dielectric_eval_masked(c->weight, p[0], p[1], p[2], p[3], wo, wi, BFO_TRANSPORT_RADIANCE, allowed, out);
    case BF_CLOSURE_SPECULAR_REFLECTION:
    case BF_CLOSURE_SPECULAR_TRANSMISSION:
        out[0] = out[1] = out[2] = out[3] = 0.0f;
        return lobe_specular_zero(c, wo, wi);
    default:
        return LOBE_INVALID;
    }
}

static inline int lobe_record_pdf(const closure_rec* c, uint32_t allowed, v3 wo, v3 wi, float* pdf) {
    if (!mf_valid_closure(c)) return LOBE_INVALID;
    const float* p = c->parameters;
    switch (c->kind) {
    case BF_CLOSURE_LAMBERTIAN:
    case BF_CLOSURE_ROUGH_DIFFUSE:
        return lobe_lambert_pdf(wo, wi, pdf);
    case BF_CLOSURE_ROUGH_CONDUCTOR:
        return mf_conductor_pdf(p[8], p[9], wo, wi, pdf);
    case BF_CLOSURE_ROUGH_DIELECTRIC:
        return mf_dielectric_pdf_masked(c->weight, p[0], p[1], p[2], p[3], wo, wi, BFO_TRANSPORT_RADIANCE, allowed, pdf);
    case BF_CLOSURE_SPECULAR_REFLECTION:
    case BF_CLOSURE_SPECULAR_TRANSMISSION:
        *pdf = 0.0f;
        return lobe_specular_zero(c, wo, wi);
    default:
        return LOBE_INVALID;
    }
}

/* sample_lambert / sample_rough_diffuse: cosine hemisphere, then eval and pdf at the sample */
static inline void lobe_diffuse_sample(const closure_rec* c, v3 wo, float u, float v, lobe_sample* out) {
    v3 wi;
    const int st = lobe_cosine_hemisphere(u, v, &wi);
    if (st != LOBE_OK) {
        out->status = st;
        return;
    }
    if (!(wo.z > 0.0f) || !(wi.z > 0.0f)) {
        out->incoming = wi;
        out->status = LOBE_OUTSIDE_SUPPORT;
        return;
    }
    float value[4], pdf;
    const int s1 = c->kind == BF_CLOSURE_LAMBERTIAN ? lobe_lambert_eval(c->weight, wo, wi, value)
                                                    : mf_rough_diffuse_eval(c->weight, c->parameters[0], wo, wi, value);
    const int s2 = lobe_lambert_pdf(wo, wi, &pdf);
    if (s1 != LOBE_OK || s2 != LOBE_OK) {
        out->status = s1 != LOBE_OK ? s1 : s2;
        return;
    }
    memcpy(out->value, value, sizeof value);
    out->incoming = wi;
    out->pdf = pdf;
    out->measure = LOBE_MEASURE_SOLID_ANGLE;
    out->lobes = BF_LOBE_DIFFUSE | BF_LOBE_REFLECTION;
    out->eta_scale = 1.0f;
    out->status = LOBE_OK;
}

static inline void lobe_record_sample(const closure_rec* c, uint32_t allowed, v3 wo, float event, float u, float v,
                                      lobe_sample* out) {
    lobe_empty_sample(out, LOBE_INVALID);
    if (!mf_valid_closure(c)) return;
    const float* p = c->parameters;
    switch (c->kind) {
    case BF_CLOSURE_LAMBERTIAN:
    case BF_CLOSURE_ROUGH_DIFFUSE:
        if (!lobe_unit_vector(wo)) return;
        lobe_diffuse_sample(c, wo, u, v, out);
        return;
    case BF_CLOSURE_ROUGH_CONDUCTOR:
        mf_conductor_sample(c->weight, p, p + 4, p[8], p[9], wo, u, v, out);
        return;
    case BF_CLOSURE_ROUGH_DIELECTRIC:
        mf_dielectric_sample_masked(c->weight, p[0], p[1], p[2], p[3], wo, event, u, v, BFO_TRANSPORT_RADIANCE, allowed, out);
        return;
    case BF_CLOSURE_SPECULAR_REFLECTION:
        lobe_specular_reflection_sample(c->weight, wo, out);
        return;
    case BF_CLOSURE_SPECULAR_TRANSMISSION:
        lobe_specular_transmission_sample(c->weight, p[0], p[1], wo, out);
        return;
    default:
        return;
    }
}

/* ---- mixtures (TransportLobesDevice.cuh:2262-2590, 2843-3080) ------------------------------ */
static inline int lobe_compensated_add(float value, float* sum, float* correction) {
    if (!bf_isfinite(value) || value < 0.0f) return 0;
    const float next = *sum + value;
    if (!bf_isfinite(next)) return 0;
    *correction += *sum >= value ? (*sum - next) + value : (value - next) + *sum;
    if (!bf_isfinite(*correction)) return 0;
    *sum = next;
    return 1;
}

static inline int lobe_mixture_has_nonblack_weight(const lobe_mixture* m) {
    for (uint32_t c = 0; c < m->active_count; ++c)
        for (int k = 0; k < 4; ++k)
            if (m->closures[c].weight[k] != 0.0f) return 1;
    return 0;
}

static inline int lobe_filtered_mixture_eval(const lobe_mixture* m, const lobe_filter_state* f, v3 wo, v3 wi, float* out) {
    if (!lobe_unit_vector(wo) || !lobe_unit_vector(wi)) return LOBE_INVALID;
    float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (uint32_t i = 0; i < m->active_count; ++i) {
        float comp[4];
        const int st = lobe_record_eval(&m->closures[i], f->allowed_directions[i], wo, wi, comp);
        if (st != LOBE_OK) return st;
        for (int k = 0; k < 4; ++k) {
            if (!bf_isfinite(comp[k]) || comp[k] < 0.0f) return LOBE_INVALID;
            acc[k] += comp[k];
            if (!bf_isfinite(acc[k]) || acc[k] < 0.0f) return LOBE_NOT_REPRESENTABLE;
        }
    }
    memcpy(out, acc, sizeof acc);
    return LOBE_OK;
}

static inline int lobe_filtered_mixture_pdf(const lobe_mixture* m, const lobe_filter_state* f, v3 wo, v3 wi, float* out) {
    if (!lobe_unit_vector(wo) || !lobe_unit_vector(wi)) return LOBE_INVALID;
    if (m->active_count == 0) {
        *out = 0.0f;
        return LOBE_OK;
    }
    float conditional[BF_MAX_CLOSURES];
    float maximum = 0.0f;
    for (uint32_t i = 0; i < m->active_count; ++i) {
        float p;
        const int st = lobe_record_pdf(&m->closures[i], f->allowed_directions[i], wo, wi, &p);
        if (st != LOBE_OK) return st;
        if (!bf_isfinite(p) || p < 0.0f) return LOBE_NOT_REPRESENTABLE;
        conditional[i] = p;
        maximum = fmaxf(maximum, p);
    }
    if (m->active_count == 1) {
        *out = conditional[0];
        return LOBE_OK;
    }
    if (maximum == 0.0f) {
        *out = 0.0f;
        return LOBE_OK;
    }
    float sum = 0.0f, correction = 0.0f;
    for (uint32_t i = 0; i < m->active_count; ++i) {
        const float ratio = conditional[i] / maximum;
        float contribution;
        if (!lobe_checked_product(m->probabilities[i], ratio, &contribution) ||
            !lobe_compensated_add(contribution, &sum, &correction))
            return LOBE_NOT_REPRESENTABLE;
    }
    const float normalized = sum + correction;
    if (!bf_isfinite(normalized) || !(normalized > 0.0f) || normalized > 1.0f) return LOBE_NOT_REPRESENTABLE;
    float mixed;
    if (!lobe_checked_product(normalized, maximum, &mixed)) return LOBE_NOT_REPRESENTABLE;
    *out = mixed;
    return LOBE_OK;
}

/* TransportLobesDevice.cuh:2618-2641 */
static inline uint32_t lobe_depth_blocked(const uint32_t* limits, const uint32_t* counters, uint32_t event) {
    uint32_t blocked = 0;
    if ((event & BF_LOBE_DIFFUSE) && counters[0] >= limits[0]) blocked |= BF_LOBE_DIFFUSE;
    if ((event & BF_LOBE_GLOSSY) && counters[1] >= limits[1]) blocked |= BF_LOBE_GLOSSY;
    if ((event & BF_LOBE_SPECULAR) && counters[2] >= limits[2]) blocked |= BF_LOBE_SPECULAR;
    if ((event & BF_LOBE_TRANSMISSION) && counters[3] >= limits[3]) blocked |= BF_LOBE_TRANSMISSION;
    return blocked;
}

/* TransportLobesDevice.cuh:2643-2718 */
static inline int lobe_build_filtered_distribution(lobe_mixture* m) {
    if (m->active_count == 0) return 1;
    if (m->active_count == 1) {
        m->probabilities[0] = 1.0f;
        m->cdf[0] = 0.0f;
        m->cdf[1] = 1.0f;
        return 1;
    }
    float total_sum = 0.0f, total_correction = 0.0f;
    for (uint32_t i = 0; i < m->active_count; ++i)
        if (!lobe_compensated_add(m->probabilities[i], &total_sum, &total_correction)) return 0;
    const float total = total_sum + total_correction;
    if (!bf_isfinite(total) || !(total > 0.0f)) return 0;
    const uint64_t scale = 1ull << 24;
    uint64_t masses[BF_MAX_CLOSURES] = {0};
    float remainders[BF_MAX_CLOSURES] = {0};
    uint64_t base_mass = 0;
    for (uint32_t i = 0; i < m->active_count; ++i) {
        const float normalized = m->probabilities[i] / total;
        if (!bf_isfinite(normalized) || !(normalized > 0.0f) || !(normalized < 1.0f)) return 0;
        const float scaled = bf_ldexpf(normalized, 24);
        const float base = floorf(scaled);
        if (!bf_isfinite(scaled) || !(scaled > 0.0f) || base < 0.0f || base > (float)scale) return 0;
        masses[i] = (uint64_t)base;
        remainders[i] = scaled - base;
        if (!bf_isfinite(remainders[i]) || remainders[i] < 0.0f || !(remainders[i] < 1.0f) || base_mass > scale - masses[i])
            return 0;
        base_mass += masses[i];
    }
    if (base_mass > scale) return 0;
    float rsum = 0.0f, rcorr = 0.0f;
    for (uint32_t i = 0; i < m->active_count; ++i)
        if (!lobe_compensated_add(remainders[i], &rsum, &rcorr)) return 0;
    const uint64_t missing = scale - base_mass;
    const float rtotal = rsum + rcorr;
    if (missing > m->active_count || fabsf(rtotal - (float)missing) > 0.5f + 0x1p-23f * 16.0f) return 0;
    int received[BF_MAX_CLOSURES] = {0};
    for (uint64_t unit = 0; unit < missing; ++unit) {
        uint32_t selected = m->active_count;
        for (uint32_t i = 0; i < m->active_count; ++i) {
            if (received[i] || !(remainders[i] > 0.0f)) continue;
            if (selected == m->active_count || remainders[i] > remainders[selected]) selected = i;
        }
        if (selected == m->active_count) return 0;
        ++masses[selected];
        received[selected] = 1;
    }
    uint64_t cumulative = 0;
    m->cdf[0] = 0.0f;
    for (uint32_t i = 0; i < m->active_count; ++i) {
        if (masses[i] == 0 || cumulative > scale - masses[i]) return 0;
        cumulative += masses[i];
        m->probabilities[i] = bf_ldexpf((float)masses[i], -24);
        m->cdf[i + 1] = bf_ldexpf((float)cumulative, -24);
    }
    return cumulative == scale && m->cdf[m->active_count] == 1.0f;
}

/* TransportLobesDevice.cuh:2720-2795 */
static inline void lobe_filter_mixture_by_depth(lobe_mixture* m, const uint32_t* limits, const uint32_t* counters,
                                                uint32_t expected_total, lobe_filter_state* r) {
    memset(r, 0, sizeof *r);
    r->source_count = m->active_count;
    const uint64_t total = (uint64_t)counters[0] + counters[1] + counters[2] + counters[4];
    const uint64_t surface = (uint64_t)counters[0] + counters[1] + counters[2];
    if (total != expected_total || counters[3] > surface || counters[0] > limits[0] || counters[1] > limits[1] ||
        counters[2] > limits[2] || counters[3] > limits[3] || counters[4] > limits[4]) {
        r->status = LOBE_INVALID;
        return;
    }
    uint32_t filtered = 0;
    for (uint32_t si = 0; si < r->source_count; ++si) {
        const closure_rec src = m->closures[si];
        const float prob = m->probabilities[si];
        uint32_t allowed = 0;
        if (src.kind == BF_CLOSURE_ROUGH_DIELECTRIC) {
            const uint32_t refl = lobe_depth_blocked(limits, counters, BF_LOBE_GLOSSY | BF_LOBE_REFLECTION);
            const uint32_t trans = lobe_depth_blocked(limits, counters, BF_LOBE_GLOSSY | BF_LOBE_TRANSMISSION);
            if (refl == 0) allowed |= BF_LOBE_REFLECTION;
            else r->blocked_lobes |= refl;
            if (trans == 0) allowed |= BF_LOBE_TRANSMISSION;
            else r->blocked_lobes |= trans;
        } else {
            if (!lobe_valid_surface_event(src.lobes)) {
                r->status = LOBE_INVALID;
                return;
            }
            const uint32_t blocked = lobe_depth_blocked(limits, counters, src.lobes);
            if (blocked == 0) allowed = src.lobes & (BF_LOBE_REFLECTION | BF_LOBE_TRANSMISSION);
            else r->blocked_lobes |= blocked;
        }
        if (allowed == 0) continue;
        m->closures[filtered] = src;
        m->probabilities[filtered] = prob;
        r->source_indices[filtered] = si;
        r->allowed_directions[filtered] = allowed;
        ++filtered;
    }
    m->active_count = filtered;
    if (filtered == r->source_count) return;
    for (uint32_t i = filtered; i < BF_MAX_CLOSURES; ++i) {
        memset(&m->closures[i], 0, sizeof m->closures[i]);
        m->probabilities[i] = 0.0f;
    }
    for (uint32_t i = 0; i <= BF_MAX_CLOSURES; ++i) m->cdf[i] = 0.0f;
    if (!lobe_build_filtered_distribution(m)) r->status = LOBE_NOT_REPRESENTABLE;
}

/* TransportLobesDevice.cuh:2979-3062 */
static inline void lobe_sample_filtered_mixture(const lobe_mixture* m, const lobe_filter_state* f, v3 wo, float component,
                                                float u, float v, lobe_mixture_sample* out) {
    memset(out, 0, sizeof *out);
    out->eta_scale = 1.0f;
    out->measure = LOBE_MEASURE_SOLID_ANGLE;
    if (!lobe_unit_vector(wo) || !bf_isfinite(component) || component < 0.0f || !(component < 1.0f) || !bf_isfinite(u) ||
        u < 0.0f || !(u < 1.0f) || !bf_isfinite(v) || v < 0.0f || !(v < 1.0f)) {
        out->status = LOBE_INVALID;
        return;
    }
    if (m->active_count == 0) {
        out->status = LOBE_OUTSIDE_SUPPORT;
        return;
    }
    uint32_t selected = m->active_count;
    for (uint32_t i = 0; i < m->active_count; ++i) {
        if (component < m->cdf[i + 1]) {
            selected = i;
            break;
        }
    }
    if (selected == m->active_count) {
        out->status = LOBE_NOT_REPRESENTABLE;
        return;
    }
    const float event = (component - m->cdf[selected]) / m->probabilities[selected];
    if (!bf_isfinite(event) || event < 0.0f || !(event < 1.0f)) {
        out->status = LOBE_NOT_REPRESENTABLE;
        return;
    }
    lobe_sample s;
    lobe_record_sample(&m->closures[selected], f->allowed_directions[selected], wo, event, u, v, &s);
    if (s.status != LOBE_OK) {
        out->status = s.status;
        return;
    }
    if (m->active_count == 1) {
        memcpy(out->value, s.value, sizeof s.value);
        out->incoming = s.incoming;
        out->pdf = s.pdf;
        out->measure = s.measure;
        out->selected = f->source_indices[0];
        out->lobes = s.lobes;
        out->eta_scale = s.eta_scale;
        out->status = LOBE_OK;
        return;
    }
    if (s.measure == LOBE_MEASURE_SOLID_ANGLE) {
        float value[4], pdf;
        const int st1 = lobe_filtered_mixture_eval(m, f, wo, s.incoming, value);
        const int st2 = lobe_filtered_mixture_pdf(m, f, wo, s.incoming, &pdf);
        if (st1 != LOBE_OK || st2 != LOBE_OK) {
            out->status = st1 != LOBE_OK ? st1 : st2;
            return;
        }
        memcpy(out->value, value, sizeof value);
        out->pdf = pdf;
        out->measure = LOBE_MEASURE_SOLID_ANGLE;
    } else {
        /* aggregate_delta_atom (:2394-2480) over the filtered mixture */
        float sums[4] = {0.0f, 0.0f, 0.0f, 0.0f}, corrections[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        float psum = 0.0f, pcorrection = 0.0f;
        uint32_t matching = 0;
        for (uint32_t i = 0; i < m->active_count; ++i) {
            const uint32_t kind = m->closures[i].kind;
            if (kind != BF_CLOSURE_SPECULAR_REFLECTION && kind != BF_CLOSURE_SPECULAR_TRANSMISSION) continue;
            lobe_sample c;
            lobe_record_sample(&m->closures[i], BF_LOBE_REFLECTION | BF_LOBE_TRANSMISSION, wo, 0.0f, 0.0f, 0.0f, &c);
            if (c.status == LOBE_OUTSIDE_SUPPORT) continue;
            if (c.status != LOBE_OK) {
                out->status = c.status;
                return;
            }
            if (c.incoming.x != s.incoming.x || c.incoming.y != s.incoming.y || c.incoming.z != s.incoming.z) continue;
            float weighted;
            if (c.measure != LOBE_MEASURE_DISCRETE || !bf_isfinite(c.pdf) || !(c.pdf > 0.0f) || c.pdf > 1.0f) {
                out->status = LOBE_NOT_REPRESENTABLE;
                return;
            }
            for (int k = 0; k < 4; ++k)
                if (!lobe_compensated_add(c.value[k], &sums[k], &corrections[k])) {
                    out->status = LOBE_NOT_REPRESENTABLE;
                    return;
                }
            if (!lobe_checked_product(m->probabilities[i], c.pdf, &weighted) ||
                !lobe_compensated_add(weighted, &psum, &pcorrection)) {
                out->status = LOBE_NOT_REPRESENTABLE;
                return;
            }
            ++matching;
        }
        const float probability = psum + pcorrection;
        if (matching == 0 || !bf_isfinite(probability) || !(probability > 0.0f) || probability > 1.0f) {
            out->status = LOBE_NOT_REPRESENTABLE;
            return;
        }
        for (int k = 0; k < 4; ++k) {
            out->value[k] = sums[k] + corrections[k];
            if (!bf_isfinite(out->value[k]) || out->value[k] < 0.0f) {
                out->status = LOBE_NOT_REPRESENTABLE;
                return;
            }
        }
        out->pdf = probability;
        out->measure = LOBE_MEASURE_DISCRETE;
    }
    out->incoming = s.incoming;
    out->selected = f->source_indices[selected];
    out->lobes = s.lobes;
    out->eta_scale = s.eta_scale;
    out->status = LOBE_OK;
}

#endif
