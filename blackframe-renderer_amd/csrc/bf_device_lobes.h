/*
 * Closure library on the device: records resolved at the path's wavelengths, depth filtering,
 * mixture sampling / evaluation / density. Arithmetic follows
 * Blackframe/XPU/CUDA/Include/Blackframe/XPU/CUDA/TransportLobesDevice.cuh and TransportDevice.cuh.
 */
#ifndef BF_DEVICE_LOBES_H
#define BF_DEVICE_LOBES_H

#include "bf_device_math.h"

struct closure_t {
    uint32_t kind, lobes;
    float weight[4];
    float parameters[10];
};

struct mixture_t {
    uint32_t active_count;
    closure_t closures[BF_MAX_CLOSURES];
    float probabilities[BF_MAX_CLOSURES];
    float cdf[BF_MAX_CLOSURES + 1];
};

struct filter_t {
    uint32_t source_indices[BF_MAX_CLOSURES];
    uint32_t allowed[BF_MAX_CLOSURES];
    uint32_t source_count;
    uint32_t blocked;
    int status;
};

struct lobe_sample_t {
    float value[4];
    v3f incoming;
    float pdf;
    uint32_t discrete;
    uint32_t lobes;
    float eta_scale;
    int status;
};

/* ---- Lambert (TransportDevice.cuh:335-478) ------------------------------------------------- */
BF_D bool reflectance_ok(const float* w) {
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (!bf_isfinite(w[k]) || w[k] < 0.0f || w[k] > 1.0f) return false;
    return true;
}

BF_D int lambert_eval(const float* r, v3f wo, v3f wi, float* out) {
    if (!reflectance_ok(r) || !unit_vector(wo) || !unit_vector(wi)) return BF_L_INVALID;
    if (!(wo.z > 0.0f) || !(wi.z > 0.0f)) {
        out[0] = out[1] = out[2] = out[3] = 0.0f;
        return BF_L_OK;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) out[k] = r[k] * BF_INV_PI;
    return BF_L_OK;
}

BF_D int lambert_pdf(v3f wo, v3f wi, float& pdf) {
    if (!unit_vector(wo) || !unit_vector(wi)) return BF_L_INVALID;
    if (!(wo.z > 0.0f) || !(wi.z > 0.0f)) {
        pdf = 0.0f;
        return BF_L_OK;
    }
    float d;
    if (!checked_product(wi.z, BF_INV_PI, d) || !(d > 0.0f)) return BF_L_NOT_REPRESENTABLE;
    pdf = d;
    return BF_L_OK;
}

/* TransportDevice.cuh:388-427 */
BF_D int cosine_hemisphere(float u, float v, v3f& out) {
    if (!bf_isfinite(u) || u < 0.0f || !(u < 1.0f) || !bf_isfinite(v) || v < 0.0f || !(v < 1.0f)) return BF_L_INVALID;
    const float ox = 2.0f * u - 1.0f;
    const float oy = 2.0f * v - 1.0f;
    if (ox == 0.0f && oy == 0.0f) {
        out = mk(0.0f, 0.0f, 1.0f);
        return BF_L_OK;
    }
    float r, phi;
    if (fabsf(ox) > fabsf(oy)) {
        r = ox;
        phi = BF_QUARTER_PI * (oy / ox);
    } else {
        r = oy;
        phi = 2.0f * BF_QUARTER_PI - BF_QUARTER_PI * (ox / oy);
    }
    const float radius = fabsf(r);
    float s, c;
    bf_sincosf(phi, &s, &c);
    out = mk(r * c, r * s, sqrtf((1.0f - radius) * (1.0f + radius)));
    return unit_vector(out) ? BF_L_OK : BF_L_NOT_REPRESENTABLE;
}

BF_D void empty_sample(lobe_sample_t& s, int status) {
    s.value[0] = s.value[1] = s.value[2] = s.value[3] = 0.0f;
    s.incoming = mk(0.0f, 0.0f, 0.0f);
    s.pdf = 0.0f;
    s.discrete = 0u;
    s.lobes = 0u;
    s.eta_scale = 1.0f;
    s.status = status;
}

#include "bf_device_microfacet.h"

/* ---- specular delta closures (TransportLobesDevice.cuh:1689-1960) ---------------------------
 * Their value and density toward any given direction are zero; sampling returns the one mirror
 * or refracted direction with a discrete probability of one. */
BF_D int specular_zero(const closure_t& c, v3f wo, v3f wi) {
    if (c.kind == BF_CLOSURE_SPECULAR_TRANSMISSION &&
        (!bf_isfinite(c.parameters[0]) || !(c.parameters[0] > 0.0f) || !bf_isfinite(c.parameters[1]) ||
         !(c.parameters[1] > 0.0f)))
        return BF_L_INVALID;
    return reflectance_ok(c.weight) && unit_vector(wo) && unit_vector(wi) ? BF_L_OK : BF_L_INVALID;
}

/* checked_positive_product_quotient (:1756-1782): left * right / denominator on significands */
BF_D int positive_product_quotient(float left, float right, float denominator, float& out) {
    out = 0.0f;
    if (left == 0.0f) return BF_L_OK;
    if (!bf_isfinite(left) || !(left > 0.0f) || !bf_isfinite(right) || !(right > 0.0f) || !bf_isfinite(denominator) ||
        !(denominator > 0.0f))
        return BF_L_INVALID;
    int le, re, de, ne;
    const float ls = bf_frexpf(left, &le), rs = bf_frexpf(right, &re), ds = bf_frexpf(denominator, &de);
    const float ns = bf_frexpf((ls * rs) / ds, &ne);
    const float r = bf_ldexpf(ns, le + re - de + ne);
    if (!bf_isfinite(r) || !(r > 0.0f)) return BF_L_NOT_REPRESENTABLE;
    out = r;
    return BF_L_OK;
}

/* scaled_delta_value (:1784-1797) */
BF_D int delta_value(const float* coefficient, float factor, float cosine, float* out) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int st = positive_product_quotient(coefficient[k], factor, cosine, out[k]);
        if (st != BF_L_OK) return st;
    }
    return BF_L_OK;
}

/* checked_squared_ratio (:1799-1819) */
BF_D int squared_ratio(float numerator, float denominator, float& out) {
    out = 0.0f;
    if (!bf_isfinite(numerator) || !(numerator > 0.0f) || !bf_isfinite(denominator) || !(denominator > 0.0f))
        return BF_L_INVALID;
    int ne, de, re;
    const float ns = bf_frexpf(numerator, &ne), ds = bf_frexpf(denominator, &de);
    const float q = ns / ds;
    const float rs = bf_frexpf(q * q, &re);
    const float r = bf_ldexpf(rs, 2 * (ne - de) + re);
    if (!bf_isfinite(r) || !(r > 0.0f)) return BF_L_NOT_REPRESENTABLE;
    out = r;
    return BF_L_OK;
}

BF_D void delta_sample(lobe_sample_t& out, v3f wi, const float* value, uint32_t lobes, float eta_scale) {
    copy4(out.value, value);
    out.incoming = wi;
    out.pdf = 1.0f;
    out.discrete = 1u;
    out.lobes = lobes;
    out.eta_scale = eta_scale;
    out.status = BF_L_OK;
}

/* sample_specular_reflection (:1821-1845) */
BF_D void specular_reflection_sample(const float* w, v3f wo, lobe_sample_t& out) {
    empty_sample(out, BF_L_INVALID);
    if (!reflectance_ok(w) || !unit_vector(wo)) return;
    if (wo.z == 0.0f) {
        out.status = BF_L_OUTSIDE;
        return;
    }
    const v3f wi = mk(-wo.x, -wo.y, wo.z);
    float value[4];
    const int st = delta_value(w, 1.0f, fabsf(wi.z), value);
    if (st != BF_L_OK) {
        out.status = st;
        return;
    }
    delta_sample(out, wi, value, BF_LOBE_SPECULAR | BF_LOBE_REFLECTION, 1.0f);
}

/* sample_specular_transmission (:1847-1958), radiance transport mode */
BF_D void specular_transmission_sample(const float* w, float exterior, float interior, v3f wo, lobe_sample_t& out) {
    empty_sample(out, BF_L_INVALID);
    if (!reflectance_ok(w) || !bf_isfinite(exterior) || !(exterior > 0.0f) || !bf_isfinite(interior) || !(interior > 0.0f) ||
        !unit_vector(wo))
        return;
    if (wo.z == 0.0f) {
        out.status = BF_L_OUTSIDE;
        return;
    }
    const bool entering = wo.z > 0.0f;
    const float eta_i = entering ? exterior : interior, eta_t = entering ? interior : exterior;
    float value[4];
    if (eta_i == eta_t) {
        const v3f wi = mk(-wo.x, -wo.y, -wo.z);
        const int st = delta_value(w, 1.0f, fabsf(wi.z), value);
        if (st != BF_L_OK) {
            out.status = st;
            return;
        }
        delta_sample(out, wi, value, BF_LOBE_SPECULAR | BF_LOBE_TRANSMISSION, 1.0f);
        return;
    }
    out.status = BF_L_NOT_REPRESENTABLE;
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
            out.status = BF_L_OUTSIDE;
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
    const v3f wi = mk(ix, iy, entering ? -t_cosine : t_cosine);
    if (!unit_vector(wi)) return;
    float factor, eta_scale;
    int st = squared_ratio(eta_i, eta_t, factor);
    const int st2 = squared_ratio(eta_t, eta_i, eta_scale);
    if (st != BF_L_OK || st2 != BF_L_OK) {
        out.status = st != BF_L_OK ? st : st2;
        return;
    }
    st = delta_value(w, factor, t_cosine, value);
    if (st != BF_L_OK) {
        out.status = st;
        return;
    }
    delta_sample(out, wi, value, BF_LOBE_SPECULAR | BF_LOBE_TRANSMISSION, eta_scale);
}

/* ---- record dispatch (TransportLobesDevice.cuh:2092-2260); transport mode radiance ---------- */
BF_D int record_eval(const closure_t& c, uint32_t allowed, v3f wo, v3f wi, float* out) {
    if (!mf_valid_closure(&c)) return BF_L_INVALID;
    const float* p = c.parameters;
    switch (c.kind) {
    case BF_CLOSURE_LAMBERTIAN:
        return lambert_eval(c.weight, wo, wi, out);
    case BF_CLOSURE_ROUGH_DIFFUSE:
        return mf_rough_diffuse_eval(c.weight, p[0], wo, wi, out);
    case BF_CLOSURE_ROUGH_CONDUCTOR:
        return mf_conductor_eval(c.weight, p, p + 4, p[8], p[9], wo, wi, out);
    case BF_CLOSURE_ROUGH_DIELECTRIC:
        return mf_dielectric_eval_masked(c.weight, p[0], p[1], p[2], p[3], wo, wi, BF_TRANSPORT_RADIANCE, allowed, out);
    case BF_CLOSURE_SPECULAR_REFLECTION:
    case BF_CLOSURE_SPECULAR_TRANSMISSION:
        out[0] = out[1] = out[2] = out[3] = 0.0f;
        return specular_zero(c, wo, wi);
    default:
        return BF_L_INVALID;
    }
}

BF_D int record_pdf(const closure_t& c, uint32_t allowed, v3f wo, v3f wi, float& pdf) {
    if (!mf_valid_closure(&c)) return BF_L_INVALID;
    const float* p = c.parameters;
    switch (c.kind) {
    case BF_CLOSURE_LAMBERTIAN:
    case BF_CLOSURE_ROUGH_DIFFUSE:
        return lambert_pdf(wo, wi, pdf);
    case BF_CLOSURE_ROUGH_CONDUCTOR:
        return mf_conductor_pdf(p[8], p[9], wo, wi, &pdf);
    case BF_CLOSURE_ROUGH_DIELECTRIC:
        return mf_dielectric_pdf_masked(c.weight, p[0], p[1], p[2], p[3], wo, wi, BF_TRANSPORT_RADIANCE, allowed, &pdf);
    case BF_CLOSURE_SPECULAR_REFLECTION:
    case BF_CLOSURE_SPECULAR_TRANSMISSION:
        pdf = 0.0f;
        return specular_zero(c, wo, wi);
    default:
        return BF_L_INVALID;
    }
}

/* sample_lambert / sample_rough_diffuse: cosine hemisphere, then eval and pdf at the sample */
BF_D void diffuse_sample(const closure_t& c, v3f wo, float u, float v, lobe_sample_t& out) {
    v3f wi;
    const int st = cosine_hemisphere(u, v, wi);
    if (st != BF_L_OK) {
        out.status = st;
        return;
    }
    if (!(wo.z > 0.0f) || !(wi.z > 0.0f)) {
        out.incoming = wi;
        out.status = BF_L_OUTSIDE;
        return;
    }
    float value[4], pdf = 0.0f;
    const int s1 = c.kind == BF_CLOSURE_LAMBERTIAN ? lambert_eval(c.weight, wo, wi, value)
                                                   : mf_rough_diffuse_eval(c.weight, c.parameters[0], wo, wi, value);
    const int s2 = lambert_pdf(wo, wi, pdf);
    if (s1 != BF_L_OK || s2 != BF_L_OK) {
        out.status = s1 != BF_L_OK ? s1 : s2;
        return;
    }
    copy4(out.value, value);
    out.incoming = wi;
    out.pdf = pdf;
    out.discrete = 0u;
    out.lobes = BF_LOBE_DIFFUSE | BF_LOBE_REFLECTION;
    out.eta_scale = 1.0f;
    out.status = BF_L_OK;
}

BF_D void record_sample(const closure_t& c, uint32_t allowed, v3f wo, float event, float u, float v, lobe_sample_t& out) {
    empty_sample(out, BF_L_INVALID);
    if (!mf_valid_closure(&c)) return;
    const float* p = c.parameters;
    switch (c.kind) {
    case BF_CLOSURE_LAMBERTIAN:
    case BF_CLOSURE_ROUGH_DIFFUSE:
        if (!unit_vector(wo)) return;
        diffuse_sample(c, wo, u, v, out);
        return;
    case BF_CLOSURE_ROUGH_CONDUCTOR:
        mf_conductor_sample(c.weight, p, p + 4, p[8], p[9], wo, u, v, &out);
        return;
    case BF_CLOSURE_ROUGH_DIELECTRIC:
        mf_dielectric_sample_masked(c.weight, p[0], p[1], p[2], p[3], wo, event, u, v, BF_TRANSPORT_RADIANCE, allowed,
                                    &out);
        return;
    case BF_CLOSURE_SPECULAR_REFLECTION:
        specular_reflection_sample(c.weight, wo, out);
        return;
    case BF_CLOSURE_SPECULAR_TRANSMISSION:
        specular_transmission_sample(c.weight, p[0], p[1], wo, out);
        return;
    default:
        return;
    }
}

/* ---- mixtures ------------------------------------------------------------------------------- */
BF_D bool compensated_add(float value, float& sum, float& correction) {
    if (!bf_isfinite(value) || value < 0.0f) return false;
    const float next = sum + value;
    if (!bf_isfinite(next)) return false;
    correction += sum >= value ? (sum - next) + value : (value - next) + sum;
    if (!bf_isfinite(correction)) return false;
    sum = next;
    return true;
}

BF_D bool mixture_nonblack(const mixture_t& m) {
    for (uint32_t c = 0; c < m.active_count; ++c)
        for (int k = 0; k < 4; ++k)
            if (m.closures[c].weight[k] != 0.0f) return true;
    return false;
}

/* TransportLobesDevice.cuh:2863-2887 */
BF_D int mixture_eval(const mixture_t& m, const filter_t& f, v3f wo, v3f wi, float* out) {
    if (!unit_vector(wo) || !unit_vector(wi)) return BF_L_INVALID;
    float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (uint32_t i = 0; i < m.active_count; ++i) {
        float comp[4];
        const int st = record_eval(m.closures[i], f.allowed[i], wo, wi, comp);
        if (st != BF_L_OK) return st;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (!bf_isfinite(comp[k]) || comp[k] < 0.0f) return BF_L_INVALID;
            acc[k] += comp[k];
            if (!bf_isfinite(acc[k]) || acc[k] < 0.0f) return BF_L_NOT_REPRESENTABLE;
        }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) out[k] = acc[k];
    return BF_L_OK;
}

/* TransportLobesDevice.cuh:2889-2958 */
BF_D int mixture_pdf(const mixture_t& m, const filter_t& f, v3f wo, v3f wi, float& out) {
    if (!unit_vector(wo) || !unit_vector(wi)) return BF_L_INVALID;
    if (m.active_count == 0) {
        out = 0.0f;
        return BF_L_OK;
    }
    float conditional[BF_MAX_CLOSURES];
    float maximum = 0.0f;
    for (uint32_t i = 0; i < m.active_count; ++i) {
        float p;
        const int st = record_pdf(m.closures[i], f.allowed[i], wo, wi, p);
        if (st != BF_L_OK) return st;
        if (!bf_isfinite(p) || p < 0.0f) return BF_L_NOT_REPRESENTABLE;
        conditional[i] = p;
        maximum = fmaxf(maximum, p);
    }
    if (m.active_count == 1) {
        out = conditional[0];
        return BF_L_OK;
    }
    if (maximum == 0.0f) {
        out = 0.0f;
        return BF_L_OK;
    }
    float sum = 0.0f, correction = 0.0f;
    for (uint32_t i = 0; i < m.active_count; ++i) {
        const float ratio = conditional[i] / maximum;
        float contribution;
        if (!checked_product(m.probabilities[i], ratio, contribution) || !compensated_add(contribution, sum, correction))
            return BF_L_NOT_REPRESENTABLE;
    }
    const float normalized = sum + correction;
    if (!bf_isfinite(normalized) || !(normalized > 0.0f) || normalized > 1.0f) return BF_L_NOT_REPRESENTABLE;
    float mixed;
    if (!checked_product(normalized, maximum, mixed)) return BF_L_NOT_REPRESENTABLE;
    out = mixed;
    return BF_L_OK;
}

/* TransportLobesDevice.cuh:2618-2641 */
BF_D uint32_t depth_blocked(const uint32_t* limits, const uint32_t* counters, uint32_t event) {
    uint32_t b = 0;
    if ((event & BF_LOBE_DIFFUSE) && counters[0] >= limits[0]) b |= BF_LOBE_DIFFUSE;
    if ((event & BF_LOBE_GLOSSY) && counters[1] >= limits[1]) b |= BF_LOBE_GLOSSY;
    if ((event & BF_LOBE_SPECULAR) && counters[2] >= limits[2]) b |= BF_LOBE_SPECULAR;
    if ((event & BF_LOBE_TRANSMISSION) && counters[3] >= limits[3]) b |= BF_LOBE_TRANSMISSION;
    return b;
}

/* TransportLobesDevice.cuh:2643-2718 */
BF_D bool build_filtered_distribution(mixture_t& m) {
    if (m.active_count == 0) return true;
    if (m.active_count == 1) {
        m.probabilities[0] = 1.0f;
        m.cdf[0] = 0.0f;
        m.cdf[1] = 1.0f;
        return true;
    }
    float ts = 0.0f, tc = 0.0f;
    for (uint32_t i = 0; i < m.active_count; ++i)
        if (!compensated_add(m.probabilities[i], ts, tc)) return false;
    const float total = ts + tc;
    if (!bf_isfinite(total) || !(total > 0.0f)) return false;
    const uint64_t scale = 1ull << 24;
    uint64_t masses[BF_MAX_CLOSURES];
    float rem[BF_MAX_CLOSURES];
    uint64_t base_mass = 0;
    for (uint32_t i = 0; i < m.active_count; ++i) {
        const float normalized = m.probabilities[i] / total;
        if (!bf_isfinite(normalized) || !(normalized > 0.0f) || !(normalized < 1.0f)) return false;
        const float scaled = bf_ldexpf(normalized, 24);
        const float base = floorf(scaled);
        if (!bf_isfinite(scaled) || !(scaled > 0.0f) || base < 0.0f || base > (float)scale) return false;
        masses[i] = (uint64_t)base;
        rem[i] = scaled - base;
        if (!bf_isfinite(rem[i]) || rem[i] < 0.0f || !(rem[i] < 1.0f) || base_mass > scale - masses[i]) return false;
        base_mass += masses[i];
    }
    if (base_mass > scale) return false;
    float rs = 0.0f, rc = 0.0f;
    for (uint32_t i = 0; i < m.active_count; ++i)
        if (!compensated_add(rem[i], rs, rc)) return false;
    const uint64_t missing = scale - base_mass;
    const float rtotal = rs + rc;
    if (missing > m.active_count || fabsf(rtotal - (float)missing) > 0.5f + 0x1p-23f * 16.0f) return false;
    bool received[BF_MAX_CLOSURES];
    for (uint32_t i = 0; i < BF_MAX_CLOSURES; ++i) received[i] = false;
    for (uint64_t unit = 0; unit < missing; ++unit) {
        uint32_t sel = m.active_count;
        for (uint32_t i = 0; i < m.active_count; ++i) {
            if (received[i] || !(rem[i] > 0.0f)) continue;
            if (sel == m.active_count || rem[i] > rem[sel]) sel = i;
        }
        if (sel == m.active_count) return false;
        ++masses[sel];
        received[sel] = true;
    }
    uint64_t cumulative = 0;
    m.cdf[0] = 0.0f;
    for (uint32_t i = 0; i < m.active_count; ++i) {
        if (masses[i] == 0 || cumulative > scale - masses[i]) return false;
        cumulative += masses[i];
        m.probabilities[i] = bf_ldexpf((float)masses[i], -24);
        m.cdf[i + 1] = bf_ldexpf((float)cumulative, -24);
    }
    return cumulative == scale && m.cdf[m.active_count] == 1.0f;
}

/* TransportLobesDevice.cuh:2720-2795 */
BF_D void filter_by_depth(mixture_t& m, const uint32_t* limits, const uint32_t* counters, uint32_t expected_total,
                          filter_t& r) {
    r.source_count = m.active_count;
    r.blocked = 0;
    r.status = BF_L_OK;
    const uint64_t total = (uint64_t)counters[0] + counters[1] + counters[2] + counters[4];
    const uint64_t surface = (uint64_t)counters[0] + counters[1] + counters[2];
    if (total != expected_total || counters[3] > surface || counters[0] > limits[0] || counters[1] > limits[1] ||
        counters[2] > limits[2] || counters[3] > limits[3] || counters[4] > limits[4]) {
        r.status = BF_L_INVALID;
        return;
    }
    uint32_t filtered = 0;
    for (uint32_t si = 0; si < r.source_count; ++si) {
        const closure_t src = m.closures[si];
        const float prob = m.probabilities[si];
        uint32_t allowed = 0;
        if (src.kind == BF_CLOSURE_ROUGH_DIELECTRIC) {
            const uint32_t refl = depth_blocked(limits, counters, BF_LOBE_GLOSSY | BF_LOBE_REFLECTION);
            const uint32_t trans = depth_blocked(limits, counters, BF_LOBE_GLOSSY | BF_LOBE_TRANSMISSION);
            if (refl == 0) allowed |= BF_LOBE_REFLECTION;
            else r.blocked |= refl;
            if (trans == 0) allowed |= BF_LOBE_TRANSMISSION;
            else r.blocked |= trans;
        } else {
            if (!valid_surface_event(src.lobes)) {
                r.status = BF_L_INVALID;
                return;
            }
            const uint32_t blocked = depth_blocked(limits, counters, src.lobes);
            if (blocked == 0) allowed = src.lobes & (BF_LOBE_REFLECTION | BF_LOBE_TRANSMISSION);
            else r.blocked |= blocked;
        }
        if (allowed == 0) continue;
        m.closures[filtered] = src;
        m.probabilities[filtered] = prob;
        r.source_indices[filtered] = si;
        r.allowed[filtered] = allowed;
        ++filtered;
    }
    m.active_count = filtered;
    if (filtered == r.source_count) return;
    for (uint32_t i = filtered; i < BF_MAX_CLOSURES; ++i) m.probabilities[i] = 0.0f;
    for (uint32_t i = 0; i <= BF_MAX_CLOSURES; ++i) m.cdf[i] = 0.0f;
    if (!build_filtered_distribution(m)) r.status = BF_L_NOT_REPRESENTABLE;
}

struct mixture_sample_t {
    float value[4];
    v3f incoming;
    float pdf;
    uint32_t discrete;
    uint32_t selected;
    uint32_t lobes;
    float eta_scale;
    int status;
};

/* TransportLobesDevice.cuh:2960-3062 */
BF_D void sample_mixture(const mixture_t& m, const filter_t& f, v3f wo, float component, float u, float v,
                         mixture_sample_t& out) {
    out.eta_scale = 1.0f;
    out.discrete = 0u;
    out.selected = 0u;
    out.lobes = 0u;
    out.pdf = 0.0f;
    if (!unit_vector(wo) || !bf_isfinite(component) || component < 0.0f || !(component < 1.0f) || !bf_isfinite(u) ||
        u < 0.0f || !(u < 1.0f) || !bf_isfinite(v) || v < 0.0f || !(v < 1.0f)) {
        out.status = BF_L_INVALID;
        return;
    }
    if (m.active_count == 0) {
        out.status = BF_L_OUTSIDE;
        return;
    }
    uint32_t sel = m.active_count;
    for (uint32_t i = 0; i < m.active_count; ++i) {
        if (component < m.cdf[i + 1]) {
            sel = i;
            break;
        }
    }
    if (sel == m.active_count) {
        out.status = BF_L_NOT_REPRESENTABLE;
        return;
    }
    const float event = (component - m.cdf[sel]) / m.probabilities[sel];
    if (!bf_isfinite(event) || event < 0.0f || !(event < 1.0f)) {
        out.status = BF_L_NOT_REPRESENTABLE;
        return;
    }
    lobe_sample_t s;
    record_sample(m.closures[sel], f.allowed[sel], wo, event, u, v, s);
    if (s.status != BF_L_OK) {
        out.status = s.status;
        return;
    }
    out.incoming = s.incoming;
    out.lobes = s.lobes;
    out.eta_scale = s.eta_scale;
    if (m.active_count == 1) {
#pragma unroll
        for (int k = 0; k < 4; ++k) out.value[k] = s.value[k];
        out.pdf = s.pdf;
        out.discrete = s.discrete;
        out.selected = f.source_indices[0];
        out.status = BF_L_OK;
        return;
    }
    if (!s.discrete) {
        float value[4], pdf;
        const int st1 = mixture_eval(m, f, wo, s.incoming, value);
        const int st2 = mixture_pdf(m, f, wo, s.incoming, pdf);
        if (st1 != BF_L_OK || st2 != BF_L_OK) {
            out.status = st1 != BF_L_OK ? st1 : st2;
            return;
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) out.value[k] = value[k];
        out.pdf = pdf;
    } else {
        /* aggregate_delta_atom (:2394-2480): every delta closure of the (filtered) mixture that
         * sends wo to the sampled direction adds its value and its selection-weighted probability */
        float sums[4] = {0.0f, 0.0f, 0.0f, 0.0f}, corrections[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        float psum = 0.0f, pcorrection = 0.0f;
        uint32_t matching = 0;
        for (uint32_t i = 0; i < m.active_count; ++i) {
            const uint32_t kind = m.closures[i].kind;
            if (kind != BF_CLOSURE_SPECULAR_REFLECTION && kind != BF_CLOSURE_SPECULAR_TRANSMISSION) continue;
            /* record_sample restricted to the two delta kinds: the microfacet samplers are not
             * inlined a second time here */
            lobe_sample_t c;
            empty_sample(c, BF_L_INVALID);
            if (mf_valid_closure(&m.closures[i])) {
                if (kind == BF_CLOSURE_SPECULAR_REFLECTION) specular_reflection_sample(m.closures[i].weight, wo, c);
                else specular_transmission_sample(m.closures[i].weight, m.closures[i].parameters[0], m.closures[i].parameters[1], wo, c);
            }
            if (c.status == BF_L_OUTSIDE) continue;
            if (c.status != BF_L_OK) {
                out.status = c.status;
                return;
            }
            if (c.incoming.x != s.incoming.x || c.incoming.y != s.incoming.y || c.incoming.z != s.incoming.z) continue;
            if (!c.discrete || !bf_isfinite(c.pdf) || !(c.pdf > 0.0f) || c.pdf > 1.0f) {
                out.status = BF_L_NOT_REPRESENTABLE;
                return;
            }
            for (int k = 0; k < 4; ++k)
                if (!compensated_add(c.value[k], sums[k], corrections[k])) {
                    out.status = BF_L_NOT_REPRESENTABLE;
                    return;
                }
            float weighted;
            if (!checked_product(m.probabilities[i], c.pdf, weighted)) {
                out.status = BF_L_NOT_REPRESENTABLE;
                return;
            }
            if (!compensated_add(weighted, psum, pcorrection)) {
                out.status = BF_L_NOT_REPRESENTABLE;
                return;
            }
            ++matching;
        }
        const float probability = psum + pcorrection;
        if (matching == 0 || !bf_isfinite(probability) || !(probability > 0.0f) || probability > 1.0f) {
            out.status = BF_L_NOT_REPRESENTABLE;
            return;
        }
        for (int k = 0; k < 4; ++k) {
            out.value[k] = sums[k] + corrections[k];
            if (!bf_isfinite(out.value[k]) || out.value[k] < 0.0f) {
                out.status = BF_L_NOT_REPRESENTABLE;
                return;
            }
        }
        out.pdf = probability;
        out.discrete = 1u;
    }
    out.selected = f.source_indices[sel];
    out.status = BF_L_OK;
}

#endif
