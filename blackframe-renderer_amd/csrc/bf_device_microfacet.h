/*
 * Microfacet and rough-diffuse lobes on the device: energy-preserving Oren–Nayar, GGX (Smith
 * masking, visible-normal sampling), rough conductor with complex-IOR Fresnel, rough dielectric
 * with direction masks. Arithmetic follows
 * Blackframe/XPU/CUDA/Include/Blackframe/XPU/CUDA/TransportLobesDevice.cuh:220-2059 operation for
 * operation (the same binary32 results as the CPU restatement); transcendentals come from
 * include/blackframe_math.h. Transport mode is radiance (camera paths).
 */
#ifndef BF_DEVICE_MICROFACET_H
#define BF_DEVICE_MICROFACET_H

#include <float.h>

#define BF_TWO_PI 6.28318530717958647692f
#define BF_TRANSPORT_RADIANCE 0u

BF_D void copy4(float* dst, const float* src) {
#pragma unroll
    for (int k = 0; k < 4; ++k) dst[k] = src[k];
}

/* ---- spectra and vectors (TransportLobesDevice.cuh:174-218, TransportDevice.cuh) ------------- */
BF_D int mf_spectrum_finite_nonnegative(const float* w) {
    for (int k = 0; k < 4; ++k)
        if (!bf_isfinite(w[k]) || w[k] < 0.0f) return 0;
    return 1;
}
BF_D int mf_spectrum_zero(const float* w) {
    return w[0] == 0.0f && w[1] == 0.0f && w[2] == 0.0f && w[3] == 0.0f;
}
/* :933-941 */
BF_D int mf_positive_spectrum(const float* w) {
    for (int k = 0; k < 4; ++k)
        if (!bf_isfinite(w[k]) || !(w[k] > 0.0f)) return 0;
    return 1;
}
/* :188-190 */
BF_D float mf_length(v3f v) { return bf_hypotf(bf_hypotf(v.x, v.y), v.z); }
/* :192-207 */
BF_D int mf_normalized(v3f v, v3f* out) {
    const float m = mf_length(v);
    if (!bf_isfinite(m) || !(m > 0.0f)) return BF_L_NOT_REPRESENTABLE;
    const v3f r = mk(v.x / m, v.y / m, v.z / m);
    if (!unit_vector(r)) return BF_L_NOT_REPRESENTABLE;
    *out = r;
    return BF_L_OK;
}
BF_D v3f mf_scaled(v3f d, float s) { return mk(s * d.x, s * d.y, s * d.z); }

/* checked_lobe_product_ratio (:1187-1226): significands renormalised after every factor. */
BF_D int mf_product_ratio(const float* num, int nn, const float* den, int nd, float* out) {
    int has_zero = 0;
    for (int i = 0; i < nn; ++i) {
        if (!bf_isfinite(num[i]) || num[i] < 0.0f) return BF_L_INVALID;
        has_zero = has_zero || num[i] == 0.0f;
    }
    for (int i = 0; i < nd; ++i)
        if (!bf_isfinite(den[i]) || !(den[i] > 0.0f)) return BF_L_INVALID;
    if (has_zero) {
        *out = 0.0f;
        return BF_L_OK;
    }
    float significand = 1.0f;
    int exponent = 0;
    for (int i = 0; i < nn; ++i) {
        int ve, ne;
        const float vs = bf_frexpf(num[i], &ve);
        significand = bf_frexpf(significand * vs, &ne);
        exponent += ve + ne;
    }
    for (int i = 0; i < nd; ++i) {
        int ve, ne;
        const float vs = bf_frexpf(den[i], &ve);
        significand = bf_frexpf(significand / vs, &ne);
        exponent += ne - ve;
    }
    const float r = bf_ldexpf(significand, exponent);
    if (!bf_isfinite(r) || !(r > 0.0f)) return BF_L_NOT_REPRESENTABLE;
    *out = r;
    return BF_L_OK;
}

/* scale_spectrum (:1228-1241) */
BF_D int mf_scale_spectrum(const float* coefficient, float scale, float* out) {
    float r[4];
    for (int k = 0; k < 4; ++k) {
        const float num[2] = {coefficient[k], scale};
        const float den[1] = {1.0f};
        const int st = mf_product_ratio(num, 2, den, 1, &r[k]);
        if (st != BF_L_OK) return st;
    }
    copy4(out, r);
    return BF_L_OK;
}

/* ---- energy-preserving Oren–Nayar (:220-296, 471-520) --------------------------------------- */
BF_D float mf_fon_coefficient(void) { return 0.5f - 2.0f * BF_INV_PI / 3.0f; }
BF_D float mf_fon_average_coefficient(void) { return 2.0f / 3.0f - 28.0f * BF_INV_PI / 15.0f; }

BF_D int mf_rough_diffuse_loss_shape(v3f d, float* out) {
    const float sine = bf_hypotf(d.x, d.y);
    const float cosine = d.z;
    const float angle = bf_atan2f(sine, cosine);
    const float stable_ratio = sine * cosine * (1.0f + sine + sine * sine) / (1.0f + sine);
    const float g = sine * (angle - sine * cosine) + (2.0f / 3.0f) * (stable_ratio - sine);
    const float loss = mf_fon_coefficient() - g * BF_INV_PI;
    if (!bf_isfinite(loss) || loss < 0.0f) return BF_L_NOT_REPRESENTABLE;
    *out = loss;
    return BF_L_OK;
}

BF_D int mf_rough_diffuse_eval(const float* rho, float roughness, v3f wo, v3f wi, float* out) {
    if (!reflectance_ok(rho) || !bf_isfinite(roughness) || roughness < 0.0f || roughness > 1.0f ||
        !unit_vector(wo) || !unit_vector(wi))
        return BF_L_INVALID;
    if (!(wo.z > 0.0f) || !(wi.z > 0.0f) || mf_spectrum_zero(rho)) {
        out[0] = out[1] = out[2] = out[3] = 0.0f;
        return BF_L_OK;
    }
    if (roughness == 0.0f) return lambert_eval(rho, wo, wi, out);
    const float fon = mf_fon_coefficient();
    const float tangent_dot = fmaf(wo.x, wi.x, wo.y * wi.y);
    const float s_over_t = tangent_dot > 0.0f ? tangent_dot / fmaxf(wo.z, wi.z) : tangent_dot;
    const float a = 1.0f / (1.0f + fon * roughness);
    const float single = BF_INV_PI * a * (1.0f + roughness * s_over_t);
    if (!bf_isfinite(single) || single < 0.0f) return BF_L_NOT_REPRESENTABLE;
    float lo, li;
    int st = mf_rough_diffuse_loss_shape(wo, &lo);
    if (st != BF_L_OK) return st;
    st = mf_rough_diffuse_loss_shape(wi, &li);
    if (st != BF_L_OK) return st;
    const float average_albedo = a * (1.0f + mf_fon_average_coefficient() * roughness);
    const float average_loss = fon - mf_fon_average_coefficient();
    const float loss_over_average = a * roughness * li * lo / average_loss;
    if (!bf_isfinite(average_albedo) || !(average_albedo > 0.0f) || average_albedo > 1.0f ||
        !bf_isfinite(loss_over_average) || loss_over_average < 0.0f)
        return BF_L_NOT_REPRESENTABLE;
    float r[4];
    for (int k = 0; k < 4; ++k) {
        const float p = rho[k];
        const float denominator = (1.0f - p) + p * average_albedo;
        if (!bf_isfinite(denominator) || !(denominator > 0.0f)) return BF_L_NOT_REPRESENTABLE;
        const float ms = p * p * average_albedo / denominator;
        const float value = p * single + ms * BF_INV_PI * loss_over_average;
        if (!bf_isfinite(value) || value < 0.0f) return BF_L_NOT_REPRESENTABLE;
        r[k] = value;
    }
    copy4(out, r);
    return BF_L_OK;
}

/* ---- GGX width pair, Smith terms, distribution (:298-470) ----------------------------------- */
BF_D int mf_ggx_widths(float ax, float ay) {
    if (!bf_isfinite(ax) || !(ax > 0.0f) || !bf_isfinite(ay) || !(ay > 0.0f)) return 0;
    const float sqrt_limit = sqrtf(FLT_MAX);
    const float sqrt_pi = sqrtf(BF_PI);
    const float maximum = sqrt_limit * sqrt_pi;
    const float minimum = 1.0f / maximum;
    if (ax < minimum || ax > maximum || ay < minimum || ay > maximum) return 0;
    if (ax == ay) return 1;
    const double inverse_pi = 0.31830988618379067154;
    const double wx = (double)ax, wy = (double)ay;
    const double ix = 1.0 / wx, iy = 1.0 / wy;
    const double min_q = fmin(1.0, fmin(ix, iy));
    const double max_q = fmax(1.0, fmax(ix, iy));
    const double scale = inverse_pi / (wx * wy);
    const double min_q2 = min_q * min_q, max_q2 = max_q * max_q;
    const double max_density = scale / (min_q2 * min_q2);
    const double min_density = scale / (max_q2 * max_q2);
    return bf_isfinited(max_density) && bf_isfinited(min_density) && max_density <= (double)FLT_MAX &&
           min_density >= (double)BF_DENORM_MIN;
}

typedef struct {
    float normal, tangent, radius, ratio;
} mf_smith;

BF_D int mf_smith_terms(float ax, float ay, v3f d, mf_smith* s) {
    float normal, tangent;
    if (ax == ay) {
        const float sine = bf_hypotf(d.x, d.y);
        const float inverse_alpha = 1.0f / ax;
        normal = ax <= 1.0f ? d.z : d.z * inverse_alpha;
        tangent = ax <= 1.0f ? ax * sine : sine;
        const float radius = bf_hypotf(normal, tangent);
        if (!bf_isfinite(normal) || !bf_isfinite(tangent) || !bf_isfinite(radius) || !(normal > 0.0f) || !(radius > 0.0f))
            return BF_L_NOT_REPRESENTABLE;
        const float ratio = normal / radius;
        if (!bf_isfinite(ratio) || !(ratio > 0.0f) || ratio > 1.0f) return BF_L_NOT_REPRESENTABLE;
        s->normal = normal;
        s->tangent = tangent;
        s->radius = radius;
        s->ratio = ratio;
        return BF_L_OK;
    }
    const float common = fmaxf(1.0f, fmaxf(ax, ay));
    normal = d.z / common;
    const float tx = (ax / common) * d.x;
    const float ty = (ay / common) * d.y;
    tangent = bf_hypotf(tx, ty);
    const float radius = bf_hypotf(normal, tangent);
    if (!bf_isfinite(normal) || !bf_isfinite(tx) || !bf_isfinite(ty) || !bf_isfinite(tangent) || !bf_isfinite(radius) ||
        !(normal > 0.0f) || !(radius > 0.0f))
        return BF_L_NOT_REPRESENTABLE;
    const float ratio = normal / radius;
    if (!bf_isfinite(ratio) || !(ratio > 0.0f) || ratio > 1.0f) return BF_L_NOT_REPRESENTABLE;
    s->normal = normal;
    s->tangent = tangent;
    s->radius = radius;
    s->ratio = ratio;
    return BF_L_OK;
}

/* ggx_smith_lambda (:313-338) */
BF_D int mf_smith_lambda(float ax, float ay, v3f d, float* out) {
    if (!mf_ggx_widths(ax, ay) || !unit_vector(d) || !(d.z > 0.0f)) return BF_L_INVALID;
    mf_smith t;
    const int st = mf_smith_terms(ax, ay, d, &t);
    if (st != BF_L_OK) return st;
    if (t.tangent == 0.0f) {
        *out = 0.0f;
        return BF_L_OK;
    }
    float value;
    if (t.tangent <= t.normal) {
        const float slope = t.tangent / t.normal;
        const float root = bf_hypotf(1.0f, slope);
        value = (slope * 0.5f) * (slope / (root + 1.0f));
    } else {
        value = (1.0f - t.ratio) / (2.0f * t.ratio);
    }
    if (!bf_isfinite(value) || !(value > 0.0f)) return BF_L_NOT_REPRESENTABLE;
    *out = value;
    return BF_L_OK;
}

/* ggx_smith_g1 (:340-357) */
BF_D int mf_smith_g1(float ax, float ay, v3f d, float* out) {
    if (!mf_ggx_widths(ax, ay) || !unit_vector(d)) return BF_L_INVALID;
    if (!(d.z > 0.0f)) {
        *out = 0.0f;
        return BF_L_OK;
    }
    mf_smith t;
    const int st = mf_smith_terms(ax, ay, d, &t);
    if (st != BF_L_OK) return st;
    const float value = 2.0f * t.ratio / (1.0f + t.ratio);
    if (!bf_isfinite(value) || !(value > 0.0f) || value > 1.0f) return BF_L_NOT_REPRESENTABLE;
    *out = value;
    return BF_L_OK;
}

/* ggx_smith_g2 (:359-387), height-correlated */
BF_D int mf_smith_g2(float ax, float ay, v3f wo, v3f wi, float* out) {
    if (!mf_ggx_widths(ax, ay) || !unit_vector(wo) || !unit_vector(wi)) return BF_L_INVALID;
    if (!(wo.z > 0.0f) || !(wi.z > 0.0f)) {
        *out = 0.0f;
        return BF_L_OK;
    }
    mf_smith to, ti;
    int st = mf_smith_terms(ax, ay, wo, &to);
    if (st != BF_L_OK) return st;
    st = mf_smith_terms(ax, ay, wi, &ti);
    if (st != BF_L_OK) return st;
    const float smaller = fminf(to.ratio, ti.ratio);
    const float larger = fmaxf(to.ratio, ti.ratio);
    const float value = 2.0f * smaller / (1.0f + smaller / larger);
    if (!bf_isfinite(value) || !(value > 0.0f) || value > 1.0f) return BF_L_NOT_REPRESENTABLE;
    *out = value;
    return BF_L_OK;
}

/* ggx_normal_distribution (:471-507) */
BF_D int mf_ggx_d(float ax, float ay, v3f m, float* out) {
    if (!mf_ggx_widths(ax, ay) || !unit_vector(m)) return BF_L_INVALID;
    if (!(m.z > 0.0f)) {
        *out = 0.0f;
        return BF_L_OK;
    }
    float denominator_root, density_root;
    if (ax == ay) {
        const float radial = bf_hypotf(m.x, m.y);
        if (ax <= 1.0f) {
            denominator_root = bf_hypotf(radial, ax * m.z);
            density_root = (ax / denominator_root) / denominator_root;
        } else {
            const float inverse_alpha = 1.0f / ax;
            denominator_root = bf_hypotf(radial * inverse_alpha, m.z);
            density_root = (inverse_alpha / denominator_root) / denominator_root;
        }
        density_root *= sqrtf(BF_INV_PI);
    } else {
        denominator_root = bf_hypotf(bf_hypotf(m.x / ax, m.y / ay), m.z);
        const float normalization_root = sqrtf(BF_INV_PI) / (sqrtf(ax) * sqrtf(ay));
        density_root = (normalization_root / denominator_root) / denominator_root;
    }
    const float value = density_root * density_root;
    if (!bf_isfinite(denominator_root) || !(denominator_root > 0.0f) || !bf_isfinite(value) || !(value > 0.0f))
        return BF_L_NOT_REPRESENTABLE;
    *out = value;
    return BF_L_OK;
}

/* ---- Fresnel (:555-718) --------------------------------------------------------------------- */
BF_D int mf_squared_magnitude_ratio(float nr, float ni, float dr, float di, float* out) {
    const float numerator = bf_hypotf(nr, ni);
    const float denominator = bf_hypotf(dr, di);
    if (!bf_isfinite(numerator) || !bf_isfinite(denominator) || !(denominator > 0.0f)) return BF_L_NOT_REPRESENTABLE;
    const float amplitude = numerator / denominator;
    const float reflectance = amplitude * amplitude;
    if (!bf_isfinite(reflectance) || reflectance < 0.0f || reflectance > 1.0f) return BF_L_NOT_REPRESENTABLE;
    *out = reflectance;
    return BF_L_OK;
}

BF_D int mf_dielectric_fresnel(float cos_i, float eta_i, float eta_t, float* out) {
    if (!bf_isfinite(cos_i) || cos_i < 0.0f || cos_i > 1.0f || !bf_isfinite(eta_i) || !(eta_i > 0.0f) ||
        !bf_isfinite(eta_t) || !(eta_t > 0.0f))
        return BF_L_INVALID;
    if (eta_i == eta_t) {
        *out = 0.0f;
        return BF_L_OK;
    }
    if (cos_i == 0.0f) {
        *out = 1.0f;
        return BF_L_OK;
    }
    const int denser = eta_i > eta_t;
    const float eta = denser ? eta_t / eta_i : eta_i / eta_t;
    if (cos_i == 1.0f) {
        const float amplitude = (1.0f - eta) / (1.0f + eta);
        const float r = amplitude * amplitude;
        if (!bf_isfinite(r) || !(r > 0.0f) || r > 1.0f) return BF_L_NOT_REPRESENTABLE;
        *out = r;
        return BF_L_OK;
    }
    const float sin_i = sqrtf((1.0f - cos_i) * (1.0f + cos_i));
    float sin_t;
    if (denser) {
        if (eta == 0.0f || sin_i >= eta) {
            *out = 1.0f;
            return BF_L_OK;
        }
        sin_t = sin_i / eta;
    } else {
        sin_t = eta * sin_i;
    }
    if (!(sin_t < 1.0f)) {
        *out = 1.0f;
        return BF_L_OK;
    }
    const float cos_t = sqrtf((1.0f - sin_t) * (1.0f + sin_t));
    const float first_eta_cos = eta * cos_t;
    const float first = (cos_i - first_eta_cos) / (cos_i + first_eta_cos);
    const float second_eta_cos = eta * cos_i;
    const float second = (second_eta_cos - cos_t) / (second_eta_cos + cos_t);
    const float r = 0.5f * fmaf(first, first, second * second);
    if (!bf_isfinite(r) || !(r > 0.0f) || r > 1.0f) return BF_L_NOT_REPRESENTABLE;
    *out = r;
    return BF_L_OK;
}

BF_D int mf_conductor_fresnel_lane(float cos_i, float eta, float k, float* out) {
    if (eta == 1.0f && k == 0.0f) {
        *out = 0.0f;
        return BF_L_OK;
    }
    if (cos_i == 0.0f) {
        *out = 1.0f;
        return BF_L_OK;
    }
    const float scale = fmaxf(1.0f, fmaxf(eta, k));
    const float ni = 1.0f / scale;
    const float ne = eta / scale;
    const float nk = k / scale;
    if (cos_i == 1.0f) {
        float r;
        const int st = mf_squared_magnitude_ratio(ne - ni, nk, ne + ni, nk, &r);
        if (st != BF_L_OK) return st;
        if (!(r > 0.0f)) return BF_L_NOT_REPRESENTABLE;
        *out = r;
        return BF_L_OK;
    }
    const float sin2 = (1.0f - cos_i) * (1.0f + cos_i);
    const float scaled_sin2 = ni * ni * sin2;
    const float x = fmaf(ne, ne, -fmaf(nk, nk, scaled_sin2));
    const float y = 2.0f * ne * nk;
    const float magnitude = bf_hypotf(x, y);
    float root_r, root_i;
    if (x >= 0.0f) {
        root_r = sqrtf(0.5f * (magnitude + x));
        root_i = root_r == 0.0f ? 0.0f : y / (2.0f * root_r);
    } else {
        root_i = sqrtf(0.5f * (magnitude - x));
        root_r = root_i == 0.0f ? 0.0f : y / (2.0f * root_i);
    }
    const float scaled_cos = ni * cos_i;
    float perpendicular, parallel;
    int st = mf_squared_magnitude_ratio(scaled_cos - root_r, -root_i, scaled_cos + root_r, root_i, &perpendicular);
    if (st != BF_L_OK) return st;
    const float sq_r = fmaf(ne, ne, -(nk * nk));
    const float sq_i = y;
    const float srr = ni * root_r, sri = ni * root_i;
    const float br = cos_i * sq_r, bi = cos_i * sq_i;
    st = mf_squared_magnitude_ratio(br - srr, bi - sri, br + srr, bi + sri, &parallel);
    if (st != BF_L_OK) return st;
    const float r = 0.5f * (perpendicular + parallel);
    if (!bf_isfinite(r) || !(r > 0.0f) || r > 1.0f) return BF_L_NOT_REPRESENTABLE;
    *out = r;
    return BF_L_OK;
}

BF_D int mf_conductor_fresnel(float cos_i, const float* eta, const float* k, float* out) {
    if (!bf_isfinite(cos_i) || cos_i < 0.0f || cos_i > 1.0f) return BF_L_INVALID;
    for (int l = 0; l < 4; ++l)
        if (!bf_isfinite(eta[l]) || !(eta[l] > 0.0f) || !bf_isfinite(k[l]) || k[l] < 0.0f) return BF_L_INVALID;
    float r[4];
    for (int l = 0; l < 4; ++l) {
        const int st = mf_conductor_fresnel_lane(cos_i, eta[l], k[l], &r[l]);
        if (st != BF_L_OK) return st;
    }
    copy4(out, r);
    return BF_L_OK;
}

/* ---- visible normals (:720-931) ------------------------------------------------------------- */
BF_D int mf_vndf_pdf(float ax, float ay, v3f wo, v3f m, float* out) {
    if (!mf_ggx_widths(ax, ay) || !unit_vector(wo) || !unit_vector(m)) return BF_L_INVALID;
    const float visible_dot = dot3(wo, m);
    if (!(wo.z > 0.0f) || !(m.z > 0.0f) || !(visible_dot > 0.0f)) {
        *out = 0.0f;
        return BF_L_OK;
    }
    float d;
    const int st = mf_ggx_d(ax, ay, m, &d);
    if (st != BF_L_OK) return st;
    const float radius = ax == ay ? bf_hypotf(wo.z, ax * bf_hypotf(wo.x, wo.y))
                                  : bf_hypotf(bf_hypotf(wo.z, ax * wo.x), ay * wo.y);
    const float denominator = wo.z + radius;
    const float value = d * (2.0f * visible_dot / denominator);
    if (!bf_isfinite(denominator) || !(denominator > 0.0f) || !bf_isfinite(value) || !(value > 0.0f))
        return BF_L_NOT_REPRESENTABLE;
    *out = value;
    return BF_L_OK;
}

BF_D int mf_nonnegative_roundoff(float value, float scale, float* out) {
    if (value >= 0.0f) {
        *out = value;
        return BF_L_OK;
    }
    const float bound = BF_EPS * 16.0f * fmaxf(1.0f, fabsf(scale));
    if (-value <= bound) {
        *out = 0.0f;
        return BF_L_OK;
    }
    return BF_L_NOT_REPRESENTABLE;
}

/* stretch (or unstretch) a direction by the GGX widths; returns 0 when a component underflows */
BF_D int mf_stretch(float ax, float ay, v3f d, v3f* out) {
    const int iso = ax == ay;
    const float inverse_alpha = 1.0f / ax;
    const float common = fmaxf(1.0f, fmaxf(ax, ay));
    const float sx = ax / common, sy = ay / common;
    v3f s;
    if (iso) s = ax <= 1.0f ? mk(ax * d.x, ax * d.y, d.z) : mk(d.x, d.y, d.z * inverse_alpha);
    else s = mk(sx * d.x, sy * d.y, d.z / common);
    if ((iso && ax <= 1.0f && ((d.x != 0.0f && s.x == 0.0f) || (d.y != 0.0f && s.y == 0.0f))) ||
        (iso && ax > 1.0f && s.z == 0.0f) ||
        (!iso && ((d.x != 0.0f && s.x == 0.0f) || (d.y != 0.0f && s.y == 0.0f) || s.z == 0.0f)))
        return 0;
    *out = s;
    return 1;
}

BF_D int mf_sample_vndf(float ax, float ay, v3f wo, float u, float v, v3f* m_out, float* pdf_out) {
    if (!mf_ggx_widths(ax, ay) || !unit_vector(wo) || !bf_isfinite(u) || u < 0.0f || !(u < 1.0f) ||
        !bf_isfinite(v) || v < 0.0f || !(v < 1.0f))
        return BF_L_INVALID;
    if (!(wo.z > 0.0f)) return BF_L_OUTSIDE;
    v3f stretched, view;
    if (!mf_stretch(ax, ay, wo, &stretched)) return BF_L_NOT_REPRESENTABLE;
    int st = mf_normalized(stretched, &view);
    if (st != BF_L_OK) return st;
    const float tangent_length = bf_hypotf(view.x, view.y);
    const v3f t1 = tangent_length > 0.0f ? mk(-view.y / tangent_length, view.x / tangent_length, 0.0f)
                                        : mk(1.0f, 0.0f, 0.0f);
    const v3f t2 = cross3(view, t1);
    const float disk_radius = sqrtf(u);
    const float azimuth = BF_TWO_PI * v;
    float sa, ca;
    bf_sincosf(azimuth, &sa, &ca);
    const float disk_x = disk_radius * ca;
    const float initial_y = disk_radius * sa;
    float limit2, hz2;
    st = mf_nonnegative_roundoff((1.0f - disk_x) * (1.0f + disk_x), 1.0f + fabsf(disk_x), &limit2);
    if (st != BF_L_OK) return st;
    const float limit = sqrtf(limit2);
    const float projection = 0.5f * (1.0f + view.z);
    const float disk_y = fmaf(projection, initial_y, (1.0f - projection) * limit);
    st = mf_nonnegative_roundoff(fmaf(-disk_y, disk_y, limit2), limit2 + fabsf(disk_y * disk_y), &hz2);
    if (st != BF_L_OK) return st;
    const float hz = sqrtf(hz2);
    const v3f h = mk(fmaf(disk_x, t1.x, fmaf(disk_y, t2.x, hz * view.x)),
                         fmaf(disk_x, t1.y, fmaf(disk_y, t2.y, hz * view.y)),
                         fmaf(disk_x, t1.z, fmaf(disk_y, t2.z, hz * view.z)));
    if (!(h.z > 0.0f)) return BF_L_NOT_REPRESENTABLE;
    v3f unstretched, m;
    if (!mf_stretch(ax, ay, h, &unstretched)) return BF_L_NOT_REPRESENTABLE;
    st = mf_normalized(unstretched, &m);
    if (st != BF_L_OK) return st;
    if (!(m.z > 0.0f) || !(dot3(wo, m) > 0.0f)) return BF_L_NOT_REPRESENTABLE;
    float pdf;
    st = mf_vndf_pdf(ax, ay, wo, m, &pdf);
    if (st != BF_L_OK) return st;
    *m_out = m;
    *pdf_out = pdf;
    return BF_L_OK;
}

/* reflection_geometry (:943-975) */
BF_D int mf_reflection_geometry(v3f wo, v3f wi, v3f* m, float* half_cos) {
    const v3f h = mk(wo.x + wi.x, wo.y + wi.y, wo.z + wi.z);
    const float magnitude = mf_length(h);
    if (!bf_isfinite(magnitude) || !(magnitude > 0.0f)) return BF_L_NOT_REPRESENTABLE;
    const v3f n = mk(h.x / magnitude, h.y / magnitude, h.z / magnitude);
    const float c = 0.5f * magnitude;
    if (!unit_vector(n) || !(n.z > 0.0f) || !bf_isfinite(c) || !(c > 0.0f) || c > 1.0f) return BF_L_NOT_REPRESENTABLE;
    *m = n;
    *half_cos = c;
    return BF_L_OK;
}

/* ---- rough conductor (:977-1098); parameters[0..3] eta, [4..7] k, [8] alpha_x, [9] alpha_y -- */
BF_D int mf_conductor_eval(const float* coefficient, const float* eta, const float* k, float ax, float ay, v3f wo,
                                    v3f wi, float* out) {
    if (!reflectance_ok(coefficient) || !mf_positive_spectrum(eta) || !mf_spectrum_finite_nonnegative(k) ||
        !mf_ggx_widths(ax, ay) || !unit_vector(wo) || !unit_vector(wi))
        return BF_L_INVALID;
    if (!(wo.z > 0.0f) || !(wi.z > 0.0f) || mf_spectrum_zero(coefficient)) {
        out[0] = out[1] = out[2] = out[3] = 0.0f;
        return BF_L_OK;
    }
    v3f m;
    float half_cos, d, g, f[4];
    int st = mf_reflection_geometry(wo, wi, &m, &half_cos);
    if (st != BF_L_OK) return st;
    const int sd = mf_ggx_d(ax, ay, m, &d);
    const int sg = mf_smith_g2(ax, ay, wo, wi, &g);
    const int sf = mf_conductor_fresnel(half_cos, eta, k, f);
    if (sd != BF_L_OK) return sd;
    if (sg != BF_L_OK) return sg;
    if (sf != BF_L_OK) return sf;
    const float denominator = 4.0f * wo.z * wi.z;
    const float scale = d * g / denominator;
    if (!bf_isfinite(denominator) || !(denominator > 0.0f) || !bf_isfinite(scale) || !(scale > 0.0f))
        return BF_L_NOT_REPRESENTABLE;
    float r[4];
    for (int l = 0; l < 4; ++l) {
        r[l] = coefficient[l] * f[l] * scale;
        if (!bf_isfinite(r[l]) || r[l] < 0.0f) return BF_L_NOT_REPRESENTABLE;
    }
    copy4(out, r);
    return BF_L_OK;
}

BF_D int mf_conductor_pdf(float ax, float ay, v3f wo, v3f wi, float* out) {
    if (!mf_ggx_widths(ax, ay) || !unit_vector(wo) || !unit_vector(wi)) return BF_L_INVALID;
    if (!(wo.z > 0.0f) || !(wi.z > 0.0f)) {
        *out = 0.0f;
        return BF_L_OK;
    }
    v3f m;
    float half_cos, visible;
    int st = mf_reflection_geometry(wo, wi, &m, &half_cos);
    if (st != BF_L_OK) return st;
    st = mf_vndf_pdf(ax, ay, wo, m, &visible);
    if (st != BF_L_OK) return st;
    const float value = visible / (4.0f * half_cos);
    if (!bf_isfinite(value) || !(value > 0.0f)) return BF_L_NOT_REPRESENTABLE;
    *out = value;
    return BF_L_OK;
}

BF_D void mf_conductor_sample(const float* coefficient, const float* eta, const float* k, float ax, float ay,
                                       v3f wo, float u, float v, lobe_sample_t* out) {
    empty_sample(*out, BF_L_INVALID);
    if (!reflectance_ok(coefficient) || !mf_positive_spectrum(eta) || !mf_spectrum_finite_nonnegative(k) ||
        !mf_ggx_widths(ax, ay) || !unit_vector(wo))
        return;
    v3f m;
    float vpdf;
    int st = mf_sample_vndf(ax, ay, wo, u, v, &m, &vpdf);
    if (st != BF_L_OK) {
        out->status = st;
        return;
    }
    const float half_cos = dot3(wo, m);
    if (!bf_isfinite(half_cos) || !(half_cos > 0.0f) || half_cos > 1.0f) {
        out->status = BF_L_NOT_REPRESENTABLE;
        return;
    }
    const v3f wi = mk(2.0f * half_cos * m.x - wo.x, 2.0f * half_cos * m.y - wo.y, 2.0f * half_cos * m.z - wo.z);
    if (!unit_vector(wi)) {
        out->status = BF_L_NOT_REPRESENTABLE;
        return;
    }
    if (!(wi.z > 0.0f)) {
        out->status = BF_L_OUTSIDE;
        return;
    }
    float value[4], pdf = 0.0f;
    const int s1 = mf_conductor_eval(coefficient, eta, k, ax, ay, wo, wi, value);
    const int s2 = mf_conductor_pdf(ax, ay, wo, wi, &pdf);
    if (s1 != BF_L_OK || s2 != BF_L_OK) {
        out->status = s1 != BF_L_OK ? s1 : s2;
        return;
    }
    copy4(out->value, value);
    out->incoming = wi;
    out->pdf = pdf;
    out->discrete = 0u;
    out->lobes = BF_LOBE_GLOSSY | BF_LOBE_REFLECTION;
    out->eta_scale = 1.0f;
    out->status = BF_L_OK;
}

/* ---- rough dielectric (:1100-1687); parameters[0] exterior eta, [1] interior, [2..3] alphas -- */
typedef struct {
    float face_sign, eta_i, eta_t, n_eta_i, n_eta_t;
} mf_interface;

BF_D mf_interface mf_dielectric_interface(v3f wo, float exterior, float interior) {
    mf_interface r;
    r.face_sign = wo.z > 0.0f ? 1.0f : -1.0f;
    r.eta_i = r.face_sign > 0.0f ? exterior : interior;
    r.eta_t = r.face_sign > 0.0f ? interior : exterior;
    const float maximum = fmaxf(r.eta_i, r.eta_t);
    r.n_eta_i = r.eta_i / maximum;
    r.n_eta_t = r.eta_t / maximum;
    return r;
}

typedef struct {
    v3f m;
    float o_dot_m, i_dot_m, half_length;
    int supported;
} mf_transmission;

BF_D int mf_transmission_geometry(v3f of, v3f inf, mf_interface f, mf_transmission* g) {
    g->supported = 0;
    v3f h = mk(fmaf(f.n_eta_i, of.x, f.n_eta_t * inf.x), fmaf(f.n_eta_i, of.y, f.n_eta_t * inf.y),
                   fmaf(f.n_eta_i, of.z, f.n_eta_t * inf.z));
    const float magnitude = mf_length(h);
    if (!bf_isfinite(magnitude) || !(magnitude > 0.0f)) return BF_L_NOT_REPRESENTABLE;
    h = mf_scaled(h, 1.0f / magnitude);
    if (h.z < 0.0f) h = mf_scaled(h, -1.0f);
    if (h.z == 0.0f) return BF_L_OK;
    if (!unit_vector(h)) return BF_L_NOT_REPRESENTABLE;
    const float od = dot3(of, h), id = dot3(inf, h);
    if (!bf_isfinite(od) || !bf_isfinite(id)) return BF_L_NOT_REPRESENTABLE;
    if (!(od > 0.0f) || !(id < 0.0f)) return BF_L_OK;
    if (od > 1.0f || id < -1.0f) return BF_L_NOT_REPRESENTABLE;
    g->m = h;
    g->o_dot_m = od;
    g->i_dot_m = id;
    g->half_length = magnitude;
    g->supported = 1;
    return BF_L_OK;
}

BF_D int mf_dielectric_valid(const float* c, float exterior, float interior, float ax, float ay) {
    return reflectance_ok(c) && bf_isfinite(exterior) && exterior > 0.0f && bf_isfinite(interior) && interior > 0.0f &&
           exterior != interior && mf_ggx_widths(ax, ay);
}

BF_D int mf_dielectric_eval(const float* c, float exterior, float interior, float ax, float ay, v3f wo, v3f wi,
                                     uint32_t mode, float* out) {
    if (!mf_dielectric_valid(c, exterior, interior, ax, ay) || mode > 1u || !unit_vector(wo) || !unit_vector(wi))
        return BF_L_INVALID;
    out[0] = out[1] = out[2] = out[3] = 0.0f;
    if (wo.z == 0.0f || wi.z == 0.0f) return BF_L_OK;
    const mf_interface f = mf_dielectric_interface(wo, exterior, interior);
    const v3f of = mf_scaled(wo, f.face_sign), inf = mf_scaled(wi, f.face_sign);
    if (inf.z > 0.0f) {
        v3f m;
        float half_cos, fr, d, g, scale;
        int st = mf_reflection_geometry(of, inf, &m, &half_cos);
        if (st != BF_L_OK) return st;
        const int s1 = mf_dielectric_fresnel(half_cos, f.eta_i, f.eta_t, &fr);
        const int s2 = mf_ggx_d(ax, ay, m, &d);
        const int s3 = mf_smith_g2(ax, ay, of, inf, &g);
        if (s1 != BF_L_OK) return s1;
        if (s2 != BF_L_OK) return s2;
        if (s3 != BF_L_OK) return s3;
        const float num[3] = {fr, d, g};
        const float den[3] = {4.0f, of.z, inf.z};
        st = mf_product_ratio(num, 3, den, 3, &scale);
        if (st != BF_L_OK) return st;
        return mf_scale_spectrum(c, scale, out);
    }
    mf_transmission t;
    int st = mf_transmission_geometry(of, inf, f, &t);
    if (st != BF_L_OK || !t.supported) return st;
    float fr, d, g, scale;
    st = mf_dielectric_fresnel(t.o_dot_m, f.eta_i, f.eta_t, &fr);
    if (st != BF_L_OK) return st;
    if (fr == 1.0f) return BF_L_OK;
    const int s2 = mf_ggx_d(ax, ay, t.m, &d);
    const int s3 = mf_smith_g2(ax, ay, of, mf_scaled(inf, -1.0f), &g);
    if (s2 != BF_L_OK) return s2;
    if (s3 != BF_L_OK) return s3;
    const float transport_eta = mode == BF_TRANSPORT_RADIANCE ? f.n_eta_i : f.n_eta_t;
    const float num[7] = {1.0f - fr, d, g, transport_eta, transport_eta, t.o_dot_m, -t.i_dot_m};
    const float den[4] = {of.z, -inf.z, t.half_length, t.half_length};
    st = mf_product_ratio(num, 7, den, 4, &scale);
    if (st != BF_L_OK) return st;
    return mf_scale_spectrum(c, scale, out);
}

BF_D int mf_dielectric_pdf(const float* c, float exterior, float interior, float ax, float ay, v3f wo, v3f wi,
                                    uint32_t mode, float* out) {
    if (!mf_dielectric_valid(c, exterior, interior, ax, ay) || mode > 1u || !unit_vector(wo) || !unit_vector(wi))
        return BF_L_INVALID;
    *out = 0.0f;
    if (wo.z == 0.0f || wi.z == 0.0f) return BF_L_OK;
    const mf_interface f = mf_dielectric_interface(wo, exterior, interior);
    const v3f of = mf_scaled(wo, f.face_sign), inf = mf_scaled(wi, f.face_sign);
    if (inf.z > 0.0f) {
        v3f m;
        float half_cos, fr, visible;
        int st = mf_reflection_geometry(of, inf, &m, &half_cos);
        if (st != BF_L_OK) return st;
        const int s1 = mf_dielectric_fresnel(half_cos, f.eta_i, f.eta_t, &fr);
        const int s2 = mf_vndf_pdf(ax, ay, of, m, &visible);
        if (s1 != BF_L_OK) return s1;
        if (s2 != BF_L_OK) return s2;
        const float num[2] = {fr, visible};
        const float den[2] = {4.0f, half_cos};
        return mf_product_ratio(num, 2, den, 2, out);
    }
    mf_transmission t;
    int st = mf_transmission_geometry(of, inf, f, &t);
    if (st != BF_L_OK || !t.supported) return st;
    float fr, visible;
    st = mf_dielectric_fresnel(t.o_dot_m, f.eta_i, f.eta_t, &fr);
    if (st != BF_L_OK) return st;
    if (fr == 1.0f) return BF_L_OK;
    st = mf_vndf_pdf(ax, ay, of, t.m, &visible);
    if (st != BF_L_OK) return st;
    const float num[5] = {1.0f - fr, visible, f.n_eta_t, f.n_eta_t, -t.i_dot_m};
    const float den[2] = {t.half_length, t.half_length};
    return mf_product_ratio(num, 5, den, 2, out);
}

BF_D void mf_dielectric_sample(const float* c, float exterior, float interior, float ax, float ay, v3f wo,
                                        float event, float u, float v, uint32_t mode, lobe_sample_t* out) {
    empty_sample(*out, BF_L_INVALID);
    if (!mf_dielectric_valid(c, exterior, interior, ax, ay) || mode > 1u || !unit_vector(wo) || !bf_isfinite(event) ||
        event < 0.0f || !(event < 1.0f))
        return;
    if (wo.z == 0.0f) {
        out->status = BF_L_OUTSIDE;
        return;
    }
    const mf_interface f = mf_dielectric_interface(wo, exterior, interior);
    const v3f of = mf_scaled(wo, f.face_sign);
    v3f m;
    float vpdf, fr;
    int st = mf_sample_vndf(ax, ay, of, u, v, &m, &vpdf);
    if (st != BF_L_OK) {
        out->status = st;
        return;
    }
    const float od = dot3(of, m);
    if (!bf_isfinite(od) || !(od > 0.0f) || od > 1.0f) {
        out->status = BF_L_NOT_REPRESENTABLE;
        return;
    }
    st = mf_dielectric_fresnel(od, f.eta_i, f.eta_t, &fr);
    if (st != BF_L_OK) {
        out->status = st;
        return;
    }
    v3f inf;
    uint32_t lobes;
    float eta_scale = 1.0f;
    if (event < fr) {
        inf = mk(2.0f * od * m.x - of.x, 2.0f * od * m.y - of.y, 2.0f * od * m.z - of.z);
        lobes = BF_LOBE_GLOSSY | BF_LOBE_REFLECTION;
        if (!unit_vector(inf)) {
            out->status = BF_L_NOT_REPRESENTABLE;
            return;
        }
        if (!(inf.z > 0.0f)) {
            out->status = BF_L_OUTSIDE;
            return;
        }
    } else {
        const v3f tangent = mk(of.x - od * m.x, of.y - od * m.y, of.z - od * m.z);
        const float sin_i = mf_length(tangent);
        if (!bf_isfinite(sin_i) || sin_i > 1.0f) {
            out->status = BF_L_NOT_REPRESENTABLE;
            return;
        }
        float cos_t = 1.0f, tangent_scale = 0.0f;
        if (sin_i > 0.0f) {
            float sin_t;
            if (f.eta_i > f.eta_t) {
                const float critical = f.eta_t / f.eta_i;
                if (critical == 0.0f || sin_i >= critical) {
                    out->status = BF_L_NOT_REPRESENTABLE;
                    return;
                }
                sin_t = sin_i / critical;
            } else {
                const float ratio = f.eta_i / f.eta_t;
                if (ratio == 0.0f) {
                    out->status = BF_L_NOT_REPRESENTABLE;
                    return;
                }
                sin_t = ratio * sin_i;
            }
            if (!bf_isfinite(sin_t) || !(sin_t < 1.0f)) {
                out->status = BF_L_NOT_REPRESENTABLE;
                return;
            }
            cos_t = sqrtf((1.0f - sin_t) * (1.0f + sin_t));
            tangent_scale = sin_t / sin_i;
            if (!bf_isfinite(cos_t) || !(cos_t > 0.0f) || !bf_isfinite(tangent_scale) || !(tangent_scale > 0.0f)) {
                out->status = BF_L_NOT_REPRESENTABLE;
                return;
            }
        }
        inf = mk(fmaf(-tangent_scale, tangent.x, -cos_t * m.x), fmaf(-tangent_scale, tangent.y, -cos_t * m.y),
                      fmaf(-tangent_scale, tangent.z, -cos_t * m.z));
        lobes = BF_LOBE_GLOSSY | BF_LOBE_TRANSMISSION;
        if (!unit_vector(inf)) {
            out->status = BF_L_NOT_REPRESENTABLE;
            return;
        }
        if (!(inf.z < 0.0f)) {
            out->status = BF_L_OUTSIDE;
            return;
        }
        if (mode == BF_TRANSPORT_RADIANCE) {
            const float num[2] = {f.n_eta_t, f.n_eta_t};
            const float den[2] = {f.n_eta_i, f.n_eta_i};
            st = mf_product_ratio(num, 2, den, 2, &eta_scale);
            if (st != BF_L_OK) {
                out->status = st;
                return;
            }
        }
    }
    const v3f wi = mf_scaled(inf, f.face_sign);
    float value[4], pdf;
    const int s1 = mf_dielectric_eval(c, exterior, interior, ax, ay, wo, wi, mode, value);
    const int s2 = mf_dielectric_pdf(c, exterior, interior, ax, ay, wo, wi, mode, &pdf);
    if (s1 != BF_L_OK || s2 != BF_L_OK) {
        out->status = s1 != BF_L_OK ? s1 : s2;
        return;
    }
    copy4(out->value, value);
    out->incoming = wi;
    out->pdf = pdf;
    out->discrete = 0u;
    out->lobes = lobes;
    out->eta_scale = eta_scale;
    out->status = BF_L_OK;
}

/* direction-masked variants (:1554-1687); allowed is REFLECTION, TRANSMISSION or both */
BF_D int mf_valid_direction_mask(uint32_t allowed) {
    return allowed == BF_LOBE_REFLECTION || allowed == BF_LOBE_TRANSMISSION ||
           allowed == (BF_LOBE_REFLECTION | BF_LOBE_TRANSMISSION);
}

BF_D int mf_dielectric_eval_masked(const float* c, float exterior, float interior, float ax, float ay, v3f wo,
                                            v3f wi, uint32_t mode, uint32_t allowed, float* out) {
    if (!mf_valid_direction_mask(allowed)) return BF_L_INVALID;
    const int st = mf_dielectric_eval(c, exterior, interior, ax, ay, wo, wi, mode, out);
    if (st != BF_L_OK || wo.z == 0.0f || wi.z == 0.0f) return st;
    const uint32_t direction = (wo.z > 0.0f) == (wi.z > 0.0f) ? BF_LOBE_REFLECTION : BF_LOBE_TRANSMISSION;
    if (!(allowed & direction)) out[0] = out[1] = out[2] = out[3] = 0.0f;
    return BF_L_OK;
}

BF_D int mf_dielectric_pdf_masked(const float* c, float exterior, float interior, float ax, float ay, v3f wo,
                                           v3f wi, uint32_t mode, uint32_t allowed, float* out) {
    if (!mf_valid_direction_mask(allowed)) {
        *out = 0.0f;
        return BF_L_INVALID;
    }
    float base;
    int st = mf_dielectric_pdf(c, exterior, interior, ax, ay, wo, wi, mode, &base);
    if (st != BF_L_OK || wo.z == 0.0f || wi.z == 0.0f || base == 0.0f) {
        *out = st == BF_L_OK ? base : 0.0f;
        return st;
    }
    const int reflection = (wo.z > 0.0f) == (wi.z > 0.0f);
    const uint32_t direction = reflection ? BF_LOBE_REFLECTION : BF_LOBE_TRANSMISSION;
    *out = 0.0f;
    if (!(allowed & direction)) return BF_L_OK;
    if (allowed == (BF_LOBE_REFLECTION | BF_LOBE_TRANSMISSION)) {
        *out = base;
        return BF_L_OK;
    }
    const mf_interface f = mf_dielectric_interface(wo, exterior, interior);
    const v3f of = mf_scaled(wo, f.face_sign), inf = mf_scaled(wi, f.face_sign);
    float branch, fr;
    if (reflection) {
        v3f m;
        float half_cos;
        st = mf_reflection_geometry(of, inf, &m, &half_cos);
        if (st != BF_L_OK) return st;
        st = mf_dielectric_fresnel(half_cos, f.eta_i, f.eta_t, &fr);
        if (st != BF_L_OK) return st;
        branch = fr;
    } else {
        mf_transmission t;
        st = mf_transmission_geometry(of, inf, f, &t);
        if (st != BF_L_OK || !t.supported) return st;
        st = mf_dielectric_fresnel(t.o_dot_m, f.eta_i, f.eta_t, &fr);
        if (st != BF_L_OK) return st;
        branch = 1.0f - fr;
    }
    if (!bf_isfinite(branch) || !(branch > 0.0f) || branch > 1.0f) return BF_L_NOT_REPRESENTABLE;
    const float num[1] = {base};
    const float den[1] = {branch};
    return mf_product_ratio(num, 1, den, 1, out);
}

BF_D void mf_dielectric_sample_masked(const float* c, float exterior, float interior, float ax, float ay, v3f wo,
                                               float event, float u, float v, uint32_t mode, uint32_t allowed,
                                               lobe_sample_t* out) {
    if (!mf_valid_direction_mask(allowed) || !bf_isfinite(event) || event < 0.0f || !(event < 1.0f)) {
        empty_sample(*out, BF_L_INVALID);
        return;
    }
    float selected = event;
    if (allowed == BF_LOBE_REFLECTION) selected = 0.0f;
    else if (allowed == BF_LOBE_TRANSMISSION) selected = bf_next_toward_zero_f(1.0f);
    mf_dielectric_sample(c, exterior, interior, ax, ay, wo, selected, u, v, mode, out);
    if (out->status != BF_L_OK) return;
    const uint32_t direction = out->lobes & (BF_LOBE_REFLECTION | BF_LOBE_TRANSMISSION);
    if ((allowed & direction) != direction) {
        empty_sample(*out, BF_L_OUTSIDE);
        return;
    }
    float pdf = 0.0f;
    const int st = mf_dielectric_pdf_masked(c, exterior, interior, ax, ay, wo, out->incoming, mode, allowed, &pdf);
    if (st != BF_L_OK) {
        empty_sample(*out, st);
        return;
    }
    out->pdf = pdf;
}

/* ---- closure validation (:1960-2059) -------------------------------------------------------- */
BF_D int mf_valid_closure(const closure_t* c) {
    if (!reflectance_ok(c->weight)) return 0;
    const float* p = c->parameters;
    switch (c->kind) {
    case BF_CLOSURE_LAMBERTIAN:
        if (c->lobes != (BF_LOBE_DIFFUSE | BF_LOBE_REFLECTION)) return 0;
        for (int i = 0; i < 10; ++i)
            if (p[i] != 0.0f) return 0;
        return 1;
    case BF_CLOSURE_ROUGH_DIFFUSE:
        if (c->lobes != (BF_LOBE_DIFFUSE | BF_LOBE_REFLECTION) || !bf_isfinite(p[0]) || p[0] < 0.0f || p[0] > 1.0f) return 0;
        for (int i = 1; i < 10; ++i)
            if (p[i] != 0.0f) return 0;
        return 1;
    case BF_CLOSURE_ROUGH_CONDUCTOR:
        if (c->lobes != (BF_LOBE_GLOSSY | BF_LOBE_REFLECTION)) return 0;
        return mf_positive_spectrum(p) && mf_spectrum_finite_nonnegative(p + 4) && mf_ggx_widths(p[8], p[9]);
    case BF_CLOSURE_ROUGH_DIELECTRIC:
        if (c->lobes != (BF_LOBE_GLOSSY | BF_LOBE_REFLECTION | BF_LOBE_TRANSMISSION) ||
            !mf_dielectric_valid(c->weight, p[0], p[1], p[2], p[3]))
            return 0;
        for (int i = 4; i < 10; ++i)
            if (p[i] != 0.0f) return 0;
        return 1;
    case BF_CLOSURE_SPECULAR_REFLECTION: /* TransportLobesDevice.cuh:2033-2042 */
        if (c->lobes != (BF_LOBE_SPECULAR | BF_LOBE_REFLECTION)) return 0;
        for (int i = 0; i < 10; ++i)
            if (p[i] != 0.0f) return 0;
        return 1;
    case BF_CLOSURE_SPECULAR_TRANSMISSION: /* :2043-2055: exterior and interior eta */
        if (c->lobes != (BF_LOBE_SPECULAR | BF_LOBE_TRANSMISSION) || !bf_isfinite(p[0]) || !(p[0] > 0.0f) ||
            !bf_isfinite(p[1]) || !(p[1] > 0.0f))
            return 0;
        for (int i = 2; i < 10; ++i)
            if (p[i] != 0.0f) return 0;
        return 1;
    default:
        return 0;
    }
}

#endif
