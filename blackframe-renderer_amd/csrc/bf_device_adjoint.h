/*
 * Closure derivatives for the path-replay adjoint (no reference counterpart: the reference
 * renders forward only). Two families of parameters are differentiable:
 *  - the closure weight (reflectance) per wavelength lane: d f / d w;
 *  - the GGX widths of rough conductors and dielectrics: d ln f / d alpha_x, d ln f / d alpha_y.
 * The GGX factor that depends on the widths is D(m) G2(wo, wi) with
 *   D   = 1 / (pi ax ay q^2),  q = (mx/ax)^2 + (my/ay)^2 + mz^2
 *   G2  = 2 r_o r_i / (r_o + r_i)   (height-correlated Smith, the reference's ggx_smith_g2),
 *   r(d)= 1 / sqrt(1 + (ax^2 dx^2 + ay^2 dy^2) / dz^2)
 * so, hand-differentiated,
 *   d ln D / d ax = -1/ax + 4 mx^2 / (ax^3 q)
 *   d r / d ax    = -r^3 ax dx^2 / dz^2
 *   d ln G2 / d ax = r_o'/r_o + r_i'/r_i - (r_o' + r_i') / (r_o + r_i)
 * (likewise for ay). Fresnel terms and Jacobians do not depend on the widths.
 */
#ifndef BF_DEVICE_ADJOINT_H
#define BF_DEVICE_ADJOINT_H

#include "bf_device_lobes.h"

struct alpha_gradient {
    float x, y;
};

BF_D float smith_ratio(float ax, float ay, v3f d, float& d_ax, float& d_ay) {
    const float z2 = d.z * d.z;
    const float a2 = (ax * ax * d.x * d.x + ay * ay * d.y * d.y) / z2;
    const float r = 1.0f / sqrtf(1.0f + a2);
    const float r3 = r * r * r;
    d_ax = -r3 * ax * (d.x * d.x / z2);
    d_ay = -r3 * ay * (d.y * d.y / z2);
    return r;
}

/* d ln(D(m) G2(a, b)) / d(ax, ay) with a, b in the upper hemisphere and m.z > 0. */
BF_D bool ggx_alpha_log_gradient(float ax, float ay, v3f m, v3f a, v3f b, alpha_gradient& g) {
    if (!(m.z > 0.0f) || !(a.z > 0.0f) || !(b.z > 0.0f)) return false;
    const float mx = m.x / ax, my = m.y / ay;
    const float q = mx * mx + my * my + m.z * m.z;
    const float dDx = -1.0f / ax + 4.0f * (mx * mx) / (ax * q);
    const float dDy = -1.0f / ay + 4.0f * (my * my) / (ay * q);
    float rax, ray, rbx, rby;
    const float ra = smith_ratio(ax, ay, a, rax, ray);
    const float rb = smith_ratio(ax, ay, b, rbx, rby);
    const float sum = ra + rb;
    g.x = dDx + rax / ra + rbx / rb - (rax + rbx) / sum;
    g.y = dDy + ray / ra + rby / rb - (ray + rby) / sum;
    return bf_isfinite(g.x) && bf_isfinite(g.y);
}

/* d ln f / d(alpha_x, alpha_y) of a rough conductor or dielectric closure at (wo, wi) in the
 * closure frame; false where the closure value is zero (no gradient flows there). */
BF_D bool closure_alpha_log_gradient(const closure_t& c, v3f wo, v3f wi, alpha_gradient& g) {
    const float* p = c.parameters;
    if (c.kind == BF_CLOSURE_ROUGH_CONDUCTOR) {
        if (!(wo.z > 0.0f) || !(wi.z > 0.0f)) return false;
        v3f m;
        float half_cos;
        if (mf_reflection_geometry(wo, wi, &m, &half_cos) != BF_L_OK) return false;
        return ggx_alpha_log_gradient(p[8], p[9], m, wo, wi, g);
    }
    if (c.kind == BF_CLOSURE_ROUGH_DIELECTRIC) {
        if (wo.z == 0.0f || wi.z == 0.0f) return false;
        const mf_interface f = mf_dielectric_interface(wo, p[0], p[1]);
        const v3f of = mf_scaled(wo, f.face_sign), inf = mf_scaled(wi, f.face_sign);
        if (inf.z > 0.0f) {
            v3f m;
            float half_cos;
            if (mf_reflection_geometry(of, inf, &m, &half_cos) != BF_L_OK) return false;
            return ggx_alpha_log_gradient(p[2], p[3], m, of, inf, g);
        }
        mf_transmission t;
        if (mf_transmission_geometry(of, inf, f, &t) != BF_L_OK || !t.supported) return false;
        return ggx_alpha_log_gradient(p[2], p[3], t.m, of, mf_scaled(inf, -1.0f), g);
    }
    return false;
}

/* d f / d w per lane, where w is the closure's weight (reflectance) at that lane. Lambert and the
 * GGX closures are linear in w (the value at unit weight); the energy-preserving Oren-Nayar
 * multiple-scattering term w^2 A / (1 - w + w A) is not:
 *   d/dw [w s + w^2 A L / (pi (1 - w + w A))] = s + (A L / pi) w (2 - w + w A) / (1 - w + w A)^2. */
BF_D int closure_weight_derivative(const closure_t& c, uint32_t allowed, v3f wo, v3f wi, float* out) {
    closure_t unit = c;
#pragma unroll
    for (int k = 0; k < 4; ++k) unit.weight[k] = 1.0f;
    if (c.kind != BF_CLOSURE_ROUGH_DIFFUSE || c.parameters[0] == 0.0f) return record_eval(unit, allowed, wo, wi, out);
    /* the terms of mf_rough_diffuse_eval, differentiated in w */
    const float roughness = c.parameters[0];
    if (!(wo.z > 0.0f) || !(wi.z > 0.0f)) {
        out[0] = out[1] = out[2] = out[3] = 0.0f;
        return BF_L_OK;
    }
    const float fon = mf_fon_coefficient();
    const float tangent_dot = fmaf(wo.x, wi.x, wo.y * wi.y);
    const float s_over_t = tangent_dot > 0.0f ? tangent_dot / fmaxf(wo.z, wi.z) : tangent_dot;
    const float a = 1.0f / (1.0f + fon * roughness);
    const float single = BF_INV_PI * a * (1.0f + roughness * s_over_t);
    float lo, li;
    if (mf_rough_diffuse_loss_shape(wo, &lo) != BF_L_OK || mf_rough_diffuse_loss_shape(wi, &li) != BF_L_OK)
        return BF_L_NOT_REPRESENTABLE;
    const float average_albedo = a * (1.0f + mf_fon_average_coefficient() * roughness);
    const float average_loss = fon - mf_fon_average_coefficient();
    const float loss_over_average = a * roughness * li * lo / average_loss;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float w = c.weight[k];
        const float den = (1.0f - w) + w * average_albedo;
        out[k] = single + average_albedo * BF_INV_PI * loss_over_average * w * (2.0f - w + w * average_albedo) /
                              (den * den);
    }
    return BF_L_OK;
}

/* d f / d w per lane of a specular (delta) closure along the delta direction `wi`: its atom's value
 * at unit weight, which the atom scales linearly (scaled_delta_value, TransportLobesDevice.cuh:
 * 1784-1797). False when the closure does not send wo to wi, i.e. when it is not part of the
 * sampled delta atom: then the continuation's throughput does not depend on its weight. */
BF_D bool delta_weight_derivative(const closure_t& c, v3f wo, v3f wi, float* out) {
    if (c.kind != BF_CLOSURE_SPECULAR_REFLECTION && c.kind != BF_CLOSURE_SPECULAR_TRANSMISSION) return false;
    const float unit[4] = {1.0f, 1.0f, 1.0f, 1.0f};
    lobe_sample_t s;
    empty_sample(s, BF_L_INVALID);
    if (c.kind == BF_CLOSURE_SPECULAR_REFLECTION) specular_reflection_sample(unit, wo, s);
    else specular_transmission_sample(unit, c.parameters[0], c.parameters[1], wo, s);
    if (s.status != BF_L_OK || !s.discrete || s.incoming.x != wi.x || s.incoming.y != wi.y || s.incoming.z != wi.z)
        return false;
#pragma unroll
    for (int k = 0; k < 4; ++k) out[k] = s.value[k];
    return true;
}

#endif
