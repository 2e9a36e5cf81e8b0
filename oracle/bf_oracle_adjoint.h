/*
 * ORACLE — test infrastructure only (see bf_oracle.h).
 * Binary64 closure derivatives for the reference adjoint (the reference itself renders forward
 * only, so these restate mathematics, not reference code):
 *  - d f / d w per lane for the closure weight w: Lambert and GGX closures are linear in w; the
 *    energy-preserving Oren-Nayar multiple-scattering term w^2 A L / (pi (1 - w + w A)) is not,
 *    d/dw = s + (A L / pi) w (2 - w + w A) / (1 - w + w A)^2;
 *  - d ln f / d(alpha_x, alpha_y) of GGX closures: only D(m) and the height-correlated Smith
 *    G2 = 2 r_o r_i / (r_o + r_i) depend on the widths, with
 *      D = 1 / (pi ax ay q^2), q = (mx/ax)^2 + (my/ay)^2 + mz^2,
 *      r(d) = 1 / sqrt(1 + (ax^2 dx^2 + ay^2 dy^2) / dz^2),
 *      d ln D / d ax = -1/ax + 4 mx^2 / (ax^3 q),  d r / d ax = -r^3 ax dx^2 / dz^2,
 *      d ln G2 / d ax = r_o'/r_o + r_i'/r_i - (r_o' + r_i') / (r_o + r_i).
 * The microfacet normal comes from the same float geometry the closure evaluation uses.
 */
#ifndef BF_ORACLE_ADJOINT_H
#define BF_ORACLE_ADJOINT_H

#include "bf_oracle_lobes.h"

static double oa_smith_ratio(double ax, double ay, v3 d, double* d_ax, double* d_ay) {
    const double x = d.x, y = d.y, z = d.z;
    const double z2 = z * z;
    const double r = 1.0 / sqrt(1.0 + (ax * ax * x * x + ay * ay * y * y) / z2);
    const double r3 = r * r * r;
    *d_ax = -r3 * ax * x * x / z2;
    *d_ay = -r3 * ay * y * y / z2;
    return r;
}

static int oa_ggx_alpha_log_gradient(double ax, double ay, v3 m, v3 a, v3 b, double g[2]) {
    if (!(m.z > 0.0f) || !(a.z > 0.0f) || !(b.z > 0.0f)) return 0;
    const double mx = m.x, my = m.y, mz = m.z;
    const double q = mx * mx / (ax * ax) + my * my / (ay * ay) + mz * mz;
    const double dDx = -1.0 / ax + 4.0 * mx * mx / (ax * ax * ax * q);
    const double dDy = -1.0 / ay + 4.0 * my * my / (ay * ay * ay * q);
    double rax, ray, rbx, rby;
    const double ra = oa_smith_ratio(ax, ay, a, &rax, &ray);
    const double rb = oa_smith_ratio(ax, ay, b, &rbx, &rby);
    g[0] = dDx + rax / ra + rbx / rb - (rax + rbx) / (ra + rb);
    g[1] = dDy + ray / ra + rby / rb - (ray + rby) / (ra + rb);
    return isfinite(g[0]) && isfinite(g[1]);
}

static int oa_closure_alpha_log_gradient(const closure_rec* c, v3 wo, v3 wi, double g[2]) {
    const float* p = c->parameters;
    if (c->kind == BF_CLOSURE_ROUGH_CONDUCTOR) {
        if (!(wo.z > 0.0f) || !(wi.z > 0.0f)) return 0;
        v3 m;
        float half_cos;
        if (mf_reflection_geometry(wo, wi, &m, &half_cos) != LOBE_OK) return 0;
        return oa_ggx_alpha_log_gradient(p[8], p[9], m, wo, wi, g);
    }
    if (c->kind == BF_CLOSURE_ROUGH_DIELECTRIC) {
        if (wo.z == 0.0f || wi.z == 0.0f) return 0;
        const mf_interface f = mf_dielectric_interface(wo, p[0], p[1]);
        const v3 of = v3_mul(wo, f.face_sign), inf = v3_mul(wi, f.face_sign);
        if (inf.z > 0.0f) {
            v3 m;
            float half_cos;
            if (mf_reflection_geometry(of, inf, &m, &half_cos) != LOBE_OK) return 0;
            return oa_ggx_alpha_log_gradient(p[2], p[3], m, of, inf, g);
        }
        mf_transmission t;
        if (mf_transmission_geometry(of, inf, f, &t) != LOBE_OK || !t.supported) return 0;
        return oa_ggx_alpha_log_gradient(p[2], p[3], t.m, of, v3_mul(inf, -1.0f), g);
    }
    return 0;
}

static int oa_closure_weight_derivative(const closure_rec* c, uint32_t allowed, v3 wo, v3 wi, double out[4]) {
    if (c->kind != BF_CLOSURE_ROUGH_DIFFUSE || c->parameters[0] == 0.0f) {
        closure_rec unit = *c;
        for (int k = 0; k < 4; ++k) unit.weight[k] = 1.0f;
        float v[4];
        if (lobe_record_eval(&unit, allowed, wo, wi, v) != LOBE_OK) return 0;
        for (int k = 0; k < 4; ++k) out[k] = v[k];
        return 1;
    }
    if (!(wo.z > 0.0f) || !(wi.z > 0.0f)) {
        for (int k = 0; k < 4; ++k) out[k] = 0.0;
        return 1;
    }
    const double inv_pi = 0.31830988618379067154;
    const double roughness = c->parameters[0];
    const double fon = 0.5 - 2.0 * inv_pi / 3.0, fon_average = 2.0 / 3.0 - 28.0 * inv_pi / 15.0;
    const double tangent_dot = (double)wo.x * wi.x + (double)wo.y * wi.y;
    const double s_over_t = tangent_dot > 0.0 ? tangent_dot / fmax(wo.z, wi.z) : tangent_dot;
    const double a = 1.0 / (1.0 + fon * roughness);
    const double single = inv_pi * a * (1.0 + roughness * s_over_t);
    float lo, li;
    if (mf_rough_diffuse_loss_shape(wo, &lo) != LOBE_OK || mf_rough_diffuse_loss_shape(wi, &li) != LOBE_OK) return 0;
    const double average_albedo = a * (1.0 + fon_average * roughness);
    const double loss_over_average = a * roughness * (double)li * lo / (fon - fon_average);
    for (int k = 0; k < 4; ++k) {
        const double w = c->weight[k];
        const double den = (1.0 - w) + w * average_albedo;
        out[k] = single + average_albedo * inv_pi * loss_over_average * w * (2.0 - w + w * average_albedo) / (den * den);
    }
    return 1;
}

/* d f / d w per lane of a specular (delta) closure along the delta direction wi: the value of its
 * atom at unit weight (the atom is linear in the weight, scaled_delta_value,
 * TransportLobesDevice.cuh:1784-1797); 0 when the closure's atom does not point along wi. */
static int oa_delta_weight_derivative(const closure_rec* c, v3 wo, v3 wi, double out[4]) {
    if (c->kind != BF_CLOSURE_SPECULAR_REFLECTION && c->kind != BF_CLOSURE_SPECULAR_TRANSMISSION) return 0;
    const float unit[4] = {1.0f, 1.0f, 1.0f, 1.0f};
    lobe_sample s;
    if (c->kind == BF_CLOSURE_SPECULAR_REFLECTION) lobe_specular_reflection_sample(unit, wo, &s);
    else lobe_specular_transmission_sample(unit, c->parameters[0], c->parameters[1], wo, &s);
    if (s.status != LOBE_OK || s.measure != LOBE_MEASURE_DISCRETE || s.incoming.x != wi.x || s.incoming.y != wi.y ||
        s.incoming.z != wi.z)
        return 0;
    for (int k = 0; k < 4; ++k) out[k] = s.value[k];
    return 1;
}

#endif
