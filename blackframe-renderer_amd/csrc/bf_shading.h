/*
 * Surface reconstruction, light sampling and path-vertex helpers of the shade stage.
 * Each function reproduces the arithmetic of the named reference function in
 * Blackframe/XPU/CUDA/Source/CudaWavefrontStages.cu.
 */
#ifndef BF_SHADING_H
#define BF_SHADING_H

#include "bf_device_lobes.h"
#include "bf_device_scene.h"

struct segment_t {
    v3f scaled, direction;
    float scale, cubed, length;
};

/* :351-388 */
BF_D bool scaled_segment(v3f disp, segment_t& s) {
    if (!fin3(disp)) return false;
    const float scale = fmaxf(fabsf(disp.x), fmaxf(fabsf(disp.y), fabsf(disp.z)));
    if (!bf_isfinite(scale) || !(scale > 0.0f)) return false;
    const v3f scaled = mk(disp.x / scale, disp.y / scale, disp.z / scale);
    const float sq = dot3(scaled, scaled);
    const float sl = sqrtf(sq);
    const float cubed = sq * sl;
    const float length = scale * sl;
    if (!fin3(scaled) || !bf_isfinite(sq) || !(sq > 0.0f) || !bf_isfinite(sl) || !(sl > 0.0f) || !bf_isfinite(cubed) ||
        !(cubed > 0.0f) || !bf_isfinite(length) || !(length > 0.0f))
        return false;
    const v3f dir = mul3(scaled, 1.0f / sl);
    if (!fin3(dir) || !unit_vector(dir)) return false;
    s.scaled = scaled;
    s.direction = dir;
    s.scale = scale;
    s.cubed = cubed;
    s.length = length;
    return true;
}

/* :390-419 */
BF_D bool light_surface_alignment(v3f n, v3f scaled_to_light, float& absolute, bool& supported) {
    const v3f out = mul3(scaled_to_light, -1.0f);
    const float sa = dot3(n, out);
    const bool ortho = (n.x == 0.0f || out.x == 0.0f) && (n.y == 0.0f || out.y == 0.0f) && (n.z == 0.0f || out.z == 0.0f);
    if (sa == 0.0f && ortho) {
        absolute = 0.0f;
        supported = false;
        return true;
    }
    const float sum = fmaf(fabsf(n.x), fabsf(out.x), fmaf(fabsf(n.y), fabsf(out.y), fabsf(n.z) * fabsf(out.z)));
    const float uncertainty = fmaf(4.0f * BF_EPS, sum, 4.0f * BF_DENORM_MIN);
    absolute = fabsf(sa);
    if (!bf_isfinite(sa) || !bf_isfinite(absolute) || !bf_isfinite(sum) || !bf_isfinite(uncertainty) || absolute <= uncertainty)
        return false;
    supported = sa > 0.0f;
    return true;
}

/* :421-465 */
BF_D bool one_sided_emission_support(v3f n, v3f wo, bool& supported) {
    const float sq = dot3(n, n);
    if (!fin3(n) || !bf_isfinite(sq) || fabsf(sq - 1.0f) > 128.0f * BF_EPS || !fin3(wo)) return false;
    const float m = fmaxf(fabsf(wo.x), fmaxf(fabsf(wo.y), fabsf(wo.z)));
    if (!bf_isfinite(m) || m == 0.0f) return false;
    const v3f s = mk(wo.x / m, wo.y / m, wo.z / m);
    const float al = dot3(n, s);
    if (!bf_isfinite(al)) return false;
    const bool ortho = (n.x == 0.0f || wo.x == 0.0f) && (n.y == 0.0f || wo.y == 0.0f) && (n.z == 0.0f || wo.z == 0.0f);
    if (al == 0.0f && ortho) {
        supported = false;
        return true;
    }
    const float sum = fmaf(fabsf(n.x), fabsf(s.x), fmaf(fabsf(n.y), fabsf(s.y), fabsf(n.z) * fabsf(s.z)));
    const float uncertainty = fmaf(4.0f * BF_EPS, sum, 4.0f * BF_DENORM_MIN);
    if (!bf_isfinite(sum) || !bf_isfinite(uncertainty) || fabsf(al) <= uncertainty) return false;
    supported = al > 0.0f;
    return true;
}

struct positioned_t {
    v3f position, error;
};

/* :1078-1110 */
BF_D bool interpolate_position_with_error(const v3f* v, v3f b, positioned_t& out) {
    float scale = 0.0f;
#pragma unroll
    for (uint32_t a = 0; a < 3; ++a) {
        const float first = comp3(v[0], a);
        scale = fmaxf(scale, fmaxf(fabsf(comp3(v[1], a) - first), fabsf(comp3(v[2], a) - first)));
    }
    const float floor_value = BF_EPS * scale;
    if (!bf_isfinite(scale) || !bf_isfinite(floor_value) || !(floor_value > 0.0f)) return false;
    const float gamma7 = (7.0f * BF_EPS) / (1.0f - 7.0f * BF_EPS);
#pragma unroll
    for (uint32_t a = 0; a < 3; ++a) {
        const float p0 = b.x * comp3(v[0], a);
        const float p1 = b.y * comp3(v[1], a);
        const float p2 = b.z * comp3(v[2], a);
        const float position = fmaf(b.x, comp3(v[0], a), fmaf(b.y, comp3(v[1], a), p2));
        const float magnitude = fabsf(p0) + fabsf(p1) + fabsf(p2);
        const float error = fmaxf(gamma7 * magnitude, floor_value);
        if (!bf_isfinite(position) || !bf_isfinite(magnitude) || !bf_isfinite(error) || error < 0.0f) return false;
        set3(out.position, a, position);
        set3(out.error, a, error);
    }
    return true;
}

/* :1112-1150 */
BF_D bool transform_position_with_error(const float* m, const positioned_t& obj, v3f traversal, positioned_t& out) {
    const double eps = (double)BF_EPS;
    const double gamma7 = (7.0 * eps) / (1.0 - 7.0 * eps);
    out.position = xform_point(m, obj.position);
    if (!fin3(out.position) || !fin3(traversal)) return false;
#pragma unroll
    for (uint32_t row = 0; row < 3; ++row) {
        double mag = fabs((double)m[row * 4 + 3]);
        double prop = 0.0;
#pragma unroll
        for (uint32_t col = 0; col < 3; ++col) {
            const double c = fabs((double)m[row * 4 + col]);
            mag += c * fabs((double)comp3(obj.position, col));
            prop += c * (double)comp3(obj.error, col);
        }
        const double discrepancy = fabs((double)comp3(out.position, row) - (double)comp3(traversal, row));
        const double bound = gamma7 * mag + (1.0 + gamma7) * prop + discrepancy;
        if (!bf_isfinited(bound) || bound < 0.0 || bound > (double)3.40282346638528859812e+38f) return false;
        float r = (float)bound;
        if ((double)r < bound) r = bf_next_up_f(r);
        if (!bf_isfinite(r) || r < 0.0f || (bound > 0.0 && r == 0.0f)) return false;
        set3(out.error, row, r);
    }
    return true;
}

/* :2596-2629 */
BF_D bool offset_point(v3f p, v3f err, v3f n, v3f w, v3f& out) {
    if (!fin3(p) || !fin3(err) || err.x < 0.0f || err.y < 0.0f || err.z < 0.0f || !fin3(n) || !fin3(w)) return false;
    const v3f oriented = dot3(n, w) >= 0.0f ? n : mul3(n, -1.0f);
    const float distance = fabsf(n.x) * err.x + fabsf(n.y) * err.y + fabsf(n.z) * err.z;
    if (!bf_isfinite(distance) || !(distance > 0.0f)) return false;
    const v3f disp = mul3(oriented, distance);
    v3f o = add3(p, disp);
    if (disp.x != 0.0f) o.x = disp.x > 0.0f ? bf_next_up_f(o.x) : bf_next_down_f(o.x);
    if (disp.y != 0.0f) o.y = disp.y > 0.0f ? bf_next_up_f(o.y) : bf_next_down_f(o.y);
    if (disp.z != 0.0f) o.z = disp.z > 0.0f ? bf_next_up_f(o.z) : bf_next_down_f(o.z);
    out = o;
    return fin3(o) && (o.x != p.x || o.y != p.y || o.z != p.z);
}

struct frame3 {
    v3f t, b, n;
};

/* :2692-2703 */
BF_D bool orthonormal_frame(const frame3& f) {
    const float tol = 128.0f * BF_EPS;
    return fin3(f.t) && fin3(f.b) && fin3(f.n) && fabsf(dot3(f.t, f.t) - 1.0f) <= tol &&
           fabsf(dot3(f.b, f.b) - 1.0f) <= tol && fabsf(dot3(f.n, f.n) - 1.0f) <= tol && fabsf(dot3(f.t, f.b)) <= tol &&
           fabsf(dot3(f.t, f.n)) <= tol && fabsf(dot3(f.b, f.n)) <= tol &&
           fabsf(dot3(cross3(f.t, f.b), f.n) - 1.0f) <= tol;
}

/* :2705-2732 */
BF_D bool frame_from_tangent(v3f normal, v3f seed, frame3& f) {
    v3f un, t, b;
    if (!normalize3(normal, un) || !normalize3(seed, t)) return false;
    seed = sub3(t, mul3(un, dot3(t, un)));
    if (!normalize3(seed, t)) return false;
    if (!normalize3(cross3(un, t), b) || !normalize3(cross3(b, un), t)) return false;
    const frame3 c{t, b, un};
    if (!orthonormal_frame(c)) return false;
    f = c;
    return true;
}

/* :2734-2748 */
BF_D bool make_local_frame(v3f normal, frame3& f) {
    v3f n;
    if (!normalize3(normal, n)) return false;
    const float sign = copysignf(1.0f, n.z);
    const float coefficient = -1.0f / (sign + n.z);
    const float product = n.x * n.y * coefficient;
    const v3f seed = mk(1.0f + sign * n.x * n.x * coefficient, sign * product, -sign * n.x);
    return frame_from_tangent(normal, seed, f);
}

/* :2750-2766 */
BF_D bool rotate_frame(const frame3& src, float angle, frame3& out) {
    if (!orthonormal_frame(src) || !bf_isfinite(angle) || angle < -BF_PI || !(angle < BF_PI)) return false;
    if (angle == 0.0f) {
        out = src;
        return true;
    }
    float s, c;
    bf_sincosf(angle, &s, &c);
    const v3f t = add3(mul3(src.t, c), mul3(src.b, s));
    return bf_isfinite(c) && bf_isfinite(s) && frame_from_tangent(src.n, t, out);
}

BF_D v3f to_local(const frame3& f, v3f d) { return mk(dot3(d, f.t), dot3(d, f.b), dot3(d, f.n)); }
BF_D bool to_world(const frame3& f, v3f l, v3f& out) {
    return normalize3(add3(add3(mul3(f.t, l.x), mul3(f.b, l.y)), mul3(f.n, l.z)), out);
}

/* :3788-3818 */
BF_D bool shading_correction(v3f ng, v3f ns, v3f wo, v3f wi, float& c) {
    const float og = dot3(ng, wo), os = dot3(ns, wo), ig = dot3(ng, wi), is = dot3(ns, wi);
    if (!unit_vector(ng) || !unit_vector(ns) || !unit_vector(wo) || !unit_vector(wi) || !(dot3(ng, ns) > 0.0f) ||
        !bf_isfinite(og) || !bf_isfinite(os) || !bf_isfinite(ig) || !bf_isfinite(is))
        return false;
    if (og == 0.0f || os == 0.0f || ig == 0.0f || is == 0.0f || bf_signbit(og) != bf_signbit(os) ||
        bf_signbit(ig) != bf_signbit(is)) {
        c = 0.0f;
        return true;
    }
    c = 1.0f;
    return true;
}

/* :3820-3840 */
BF_D bool geometric_event_support(v3f ng, v3f wo, v3f wi, uint32_t lobes, bool& supported) {
    if (!valid_surface_event(lobes)) return false;
    const float o = dot3(ng, wo), i = dot3(ng, wi);
    if (!bf_isfinite(o) || !bf_isfinite(i)) return false;
    if (o == 0.0f || i == 0.0f) {
        supported = false;
        return true;
    }
    const bool same = bf_signbit(o) == bf_signbit(i);
    supported = (lobes & BF_LOBE_REFLECTION) ? same : !same;
    return true;
}

/* uniform selection: TransportDevice.cuh:479-551 */
BF_D bool uniform_light_selection(uint32_t count, float u, uint32_t& index, float& probability) {
    if (count == 0 || !bf_isfinite(u) || u < 0.0f || !(u < 1.0f) || count > (1u << 24)) return false;
    const float sc = (float)count;
    uint32_t first = 1, last = count + 1;
    while (first < last) {
        const uint32_t mid = first + (last - first) / 2;
        const float boundary = mid == count ? 1.0f : (float)mid / sc;
        if (!(u < boundary)) first = mid + 1;
        else last = mid;
    }
    if (first == 0 || first > count) return false;
    const uint32_t i = first - 1;
    const float lower = i == 0 ? 0.0f : (float)i / sc;
    const float upper = first == count ? 1.0f : (float)first / sc;
    const float interval = upper - lower;
    if (!(lower <= u) || !(u < upper) || !(interval > 0.0f) || interval > 1.0f) return false;
    index = i;
    probability = interval;
    return true;
}

/* :2216-2230 */
BF_D bool uniform_light_probability(uint32_t count, uint32_t i, float& probability) {
    if (count == 0 || count > (1u << 24) || i >= count) return false;
    const float sc = (float)count;
    const float lower = i == 0 ? 0.0f : (float)i / sc;
    const float upper = i + 1 == count ? 1.0f : (float)(i + 1) / sc;
    probability = upper - lower;
    return probability > 0.0f && probability <= 1.0f;
}

/* Power-weighted selection over the host-built CDF (LightSampler::sample, Renderer/Source/
 * LightSampler.cpp:1023-1057): the first boundary above u closes the selected light's interval. */
BF_D bool cdf_light_selection(const float* cdf, uint32_t count, float u, uint32_t& index, float& probability) {
    if (count == 0 || !bf_isfinite(u) || u < 0.0f || !(u < 1.0f)) return false;
    uint32_t first = 1, last = count + 1;
    while (first < last) {
        const uint32_t mid = first + (last - first) / 2;
        if (!(u < cdf[mid])) first = mid + 1;
        else last = mid;
    }
    if (first > count) return false;
    const float p = cdf[first] - cdf[first - 1];
    if (!bf_isfinite(p) || !(p > 0.0f) || p > 1.0f) return false;
    index = first - 1;
    probability = p;
    return true;
}

/* LightSampler::probability (:1133-1151): a light without power has probability 0 */
BF_D bool cdf_light_probability(const float* cdf, uint32_t count, uint32_t i, float& probability) {
    if (i >= count) return false;
    const float p = cdf[i + 1] - cdf[i];
    if (!bf_isfinite(p) || p < 0.0f || p > 1.0f) return false;
    probability = p;
    return true;
}

#endif
