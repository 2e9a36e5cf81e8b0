/*
 * Normal and bump maps on the device: EWA filtering of mip-mapped image textures over a ray cone's
 * footprint, and the shading-normal perturbation each map makes. Follows
 * Blackframe/XPU/CUDA/Source/CudaWavefrontStages.cu:2992-3673 (wrap_texture_index,
 * cone_texture_differentials, make_ewa_footprint, make_ewa_ellipse, sample_ewa_ellipse,
 * filter_image_ewa, projected_tangent_frame, apply_normal_map, apply_bump_map). Image tables and
 * map parameters are validated when the scene is created, so the checks here are the numerical
 * ones the reference makes per lookup. Transcendentals come from blackframe_math.h.
 */
#ifndef BF_DEVICE_MAPS_H
#define BF_DEVICE_MAPS_H

#include "bf_device_math.h"
#include "bf_device_scene.h"

enum bf_map_status { BF_MAP_OK = 0, BF_MAP_INVALID = 1, BF_MAP_NUMERICAL = 2, BF_MAP_UNSUPPORTED = 3 };

#define BF_EXACT_TEXEL_LIMIT 8388608.0f

struct map_differentials {
    float dudx, dvdx, dudy, dvdy;
};

struct map_mip {
    uint32_t width, height;
    uint64_t texel_base; /* index of the level's first value in the scene's texel array */
};

struct ewa_footprint {
    float major_x, major_y, minor_x, minor_y, major_radius, minor_radius, minor_length;
};

struct ewa_ellipse {
    float center_x, center_y, major_x, major_y, minor_x, minor_y, inv_major, inv_minor;
    int64_t x0, x1, y0, y1;
    uint64_t visits;
};

/* wrap_texture_index (:2992-3034): repeat, clamp, mirror, or border (present = false outside) */
BF_D void wrap_index(int64_t index, uint32_t extent, uint32_t mode, bool& present, uint32_t& wrapped) {
    const int64_t n = (int64_t)extent;
    present = true;
    if (mode == BF_TEXTURE_WRAP_REPEAT) {
        int64_t r = index % n;
        wrapped = (uint32_t)(r < 0 ? r + n : r);
    } else if (mode == BF_TEXTURE_WRAP_CLAMP) {
        wrapped = index < 0 ? 0u : (index >= n ? extent - 1u : (uint32_t)index);
    } else if (mode == BF_TEXTURE_WRAP_MIRROR) {
        int64_t phase = index % (2 * n);
        if (phase < 0) phase += 2 * n;
        wrapped = (uint32_t)(phase < n ? phase : 2 * n - 1 - phase);
    } else {
        present = index >= 0 && index < n;
        wrapped = present ? (uint32_t)index : 0u;
    }
}

/* cone_texture_differentials (:3036-3088): the cone's circular cross-section at the hit, projected
 * onto the surface and pulled back to (u, v) through the metric of dp/du, dp/dv, as the lower
 * triangular factor of its covariance */
BF_D bool cone_differentials(float2 cone, v3f ray_direction, v3f ng, v3f dpdu, v3f dpdv, map_differentials& out) {
    v3f d;
    if (!bf_isfinite(cone.x) || cone.x < 0.0f || !bf_isfinite(cone.y) || cone.y < 0.0f || !normalize3(ray_direction, d))
        return false;
    const float cosine = fabsf(dot3(d, ng));
    if (!bf_isfinite(cosine) || !(cosine > 0.0f)) return false;
    const float radius = cone.x / cosine;
    const float guu = dot3(dpdu, dpdu), guv = dot3(dpdu, dpdv), gvv = dot3(dpdv, dpdv);
    const float det = fmaf(guu, gvv, -guv * guv);
    if (!bf_isfinite(radius) || radius < 0.0f || !bf_isfinite(guu) || !bf_isfinite(guv) || !bf_isfinite(gvv) ||
        !(guu > 0.0f) || !(gvv > 0.0f) || !bf_isfinite(det) || !(det > 0.0f) || radius == 0.0f)
        return false;
    const float r2 = radius * radius;
    const float cuu = r2 * gvv / det, cuv = -r2 * guv / det, cvv = r2 * guu / det;
    if (!bf_isfinite(r2) || !(r2 > 0.0f) || !bf_isfinite(cuu) || !bf_isfinite(cuv) || !bf_isfinite(cvv) ||
        !(cuu > 0.0f) || !(cvv > 0.0f))
        return false;
    const float dudx = sqrtf(cuu);
    const float dvdx = cuv / dudx;
    float rest = fmaf(-dvdx, dvdx, cvv);
    if (rest < 0.0f && rest >= -(64.0f * BF_EPS * cvv)) rest = 0.0f;
    const float dvdy = sqrtf(rest);
    out = map_differentials{dudx, dvdx, 0.0f, dvdy};
    return bf_isfinite(dudx) && dudx > 0.0f && bf_isfinite(dvdx) && bf_isfinite(dvdy) && dvdy > 0.0f;
}

/* make_ewa_footprint (:3090-3141): principal axes of the footprint's texel-space covariance, the
 * minor axis widened to at most max_anisotropy : 1, radii grown by one texel (hypot(1, length)) */
BF_D bool ewa_make_footprint(map_mip mip, const map_differentials& g, uint32_t max_anisotropy, ewa_footprint& f) {
    const float w = (float)mip.width, h = (float)mip.height;
    const float dudx = g.dudx * w, dvdx = g.dvdx * h, dudy = g.dudy * w, dvdy = g.dvdy * h;
    if (!bf_isfinite(dudx) || !bf_isfinite(dvdx) || !bf_isfinite(dudy) || !bf_isfinite(dvdy) ||
        fabsf(dudx) >= BF_EXACT_TEXEL_LIMIT || fabsf(dvdx) >= BF_EXACT_TEXEL_LIMIT ||
        fabsf(dudy) >= BF_EXACT_TEXEL_LIMIT || fabsf(dvdy) >= BF_EXACT_TEXEL_LIMIT)
        return false;
    const float cxx = fmaf(dudx, dudx, dudy * dudy);
    const float cxy = fmaf(dudx, dvdx, dudy * dvdy);
    const float cyy = fmaf(dvdx, dvdx, dvdy * dvdy);
    const float trace = cxx + cyy;
    const float disc = bf_hypotf(cxx - cyy, 2.0f * cxy);
    const float lmax = 0.5f * (trace + disc);
    const float jac = fmaf(dudx, dvdy, -dudy * dvdx);
    float lmin = lmax > 0.0f ? (jac * jac) / lmax : 0.0f;
    const float a = (float)max_anisotropy;
    lmin = fmaxf(lmin, lmax / (a * a));
    const float angle = lmax > 0.0f ? 0.5f * bf_atan2f(2.0f * cxy, cxx - cyy) : 0.0f;
    float sn, cs;
    bf_sincosf(angle, &sn, &cs);
    const float major_len = sqrtf(lmax), minor_len = sqrtf(lmin);
    f = ewa_footprint{cs, sn, -sn, cs, bf_hypotf(1.0f, major_len), bf_hypotf(1.0f, minor_len), minor_len};
    return bf_isfinite(cxx) && bf_isfinite(cxy) && bf_isfinite(cyy) && bf_isfinite(lmax) && lmax >= 0.0f &&
           bf_isfinite(lmin) && lmin >= 0.0f && bf_isfinite(f.major_x) && bf_isfinite(f.major_y) &&
           bf_isfinite(f.major_radius) && bf_isfinite(f.minor_radius) && bf_isfinite(f.minor_length);
}

/* make_ewa_ellipse (:3143-3196): the footprint centred on the lookup in texel space (texel centres
 * at integers) and the integer bounding box of texels it covers */
BF_D bool ewa_make_ellipse(map_mip mip, float u, float v, const ewa_footprint& f, ewa_ellipse& e) {
    const float sx = u * (float)mip.width, sy = v * (float)mip.height;
    if (!bf_isfinite(u) || !bf_isfinite(v) || !bf_isfinite(sx) || !bf_isfinite(sy) || sx <= -BF_EXACT_TEXEL_LIMIT ||
        sx >= BF_EXACT_TEXEL_LIMIT || sy <= -BF_EXACT_TEXEL_LIMIT || sy >= BF_EXACT_TEXEL_LIMIT)
        return false;
    e.center_x = sx - 0.5f;
    e.center_y = sy - 0.5f;
    e.major_x = f.major_x;
    e.major_y = f.major_y;
    e.minor_x = f.minor_x;
    e.minor_y = f.minor_y;
    e.inv_major = 1.0f / f.major_radius;
    e.inv_minor = 1.0f / f.minor_radius;
    const float rx = bf_hypotf(f.major_radius * f.major_x, f.minor_radius * f.minor_x);
    const float ry = bf_hypotf(f.major_radius * f.major_y, f.minor_radius * f.minor_y);
    const float x0 = ceilf(e.center_x - rx), x1 = floorf(e.center_x + rx);
    const float y0 = ceilf(e.center_y - ry), y1 = floorf(e.center_y + ry);
    if (!bf_isfinite(rx) || !bf_isfinite(ry) || !bf_isfinite(x0) || !bf_isfinite(x1) || !bf_isfinite(y0) ||
        !bf_isfinite(y1) || x0 <= -BF_EXACT_TEXEL_LIMIT || x1 >= BF_EXACT_TEXEL_LIMIT || y0 <= -BF_EXACT_TEXEL_LIMIT ||
        y1 >= BF_EXACT_TEXEL_LIMIT)
        return false;
    e.x0 = (int64_t)x0;
    e.x1 = (int64_t)x1;
    e.y0 = (int64_t)y0;
    e.y1 = (int64_t)y1;
    if (e.x1 < e.x0 || e.y1 < e.y0) return false;
    e.visits = (uint64_t)(e.x1 - e.x0 + 1) * (uint64_t)(e.y1 - e.y0 + 1);
    return true;
}

/* sample_ewa_ellipse (:3198-3269): Gaussian weights exp(-alpha r^2) - exp(-alpha) (alpha = 2,
 * written as exp(-alpha) expm1(alpha (1 - r^2))) inside the unit ellipse, accumulated as a running
 * weighted mean; border texels outside the image read 0 */
BF_D int ewa_sample(const float* texels, uint32_t channels, map_mip mip, const ewa_ellipse& e, uint32_t channel,
                    uint32_t u_wrap, uint32_t v_wrap, float& out) {
    const float alpha = 2.0f;
    const float edge = bf_expf(-alpha);
    float total = 0.0f;
    out = 0.0f;
    for (int64_t y = e.y0; y <= e.y1; ++y) {
        bool y_in;
        uint32_t wy;
        wrap_index(y, mip.height, v_wrap, y_in, wy);
        const float oy = (float)y - e.center_y;
        for (int64_t x = e.x0; x <= e.x1; ++x) {
            const float ox = (float)x - e.center_x;
            const float a = fmaf(ox, e.major_x, oy * e.major_y) * e.inv_major;
            const float b = fmaf(ox, e.minor_x, oy * e.minor_y) * e.inv_minor;
            const float r2 = fmaf(a, a, b * b);
            if (!bf_isfinite(r2) || r2 < 0.0f) return BF_MAP_NUMERICAL;
            if (r2 >= 1.0f) continue;
            const float weight = edge * bf_expm1f(alpha * (1.0f - r2));
            if (!bf_isfinite(weight) || !(weight > 0.0f)) return BF_MAP_NUMERICAL;
            bool x_in;
            uint32_t wx;
            wrap_index(x, mip.width, u_wrap, x_in, wx);
            float value = 0.0f;
            if (x_in && y_in) {
                value = texels[mip.texel_base + ((uint64_t)wy * mip.width + wx) * channels + channel];
                if (!bf_isfinite(value)) return BF_MAP_NUMERICAL;
            }
            const float next = total + weight;
            if (!bf_isfinite(next) || !(next > 0.0f)) return BF_MAP_NUMERICAL;
            if (total == 0.0f) out = value;
            else out += (value - out) * (weight / next);
            total = next;
        }
    }
    return total > 0.0f && bf_isfinite(out) ? BF_MAP_OK : BF_MAP_NUMERICAL;
}

BF_D map_mip image_level(const bf_dev_image& img, uint32_t level) {
    map_mip m{img.width, img.height, img.texel_base};
    for (uint32_t k = 0; k < level; ++k) {
        m.texel_base += (uint64_t)m.width * m.height * img.channels;
        m.width = m.width > 1u ? m.width / 2u : 1u;
        m.height = m.height > 1u ? m.height / 2u : 1u;
    }
    return m;
}

/* filter_image_ewa (:3271-3355): the finest level whose footprint's minor axis is at most one texel
 * (blended in log(minor length) with the level below it), or the coarsest level filtered as a
 * point when none is; refused when the ellipses would read more than max_visits texels */
BF_D int ewa_filter(const bf_dev_scene& S, uint32_t image, float u, float v, const map_differentials& g,
                    uint32_t channel, uint32_t u_wrap, uint32_t v_wrap, uint32_t max_anisotropy, uint32_t max_visits,
                    float& out) {
    const bf_dev_image img = S.images[image];
    if (!bf_isfinite(u) || !bf_isfinite(v)) return BF_MAP_INVALID;
    map_mip lower = image_level(img, 0), upper{};
    ewa_footprint lf, uf{};
    if (!ewa_make_footprint(lower, g, max_anisotropy, lf)) return BF_MAP_NUMERICAL;
    bool has_upper = false;
    float fraction = 0.0f;
    if (lf.minor_length > 1.0f) {
        for (uint32_t level = 1; level < img.mips; ++level) {
            const map_mip cm = image_level(img, level);
            ewa_footprint cf;
            if (!ewa_make_footprint(cm, g, max_anisotropy, cf)) return BF_MAP_NUMERICAL;
            if (cf.minor_length <= 1.0f) {
                const float ll = bf_logf(lf.minor_length), ul = bf_logf(cf.minor_length);
                const float den = ll - ul;
                const float blend = ll / den;
                if (!bf_isfinite(blend) || !(den > 0.0f) || blend < 0.0f || blend > 1.0f) return BF_MAP_NUMERICAL;
                if (blend == 1.0f) {
                    lower = cm;
                    lf = cf;
                } else {
                    upper = cm;
                    uf = cf;
                    has_upper = true;
                    fraction = blend;
                }
                break;
            }
            lower = cm;
            lf = cf;
        }
    }
    /* a footprint wider than the coarsest level: a one-texel Gaussian there (EwaFootprint's defaults) */
    if (!has_upper && lf.minor_length > 1.0f) lf = ewa_footprint{1.0f, 0.0f, 0.0f, 1.0f, 1.0f, 1.0f, 0.0f};
    ewa_ellipse le, ue;
    if (!ewa_make_ellipse(lower, u, v, lf, le)) return BF_MAP_NUMERICAL;
    uint64_t visits = le.visits;
    if (has_upper) {
        if (!ewa_make_ellipse(upper, u, v, uf, ue) || visits > ~0ull - ue.visits) return BF_MAP_NUMERICAL;
        visits += ue.visits;
    }
    if (visits > max_visits) return BF_MAP_UNSUPPORTED;
    float lo = 0.0f;
    int st = ewa_sample(S.image_texels, img.channels, lower, le, channel, u_wrap, v_wrap, lo);
    if (st != BF_MAP_OK) return st;
    if (!has_upper) {
        out = lo;
        return BF_MAP_OK;
    }
    float hi = 0.0f;
    st = ewa_sample(S.image_texels, img.channels, upper, ue, channel, u_wrap, v_wrap, hi);
    if (st != BF_MAP_OK) return st;
    out = lo + fraction * (hi - lo);
    return bf_isfinite(out) ? BF_MAP_OK : BF_MAP_NUMERICAL;
}

/* projected_tangent_frame (:3357-3381): dp/du made orthogonal to the shading normal, and the
 * bitangent n x t signed by dp/dv */
BF_D bool projected_frame(v3f ns, v3f dpdu, v3f dpdv, v3f& tangent, v3f& bitangent) {
    const float along = dot3(ns, dpdu);
    const v3f projected = mk(fmaf(-along, ns.x, dpdu.x), fmaf(-along, ns.y, dpdu.y), fmaf(-along, ns.z, dpdu.z));
    v3f base;
    if (!normalize3(projected, tangent) || !normalize3(cross3(ns, tangent), base)) return false;
    const float scale = fmaxf(fabsf(dpdv.x), fmaxf(fabsf(dpdv.y), fabsf(dpdv.z)));
    if (!bf_isfinite(scale) || scale == 0.0f) return false;
    const float hand = dot3(base, mul3(dpdv, 1.0f / scale));
    if (!bf_isfinite(hand) || hand == 0.0f) return false;
    bitangent = hand > 0.0f ? base : mul3(base, -1.0f);
    return true;
}

/* normalized_cross (:3383-3393): the cross product of the two vectors, each scaled to a largest
 * component of 1 first */
BF_D bool scaled_cross(v3f a, v3f b, v3f& out) {
    const float sa = fmaxf(fabsf(a.x), fmaxf(fabsf(a.y), fabsf(a.z)));
    const float sb = fmaxf(fabsf(b.x), fmaxf(fabsf(b.y), fabsf(b.z)));
    return bf_isfinite(sa) && bf_isfinite(sb) && sa > 0.0f && sb > 0.0f &&
           normalize3(cross3(mul3(a, 1.0f / sa), mul3(b, 1.0f / sb)), out);
}

/* apply_normal_map (:3395-3494) */
BF_D int apply_normal_map(const bf_dev_scene& S, const bf_surface_maps_desc& m, const map_differentials& g, float u,
                          float v, v3f ng, v3f dpdu, v3f dpdv, v3f& ns) {
    float enc[3];
    for (uint32_t c = 0; c < 3u; ++c) {
        const int st = ewa_filter(S, m.normal_texture, u, v, g, m.normal_channels[c], m.normal_u_wrap, m.normal_v_wrap,
                                  m.normal_max_anisotropy, m.normal_max_texel_visits, enc[c]);
        if (st != BF_MAP_OK) return st;
        if (enc[c] < 0.0f || enc[c] > 1.0f) return BF_MAP_NUMERICAL;
    }
    v3f mapped;
    if (!normalize3(mk(fmaf(2.0f, enc[0], -1.0f), fmaf(2.0f, enc[1], -1.0f), fmaf(2.0f, enc[2], -1.0f)), mapped) ||
        !(mapped.z > 0.0f))
        return BF_MAP_NUMERICAL;
    v3f t, b;
    if (!projected_frame(ns, dpdu, dpdv, t, b)) return BF_MAP_NUMERICAL;
    const float my = m.normal_y_convention == 0u ? mapped.y : -mapped.y;
    const v3f n = mk(fmaf(t.x, mapped.x, fmaf(b.x, my, ns.x * mapped.z)), fmaf(t.y, mapped.x, fmaf(b.y, my, ns.y * mapped.z)),
                     fmaf(t.z, mapped.x, fmaf(b.z, my, ns.z * mapped.z)));
    v3f out;
    if (!normalize3(n, out) || !(dot3(ng, out) > 0.0f)) return BF_MAP_NUMERICAL;
    ns = out;
    return BF_MAP_OK;
}

/* apply_bump_map (:3496-3620): central differences of the filtered height one footprint (at least
 * one texel) either side in u and in v, applied along the shading normal to the projected dp/du
 * and dp/dv */
BF_D int apply_bump_map(const bf_dev_scene& S, const bf_surface_maps_desc& m, const map_differentials& g, float u,
                        float v, v3f ng, v3f dpdu, v3f dpdv, v3f& ns) {
    const bf_dev_image img = S.images[m.bump_texture];
    const float au = dot3(ns, dpdu), av = dot3(ns, dpdv);
    const v3f pu = mk(fmaf(-au, ns.x, dpdu.x), fmaf(-au, ns.y, dpdu.y), fmaf(-au, ns.z, dpdu.z));
    const v3f pv = mk(fmaf(-av, ns.x, dpdv.x), fmaf(-av, ns.y, dpdv.y), fmaf(-av, ns.z, dpdv.z));
    v3f orientation;
    if (!bf_isfinite(au) || !bf_isfinite(av) || !fin3(pu) || !fin3(pv) || !scaled_cross(pu, pv, orientation))
        return BF_MAP_NUMERICAL;
    const float hand = dot3(ns, orientation);
    if (!bf_isfinite(hand) || hand == 0.0f) return BF_MAP_NUMERICAL;
    const float du = fmaxf(1.0f / (float)img.width, bf_hypotf(g.dudx, g.dudy));
    const float dv = fmaxf(1.0f / (float)img.height, bf_hypotf(g.dvdx, g.dvdy));
    if (!bf_isfinite(du) || !(du > 0.0f) || !bf_isfinite(dv) || !(dv > 0.0f)) return BF_MAP_NUMERICAL;
    const float at[4][2] = {{u + du, v}, {u - du, v}, {u, v + dv}, {u, v - dv}};
    float hgt[4];
    for (int k = 0; k < 4; ++k) {
        const int st = ewa_filter(S, m.bump_texture, at[k][0], at[k][1], g, m.bump_channel, m.bump_u_wrap, m.bump_v_wrap,
                                  m.bump_max_anisotropy, m.bump_max_texel_visits, hgt[k]);
        if (st != BF_MAP_OK) return st;
    }
    const float dhdu = (hgt[0] - hgt[1]) / (2.0f * du);
    const float dhdv = (hgt[2] - hgt[3]) / (2.0f * dv);
    const float su = m.bump_scale * dhdu, sv = m.bump_scale * dhdv;
    const v3f tu = mk(fmaf(ns.x, su, pu.x), fmaf(ns.y, su, pu.y), fmaf(ns.z, su, pu.z));
    const v3f tv = mk(fmaf(ns.x, sv, pv.x), fmaf(ns.y, sv, pv.y), fmaf(ns.z, sv, pv.z));
    v3f n;
    if (!bf_isfinite(dhdu) || !bf_isfinite(dhdv) || !bf_isfinite(su) || !bf_isfinite(sv) || !fin3(tu) || !fin3(tv) ||
        !scaled_cross(tu, tv, n))
        return BF_MAP_NUMERICAL;
    if (dot3(n, ns) < 0.0f) n = mul3(n, -1.0f);
    if (!(dot3(n, ns) > 0.0f) || !(dot3(ng, n) > 0.0f)) return BF_MAP_NUMERICAL;
    ns = n;
    return BF_MAP_OK;
}

/* apply_surface_maps (:3622-3643): the cone's footprint, then the normal map, then the bump map */
BF_D int apply_surface_maps(const bf_dev_scene& S, uint32_t material, float2 cone, v3f ray_direction, float u, float v,
                            v3f ng, v3f dpdu, v3f dpdv, v3f& ns) {
    const bf_surface_maps_desc m = S.maps[material];
    map_differentials g;
    if (!cone_differentials(cone, ray_direction, ng, dpdu, dpdv, g)) return BF_MAP_NUMERICAL;
    if (m.normal_texture != BF_INVALID_INDEX) {
        const int st = apply_normal_map(S, m, g, u, v, ng, dpdu, dpdv, ns);
        if (st != BF_MAP_OK) return st;
    }
    if (m.bump_texture != BF_INVALID_INDEX) return apply_bump_map(S, m, g, u, v, ng, dpdu, dpdv, ns);
    return BF_MAP_OK;
}

#endif
