/* ORACLE — test infrastructure only (see bf_oracle.h). Normal and bump maps: EWA filtering of
 * mip-mapped image textures over a ray cone's footprint and the shading-normal perturbation of
 * each map, restated from Blackframe/XPU/CUDA/Source/CudaWavefrontStages.cu:2992-3673
 * (wrap_texture_index, cone_texture_differentials, make_ewa_footprint, make_ewa_ellipse,
 * sample_ewa_ellipse, filter_image_ewa, projected_tangent_frame, normalized_cross,
 * apply_normal_map, apply_bump_map, apply_surface_maps). Image tables and map parameters are
 * validated when the scene is built; transcendentals come from blackframe_math.h.
 * Pins: tests/test_oracle_maps.py (closed forms of constant, ramp and flat maps, wrap modes, and
 * the filter against a binary64 evaluation of the same weights). */
#ifndef BF_ORACLE_MAPS_H
#define BF_ORACLE_MAPS_H

#include "bf_oracle_internal.h"

enum { MAP_OK = 0, MAP_INVALID = 1, MAP_NUMERICAL = 2, MAP_UNSUPPORTED = 3 };
#define MAP_TEXEL_LIMIT 8388608.0f

typedef struct {
    float dudx, dvdx, dudy, dvdy;
} map_diff;
typedef struct {
    uint32_t width, height;
    uint64_t base; /* first value of the level in the image's texels */
} map_level;
typedef struct {
    float ax, ay, bx, by, ra, rb, minor; /* major axis, minor axis, radii, minor length */
} map_footprint;
typedef struct {
    float cx, cy, ax, ay, bx, by, inv_a, inv_b;
    int64_t x0, x1, y0, y1;
    uint64_t visits;
} map_ellipse;

/* :2992-3034 */
static int map_wrap(int64_t i, uint32_t extent, uint32_t mode, uint32_t* out) {
    const int64_t n = (int64_t)extent;
    switch (mode) {
    case BF_TEXTURE_WRAP_REPEAT: {
        int64_t r = i % n;
        if (r < 0) r += n;
        *out = (uint32_t)r;
        return 1;
    }
    case BF_TEXTURE_WRAP_CLAMP:
        *out = i < 0 ? 0u : (i >= n ? extent - 1u : (uint32_t)i);
        return 1;
    case BF_TEXTURE_WRAP_MIRROR: {
        int64_t p = i % (2 * n);
        if (p < 0) p += 2 * n;
        *out = (uint32_t)(p < n ? p : 2 * n - 1 - p);
        return 1;
    }
    default: /* border: absent outside */
        *out = i >= 0 && i < n ? (uint32_t)i : 0u;
        return i >= 0 && i < n;
    }
}

/* :3036-3088 */
static int map_cone_diff(bf_ray_cone cone, v3 dir, v3 ng, v3 dpdu, v3 dpdv, map_diff* g) {
    v3 d;
    if (!bf_isfinite(cone.width) || cone.width < 0.0f || !bf_isfinite(cone.spread) || cone.spread < 0.0f ||
        !normalize3(dir, &d))
        return 0;
    const float cosine = fabsf(v3_dot(d, ng));
    if (!bf_isfinite(cosine) || !(cosine > 0.0f)) return 0;
    const float radius = cone.width / cosine;
    const float guu = v3_dot(dpdu, dpdu), guv = v3_dot(dpdu, dpdv), gvv = v3_dot(dpdv, dpdv);
    const float det = fmaf(guu, gvv, -guv * guv);
    if (!bf_isfinite(radius) || radius < 0.0f || !bf_isfinite(guu) || !bf_isfinite(guv) || !bf_isfinite(gvv) ||
        !(guu > 0.0f) || !(gvv > 0.0f) || !bf_isfinite(det) || !(det > 0.0f) || radius == 0.0f)
        return 0;
    const float rr = radius * radius;
    const float cuu = rr * gvv / det, cuv = -rr * guv / det, cvv = rr * guu / det;
    if (!bf_isfinite(rr) || !(rr > 0.0f) || !bf_isfinite(cuu) || !bf_isfinite(cuv) || !bf_isfinite(cvv) ||
        !(cuu > 0.0f) || !(cvv > 0.0f))
        return 0;
    g->dudx = sqrtf(cuu);
    g->dvdx = cuv / g->dudx;
    float rest = fmaf(-g->dvdx, g->dvdx, cvv);
    if (rest < 0.0f && rest >= -(64.0f * BFO_FLOAT_EPSILON * cvv)) rest = 0.0f;
    g->dudy = 0.0f;
    g->dvdy = sqrtf(rest);
    return bf_isfinite(g->dudx) && g->dudx > 0.0f && bf_isfinite(g->dvdx) && bf_isfinite(g->dvdy) && g->dvdy > 0.0f;
}

/* :3090-3141 */
static int map_footprint_at(map_level lv, const map_diff* g, uint32_t anisotropy, map_footprint* f) {
    const float w = (float)lv.width, h = (float)lv.height;
    const float ux = g->dudx * w, vx = g->dvdx * h, uy = g->dudy * w, vy = g->dvdy * h;
    if (!bf_isfinite(ux) || !bf_isfinite(vx) || !bf_isfinite(uy) || !bf_isfinite(vy) || fabsf(ux) >= MAP_TEXEL_LIMIT ||
        fabsf(vx) >= MAP_TEXEL_LIMIT || fabsf(uy) >= MAP_TEXEL_LIMIT || fabsf(vy) >= MAP_TEXEL_LIMIT)
        return 0;
    const float xx = fmaf(ux, ux, uy * uy), xy = fmaf(ux, vx, uy * vy), yy = fmaf(vx, vx, vy * vy);
    const float big = 0.5f * ((xx + yy) + bf_hypotf(xx - yy, 2.0f * xy));
    const float j = fmaf(ux, vy, -uy * vx);
    float small = big > 0.0f ? (j * j) / big : 0.0f;
    const float a = (float)anisotropy;
    const float floor_value = big / (a * a);
    if (floor_value > small) small = floor_value;
    const float theta = big > 0.0f ? 0.5f * bf_atan2f(2.0f * xy, xx - yy) : 0.0f;
    float sn, cs;
    bf_sincosf(theta, &sn, &cs);
    const float la = sqrtf(big), lb = sqrtf(small);
    f->ax = cs;
    f->ay = sn;
    f->bx = -sn;
    f->by = cs;
    f->ra = bf_hypotf(1.0f, la);
    f->rb = bf_hypotf(1.0f, lb);
    f->minor = lb;
    return bf_isfinite(xx) && bf_isfinite(xy) && bf_isfinite(yy) && bf_isfinite(big) && big >= 0.0f &&
           bf_isfinite(small) && small >= 0.0f && bf_isfinite(f->ax) && bf_isfinite(f->ay) && bf_isfinite(f->ra) &&
           bf_isfinite(f->rb) && bf_isfinite(f->minor);
}

/* :3143-3196 */
static int map_ellipse_at(map_level lv, float u, float v, const map_footprint* f, map_ellipse* e) {
    const float sx = u * (float)lv.width, sy = v * (float)lv.height;
    if (!bf_isfinite(u) || !bf_isfinite(v) || !bf_isfinite(sx) || !bf_isfinite(sy) || fabsf(sx) >= MAP_TEXEL_LIMIT ||
        fabsf(sy) >= MAP_TEXEL_LIMIT)
        return 0;
    e->cx = sx - 0.5f;
    e->cy = sy - 0.5f;
    e->ax = f->ax;
    e->ay = f->ay;
    e->bx = f->bx;
    e->by = f->by;
    e->inv_a = 1.0f / f->ra;
    e->inv_b = 1.0f / f->rb;
    const float rx = bf_hypotf(f->ra * f->ax, f->rb * f->bx);
    const float ry = bf_hypotf(f->ra * f->ay, f->rb * f->by);
    const float x0 = ceilf(e->cx - rx), x1 = floorf(e->cx + rx), y0 = ceilf(e->cy - ry), y1 = floorf(e->cy + ry);
    if (!bf_isfinite(rx) || !bf_isfinite(ry) || !bf_isfinite(x0) || !bf_isfinite(x1) || !bf_isfinite(y0) ||
        !bf_isfinite(y1) || x0 <= -MAP_TEXEL_LIMIT || x1 >= MAP_TEXEL_LIMIT || y0 <= -MAP_TEXEL_LIMIT ||
        y1 >= MAP_TEXEL_LIMIT)
        return 0;
    e->x0 = (int64_t)x0;
    e->x1 = (int64_t)x1;
    e->y0 = (int64_t)y0;
    e->y1 = (int64_t)y1;
    if (e->x1 < e->x0 || e->y1 < e->y0) return 0;
    e->visits = (uint64_t)(e->x1 - e->x0 + 1) * (uint64_t)(e->y1 - e->y0 + 1);
    return 1;
}

/* :3198-3269 */
static int map_sample(const bf_image_texture_desc* img, map_level lv, const map_ellipse* e, uint32_t channel,
                      uint32_t u_wrap, uint32_t v_wrap, float* out) {
    const float edge = bf_expf(-2.0f);
    float weight_sum = 0.0f, mean = 0.0f;
    for (int64_t y = e->y0; y <= e->y1; ++y) {
        uint32_t wy;
        const int y_in = map_wrap(y, lv.height, v_wrap, &wy);
        const float dy = (float)y - e->cy;
        for (int64_t x = e->x0; x <= e->x1; ++x) {
            const float dx = (float)x - e->cx;
            const float p = fmaf(dx, e->ax, dy * e->ay) * e->inv_a;
            const float q = fmaf(dx, e->bx, dy * e->by) * e->inv_b;
            const float r2 = fmaf(p, p, q * q);
            if (!bf_isfinite(r2) || r2 < 0.0f) return MAP_NUMERICAL;
            if (r2 >= 1.0f) continue;
            const float weight = edge * bf_expm1f(2.0f * (1.0f - r2));
            if (!bf_isfinite(weight) || !(weight > 0.0f)) return MAP_NUMERICAL;
            uint32_t wx;
            const int x_in = map_wrap(x, lv.width, u_wrap, &wx);
            float value = 0.0f;
            if (x_in && y_in) {
                value = img->texels[lv.base + ((uint64_t)wy * lv.width + wx) * img->channel_count + channel];
                if (!bf_isfinite(value)) return MAP_NUMERICAL;
            }
            const float next = weight_sum + weight;
            if (!bf_isfinite(next) || !(next > 0.0f)) return MAP_NUMERICAL;
            mean = weight_sum == 0.0f ? value : mean + (value - mean) * (weight / next);
            weight_sum = next;
        }
    }
    *out = mean;
    return weight_sum > 0.0f && bf_isfinite(mean) ? MAP_OK : MAP_NUMERICAL;
}

static map_level map_level_of(const bf_image_texture_desc* img, uint32_t level) {
    map_level lv = {img->width, img->height, 0};
    for (uint32_t k = 0; k < level; ++k) {
        lv.base += (uint64_t)lv.width * lv.height * img->channel_count;
        lv.width = lv.width > 1u ? lv.width / 2u : 1u;
        lv.height = lv.height > 1u ? lv.height / 2u : 1u;
    }
    return lv;
}

/* :3271-3355 */
static int map_filter(const bf_image_texture_desc* img, float u, float v, const map_diff* g, uint32_t channel,
                      uint32_t u_wrap, uint32_t v_wrap, uint32_t anisotropy, uint32_t visits_max, float* out) {
    if (!bf_isfinite(u) || !bf_isfinite(v)) return MAP_INVALID;
    map_level lo = map_level_of(img, 0), hi = lo;
    map_footprint flo, fhi;
    memset(&fhi, 0, sizeof fhi);
    if (!map_footprint_at(lo, g, anisotropy, &flo)) return MAP_NUMERICAL;
    int two = 0;
    float t = 0.0f;
    if (flo.minor > 1.0f) {
        for (uint32_t k = 1; k < img->mip_count; ++k) {
            const map_level cand = map_level_of(img, k);
            map_footprint fc;
            if (!map_footprint_at(cand, g, anisotropy, &fc)) return MAP_NUMERICAL;
            if (fc.minor <= 1.0f) {
                const float a = bf_logf(flo.minor), b = bf_logf(fc.minor);
                const float den = a - b;
                const float blend = a / den;
                if (!bf_isfinite(blend) || !(den > 0.0f) || blend < 0.0f || blend > 1.0f) return MAP_NUMERICAL;
                if (blend == 1.0f) {
                    lo = cand;
                    flo = fc;
                } else {
                    hi = cand;
                    fhi = fc;
                    two = 1;
                    t = blend;
                }
                break;
            }
            lo = cand;
            flo = fc;
        }
    }
    if (!two && flo.minor > 1.0f) { /* wider than the coarsest level: the unit footprint there */
        const map_footprint unit = {1.0f, 0.0f, 0.0f, 1.0f, 1.0f, 1.0f, 0.0f};
        flo = unit;
    }
    map_ellipse elo, ehi;
    if (!map_ellipse_at(lo, u, v, &flo, &elo)) return MAP_NUMERICAL;
    uint64_t visits = elo.visits;
    if (two) {
        if (!map_ellipse_at(hi, u, v, &fhi, &ehi) || visits > UINT64_MAX - ehi.visits) return MAP_NUMERICAL;
        visits += ehi.visits;
    }
    if (visits > visits_max) return MAP_UNSUPPORTED;
    float a = 0.0f, b = 0.0f;
    int st = map_sample(img, lo, &elo, channel, u_wrap, v_wrap, &a);
    if (st != MAP_OK) return st;
    if (!two) {
        *out = a;
        return MAP_OK;
    }
    st = map_sample(img, hi, &ehi, channel, u_wrap, v_wrap, &b);
    if (st != MAP_OK) return st;
    *out = a + t * (b - a);
    return bf_isfinite(*out) ? MAP_OK : MAP_NUMERICAL;
}

static float map_max_abs(v3 a) { return fmaxf(fabsf(a.x), fmaxf(fabsf(a.y), fabsf(a.z))); }

/* :3357-3381 */
static int map_tangent_frame(v3 n, v3 dpdu, v3 dpdv, v3* t, v3* b) {
    const float along = v3_dot(n, dpdu);
    const v3 p = v3_make(fmaf(-along, n.x, dpdu.x), fmaf(-along, n.y, dpdu.y), fmaf(-along, n.z, dpdu.z));
    v3 base;
    if (!normalize3(p, t) || !normalize3(v3_cross(n, *t), &base)) return 0;
    const float scale = map_max_abs(dpdv);
    if (!bf_isfinite(scale) || scale == 0.0f) return 0;
    const float side = v3_dot(base, v3_mul(dpdv, 1.0f / scale));
    if (!bf_isfinite(side) || side == 0.0f) return 0;
    *b = side > 0.0f ? base : v3_mul(base, -1.0f);
    return 1;
}

/* :3383-3393 */
static int map_cross(v3 a, v3 b, v3* out) {
    const float sa = map_max_abs(a), sb = map_max_abs(b);
    if (!bf_isfinite(sa) || !bf_isfinite(sb) || !(sa > 0.0f) || !(sb > 0.0f)) return 0;
    return normalize3(v3_cross(v3_mul(a, 1.0f / sa), v3_mul(b, 1.0f / sb)), out);
}

/* :3395-3494 */
static int map_normal(const bfo_scene* s, const bf_surface_maps_desc* m, const map_diff* g, float u, float v, v3 ng,
                      v3 dpdu, v3 dpdv, v3* ns) {
    const bf_image_texture_desc* img = &s->image_textures[m->normal_texture];
    float c[3];
    for (int k = 0; k < 3; ++k) {
        const int st = map_filter(img, u, v, g, m->normal_channels[k], m->normal_u_wrap, m->normal_v_wrap,
                                  m->normal_max_anisotropy, m->normal_max_texel_visits, &c[k]);
        if (st != MAP_OK) return st;
        if (c[k] < 0.0f || c[k] > 1.0f) return MAP_NUMERICAL;
    }
    v3 tn;
    if (!normalize3(v3_make(fmaf(2.0f, c[0], -1.0f), fmaf(2.0f, c[1], -1.0f), fmaf(2.0f, c[2], -1.0f)), &tn) ||
        !(tn.z > 0.0f))
        return MAP_NUMERICAL;
    v3 t, b;
    if (!map_tangent_frame(*ns, dpdu, dpdv, &t, &b)) return MAP_NUMERICAL;
    const float ty = m->normal_y_convention ? -tn.y : tn.y;
    const v3 n = v3_make(fmaf(t.x, tn.x, fmaf(b.x, ty, ns->x * tn.z)), fmaf(t.y, tn.x, fmaf(b.y, ty, ns->y * tn.z)),
                         fmaf(t.z, tn.x, fmaf(b.z, ty, ns->z * tn.z)));
    v3 r;
    if (!normalize3(n, &r) || !(v3_dot(ng, r) > 0.0f)) return MAP_NUMERICAL;
    *ns = r;
    return MAP_OK;
}

/* :3496-3620 */
static int map_bump(const bfo_scene* s, const bf_surface_maps_desc* m, const map_diff* g, float u, float v, v3 ng,
                    v3 dpdu, v3 dpdv, v3* ns) {
    const bf_image_texture_desc* img = &s->image_textures[m->bump_texture];
    const v3 n = *ns;
    const float au = v3_dot(n, dpdu), av = v3_dot(n, dpdv);
    const v3 pu = v3_make(fmaf(-au, n.x, dpdu.x), fmaf(-au, n.y, dpdu.y), fmaf(-au, n.z, dpdu.z));
    const v3 pv = v3_make(fmaf(-av, n.x, dpdv.x), fmaf(-av, n.y, dpdv.y), fmaf(-av, n.z, dpdv.z));
    v3 orient;
    if (!bf_isfinite(au) || !bf_isfinite(av) || !v3_finite(pu) || !v3_finite(pv) || !map_cross(pu, pv, &orient))
        return MAP_NUMERICAL;
    const float side = v3_dot(n, orient);
    if (!bf_isfinite(side) || side == 0.0f) return MAP_NUMERICAL;
    const float du = fmaxf(1.0f / (float)img->width, bf_hypotf(g->dudx, g->dudy));
    const float dv = fmaxf(1.0f / (float)img->height, bf_hypotf(g->dvdx, g->dvdy));
    if (!bf_isfinite(du) || !(du > 0.0f) || !bf_isfinite(dv) || !(dv > 0.0f)) return MAP_NUMERICAL;
    const float su[4] = {u + du, u - du, u, u}, sv[4] = {v, v, v + dv, v - dv};
    float h[4];
    for (int k = 0; k < 4; ++k) {
        const int st = map_filter(img, su[k], sv[k], g, m->bump_channel, m->bump_u_wrap, m->bump_v_wrap,
                                  m->bump_max_anisotropy, m->bump_max_texel_visits, &h[k]);
        if (st != MAP_OK) return st;
    }
    const float hu = (h[0] - h[1]) / (2.0f * du), hv = (h[2] - h[3]) / (2.0f * dv);
    const float ku = m->bump_scale * hu, kv = m->bump_scale * hv;
    const v3 tu = v3_make(fmaf(n.x, ku, pu.x), fmaf(n.y, ku, pu.y), fmaf(n.z, ku, pu.z));
    const v3 tv = v3_make(fmaf(n.x, kv, pv.x), fmaf(n.y, kv, pv.y), fmaf(n.z, kv, pv.z));
    v3 r;
    if (!bf_isfinite(hu) || !bf_isfinite(hv) || !bf_isfinite(ku) || !bf_isfinite(kv) || !v3_finite(tu) ||
        !v3_finite(tv) || !map_cross(tu, tv, &r))
        return MAP_NUMERICAL;
    if (v3_dot(r, n) < 0.0f) r = v3_mul(r, -1.0f);
    if (!(v3_dot(r, n) > 0.0f) || !(v3_dot(ng, r) > 0.0f)) return MAP_NUMERICAL;
    *ns = r;
    return MAP_OK;
}

/* :3622-3643 */
static int map_apply(const bfo_scene* s, uint32_t material, bf_ray_cone cone, v3 dir, float u, float v, v3 ng, v3 dpdu,
                     v3 dpdv, v3* ns) {
    const bf_surface_maps_desc* m = &s->surface_maps[material];
    map_diff g;
    if (!map_cone_diff(cone, dir, ng, dpdu, dpdv, &g)) return MAP_NUMERICAL;
    if (m->normal_texture != BF_INVALID_INDEX) {
        const int st = map_normal(s, m, &g, u, v, ng, dpdu, dpdv, ns);
        if (st != MAP_OK) return st;
    }
    return m->bump_texture != BF_INVALID_INDEX ? map_bump(s, m, &g, u, v, ng, dpdu, dpdv, ns) : MAP_OK;
}

#endif
