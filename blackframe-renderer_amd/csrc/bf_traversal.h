/*
 * BVH2 closest-hit and any-hit traversal over 64-byte wide nodes with a per-lane LDS stack.
 *
 * Decisions are bit-identical to the reference traversal (CudaSceneClosestHit.cu:407-1186):
 *  - the slab test is the binary64 quotient widened by one ulp (:407-455);
 *  - children are pushed far-then-near with the first (lower canonical index) child winning ties
 *    (:466-512); the first child of an internal node always has the lower pre-order index;
 *  - triangles use the watertight binary32 test with binary64 re-decision of ambiguous edges
 *    (:583-824); candidates are ordered by (t, instance id, primitive id) (:873-881).
 * The reference re-tests a node's own box with the current t_max when it is popped. A node's box
 * passed its push-time test, so that pop test reduces to near <= t_max; the stack stores near
 * rounded UP to binary32 (exactly equivalent for a binary32 t_max), which lets a popped leaf go
 * straight to its triangle without fetching a node. The reference keeps one 64-entry stack per
 * structure; here TLAS and BLAS share one stack of BF_STACK_DEPTH entries, each part held to the
 * reference's 64, so a lane reports stack overflow exactly where the reference's would. Slab quotients are formed from a per-ray
 * binary64 reciprocal with one correction step, which yields the correctly rounded quotient.
 */
#ifndef BF_TRAVERSAL_H
#define BF_TRAVERSAL_H

#include "bf_device_math.h"
#include "bf_device_scene.h"

struct trav_ray {
    v3f o, d;
    float t_min, t_max;
    uint32_t mask;
};

struct trav_hit {
    bool present;
    float t, b0, b1, b2;
    uint32_t instance_id, primitive, scene_instance, leaf;
};

struct trav_counters {
    uint32_t nodes, triangles, instances;
};

/* Shear of the watertight triangle test for one direction (CudaSceneClosestHit.cu:694-740): the
 * dominant axis kz, the other two in winding order, and rz = 1 / d_kz, sx = d_kx rz, sy = d_ky rz.
 * `ok` is false when one of them is not finite (every triangle test then fails numerically). */
struct trav_shear {
    uint32_t kx, ky, kz;
    float rz, sx, sy;
    bool ok;
};

BF_D void make_shear(v3f d, trav_shear& s) {
    uint32_t kz = 0;
    float maxd = fabsf(d.x);
    if (fabsf(d.y) > maxd) {
        kz = 1;
        maxd = fabsf(d.y);
    }
    if (fabsf(d.z) > maxd) kz = 2;
    uint32_t kx = kz + 1;
    if (kx == 3) kx = 0;
    uint32_t ky = kx + 1;
    if (ky == 3) ky = 0;
    if (comp3(d, kz) < 0.0f) {
        const uint32_t t = kx;
        kx = ky;
        ky = t;
    }
    s.kx = kx;
    s.ky = ky;
    s.kz = kz;
    s.rz = 1.0f / comp3(d, kz);
    s.sx = comp3(d, kx) * s.rz;
    s.sy = comp3(d, ky) * s.rz;
    s.ok = bf_isfinite(s.rz) && bf_isfinite(s.sx) && bf_isfinite(s.sy);
}

/* A ray expressed in one space (world for the TLAS, an instance's object space for its BLAS),
 * with the correctly rounded binary64 reciprocal of every non-zero direction component and the
 * binary32 roundings the slab filter uses (finv, and efloor = 2^-147 max |finv|), formed once
 * instead of per node. (Keeping the triangle test's shear here as well measured slower: the
 * extra live registers spill in the traversal kernels.) */
struct trav_space {
    v3f o, d;
    double inv[3];
    float finv[3];
    float efloor;
};

/* The derived binary32 constants of a space whose o, d and inv are set. */
BF_D void finish_space(trav_space& s) {
    float inv_max = 0.0f;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        s.finv[a] = (float)s.inv[a];
        inv_max = fmaxf(inv_max, fabsf(s.finv[a]));
    }
    s.efloor = 0x1p-147f * inv_max;
}

BF_D void make_space(v3f o, v3f d, trav_space& s) {
    s.o = o;
    s.d = d;
    s.inv[0] = d.x != 0.0f ? 1.0 / (double)d.x : 0.0;
    s.inv[1] = d.y != 0.0f ? 1.0 / (double)d.y : 0.0;
    s.inv[2] = d.z != 0.0f ? 1.0 / (double)d.z : 0.0;
    finish_space(s);
}

/* a / b rounded to nearest from y = RN(1/b): q0 = RN(a y), r = a - b q0 (exact through the fma),
 * q0 + r y rounded once is RN(a / b) (Markstein's quotient correction), i.e. the same binary64
 * quotient the reference's division returns, for three fused operations instead of a division. */
BF_D double quotient(double a, double b, double y) {
    const double q0 = a * y;
    const double r = fma(-q0, b, a);
    return fma(r, y, q0);
}

/* CudaSceneClosestHit.cu:407-455 (boxes are validated finite and ordered by the host builder):
 * per axis the binary64 plane distances (lo - o) / d and (hi - o) / d, widened by one ulp. */
BF_D uint32_t slab_exact(float lx, float ly, float lz, float hx, float hy, float hz, const trav_space& s, float t_min,
                         float t_max, bool& hit, double& near_value) {
    double nv = (double)t_min;
    double fv = (double)t_max;
    const float lo[3] = {lx, ly, lz};
    const float hi[3] = {hx, hy, hz};
    const float oo[3] = {s.o.x, s.o.y, s.o.z};
    const float dd[3] = {s.d.x, s.d.y, s.d.z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (dd[a] == 0.0f) {
            if (oo[a] < lo[a] || oo[a] > hi[a]) {
                hit = false;
                return BF_HIT_MISS;
            }
            continue;
        }
        const double db = (double)dd[a];
        const double first = quotient((double)lo[a] - (double)oo[a], db, s.inv[a]);
        const double second = quotient((double)hi[a] - (double)oo[a], db, s.inv[a]);
        if (bf_isnand(first) || bf_isnand(second)) return BF_HIT_NUMERICAL_FAILURE;
        nv = fmax(nv, bf_next_down(fmin(first, second)));
        fv = fmin(fv, bf_next_up(fmax(first, second)));
        if (nv > fv) {
            hit = false;
            return BF_HIT_MISS;
        }
    }
    hit = true;
    near_value = nv;
    return BF_HIT_MISS;
}

/* A child a node step pushes: its stack entry and bounds [key_lo, key_hi] of the binary32 rounding
 * of its binary64 near value (the reference orders siblings by that rounding). */
struct child_near {
    uint2 entry;
    float key_lo, key_hi;
    uint32_t tie_axis; /* 0-2: that axis's near plane alone sets the binary64 near value; 3: t_min
                        * does; 4: not known */
    float tie_plane;   /* that plane's coordinate */
};

BF_D float round_up_float(double v) {
    float f = (float)v;
    if ((double)f < v) f = bf_next_up_f(f);
    return f;
}

/* An entry whose near word is the exact binary64 near value rounded up to binary32: visited iff it
 * is <= the current t_max, exactly the reference's pop test. */
BF_D void exact_child(uint32_t ref, double v, child_near& c) {
    c.entry = make_uint2(ref | BF_NEAR_EXACT, __float_as_uint(round_up_float(v)));
    c.key_lo = c.key_hi = (float)v;
    c.tie_axis = 4u;
    c.tie_plane = 0.0f;
}

/*
 * An entry that keeps bounds lo <= near <= hi of the binary64 near value (0 <= lo): bits 8-31 of
 * the near word hold hi rounded up to a multiple of 256 ulps (still an upper bound) and bits 0-7
 * k with 2^(k-127) >= hi' - lo, so that hi' - 2^(k-127) is a lower bound (k = 255: none kept).
 * With E the exponent field of RN(hi' - lo), 2^(E+1-127) exceeds hi' - lo itself.
 */
BF_D void bounded_child(uint32_t ref, float lo, float hi, child_near& c) {
    const uint32_t hb = (__float_as_uint(hi) + 0xFFu) & ~0xFFu;
    const uint32_t k = min((__float_as_uint(__uint_as_float(hb) - lo) >> 23) + 1u, 255u);
    c.entry = make_uint2(ref, hb | k);
    c.key_lo = lo;
    c.key_hi = hi;
}

/*
 * The same decision with a binary32 filter in front. Each binary32 plane distance
 * q~ = RN32(RN32(plane - o) * RN32(1/d)) is within 3 binary32 roundings of the exact quotient
 * (relative 3u, u = 2^-24), and the binary64 quotient and its one-ulp widening add under 2^-51
 * relative; subnormal differences add at most |1/d| 2^-149 absolutely. So every binary64
 * interval end lies within e = 4u max|q~| + max |1/d| 2^-147 + 2^-126 of its binary32 estimate,
 * and when the estimated interval is empty (or non-empty) by more than 2e the binary64 test
 * reaches the same verdict. Only a box within that band takes binary64 arithmetic; a certain hit
 * returns its near value as the binary32 interval [nv - en, nv + en], en = 4u nv + the absolute
 * terms (widened outward by one ulp each side, and never below t_min, which bounds the binary64
 * value from below).
 */
BF_D uint32_t slab(float lx, float ly, float lz, float hx, float hy, float hz, const trav_space& s, float t_min,
                   float t_max, uint32_t ref, bool& hit, child_near& near) {
    const float lo[3] = {lx, ly, lz};
    const float hi[3] = {hx, hy, hz};
    const float oo[3] = {s.o.x, s.o.y, s.o.z};
    const float dd[3] = {s.d.x, s.d.y, s.d.z};
    /* nv: the largest near estimate (t_min included), second: the next one, axis: nv's source */
    float nv = t_min, second = -BF_INFINITY_F, fv = t_max, scale = 0.0f, plane = 0.0f;
    uint32_t axis = 3u;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (dd[a] == 0.0f) {
            if (oo[a] < lo[a] || oo[a] > hi[a]) {
                hit = false;
                return BF_HIT_MISS;
            }
            continue;
        }
        const float inv = s.finv[a];
        const float q1 = (lo[a] - oo[a]) * inv;
        const float q2 = (hi[a] - oo[a]) * inv;
        const float qn = fminf(q1, q2);
        if (qn > nv) {
            second = nv;
            nv = qn;
            axis = (uint32_t)a;
            plane = dd[a] > 0.0f ? lo[a] : hi[a];
        } else {
            second = fmaxf(second, qn);
        }
        fv = fminf(fv, fmaxf(q1, q2));
        scale = fmaxf(scale, fmaxf(fabsf(q1), fabsf(q2)));
    }
    const float e = 4.0f * 0x1p-24f * scale + s.efloor + 0x1p-126f;
    if (bf_isfinite(e) && bf_isfinite(nv) && bf_isfinite(scale)) {
        if (nv - fv > 2.0f * e) {
            hit = false;
            return BF_HIT_MISS;
        }
        /* The near value alone has a tighter bound than the interval test: an axis whose near
         * estimate lies below t_min by more than its own error contributes t_min exactly, and every
         * other axis's near quotient is at most nv (or tiny), so its error is under 4u nv plus the
         * absolute terms. A ray nearly parallel to an axis (huge quotients there) keeps tight near
         * bounds this way. */
        const float en = 4.0f * 0x1p-24f * nv + s.efloor + 0x1p-126f;
        const float up = nv + en; /* > 0: nv >= t_min >= 0, en > 0 */
        if (fv - nv > 2.0f * e && up < 0x1p127f) {
            const float low = nv - en;
            hit = true;
            /* one ulp outward on each side (both positive where they bind) */
            bounded_child(ref, low > t_min ? __uint_as_float(__float_as_uint(low) - 1u) : t_min,
                          __uint_as_float(__float_as_uint(up) + 1u), near);
            /* every estimate is within en of its binary64 term, so a lead of more than 2 en makes
             * the source of nv the source of the binary64 near value too */
            near.tie_axis = nv - second > 2.0f * en ? axis : 4u;
            near.tie_plane = plane;
            return BF_HIT_MISS;
        }
    }
    double v = 0.0;
    const uint32_t st = slab_exact(lx, ly, lz, hx, hy, hz, s, t_min, t_max, hit, v);
    exact_child(ref, v, near);
    return st;
}

/* The binary64 near value of a box the ray is known to meet: per axis the near plane (lo for
 * d > 0, hi for d < 0; rounding is monotone) widened down by one ulp, and t_min. */
BF_D double near_exact(const float* box, const trav_space& s, float t_min) {
    const float oo[3] = {s.o.x, s.o.y, s.o.z};
    const float dd[3] = {s.d.x, s.d.y, s.d.z};
    double n = (double)t_min;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (dd[a] == 0.0f) continue;
        const float plane = dd[a] > 0.0f ? box[a] : box[3 + a];
        n = fmax(n, bf_next_down(quotient((double)plane - (double)oo[a], (double)dd[a], s.inv[a])));
    }
    return n;
}

/* 1: the entry's near value is <= t (visit), 0: it is > t (skip), 2: its bounds cannot tell */
BF_D uint32_t near_decision(uint2 entry, float t) {
    if (entry.x & BF_NEAR_EXACT) return __uint_as_float(entry.y) <= t ? 1u : 0u;
    const float hi = __uint_as_float(entry.y & ~0xFFu);
    if (hi <= t) return 1u;
    const uint32_t k = entry.y & 0xFFu;
    if (k != 255u && bf_next_down_f(hi - __uint_as_float(k << 23)) > t) return 0u;
    return 2u;
}

struct tri_result {
    bool hit;
    float t, b0, b1, b2;
};

/* CudaSceneClosestHit.cu:583-692 */
template <bool BARY>
BF_D uint32_t refine_edges(v3f v0, v3f v1, v3f v2, v3f o, v3f d, uint32_t kx, uint32_t ky, uint32_t kz, float t_min,
                           float t_max, tri_result& r) {
    const double dx = (double)comp3(d, kx), dy = (double)comp3(d, ky), dz = (double)comp3(d, kz);
    if (dz == 0.0) return BF_HIT_NUMERICAL_FAILURE;
    const double p0x = (double)comp3(v0, kx) - (double)comp3(o, kx);
    const double p0y = (double)comp3(v0, ky) - (double)comp3(o, ky);
    const double p0z = (double)comp3(v0, kz) - (double)comp3(o, kz);
    const double p1x = (double)comp3(v1, kx) - (double)comp3(o, kx);
    const double p1y = (double)comp3(v1, ky) - (double)comp3(o, ky);
    const double p1z = (double)comp3(v1, kz) - (double)comp3(o, kz);
    const double p2x = (double)comp3(v2, kx) - (double)comp3(o, kx);
    const double p2y = (double)comp3(v2, ky) - (double)comp3(o, ky);
    const double p2z = (double)comp3(v2, kz) - (double)comp3(o, kz);
    const double ax = p0x * dz - dx * p0z;
    const double ay = p0y * dz - dy * p0z;
    const double bx = p1x * dz - dx * p1z;
    const double by = p1y * dz - dy * p1z;
    const double cx = p2x * dz - dx * p2z;
    const double cy = p2y * dz - dy * p2z;
    const double e0 = bx * cy - by * cx;
    const double e1 = cx * ay - cy * ax;
    const double e2 = ax * by - ay * bx;
    if (!bf_isfinited(e0) || !bf_isfinited(e1) || !bf_isfinited(e2)) return BF_HIT_NUMERICAL_FAILURE;
    const bool neg = e0 < 0.0 || e1 < 0.0 || e2 < 0.0;
    const bool pos = e0 > 0.0 || e1 > 0.0 || e2 > 0.0;
    if (neg && pos) return BF_HIT_MISS;
    const double det = e0 + e1 + e2;
    if (!bf_isfinited(det)) return BF_HIT_NUMERICAL_FAILURE;
    if (det == 0.0) return BF_HIT_MISS;
    const double num = fma(e0, p0z, fma(e1, p1z, e2 * p2z));
    const double den = det * dz;
    if (!bf_isfinited(num) || !bf_isfinited(den) || den == 0.0) return BF_HIT_NUMERICAL_FAILURE;
    const double t = num / den;
    if (!BARY) {
        if (!bf_isfinited(t)) return BF_HIT_NUMERICAL_FAILURE;
        if (t < (double)t_min || t > (double)t_max) return BF_HIT_MISS;
        const float ft = (float)t;
        if (!bf_isfinite(ft)) return BF_HIT_NUMERICAL_FAILURE;
        r.hit = true;
        r.t = ft == 0.0f ? 0.0f : ft;
        return BF_HIT_MISS;
    }
    const double b0 = e0 / det, b1 = e1 / det, b2 = e2 / det;
    if (!bf_isfinited(t) || !bf_isfinited(b0) || !bf_isfinited(b1) || !bf_isfinited(b2)) return BF_HIT_NUMERICAL_FAILURE;
    if (t < (double)t_min || t > (double)t_max) return BF_HIT_MISS;
    const float ft = (float)t, f0 = (float)b0, f1 = (float)b1, f2 = (float)b2;
    if (!bf_isfinite(ft) || !bf_isfinite(f0) || !bf_isfinite(f1) || !bf_isfinite(f2)) return BF_HIT_NUMERICAL_FAILURE;
    r.hit = true;
    r.t = ft == 0.0f ? 0.0f : ft;
    r.b0 = f0 == 0.0f ? 0.0f : f0;
    r.b1 = f1 == 0.0f ? 0.0f : f1;
    r.b2 = f2 == 0.0f ? 0.0f : f2;
    return BF_HIT_MISS;
}

/* CudaSceneClosestHit.cu:694-824, with the direction's shear formed by make_shear */
template <bool BARY>
BF_D uint32_t watertight(v3f v0, v3f v1, v3f v2, v3f o, v3f d, const trav_shear& sh, float t_min, float t_max,
                         tri_result& r) {
    const uint32_t kx = sh.kx, ky = sh.ky, kz = sh.kz;
    const v3f t0 = sub3(v0, o), t1 = sub3(v1, o), t2 = sub3(v2, o);
    if (!fin3(t0) || !fin3(t1) || !fin3(t2)) return BF_HIT_NUMERICAL_FAILURE;
    const float rz = sh.rz, sx = sh.sx, sy = sh.sy;
    if (!sh.ok) return BF_HIT_NUMERICAL_FAILURE;
    const float ax = fmaf(-sx, comp3(t0, kz), comp3(t0, kx));
    const float ay = fmaf(-sy, comp3(t0, kz), comp3(t0, ky));
    const float bx = fmaf(-sx, comp3(t1, kz), comp3(t1, kx));
    const float by = fmaf(-sy, comp3(t1, kz), comp3(t1, ky));
    const float cx = fmaf(-sx, comp3(t2, kz), comp3(t2, kx));
    const float cy = fmaf(-sy, comp3(t2, kz), comp3(t2, ky));
    const float az = comp3(t0, kz) * rz;
    const float bz = comp3(t1, kz) * rz;
    const float cz = comp3(t2, kz) * rz;
    if (!bf_isfinite(ax) || !bf_isfinite(ay) || !bf_isfinite(bx) || !bf_isfinite(by) || !bf_isfinite(cx) ||
        !bf_isfinite(cy) || !bf_isfinite(az) || !bf_isfinite(bz) || !bf_isfinite(cz))
        return BF_HIT_NUMERICAL_FAILURE;
    const float e0 = bx * cy - by * cx;
    const float e1 = cx * ay - cy * ax;
    const float e2 = ax * by - ay * bx;
    if (!bf_isfinite(e0) || !bf_isfinite(e1) || !bf_isfinite(e2)) return BF_HIT_NUMERICAL_FAILURE;
    const bool neg = e0 < 0.0f || e1 < 0.0f || e2 < 0.0f;
    const bool pos = e0 > 0.0f || e1 > 0.0f || e2 > 0.0f;
    if (neg && pos) {
        /* Binary32 filter in front of the binary64 re-decision. Each sheared coordinate is within
         * 4u (u = 2^-24) of its exact value relative to m = |t_x| + |s t_z|, so each edge function
         * is within 10.001u (m_b m_c' + m_b' m_c) of the exact one, while the binary64
         * re-decision is accurate to about 2^-50 of the same terms. When one edge is negative and
         * another positive by more than 12u of that (plus subnormal floors), the binary64 signs
         * are mixed too and the ray misses exactly as refine_edges would decide. */
        const float fl = 0x1p-120f;
        const float max_ = fabsf(comp3(t0, kx)) + fabsf(sx * comp3(t0, kz)) + fl;
        const float may = fabsf(comp3(t0, ky)) + fabsf(sy * comp3(t0, kz)) + fl;
        const float mbx = fabsf(comp3(t1, kx)) + fabsf(sx * comp3(t1, kz)) + fl;
        const float mby = fabsf(comp3(t1, ky)) + fabsf(sy * comp3(t1, kz)) + fl;
        const float mcx = fabsf(comp3(t2, kx)) + fabsf(sx * comp3(t2, kz)) + fl;
        const float mcy = fabsf(comp3(t2, ky)) + fabsf(sy * comp3(t2, kz)) + fl;
        const float k = 12.0f * 0x1p-24f;
        const float b0 = k * (mbx * mcy + mby * mcx) + 0x1p-100f;
        const float b1 = k * (mcx * may + mcy * max_) + 0x1p-100f;
        const float b2 = k * (max_ * mby + may * mbx) + 0x1p-100f;
        const bool sure_neg = e0 < -b0 || e1 < -b1 || e2 < -b2;
        const bool sure_pos = e0 > b0 || e1 > b1 || e2 > b2;
        if (sure_neg && sure_pos && bf_isfinite(b0) && bf_isfinite(b1) && bf_isfinite(b2)) return BF_HIT_MISS;
    }
    if (e0 == 0.0f || e1 == 0.0f || e2 == 0.0f || (neg && pos))
        return refine_edges<BARY>(v0, v1, v2, o, d, kx, ky, kz, t_min, t_max, r);
    const float det = e0 + e1 + e2;
    if (!bf_isfinite(det)) return BF_HIT_NUMERICAL_FAILURE;
    if (det == 0.0f) return BF_HIT_MISS;
    const float sp = fmaf(e0, az, fmaf(e1, bz, e2 * cz));
    if (!bf_isfinite(sp)) return BF_HIT_NUMERICAL_FAILURE;
    const float smin = t_min * det, smax = t_max * det;
    if (bf_isnan(smin) || bf_isnan(smax)) return BF_HIT_NUMERICAL_FAILURE;
    if ((det > 0.0f && (sp < smin || sp > smax)) || (det < 0.0f && (sp > smin || sp < smax))) return BF_HIT_MISS;
    const float inv = 1.0f / det;
    const float t = sp * inv;
    if (!BARY) {
        if (!bf_isfinite(t)) return BF_HIT_NUMERICAL_FAILURE;
        if (t < t_min || t > t_max) return BF_HIT_MISS;
        r.hit = true;
        r.t = t == 0.0f ? 0.0f : t;
        return BF_HIT_MISS;
    }
    const float b0 = e0 * inv, b1 = e1 * inv, b2 = e2 * inv;
    if (!bf_isfinite(t) || !bf_isfinite(b0) || !bf_isfinite(b1) || !bf_isfinite(b2)) return BF_HIT_NUMERICAL_FAILURE;
    if (t < t_min || t > t_max) return BF_HIT_MISS;
    r.hit = true;
    r.t = t == 0.0f ? 0.0f : t;
    r.b0 = b0 == 0.0f ? 0.0f : b0;
    r.b1 = b1 == 0.0f ? 0.0f : b1;
    r.b2 = b2 == 0.0f ? 0.0f : b2;
    return BF_HIT_MISS;
}

/* CudaSceneClosestHit.cu:873-881 */
BF_D bool precedes(float t, uint32_t inst, uint32_t prim, const trav_hit& c) {
    return !c.present || t < c.t || (t == c.t && (inst < c.instance_id || (inst == c.instance_id && prim < c.primitive)));
}

/*
 * Per-lane traversal stack of (node reference, near bits) pairs. The top S entries live in LDS,
 * laid out [entry][lane] so that a wave's pushes and pops touch consecutive banks; when a push
 * finds them full, the lower half moves to the lane's column of a global spill area and comes
 * back when the LDS part runs empty. With S = BF_STACK_DEPTH nothing ever spills.
 */
template <uint32_t S, uint32_t LANES>
struct trav_stack {
    uint2* lds;             /* entry k at lds[k * LANES] (LANES = lanes of the workgroup) */
    uint2* spill;           /* entry k at spill[k * spill_stride] */
    uint32_t spill_stride;
    uint32_t sp, gsp;       /* entries in LDS, entries spilled */
    uint32_t cap;           /* entries the lane may hold in all: S without a spill area */
};

template <uint32_t S, uint32_t LANES>
BF_D uint32_t stack_depth(const trav_stack<S, LANES>& s) { return s.sp + s.gsp; }

template <uint32_t S, uint32_t LANES>
BF_D void stack_push(trav_stack<S, LANES>& s, uint2 entry) {
    if (S < BF_STACK_DEPTH && s.sp == S) {
#pragma unroll
        for (uint32_t k = 0; k < S / 2; ++k) s.spill[(size_t)(s.gsp + k) * s.spill_stride] = s.lds[k * LANES];
#pragma unroll
        for (uint32_t k = 0; k < S / 2; ++k) s.lds[k * LANES] = s.lds[(k + S / 2) * LANES];
        s.gsp += S / 2;
        s.sp = S / 2;
    }
    s.lds[s.sp * LANES] = entry;
    ++s.sp;
}

template <uint32_t S, uint32_t LANES>
BF_D uint2 stack_pop(trav_stack<S, LANES>& s) {
    if (S < BF_STACK_DEPTH && s.sp == 0u) {
        s.gsp -= S / 2;
#pragma unroll
        for (uint32_t k = 0; k < S / 2; ++k) s.lds[k * LANES] = s.spill[(size_t)(s.gsp + k) * s.spill_stride];
        s.sp = S / 2;
    }
    --s.sp;
    return s.lds[s.sp * LANES];
}

/*
 * Traverses TLAS then BLASes depth-first in the reference's order. Returns a traversal status
 * (BF_HIT_MISS = no error). The stack must be empty on entry and is empty on a normal return.
 * `world(o, d)` re-reads the world-space ray when a BLAS is left, so that no copy of it stays
 * live through the loop. The winning candidate keeps (t, ids, leaf); its barycentrics are
 * re-derived by closest_hit_barycentrics from the same deterministic triangle test.
 */
/* The world-space ray of a traversal, re-read through its functor when a BLAS is left. A functor
 * type may provide its own overload (found by argument-dependent lookup) that keeps the
 * reciprocals instead of recomputing them. */
template <typename WorldRay>
BF_D void world_space(const WorldRay& world, trav_space& s) {
    v3f o, d;
    world(o, d);
    make_space(o, d, s);
}

/* True when the next step pops a leaf (a triangle or an instance) rather than an interior node. */
template <uint32_t S, uint32_t LANES>
BF_D bool leaf_next(const trav_stack<S, LANES>& s) {
    return s.sp != 0u && (s.lds[(s.sp - 1u) * LANES].x & BF_WIDE_LEAF) != 0u;
}

/* Loop state of one ray's traversal between steps (the stack lives in trav_stack). */
struct trav_state {
    trav_space space;
    float t_min, t_max;
    uint32_t mask;
    uint32_t in_blas, blas_base;
    uint32_t inst_id, scene_inst;
};

/* Starts a ray: empties the stack and pushes the TLAS root when the ray meets its box. On a normal
 * return an empty stack means the ray is already done. */
template <bool COUNT, uint32_t S, uint32_t LANES>
BF_D uint32_t traverse_begin(const bf_dev_scene& S_, const trav_ray& ray, trav_stack<S, LANES>& stack, trav_state& ts,
                             trav_hit& best, bool& occluded, trav_counters& ctr) {
    best.present = false;
    best.t = ray.t_max;
    occluded = false;
    stack.sp = 0u;
    stack.gsp = 0u;
    if (S_.tlas_root_ref == BF_INVALID_INDEX) return BF_HIT_MISS;
    ts.t_min = ray.t_min;
    ts.t_max = ray.t_max;
    ts.mask = ray.mask;
    ts.in_blas = 0u;
    ts.blas_base = 0u;
    ts.inst_id = 0u;
    ts.scene_inst = 0u;
    make_space(ray.o, ray.d, ts.space);
    bool h;
    child_near nv;
    const uint32_t st = slab(S_.tlas_root_box[0], S_.tlas_root_box[1], S_.tlas_root_box[2], S_.tlas_root_box[3],
                             S_.tlas_root_box[4], S_.tlas_root_box[5], ts.space, ts.t_min, ts.t_max, S_.tlas_root_ref,
                             h, nv);
    if (st != BF_HIT_MISS) return st;
    if (!h) {
        if (COUNT) ctr.nodes++;
        return BF_HIT_MISS;
    }
    stack_push(stack, nv.entry);
    return BF_HIT_MISS;
}

/* One step: pops one stack entry and processes it (a node's two child boxes, an instance entry,
 * or a triangle). Returns a traversal status (BF_HIT_MISS = no error); the ray is done when the
 * stack is empty. `world(o, d)` re-reads the world-space ray when a BLAS is left, so that no copy
 * of it stays live between steps. */
template <bool ANY, bool COUNT, uint32_t S, uint32_t LANES, typename WorldRay>
BF_D uint32_t traverse_step(const bf_dev_scene& S_, WorldRay world, trav_stack<S, LANES>& stack, trav_state& ts,
                            trav_hit& best, bool& occluded, trav_counters& ctr) {
    if (ts.in_blas && stack_depth(stack) <= ts.blas_base) {
        ts.in_blas = 0u;
        world_space(world, ts.space); /* back in world space (recomputing is exact) */
    }
    const uint2 entry = stack_pop(stack);
    const uint32_t ref = entry.x & ~BF_NEAR_EXACT;
    if (COUNT) ctr.nodes++;
    const float tcur = ANY ? ts.t_max : best.t;
    /* the reference's pop test: the entry's binary64 near value against the current t_max */
    uint32_t visit = near_decision(entry, tcur);
    if (visit == 2u) {
        const uint32_t index = ref & BF_REF_MASK;
        const float* box = (ref & BF_WIDE_LEAF) ? (ts.in_blas ? S_.blas_leaf_boxes : S_.tlas_leaf_boxes)
                                                : (ts.in_blas ? S_.blas_node_boxes : S_.tlas_node_boxes);
        visit = near_exact(box + 6 * (size_t)index, ts.space, ts.t_min) <= (double)tcur ? 1u : 0u;
    }
    if (visit == 0u) return BF_HIT_MISS;
    if (ref & BF_WIDE_LEAF) {
        const uint32_t leaf = ref & BF_REF_MASK;
        if (!ts.in_blas) {
            const bf_dev_tlas_leaf L = S_.tlas_leaves[leaf];
            if ((ts.mask & L.visibility_mask) == 0u) return BF_HIT_MISS;
            const v3f lo = xform_point(L.world_to_local, ts.space.o);
            const v3f ld = xform_vector(L.world_to_local, ts.space.d);
            if (!fin3(lo) || !fin3(ld) || (ld.x == 0.0f && ld.y == 0.0f && ld.z == 0.0f)) return BF_HIT_NUMERICAL_FAILURE;
            if (COUNT) ctr.instances++;
            const bf_dev_blas B = S_.blases[L.blas_index];
            trav_space local;
            make_space(lo, ld, local);
            bool h;
            child_near nv;
            const uint32_t st = slab(B.root_box[0], B.root_box[1], B.root_box[2], B.root_box[3], B.root_box[4],
                                     B.root_box[5], local, ts.t_min, tcur, B.root_ref, h, nv);
            if (st != BF_HIT_MISS) return st;
            if (!h) {
                if (COUNT) ctr.nodes++;
                return BF_HIT_MISS;
            }
            /* the BLAS root opens the BLAS part of the stack (a fresh stack in the reference) */
            if (stack_depth(stack) + 1u > stack.cap) return BF_HIT_STACK_OVERFLOW;
            ts.inst_id = L.instance_id;
            ts.scene_inst = L.scene_instance;
            ts.in_blas = 1u;
            ts.blas_base = stack_depth(stack);
            ts.space = local;
            stack_push(stack, nv.entry);
            return BF_HIT_MISS;
        }
        const float4 q0 = S_.leaf_triangles[3 * (size_t)leaf + 0];
        const float4 q1 = S_.leaf_triangles[3 * (size_t)leaf + 1];
        const float4 q2 = S_.leaf_triangles[3 * (size_t)leaf + 2];
        const uint32_t prim = __float_as_uint(q0.w);
        if (COUNT) ctr.triangles++;
        tri_result tr;
        tr.hit = false;
        trav_shear sh;
        make_shear(ts.space.d, sh);
        const uint32_t st = watertight<false>(mk(q0.x, q0.y, q0.z), mk(q1.x, q1.y, q1.z), mk(q2.x, q2.y, q2.z), ts.space.o,
                                              ts.space.d, sh, ts.t_min, tcur, tr);
        if (st != BF_HIT_MISS) return st;
        if (!tr.hit) return BF_HIT_MISS;
        if (ANY) {
            occluded = true;
            stack.sp = 0;
            stack.gsp = 0;
            return BF_HIT_MISS;
        }
        if (!precedes(tr.t, ts.inst_id, prim, best)) return BF_HIT_MISS;
        best.present = true;
        best.t = tr.t;
        best.instance_id = ts.inst_id;
        best.primitive = prim;
        best.scene_instance = ts.scene_inst;
        best.leaf = leaf;
        return BF_HIT_MISS;
    }
    const bf_wide_node* nodes = ts.in_blas ? S_.blas_nodes : S_.tlas_nodes;
    const bf_wide_node N = nodes[ref];
    bool h0, h1;
    child_near n0, n1;
    uint32_t st = slab(N.lo_x[0], N.lo_y[0], N.lo_z[0], N.hi_x[0], N.hi_y[0], N.hi_z[0], ts.space, ts.t_min, tcur,
                       N.child[0], h0, n0);
    if (st != BF_HIT_MISS) return st;
    st = slab(N.lo_x[1], N.lo_y[1], N.lo_z[1], N.hi_x[1], N.hi_y[1], N.hi_z[1], ts.space, ts.t_min, tcur, N.child[1], h1,
              n1);
    if (st != BF_HIT_MISS) return st;
    if (!h0 && !h1) return BF_HIT_MISS;
    /* entries of the current structure (TLAS, or the BLAS above blas_base) against its own cap */
    const uint32_t depth = stack_depth(stack), used = ts.in_blas ? depth - ts.blas_base : depth;
    const uint32_t pushes = h0 && h1 ? 2u : 1u;
    if (used + pushes > BF_STRUCTURE_STACK || depth + pushes > stack.cap) return BF_HIT_STACK_OVERFLOW;
    if (h0 && h1) {
        /* near child first: RN32 of the binary64 near values, ties to the first child (the lower
         * canonical index). Bounds of both roundings decide it unless they overlap; then both
         * near values are formed exactly. */
        bool first_near;
        if (n0.key_hi <= n1.key_lo) first_near = true;
        else if (n0.key_lo > n1.key_hi) first_near = false;
        else if (n0.tie_axis < 4u && n0.tie_axis == n1.tie_axis && (n0.tie_axis == 3u || n0.tie_plane == n1.tie_plane))
            first_near = true; /* both near values are t_min, or the same plane's quotient: a tie */
        else {
            const float b0[6] = {N.lo_x[0], N.lo_y[0], N.lo_z[0], N.hi_x[0], N.hi_y[0], N.hi_z[0]};
            const float b1[6] = {N.lo_x[1], N.lo_y[1], N.lo_z[1], N.hi_x[1], N.hi_y[1], N.hi_z[1]};
            exact_child(N.child[0], near_exact(b0, ts.space, ts.t_min), n0);
            exact_child(N.child[1], near_exact(b1, ts.space, ts.t_min), n1);
            first_near = n0.key_lo <= n1.key_lo;
        }
        stack_push(stack, first_near ? n1.entry : n0.entry);
        stack_push(stack, first_near ? n0.entry : n1.entry);
    } else {
        stack_push(stack, h0 ? n0.entry : n1.entry);
    }
    return BF_HIT_MISS;
}

/*
 * Traverses TLAS then BLASes depth-first in the reference's order. Returns a traversal status
 * (BF_HIT_MISS = no error). The stack must be empty on entry and is empty on a normal return.
 * The winning candidate keeps (t, ids, leaf); its barycentrics are re-derived by
 * closest_hit_barycentrics from the same deterministic triangle test.
 */
template <bool ANY, bool COUNT, uint32_t S, uint32_t LANES, typename WorldRay>
BF_D uint32_t traverse(const bf_dev_scene& S_, const trav_ray& ray, WorldRay world, trav_stack<S, LANES>& stack, trav_hit& best,
                       bool& occluded, trav_counters& ctr) {
    trav_state ts;
    uint32_t st = traverse_begin<COUNT>(S_, ray, stack, ts, best, occluded, ctr);
    while (st == BF_HIT_MISS && stack_depth(stack) != 0u) st = traverse_step<ANY, COUNT>(S_, world, stack, ts, best, occluded, ctr);
    return st;
}

/* Barycentrics of the winning candidate: the triangle test is a pure function of the triangle, the
 * object-space ray and the parameter range, and the winner passed it within [t_min, t_max] of the
 * ray (re-testing against its own t could fail: the range test compares sp with t_max * det
 * before the division that produces t). */
BF_D uint32_t closest_hit_barycentrics(const bf_dev_scene& S, v3f o, v3f d, float t_min, float t_max, trav_hit& best) {
    const float* m = S.instances[best.scene_instance].world_to_local;
    const v3f lo = xform_point(m, o), ld = xform_vector(m, d);
    const float4 q0 = S.leaf_triangles[3 * (size_t)best.leaf + 0];
    const float4 q1 = S.leaf_triangles[3 * (size_t)best.leaf + 1];
    const float4 q2 = S.leaf_triangles[3 * (size_t)best.leaf + 2];
    tri_result tr;
    tr.hit = false;
    trav_shear sh;
    make_shear(ld, sh);
    const uint32_t st = watertight<true>(mk(q0.x, q0.y, q0.z), mk(q1.x, q1.y, q1.z), mk(q2.x, q2.y, q2.z), lo, ld, sh,
                                         t_min, t_max, tr);
    if (st != BF_HIT_MISS) return st;
    if (!tr.hit || tr.t != best.t) return BF_HIT_NUMERICAL_FAILURE;
    best.b0 = tr.b0;
    best.b1 = tr.b1;
    best.b2 = tr.b2;
    return BF_HIT_MISS;
}

/* CudaSceneClosestHit.cu:826-871 */
BF_D bool world_normal(const float* m, v3f l0, v3f l1, v3f l2, v3f& n) {
    const v3f w0 = xform_point(m, l0), w1 = xform_point(m, l1), w2 = xform_point(m, l2);
    if (!fin3(w0) || !fin3(w1) || !fin3(w2)) return false;
    const v3f f = sub3(w1, w0), s = sub3(w2, w0);
    if (!fin3(f) || !fin3(s)) return false;
    const double cx = (double)f.y * (double)s.z - (double)f.z * (double)s.y;
    const double cy = (double)f.z * (double)s.x - (double)f.x * (double)s.z;
    const double cz = (double)f.x * (double)s.y - (double)f.y * (double)s.x;
    const double mx = fmax(fabs(cx), fmax(fabs(cy), fabs(cz)));
    if (!bf_isfinited(cx) || !bf_isfinited(cy) || !bf_isfinited(cz) || !bf_isfinited(mx) || mx == 0.0) return false;
    const double sx = cx / mx, sy = cy / mx, sz = cz / mx;
    const double len = sqrt(sx * sx + sy * sy + sz * sz);
    if (!bf_isfinited(len) || len == 0.0) return false;
    n = mk((float)(sx / len), (float)(sy / len), (float)(sz / len));
    return fin3(n);
}

/* CudaSceneClosestHit.cu:397-405 */
BF_D bool valid_traversal_ray(const bf_transport_ray& r) {
    const v3f o = mk(r.origin_x, r.origin_y, r.origin_z);
    const v3f d = mk(r.direction_x, r.direction_y, r.direction_z);
    return fin3(o) && fin3(d) && (d.x != 0.0f || d.y != 0.0f || d.z != 0.0f) && bf_isfinite(r.t_min) && r.t_min >= 0.0f &&
           !bf_isnan(r.t_max) && r.t_max >= r.t_min && bf_isfinite(r.time) && r.time >= 0.0f && r.time <= 1.0f &&
           r.reserved == 0u;
}

#endif
