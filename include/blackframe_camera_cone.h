/* Primary-ray cones of the pinhole camera: the one-pixel ray differentials of a raster sample
 * (PinholeCamera.hpp:224-252) reduced to a circular cone (RayCone.hpp:194-220), in binary32.
 *
 * The reference builds these cones on the host for every camera path of its wavefront frame; this
 * header is shared by the device camera stage and the CPU restatement so that frames whose
 * materials carry normal or bump maps can start their paths with the same footprint. Its results
 * are pinned against a binary64 evaluation and the closed form of the image-centre pixel.
 *
 * Every function returns 1 on success and 0 where the reference reports an invalid camera input
 * or an unrepresentable cone. */
#ifndef BLACKFRAME_CAMERA_CONE_H
#define BLACKFRAME_CAMERA_CONE_H

#include "blackframe_hip_abi.h"
#include "blackframe_math.h"

/* std::hypot(x, y, z) as the reference's host library forms it (libstdc++ <cmath>, __hypot3): the
 * largest magnitude a times sqrt((x/a)^2 + (y/a)^2 + (z/a)^2), every step in binary32. */
BF_HD float bf_hypot3f(float x, float y, float z) {
    x = fabsf(x);
    y = fabsf(y);
    z = fabsf(z);
    const float a = x < y ? (y < z ? z : y) : (x < z ? z : x);
    if (a == 0.0f) return 0.0f;
    const float xa = x / a, ya = y / a, za = z / a;
    return a * sqrtf(xa * xa + ya * ya + za * za);
}

BF_HD int bf_finite3(const float v[3]) { return bf_isfinite(v[0]) && bf_isfinite(v[1]) && bf_isfinite(v[2]); }

/* pinhole_camera_detail::robust_unit_direction: scaled by the largest magnitude, then normalized */
BF_HD int bf_pinhole_unit(const float v[3], float out[3]) {
    if (!bf_finite3(v)) return 0;
    const float m = fmaxf(fabsf(v[0]), fmaxf(fabsf(v[1]), fabsf(v[2])));
    if (m == 0.0f) return 0;
    const float s[3] = {v[0] / m, v[1] / m, v[2] / m};
    const float mag = sqrtf(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    for (int k = 0; k < 3; ++k) out[k] = s[k] / mag;
    return bf_finite3(out);
}

/* generate_primary_ray_from_normalized_raster (PinholeCamera.hpp:255-276): the world direction */
BF_HD int bf_pinhole_direction(const bf_camera_desc* cam, float nx, float ny, float out[3]) {
    const float v[3] = {nx * cam->half_width, ny * cam->half_height, -1.0f};
    float u[3];
    if (!bf_pinhole_unit(v, u)) return 0;
    const float w[3] = {cam->tangent[0] * u[0] + cam->bitangent[0] * u[1] + cam->normal[0] * u[2],
                        cam->tangent[1] * u[0] + cam->bitangent[1] * u[1] + cam->normal[1] * u[2],
                        cam->tangent[2] * u[0] + cam->bitangent[2] * u[1] + cam->normal[2] * u[2]};
    return bf_pinhole_unit(w, out);
}

/* ray_cone_detail::unit_direction (RayCone.hpp:27-44): scaled by the largest magnitude, divided by
 * its three-argument hypot */
BF_HD int bf_cone_unit(const float v[3], float out[3]) {
    if (!bf_finite3(v)) return 0;
    const float m = fmaxf(fabsf(v[0]), fmaxf(fabsf(v[1]), fabsf(v[2])));
    if (!(m > 0.0f)) return 0;
    const float s[3] = {v[0] / m, v[1] / m, v[2] / m};
    const float mag = bf_hypot3f(s[0], s[1], s[2]);
    for (int k = 0; k < 3; ++k) out[k] = s[k] / mag;
    return bf_isfinite(mag) && mag > 0.0f && bf_finite3(out);
}

/* ray_cone_detail::differential_cross_section (RayCone.hpp:107-135): where the auxiliary ray
 * crosses the plane through the central origin normal to the central direction (offset) and how
 * fast that crossing moves per unit distance along the central ray (slope). `delta` is the
 * auxiliary origin minus the central origin. */
BF_HD int bf_cone_cross_section(const float delta[3], const float central_dir[3], const float aux_dir[3],
                                float offset[3], float slope[3]) {
    float c[3], a[3];
    if (!bf_cone_unit(central_dir, c) || !bf_cone_unit(aux_dir, a)) return 0;
    const float den = a[0] * c[0] + a[1] * c[1] + a[2] * c[2];
    if (!bf_isfinite(den) || !(den > 0.0f)) return 0;
    const float proj = delta[0] * c[0] + delta[1] * c[1] + delta[2] * c[2];
    const float t0 = -proj / den;
    const float t1 = (1.0f - proj) / den;
    for (int k = 0; k < 3; ++k) {
        offset[k] = delta[k] + t0 * a[k];
        const float at_unit = delta[k] + t1 * a[k] - c[k];
        slope[k] = at_unit - offset[k];
    }
    return bf_isfinite(t0) && bf_isfinite(t1) && bf_finite3(offset) && bf_finite3(slope);
}

/* ray_cone_detail::maximum_singular_value (RayCone.hpp:58-82): the larger singular value of the
 * 3x2 matrix [p q], from the 2x2 Gram matrix of the columns scaled by their largest entry */
BF_HD int bf_cone_max_singular(const float p[3], const float q[3], float* out) {
    if (!bf_finite3(p) || !bf_finite3(q)) return 0;
    float scale = 0.0f;
    for (int k = 0; k < 3; ++k) scale = fmaxf(scale, fmaxf(fabsf(p[k]), fabsf(q[k])));
    if (scale == 0.0f) {
        *out = 0.0f;
        return 1;
    }
    const float x[3] = {p[0] / scale, p[1] / scale, p[2] / scale};
    const float y[3] = {q[0] / scale, q[1] / scale, q[2] / scale};
    const float a = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
    const float b = x[0] * y[0] + x[1] * y[1] + x[2] * y[2];
    const float c = y[0] * y[0] + y[1] * y[1] + y[2] * y[2];
    const float largest = 0.5f * (a + c + bf_hypotf(a - c, 2.0f * b));
    const float normalized = sqrtf(fmaxf(0.0f, largest));
    const float result = scale * normalized;
    *out = result;
    return bf_isfinite(result) && result >= 0.0f && !(normalized != 0.0f && result == 0.0f);
}

/* generate_primary_ray_cone (PinholeCamera.hpp:224-252, RayCone.hpp:194-220): the cone of the
 * primary ray through raster point (raster_x, raster_y) = pixel + sub-pixel offset, whose unit
 * world direction is `dir`. The auxiliary rays pass one pixel right and one pixel down (past the
 * film's last column and row too); both start at the camera origin, so the width is zero and the
 * spread is the larger singular value of the two unit-distance slopes. */
BF_HD int bf_camera_cone(const bf_camera_desc* cam, float raster_x, float raster_y, const float dir[3], float* width,
                         float* spread) {
    const float w = (float)cam->width, h = (float)cam->height;
    float rx[3], ry[3];
    if (!bf_pinhole_direction(cam, 2.0f * (raster_x + 1.0f) / w - 1.0f, 1.0f - 2.0f * raster_y / h, rx)) return 0;
    if (!bf_pinhole_direction(cam, 2.0f * raster_x / w - 1.0f, 1.0f - 2.0f * (raster_y + 1.0f) / h, ry)) return 0;
    if ((rx[0] == dir[0] && rx[1] == dir[1] && rx[2] == dir[2]) || (ry[0] == dir[0] && ry[1] == dir[1] && ry[2] == dir[2]))
        return 0;
    const float zero[3] = {0.0f, 0.0f, 0.0f};
    float ox[3], sx[3], oy[3], sy[3];
    if (!bf_cone_cross_section(zero, dir, rx, ox, sx) || !bf_cone_cross_section(zero, dir, ry, oy, sy)) return 0;
    if (!bf_cone_max_singular(ox, oy, width) || !bf_cone_max_singular(sx, sy, spread)) return 0;
    return bf_isfinite(*width) && *width >= 0.0f && bf_isfinite(*spread) && *spread >= 0.0f;
}

#endif
