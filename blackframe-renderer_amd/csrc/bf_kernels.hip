/*
 * HIP kernels of the MI355X wavefront path-tracing backend (gfx950, wave64).
 *
 * One bounce of the path loop is: trace_closest (BVH2 closest hit; misses resolved in place) ->
 * stable compaction into the hit queue -> shade (emission + MIS, NEE light sample, BSDF sample,
 * Russian roulette) -> stable compaction into shadow / continuation queues -> trace_shadow
 * (any-hit; visible contributions accumulated) -> stable append of shadow continuations.
 * Stage semantics follow Blackframe/XPU/CUDA/Source/CudaWavefrontStages.cu; queue routing follows
 * Blackframe/Backends/GPU/CUDA/Source/WavefrontTransport.cpp:2222-2649.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "../../include/blackframe_camera_cone.h"
#include "../../include/blackframe_texture.h"
#include "bf_device_adjoint.h"
#include "bf_device_maps.h"
#include "bf_kernels.h"
#include "bf_scene_soa.h"
#include "bf_shading.h"
#include "bf_traversal.h"

#define BF_BLOCK 256u

/* RGB basis spectra of textured reflectance, [3][BF_RGB_BASIS_SAMPLES] */
__constant__ float c_rgb_basis[3 * BF_RGB_BASIS_SAMPLES] = {
#include "data/rgb_basis.inc"
};

/* ---- path helpers ----------------------------------------------------------------------------- */
/* The frame a slot belongs to: the kernel argument in a single-frame launch, the job table entry
 * recorded for the slot by the camera stage in a batched one. */
template <bool BATCH>
BF_D const bf_dev_job& job_at(const bf_transport_params& P, uint32_t slot) {
    if constexpr (BATCH) return P.jobs[P.paths.job[slot]];
    else return P.job0;
}

BF_D uint64_t path_of(const bf_dev_job& J, uint32_t slot) { return J.path_base + (slot - J.slot_begin); }

BF_D sample_key key_for_slot(const bf_dev_job& J, uint32_t slot) {
    const uint64_t global = path_of(J, slot);
    if (J.in_samples) {
        const bf_sample_stream_index in = J.in_samples[global];
        return sample_key{in.pixel_x, in.pixel_y, in.sample_index, in.seed};
    }
    const uint64_t s = global / J.pixels;
    const uint32_t rem = (uint32_t)(global - s * J.pixels);
    sample_key k;
    k.px = rem % J.camera.width;
    k.py = rem / J.camera.width;
    k.s = (uint64_t)J.first_sample_index + s;
    k.seed = J.seed;
    return k;
}

/* Marks a path failed; the workspace keeps the failure count and the lowest (job, path, code). */
BF_D void fail_path(const bf_transport_params& P, const bf_dev_job& J, uint32_t slot, uint32_t code) {
    P.paths.status[slot] = BF_TERM_FAILED | (code << 8);
    atomicAdd(&P.failure[0], 1u);
    atomicMin((unsigned long long*)(P.failure + 2),
              ((unsigned long long)J.index << 48) | ((unsigned long long)path_of(J, slot) << 8) | code);
}

/* Adds the number of active lanes to *counter with one atomic per wave. */
BF_D void wave_count(uint32_t* counter) {
    const uint64_t active = __ballot(1);
    if ((threadIdx.x & 63u) == (uint32_t)(__ffsll((unsigned long long)active) - 1)) atomicAdd(counter, (uint32_t)__popcll(active));
}

BF_D float spectrum_at(const bf_dev_scene& S, uint32_t id, float wavelength) {
    if (id == BF_INVALID_INDEX) return 0.0f;
    const bf_spectrum_desc sp = S.spectra[id];
    const float* v = S.spectrum_values + sp.offset;
    if (sp.count & BF_SPECTRUM_SAMPLED) {
        /* four (wavelength, value) pairs; the host admits only frames whose packet is these
         * wavelengths, so one lane always matches */
#pragma unroll
        for (int lane = 0; lane < 4; ++lane)
            if (v[lane] == wavelength) return v[4 + lane];
        return __builtin_nanf("");
    }
    if (sp.count <= 1u) return v[0];
    const float step = 470.0f / (float)(sp.count - 1u);
    const float t = (wavelength - 360.0f) / step;
    uint32_t i = t <= 0.0f ? 0u : (uint32_t)t;
    if (i > sp.count - 2u) i = sp.count - 2u;
    const float f = t - (float)i;
    return fmaf(v[i + 1] - v[i], f, v[i]);
}

/* PinholeCamera.hpp:140-301 with PixelJitter.hpp:37-63 */
BF_D bool robust_unit(v3f d, v3f& out) {
    if (!fin3(d)) return false;
    const float m = fmaxf(fabsf(d.x), fmaxf(fabsf(d.y), fabsf(d.z)));
    if (m == 0.0f) return false;
    const v3f s = mk(d.x / m, d.y / m, d.z / m);
    const float mag = sqrtf(s.x * s.x + s.y * s.y + s.z * s.z);
    out = mk(s.x / mag, s.y / mag, s.z / mag);
    return fin3(out);
}

/* `cone` (optional) receives the primary ray's cone (width, spread) from its one-pixel
 * differentials, for frames whose materials carry normal or bump maps */
BF_D bool camera_ray(const bf_camera_desc& cam, const sampler_ctx& sc, const sample_key& k, v3f& origin, v3f& dir,
                     float2* cone = nullptr) {
    float ox = 0.5f, oy = 0.5f;
    if (cam.jitter_mode == BF_JITTER_UNIFORM) {
        ox = sample_1d(sc, k, BF_DIM_RASTER_X);
        oy = sample_1d(sc, k, BF_DIM_RASTER_Y);
    }
    if (k.px >= cam.width || k.py >= cam.height || !(ox >= 0.0f) || !(ox < 1.0f) || !(oy >= 0.0f) || !(oy < 1.0f))
        return false;
    const float w = (float)cam.width, h = (float)cam.height;
    const float tx = 2.0f * (float)k.px;
    const float ty = 2.0f * (float)k.py;
    const float nx = (tx + 2.0f * ox) / w - 1.0f;
    const float ny = 1.0f - (ty + 2.0f * oy) / h;
    v3f unit;
    if (!robust_unit(mk(nx * cam.half_width, ny * cam.half_height, -1.0f), unit)) return false;
    const v3f wd = mk(cam.tangent[0] * unit.x + cam.bitangent[0] * unit.y + cam.normal[0] * unit.z,
                      cam.tangent[1] * unit.x + cam.bitangent[1] * unit.y + cam.normal[1] * unit.z,
                      cam.tangent[2] * unit.x + cam.bitangent[2] * unit.y + cam.normal[2] * unit.z);
    if (!robust_unit(wd, dir)) return false;
    origin = mk(cam.origin[0], cam.origin[1], cam.origin[2]);
    if (cone) {
        const float unit_dir[3] = {dir.x, dir.y, dir.z};
        float width, spread;
        if (!bf_camera_cone(&cam, (float)k.px + ox, (float)k.py + oy, unit_dir, &width, &spread)) return false;
        *cone = make_float2(width, spread);
    }
    return true;
}

/* ---- camera stage: device-side primary rays, wavelengths and initial state ------------------- */
__constant__ double c_cie_table_rows[471][3] = {
#include "data/cie1931_cmf.inc"
};

/* color_matching_functions (Cie1931Sensor.cpp:69-93) */
BF_D void cie_cmf(float wavelength, float* response) {
    const float coordinate = wavelength - 360.0f;
    const uint32_t lower_index = (uint32_t)coordinate;
    if (lower_index + 1u >= 471u) {
        for (int c = 0; c < 3; ++c) response[c] = (float)c_cie_table_rows[470][c];
    } else {
        const float fraction = coordinate - (float)lower_index;
        for (int c = 0; c < 3; ++c) {
            const float lower = (float)c_cie_table_rows[lower_index][c];
            const float upper = (float)c_cie_table_rows[lower_index + 1][c];
            response[c] = fmaf(upper - lower, fraction, lower);
        }
    }
}

/* d(film RGB)/d(radiance lane) for the path's pixel, contracted with dLoss/d(film RGB): the film
 * adds rgb = M * sum_lanes L cmf(lambda) normalization / pdf per path (cie_xyz below). */
BF_D float4 film_adjoint(const bf_transport_params& P, const bf_dev_job& J, const sample_key& k, const float* wl) {
    const size_t pixel = (size_t)k.py * J.camera.width + k.px;
    const float d[3] = {J.film_gradient[3 * pixel], J.film_gradient[3 * pixel + 1], J.film_gradient[3 * pixel + 2]};
    const float dx = 3.2404542f * d[0] - 0.9692660f * d[1] + 0.0556434f * d[2];
    const float dy = -1.5371385f * d[0] + 1.8760108f * d[1] - 0.2040259f * d[2];
    const float dz = -0.4985314f * d[0] + 0.0415560f * d[1] + 1.0572252f * d[2];
    const float scale = P.film_normalization / P.film_pdf;
    float a[4];
#pragma unroll
    for (int lane = 0; lane < 4; ++lane) {
        float r[3];
        cie_cmf(wl[lane], r);
        a[lane] = (dx * r[0] + dy * r[1] + dz * r[2]) * scale;
    }
    return make_float4(a[0], a[1], a[2], a[3]);
}

/* camera_stage_kernel's path-input form (CudaWavefrontStages.cu:1347-1405, validity checks
 * :528-575): the slot takes its path's ray, ray cone and path state. Depth counters above 255 (kept
 * in bytes here) and participating media fail the path. */
BF_D bool valid_input_ray(const bf_transport_ray& r) {
    const float d2 = r.direction_x * r.direction_x + r.direction_y * r.direction_y + r.direction_z * r.direction_z;
    return bf_isfinite(r.origin_x) && bf_isfinite(r.origin_y) && bf_isfinite(r.origin_z) && bf_isfinite(r.t_min) &&
           r.t_min >= 0.0f && !isnan(r.t_max) && r.t_max >= r.t_min && bf_isfinite(r.time) && r.time >= 0.0f &&
           r.time <= 1.0f && bf_isfinite(r.direction_x) && bf_isfinite(r.direction_y) && bf_isfinite(r.direction_z) &&
           d2 > 0.0f && r.reserved == 0u;
}

BF_D bool valid_cone(float w, float s) { return bf_isfinite(w) && w >= 0.0f && bf_isfinite(s) && s >= 0.0f; }

BF_D bool valid_input_state(const bf_path_state_lane& s) {
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (!bf_isfinite(s.beta.values[k]) || s.beta.values[k] < 0.0f || !bf_isfinite(s.accumulated_radiance.values[k]) ||
            s.accumulated_radiance.values[k] < 0.0f)
            return false;
    if (!bf_isfinite(s.eta_scale) || !(s.eta_scale > 0.0f)) return false;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float w = s.wavelength_nanometers.values[k], pdf = s.wavelength_pdf_values.values[k];
        if (!bf_isfinite(w) || w < 360.0f || w > 830.0f || !bf_isfinite(pdf) || !(pdf > 0.0f) || s.wavelength_pdf_measures[k] != 5u)
            return false;
    }
#pragma unroll
    for (int k = 0; k < 6; ++k)
        if (s.reserved[k] != 0u) return false;
    const uint64_t total = (uint64_t)s.diffuse_depth + s.glossy_depth + s.specular_depth + s.volume_depth;
    const uint64_t surface = (uint64_t)s.diffuse_depth + s.glossy_depth + s.specular_depth;
    return total == s.depth && s.transmission_depth <= surface && (s.delta_flags & ~3u) == 0u;
}

BF_D bool input_lane(const bf_transport_params& P, const bf_dev_job& J, uint32_t slot) {
    const uint64_t path = path_of(J, slot);
    const bf_transport_ray r = J.in_rays[path];
    const bf_ray_cone c = J.in_cones[path];
    const bf_path_state_lane& s = J.in_states[path];
    P.paths.status[slot] = BF_TERM_NONE;
    P.paths.previous[slot] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    uint32_t code = 0u;
    if (!valid_input_ray(r)) code = 3u;
    else if (!valid_cone(c.width, c.spread) || !valid_input_state(s) || r.current_medium != s.current_medium) code = 4u;
    else if (s.current_medium != 0u || s.volume_depth != 0u) code = 5u;
    else if (s.diffuse_depth > 255u || s.glossy_depth > 255u || s.specular_depth > 255u || s.transmission_depth > 255u)
        code = 6u;
    if (code != 0u) {
        fail_path(P, J, slot, code);
        return false;
    }
    P.paths.beta[slot] = make_float4(s.beta.values[0], s.beta.values[1], s.beta.values[2], s.beta.values[3]);
    P.paths.radiance[slot] = make_float4(s.accumulated_radiance.values[0], s.accumulated_radiance.values[1],
                                         s.accumulated_radiance.values[2], s.accumulated_radiance.values[3]);
    P.paths.wavelengths[slot] = make_float4(s.wavelength_nanometers.values[0], s.wavelength_nanometers.values[1],
                                            s.wavelength_nanometers.values[2], s.wavelength_nanometers.values[3]);
    P.paths.depth[slot] = make_uint4(s.diffuse_depth | (s.glossy_depth << 8) | (s.specular_depth << 16) |
                                         (s.transmission_depth << 24),
                                     s.depth, s.delta_flags | BF_FIRST_RAY, __float_as_uint(s.eta_scale));
    P.paths.visibility[slot] = r.visibility_mask;
    P.paths.cone[slot] = make_float2(c.width, c.spread);
    P.paths.ray_origin[slot] = make_float4(r.origin_x, r.origin_y, r.origin_z, r.time);
    P.paths.ray_direction[slot] = make_float4(r.direction_x, r.direction_y, r.direction_z, r.t_min);
    return true;
}

/* ---- ray cones (CudaWavefrontStages.cu:2816-2867, 3675-3782) ---------------------------------- */
BF_D bool cone_product(float a, float b, float& out) {
    out = a * b;
    return bf_isfinite(out) && !(a != 0.0f && b != 0.0f && out == 0.0f);
}

BF_D bool cone_ratio(float n, float d, float& out) {
    if (!bf_isfinite(n) || !(n > 0.0f) || !bf_isfinite(d) || !(d > 0.0f)) return false;
    out = n / d;
    return bf_isfinite(out) && out > 0.0f;
}

/* the cone at the hit: width + spread x the distance travelled along the (unnormalised) ray */
BF_D bool advance_cone(float2 cone, v3f d, float t, float2& out) {
    if (!valid_cone(cone.x, cone.y) || !bf_isfinite(t) || t < 0.0f) return false;
    const float scale = fmaxf(fabsf(d.x), fmaxf(fabsf(d.y), fabsf(d.z)));
    if (!bf_isfinite(scale) || !(scale > 0.0f)) return false;
    const float inv = 1.0f / scale;
    const float sx = d.x * inv, sy = d.y * inv, sz = d.z * inv;
    const float nl = sqrtf(sx * sx + sy * sy + sz * sz);
    float length, distance, growth;
    if (!bf_isfinite(nl) || !(nl > 0.0f) || !cone_product(scale, nl, length) || !(length > 0.0f)) return false;
    if (!cone_product(t, length, distance) || distance < 0.0f) return false;
    if (!cone_product(cone.y, distance, growth)) return false;
    const float width = fmaf(cone.y, distance, cone.x);
    if (!bf_isfinite(width) || width < cone.x || (growth != 0.0f && width == cone.x)) return false;
    out = make_float2(width, cone.y);
    return valid_cone(out.x, out.y);
}

/* the cone leaving a scattering event: the spread is floored by the sampled closure's roughness
 * (1 for diffuse closures); refraction scales the spread by the eta and cosine ratios and the width
 * by the inverse cosine ratio */
BF_D bool propagate_cone(float2 cone, const closure_t& c, uint32_t event, v3f wo, v3f wi, float2& out) {
    if (!valid_cone(cone.x, cone.y) || !valid_surface_event(event) || !unit_vector(wo) || !unit_vector(wi) ||
        wo.z == 0.0f || wi.z == 0.0f)
        return false;
    const bool reflection = (event & BF_LOBE_REFLECTION) != 0u, transmission = (event & BF_LOBE_TRANSMISSION) != 0u;
    const bool same = (wo.z > 0.0f) == (wi.z > 0.0f);
    if ((reflection && !same) || (transmission && same)) return false;
    float floor_spread = 0.0f, spread_scale = 1.0f, width_scale = 1.0f;
    switch (c.kind) {
    case BF_CLOSURE_LAMBERTIAN:
    case BF_CLOSURE_ROUGH_DIFFUSE:
        if (!reflection || !(event & BF_LOBE_DIFFUSE)) return false;
        floor_spread = 1.0f;
        break;
    case BF_CLOSURE_ROUGH_CONDUCTOR:
        if (!reflection || !(event & BF_LOBE_GLOSSY)) return false;
        if (!bf_isfinite(c.parameters[8]) || c.parameters[8] < 0.0f || !bf_isfinite(c.parameters[9]) ||
            c.parameters[9] < 0.0f)
            return false;
        floor_spread = fmaxf(c.parameters[8], c.parameters[9]);
        break;
    case BF_CLOSURE_ROUGH_DIELECTRIC:
        if (!(event & BF_LOBE_GLOSSY)) return false;
        if (!bf_isfinite(c.parameters[2]) || c.parameters[2] < 0.0f || !bf_isfinite(c.parameters[3]) ||
            c.parameters[3] < 0.0f)
            return false;
        floor_spread = fmaxf(c.parameters[2], c.parameters[3]);
        break;
    case BF_CLOSURE_SPECULAR_REFLECTION:
        if (!reflection || !(event & BF_LOBE_SPECULAR)) return false;
        break;
    case BF_CLOSURE_SPECULAR_TRANSMISSION:
        if (!transmission || !(event & BF_LOBE_SPECULAR)) return false;
        break;
    default:
        return false;
    }
    if (!bf_isfinite(floor_spread) || floor_spread < 0.0f) return false;
    if (transmission) {
        if (c.kind != BF_CLOSURE_ROUGH_DIELECTRIC && c.kind != BF_CLOSURE_SPECULAR_TRANSMISSION) return false;
        const float incident = wo.z > 0.0f ? c.parameters[0] : c.parameters[1];
        const float transmitted = wo.z > 0.0f ? c.parameters[1] : c.parameters[0];
        float eta_ratio, cosine_ratio, inverse_ratio, meridional;
        if (!cone_ratio(incident, transmitted, eta_ratio) || !cone_ratio(fabsf(wo.z), fabsf(wi.z), cosine_ratio) ||
            !cone_ratio(fabsf(wi.z), fabsf(wo.z), inverse_ratio))
            return false;
        if (!cone_product(eta_ratio, cosine_ratio, meridional)) return false;
        spread_scale = fmaxf(eta_ratio, meridional);
        width_scale = fmaxf(1.0f, inverse_ratio);
    }
    float width, spread;
    if (!cone_product(cone.x, width_scale, width) || !cone_product(cone.y, spread_scale, spread)) return false;
    out = make_float2(width, fmaxf(spread, floor_spread));
    return valid_cone(out.x, out.y);
}

template <bool BATCH, bool ADJ>
__global__ __launch_bounds__(BF_BLOCK) void k_camera(bf_transport_params P, uint32_t count, uint32_t* ray_queue,
                                                     uint8_t* route) {
    const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= count) return;
    uint32_t j = 0;
    if constexpr (BATCH) {
        /* jobs are ordered by slot_begin: the last one starting at or before this slot */
        uint32_t lo = 0, hi = P.job_count;
        while (hi - lo > 1u) {
            const uint32_t mid = (lo + hi) >> 1;
            if (P.jobs[mid].slot_begin <= slot) lo = mid;
            else hi = mid;
        }
        j = lo;
        P.paths.job[slot] = (uint16_t)j;
    }
    const bf_dev_job& J = BATCH ? P.jobs[j] : P.job0;
    if (J.in_rays) {
        /* the ray queue is compacted from these routes: paths whose inputs failed never trace */
        ray_queue[slot] = slot;
        route[slot] = input_lane(P, J, slot) ? 1u : 0u;
        return;
    }
    const sample_key k = key_for_slot(J, slot);
    v3f o, d;
    ray_queue[slot] = slot;
    P.paths.status[slot] = BF_TERM_NONE;
    P.paths.beta[slot] = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
    P.paths.radiance[slot] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    P.paths.depth[slot] = make_uint4(0u, 0u, BF_FIRST_RAY, __float_as_uint(1.0f));
    P.paths.previous[slot] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    P.paths.visibility[slot] = J.camera.visibility_mask;
    float wl[4];
    if (P.wavelength_mode == BF_WAVELENGTHS_PER_PATH) {
        const float u = sample_1d(P.sampler, k, 3ull);
        if (!(u >= 0.0f) || !(u < 1.0f)) {
            fail_path(P, J, slot, 2);
            return;
        }
        /* WavelengthSampling.hpp:62-84 */
#pragma unroll
        for (int lane = 0; lane < 4; ++lane) {
            const float offset = (float)lane / 4.0f;
            const float wrap = 1.0f - offset;
            const float shifted = lane != 0 && u >= wrap ? u - wrap : u + offset;
            wl[lane] = fmaf(470.0f, shifted, 360.0f);
        }
    } else {
#pragma unroll
        for (int lane = 0; lane < 4; ++lane) wl[lane] = P.packet[lane];
    }
    P.paths.wavelengths[slot] = make_float4(wl[0], wl[1], wl[2], wl[3]);
    if (J.tape_visibility) {
        const uint64_t path = path_of(J, slot);
        if constexpr (ADJ) {
            /* adjoint lanes keep the tape in the shadow columns, which the replay never traces */
            const unsigned long long vis = J.tape_visibility[path];
            P.paths.shadow_origin[slot] = J.tape_radiance[path];
            P.paths.shadow_direction[slot] =
                make_float4(__uint_as_float((uint32_t)vis), __uint_as_float((uint32_t)(vis >> 32)), 0.0f, 0.0f);
            P.paths.shadow_contribution[slot] = film_adjoint(P, J, k, wl);
        } else {
            J.tape_visibility[path] = 0ull;
        }
    }
    float2 cone;
    if (!camera_ray(J.camera, P.sampler, k, o, d, J.frame_cones ? &cone : nullptr)) {
        fail_path(P, J, slot, 1);
        return;
    }
    if (J.frame_cones) P.paths.cone[slot] = cone;
    P.paths.ray_origin[slot] = make_float4(o.x, o.y, o.z, J.camera.time);
    P.paths.ray_direction[slot] = make_float4(d.x, d.y, d.z, J.camera.t_min);
}

/* ---- closest-hit stage: gather + traversal + classify + miss ---------------------------------- */
/* The traversal kernels are persistent: a grid sized to the device's resident capacity, each wave
 * taking 64 consecutive queue items at a time with a static stride, so that the per-lane stack
 * spill area is bounded by the grid and empty tail workgroups cost nothing. */

/* One lane's ray between traversal steps. */
struct lane_ray {
    uint32_t item, slot;
    float t_min, t_max;
    trav_state ts;
    trav_hit best;
    bool occluded;
};

/* World-space ray of a slot, re-read from the path columns (closest-hit rays). `inv` (optional)
 * holds the lane's world-space reciprocals, entry k at inv[k * stride]. */
struct closest_world {
    const bf_dev_paths* paths;
    uint32_t slot;
    double* inv;
    uint32_t stride;
    BF_D void operator()(v3f& o, v3f& d) const {
        const float4 a = paths->ray_origin[slot];
        const float4 b = paths->ray_direction[slot];
        o = mk(a.x, a.y, a.z);
        d = mk(b.x, b.y, b.z);
    }
};

BF_D void world_space(const closest_world& world, trav_space& s) {
    world(s.o, s.d);
    if (world.inv) {
        s.inv[0] = world.inv[0];
        s.inv[1] = world.inv[world.stride];
        s.inv[2] = world.inv[2u * world.stride];
        finish_space(s);
    } else {
        make_space(s.o, s.d, s);
    }
}

/* Starts queue item i; returns the traversal status of the root test. */
template <bool COUNT, bool BATCH, uint32_t BLOCK>
BF_D uint32_t closest_start(const bf_transport_params& P, const uint32_t* queue, uint32_t i,
                            trav_stack<BF_SHORT_STACK, BLOCK>& stack, lane_ray& L, trav_counters& ctr) {
    L.item = i;
    L.slot = queue[i];
    const bf_dev_job& J = job_at<BATCH>(P, L.slot);
    const float4 ro = P.paths.ray_origin[L.slot];
    const float4 rd = P.paths.ray_direction[L.slot];
    trav_ray ray;
    ray.o = mk(ro.x, ro.y, ro.z);
    ray.d = mk(rd.x, rd.y, rd.z);
    ray.t_min = rd.w;
    ray.t_max = (P.paths.depth[L.slot].z & BF_FIRST_RAY) == 0u
                    ? BF_INFINITY_F
                    : (J.in_rays ? J.in_rays[path_of(J, L.slot)].t_max : J.camera.t_max);
    ray.mask = P.paths.visibility[L.slot];
    L.t_min = ray.t_min;
    L.t_max = ray.t_max;
    bool occluded;
    return traverse_begin<COUNT>(J.scene, ray, stack, L.ts, L.best, occluded, ctr);
}

template <bool COUNT, bool BATCH, uint32_t BLOCK>
BF_D uint32_t closest_step(const bf_transport_params& P, trav_stack<BF_SHORT_STACK, BLOCK>& stack, lane_ray& L,
                           trav_counters& ctr, double* inv) {
    const bf_dev_job& J = job_at<BATCH>(P, L.slot);
    bool occluded;
    return traverse_step<false, COUNT>(J.scene, closest_world{&P.paths, L.slot, inv, BLOCK}, stack, L.ts, L.best,
                                       occluded, ctr);
}

/* Classifies a finished closest-hit query: hit (route 1, barycentrics re-derived for the winner),
 * miss (environment radiance, path escapes) or traversal failure. */
template <bool BATCH>
BF_D void closest_finish(const bf_transport_params& P, uint8_t* route, lane_ray& L, uint32_t st) {
    const uint32_t slot = L.slot;
    const bf_dev_job& J = job_at<BATCH>(P, slot);
    trav_hit& best = L.best;
    if (st == BF_HIT_MISS && best.present) {
        v3f o, d;
        closest_world{&P.paths, slot, nullptr, 0u}(o, d);
        st = closest_hit_barycentrics(J.scene, o, d, L.t_min, L.t_max, best);
    }
    uint8_t r = 0;
    if (st != BF_HIT_MISS) {
        fail_path(P, J, slot, 60u + st);
    } else if (best.present) {
        P.paths.hit[slot] = make_float4(best.t, best.b0, best.b1, best.b2);
        P.paths.hit_ids[slot] = make_uint2(best.scene_instance, best.primitive);
        r = 1;
    } else {
        /* miss stage: CudaWavefrontStages.cu:3962-4040 */
        if (J.scene.has_environment) {
            const float4 wl = P.paths.wavelengths[slot];
            const float4 beta = P.paths.beta[slot];
            float4 Lr = P.paths.radiance[slot];
            const float w4[4] = {wl.x, wl.y, wl.z, wl.w};
            const float b4[4] = {beta.x, beta.y, beta.z, beta.w};
            float l4[4] = {Lr.x, Lr.y, Lr.z, Lr.w};
            bool ok = true;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                float product;
                const float radiance = spectrum_at(J.scene, J.scene.environment_spectrum, w4[k]);
                if (!checked_product(b4[k], radiance, product)) ok = false;
                else {
                    l4[k] += product;
                    if (!bf_isfinite(l4[k]) || l4[k] < 0.0f) ok = false;
                }
            }
            if (!ok) {
                fail_path(P, J, slot, 50);
            } else {
                P.paths.radiance[slot] = make_float4(l4[0], l4[1], l4[2], l4[3]);
                P.paths.status[slot] = BF_TERM_ESCAPED;
            }
        } else {
            P.paths.status[slot] = BF_TERM_ESCAPED;
        }
    }
    route[L.item] = r;
}

/*
 * The persistent lane loop shared by both traversal stages. Every lane holds one ray; a lane whose
 * ray is done stays idle until REFILL lanes of its wave are (or all are), and then the
 * wave claims that many queue items with one atomic on the launch's fetch counter. A wave
 * therefore keeps working while any ray of the queue is unclaimed instead of waiting for its
 * slowest ray before it moves on. Within a wave, an iteration steps either the lanes whose next
 * entry is an interior node or those whose next entry is a leaf (LEAF lanes), so the wave issues
 * one kind of step code at a time instead of every kind under partial masks. Every ray's
 * traversal is the same sequence of steps in any schedule, so results do not depend on it; the
 * loop ends when the counter has passed the item count and every lane of the wave is idle.
 */
template <uint32_t REFILL, uint32_t LEAF, typename Start, typename Step, typename Finish>
BF_D void lane_loop(uint32_t n, uint32_t* fetch, Start start, Step step, Finish finish) {
    const uint32_t lane = threadIdx.x & 63u;
    bool active = false, exhausted = false;
    for (;;) {
        const uint64_t idle = __ballot(!active);
        const uint32_t want = (uint32_t)__popcll(idle);
        if (!exhausted && (want >= REFILL || want == 64u)) {
            uint32_t base = 0u;
            if (lane == 0u) base = atomicAdd(fetch, want);
            base = __shfl(base, 0, 64);
            exhausted = base >= n || n - base <= want;
            if (!active) {
                const uint32_t i = base + (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
                if (i < n && base < n) {
                    const uint32_t st = start(i);
                    if (st != BF_HIT_MISS || !step.pending()) finish(st);
                    else active = true;
                }
            }
        }
        if (__ballot(active) == 0ull) {
            if (exhausted) break;
            continue;
        }
        bool go = active;
        if (LEAF != 0u) {
            /* one kind of step per iteration: interior nodes, unless LEAF lanes (or all
             * active ones) have a leaf next */
            const bool leaf = active && step.leaf_next();
            const uint64_t leaves = __ballot(leaf);
            const bool leaf_phase = (uint32_t)__popcll(leaves) >= LEAF || leaves == __ballot(active);
            go = active && leaf == leaf_phase;
        }
        if (go) {
            const uint32_t st = step();
            if (st != BF_HIT_MISS || !step.pending()) {
                finish(st);
                active = false;
            }
        }
    }
}

BF_D void reduce_counters(const trav_counters& c, unsigned long long* counters) {
    trav_counters ctr = c;
    for (int off = 32; off > 0; off >>= 1) {
        ctr.nodes += __shfl_down(ctr.nodes, off, 64);
        ctr.triangles += __shfl_down(ctr.triangles, off, 64);
        ctr.instances += __shfl_down(ctr.instances, off, 64);
    }
    if ((threadIdx.x & 63u) == 0u) {
        atomicAdd(&counters[0], (unsigned long long)ctr.nodes);
        atomicAdd(&counters[1], (unsigned long long)ctr.triangles);
        atomicAdd(&counters[2], (unsigned long long)ctr.instances);
    }
}

template <bool COUNT, bool BATCH, uint32_t BLOCK>
__global__ __launch_bounds__(BLOCK, BF_TRACE_WAVES_PER_SIMD) void k_trace_closest(bf_transport_params P, const uint32_t* queue,
                                                                      const uint32_t* count_ptr, const uint32_t* order,
                                                                      uint8_t* route, unsigned long long* counters) {
    __shared__ uint2 s_stack[BF_SHORT_STACK * BLOCK];
    const uint32_t n = *count_ptr;
    trav_stack<BF_SHORT_STACK, BLOCK> stack{s_stack + threadIdx.x, P.spill + blockIdx.x * BLOCK + threadIdx.x,
                                            P.spill_stride, 0u, 0u, BF_STACK_DEPTH};
    trav_counters ctr{0u, 0u, 0u};
    lane_ray L;
    struct stepper {
        const bf_transport_params* P;
        trav_stack<BF_SHORT_STACK, BLOCK>* stack;
        lane_ray* L;
        trav_counters* ctr;
        double* inv;
        BF_D uint32_t operator()() const { return closest_step<COUNT, BATCH>(*P, *stack, *L, *ctr, inv); }
        BF_D bool pending() const { return stack_depth(*stack) != 0u; }
        BF_D bool leaf_next() const { return ::leaf_next(*stack); }
    };
#if BF_WORLD_INV_CACHE
    /* the lane's world-space reciprocals, kept for the return from each BLAS */
    __shared__ double s_inv[3 * BLOCK];
    double* inv = s_inv + threadIdx.x;
#else
    double* inv = nullptr;
#endif
    lane_loop<BF_TRACE_REFILL, BF_LEAF_PHASE>(
        n, P.fetch,
        [&](uint32_t i) {
            const uint32_t st = closest_start<COUNT, BATCH>(P, queue, order ? order[i] : i, stack, L, ctr);
            if (inv) {
                inv[0] = L.ts.space.inv[0];
                inv[BLOCK] = L.ts.space.inv[1];
                inv[2u * BLOCK] = L.ts.space.inv[2];
            }
            return st;
        },
        stepper{&P, &stack, &L, &ctr, inv}, [&](uint32_t st) { closest_finish<BATCH>(P, route, L, st); });
    if (COUNT) reduce_counters(ctr, counters);
}

/* ---- shade stage: CudaWavefrontStages.cu:4107-4727 -------------------------------------------- */
struct surface_t {
    v3f position, error, ng, ns, dpdu, dpdv;
    float uv[2];
    uint32_t material;
};

/* surface_data (:1523-1665) plus the traversal's world normal (CudaSceneClosestHit.cu:826-871) */
BF_D bool surface_data(const bf_dev_scene& S, uint32_t inst, uint32_t prim, float4 hit, v3f ro, v3f rd, surface_t& out) {
    const bf_dev_instance& I = S.instances[inst];
    const bf_dev_geometry G = S.geometries[I.geometry_index];
    if (prim >= G.triangle_count) return false;
    const size_t t = (size_t)G.triangle_offset + prim;
    uint32_t idx[3] = {S.indices[3 * t], S.indices[3 * t + 1], S.indices[3 * t + 2]};
    const v3f bary = mk(hit.y, hit.z, hit.w);
    v3f lv[3], ln = mk(0.0f, 0.0f, 0.0f);
    float tc[3][2];
    for (int c = 0; c < 3; ++c) {
        if (idx[c] >= G.vertex_count) return false;
        const size_t gv = (size_t)G.vertex_offset + idx[c];
        lv[c] = mk(S.positions[3 * gv], S.positions[3 * gv + 1], S.positions[3 * gv + 2]);
        tc[c][0] = S.uvs[2 * gv];
        tc[c][1] = S.uvs[2 * gv + 1];
        const float w = c == 0 ? bary.x : (c == 1 ? bary.y : bary.z);
        const v3f vn = mk(S.normals[3 * gv], S.normals[3 * gv + 1], S.normals[3 * gv + 2]);
        if (!fin3(lv[c]) || !fin3(vn) || !bf_isfinite(tc[c][0]) || !bf_isfinite(tc[c][1])) return false;
        ln = add3(ln, mul3(vn, w));
    }
    v3f ng;
    if (!world_normal(I.local_to_world, lv[0], lv[1], lv[2], ng)) return false;
    const float* inv = I.world_to_local;
    const float* m = I.local_to_world;
    const v3f tn = mk(fmaf(inv[0], ln.x, fmaf(inv[4], ln.y, inv[8] * ln.z)), fmaf(inv[1], ln.x, fmaf(inv[5], ln.y, inv[9] * ln.z)),
                      fmaf(inv[2], ln.x, fmaf(inv[6], ln.y, inv[10] * ln.z)));
    v3f ns;
    if (!normalize3(tn, ns) || !normalize3(ng, ng)) return false;
    if (dot3(ns, ng) < 0.0f) ns = mul3(ns, -1.0f);
    if (!(dot3(ns, ng) > 0.0f)) return false;
    v3f wv[3];
    for (int c = 0; c < 3; ++c) {
        wv[c] = xform_point(m, lv[c]);
        if (!fin3(wv[c])) return false;
    }
    /* surface_derivatives (:1036-1060) */
    const v3f e1 = sub3(wv[1], wv[0]), e2 = sub3(wv[2], wv[0]);
    const float du1 = tc[1][0] - tc[0][0], dv1 = tc[1][1] - tc[0][1];
    const float du2 = tc[2][0] - tc[0][0], dv2 = tc[2][1] - tc[0][1];
    const float det = fmaf(du1, dv2, -dv1 * du2);
    v3f dpdu = mk(0.0f, 0.0f, 0.0f), dpdv = mk(0.0f, 0.0f, 0.0f);
    if (!fin3(e1) || !fin3(e2) || !bf_isfinite(det)) return false;
    if (det != 0.0f) {
        const float rcp = 1.0f / det;
        dpdu = mul3(sub3(mul3(e1, dv2), mul3(e2, dv1)), rcp);
        dpdv = mul3(sub3(mul3(e2, du1), mul3(e1, du2)), rcp);
        if (!bf_isfinite(rcp) || !fin3(dpdu) || !fin3(dpdv)) return false;
    }
    /* texture coordinates at the hit, for textured closures */
    for (int k = 0; k < 2; ++k) out.uv[k] = fmaf(tc[2][k], bary.z, fmaf(tc[1][k], bary.y, tc[0][k] * bary.x));
    positioned_t obj, world;
    const v3f traversal = mk(fmaf(hit.x, rd.x, ro.x), fmaf(hit.x, rd.y, ro.y), fmaf(hit.x, rd.z, ro.z));
    if (!interpolate_position_with_error(lv, bary, obj) || !transform_position_with_error(m, obj, traversal, world))
        return false;
    out.position = world.position;
    out.error = world.error;
    out.ng = ng;
    out.ns = ns;
    out.dpdu = dpdu;
    out.dpdv = dpdv;
    out.material = I.material_index;
    return fin3(out.position) && fin3(out.error);
}

/* Filtered texel value of a scene texture at (u, v); false when the lookup is not representable. */
BF_D bool texture_value(const bf_dev_scene& S, uint32_t slot, const float* uv, float* value, bf_texel_footprint& f) {
    const bf_dev_texture T = S.textures[slot];
    if (!bf_texture_footprint(T.width, T.height, uv[0], uv[1], &f)) return false;
    bf_texture_filter(T.data, T.channels, &f, value);
    return true;
}

/* Closure records at the path's wavelengths; textured weights and GGX widths are looked up at the
 * hit's texture coordinates. False when a texture lookup is not representable. */
template <bool TEX>
BF_D bool load_mixture(const bf_dev_scene& S, const bf_material_desc& M, const float* wl, const float* uv, mixture_t& mix) {
    mix.active_count = M.closure_count;
    for (uint32_t c = 0; c < M.closure_count; ++c) {
        const bf_closure_desc D = S.closures[M.closure_offset + c];
        closure_t& r = mix.closures[c];
        r.kind = D.kind;
        r.lobes = D.lobes;
#pragma unroll
        for (int k = 0; k < 10; ++k) r.parameters[k] = D.parameters[k];
#pragma unroll
        for (int lane = 0; lane < 4; ++lane) {
            r.weight[lane] = spectrum_at(S, D.weight_spectrum, wl[lane]);
            if (D.kind == BF_CLOSURE_ROUGH_CONDUCTOR) {
                r.parameters[lane] = spectrum_at(S, D.eta_spectrum, wl[lane]);
                r.parameters[4 + lane] = spectrum_at(S, D.k_spectrum, wl[lane]);
            }
        }
        if (TEX && D.weight_texture) {
            float t[3];
            bf_texel_footprint f;
            if (!texture_value(S, D.weight_texture - 1u, uv, t, f)) return false;
            const uint32_t channels = S.textures[D.weight_texture - 1u].channels;
#pragma unroll
            for (int lane = 0; lane < 4; ++lane) r.weight[lane] *= bf_texel_reflectance(c_rgb_basis, t, channels, wl[lane]);
        }
        if (TEX && D.roughness_texture) {
            float t[3];
            bf_texel_footprint f;
            if (!texture_value(S, D.roughness_texture - 1u, uv, t, f)) return false;
            const float ay = S.textures[D.roughness_texture - 1u].channels == 2u ? t[1] : t[0];
            const int at = D.kind == BF_CLOSURE_ROUGH_CONDUCTOR ? 8 : 2;
            r.parameters[at] = t[0];
            r.parameters[at + 1] = ay;
        }
        mix.probabilities[c] = M.probabilities[c];
    }
#pragma unroll
    for (uint32_t c = 0; c <= BF_MAX_CLOSURES; ++c) mix.cdf[c] = M.cdf[c];
    return true;
}

/* IncidentLightKind of the reference: a mesh-area sample (solid-angle pdf, MIS), a point or
 * spot light (discrete, shadow ray to the contracted point) or a directional light (discrete,
 * shadow ray to infinity). */
enum : uint32_t { BF_INCIDENT_SURFACE = 0u, BF_INCIDENT_POINT = 1u, BF_INCIDENT_INFINITE = 2u };

struct incident_t {
    float radiance[4];
    v3f direction, position, error, normal;
    float conditional, selection;
    uint32_t kind;
    bool present;
};

/* sample_punctual_light (:2240-2376): the scene checks it makes per sample are made by the host. */
BF_D bool sample_punctual(const bf_dev_scene& S, uint32_t index, const float* wl, const surface_t& surf, float sel,
                          incident_t& inc) {
    const bf_dev_punctual L = S.punctuals[index];
    float spectrum[4];
    bool black = true;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        spectrum[c] = spectrum_at(S, L.spectrum, wl[c]);
        if (!bf_isfinite(spectrum[c]) || spectrum[c] < 0.0f) return false;
        if (spectrum[c] != 0.0f) black = false;
    }
    const v3f axis = mk(L.direction[0], L.direction[1], L.direction[2]);
    if (L.kind == BF_PUNCTUAL_DIRECTIONAL) {
        if (black) return true;
#pragma unroll
        for (int c = 0; c < 4; ++c) inc.radiance[c] = spectrum[c];
        inc.direction = mul3(axis, -1.0f);
        inc.position = inc.error = inc.normal = mk(0.0f, 0.0f, 0.0f);
        inc.conditional = 1.0f;
        inc.selection = sel;
        inc.kind = BF_INCIDENT_INFINITE;
        inc.present = true;
        return true;
    }
    const v3f position = mk(L.position[0], L.position[1], L.position[2]);
    segment_t seg;
    if (!scaled_segment(sub3(position, surf.position), seg)) return false;
    const float ssd = dot3(seg.scaled, seg.scaled);
    if (!bf_isfinite(ssd) || !(ssd > 0.0f)) return false;
    float transition = 1.0f, smooth = 1.0f;
    if (L.kind == BF_PUNCTUAL_SPOT) {
        const float cone = dot3(axis, mul3(seg.direction, -1.0f));
        if (!bf_isfinite(cone)) return false;
        if (cone <= L.cosine_outer) return true;
        if (L.smooth && cone < L.cosine_inner) {
            transition = (cone - L.cosine_outer) / (L.cosine_inner - L.cosine_outer);
            smooth = 3.0f - 2.0f * transition;
            if (!bf_isfinite(transition) || !(transition > 0.0f) || !(transition < 1.0f) || !bf_isfinite(smooth) ||
                !(smooth > 0.0f))
                return false;
        }
    }
    float radiance[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const float num[4] = {spectrum[c], transition, transition, smooth};
        const float den[3] = {seg.scale, seg.scale, ssd};
        if (checked_product_quotient(num, den, radiance[c]) != BF_L_OK) return false;
    }
    if (black) return true;
#pragma unroll
    for (int c = 0; c < 4; ++c) inc.radiance[c] = radiance[c];
    inc.direction = seg.direction;
    inc.position = position;
    inc.error = mk(L.position_error[0], L.position_error[1], L.position_error[2]);
    inc.normal = mk(0.0f, 0.0f, 0.0f);
    inc.conditional = 1.0f;
    inc.selection = sel;
    inc.kind = BF_INCIDENT_POINT;
    inc.present = true;
    return true;
}

/* contract_point_endpoint (:2630-2682): the shadow ray of a point light stops short of the
 * light's position by its error bound along the direction to the light. */
BF_D bool endpoint_support(float direction, float error, float& contraction) {
    const float candidate = fmaf(fabsf(direction), error, contraction);
    if (!bf_isfinite(candidate) || candidate < contraction) return false;
    if (direction != 0.0f && error != 0.0f) {
        contraction = bf_next_up_f(candidate);
        return bf_isfinite(contraction);
    }
    contraction = candidate;
    return true;
}

BF_D bool contract_point_endpoint(v3f p, v3f err, v3f to_light, v3f& out) {
    if (!fin3(p) || !fin3(err) || err.x < 0.0f || err.y < 0.0f || err.z < 0.0f || !fin3(to_light)) return false;
    float contraction = 0.0f;
    if (!endpoint_support(to_light.x, err.x, contraction) || !endpoint_support(to_light.y, err.y, contraction) ||
        !endpoint_support(to_light.z, err.z, contraction))
        return false;
    out = p;
    if (contraction == 0.0f) return true;
    out = mk(fmaf(-to_light.x, contraction, p.x), fmaf(-to_light.y, contraction, p.y), fmaf(-to_light.z, contraction, p.z));
    if (to_light.x != 0.0f) out.x = to_light.x > 0.0f ? bf_next_down_f(out.x) : bf_next_up_f(out.x);
    if (to_light.y != 0.0f) out.y = to_light.y > 0.0f ? bf_next_down_f(out.y) : bf_next_up_f(out.y);
    if (to_light.z != 0.0f) out.z = to_light.z > 0.0f ? bf_next_down_f(out.z) : bf_next_up_f(out.z);
    return fin3(out);
}

/* sample_registered_light (:2378-2460) + sample_mesh_area_light (:2044-2201); the per-light area
 * CDF of :2085-2147 is a scene constant evaluated by the host, searched here in O(log n). */
BF_D bool sample_light(const bf_transport_params& P, const bf_dev_scene& S, const sample_key& k, uint64_t first_dim,
                       const float* wl, uint32_t visibility, const surface_t& surf, incident_t& inc) {
    inc.present = false;
    const float u_sel = sample_1d(P.sampler, k, first_dim + 0);
    uint32_t li;
    float sel;
    if (S.light_cdf ? !cdf_light_selection(S.light_cdf, S.light_count, u_sel, li, sel)
                    : !uniform_light_selection(S.light_count, u_sel, li, sel))
        return false;
    if (li < S.punctual_count) return sample_punctual(S, li, wl, surf, sel, inc);
    const bf_dev_light L = S.lights[li - S.punctual_count];
    const bf_dev_instance& I = S.instances[L.instance];
    if ((I.visibility_mask & visibility) == 0u) return true;
    const float area_sample = sample_1d(P.sampler, k, first_dim + 1);
    const float triangle_sample = sample_1d(P.sampler, k, first_dim + 2);
    if (!(area_sample >= 0.0f) || !(area_sample < 1.0f) || !(triangle_sample >= 0.0f) || !(triangle_sample < 1.0f))
        return false;
    /* first triangle whose cumulative bound exceeds the sample (bounds strictly increase) */
    const float* next = S.light_next + L.table_offset;
    uint32_t lo = 0, hi = L.triangle_count;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (area_sample < next[mid]) hi = mid;
        else lo = mid + 1;
    }
    const uint32_t t = lo;
    if (t >= L.triangle_count) return false;
    const float prob = S.light_probabilities[L.table_offset + t];
    const float area = S.light_areas[L.table_offset + t];
    const float remapped = (area_sample - S.light_cumulative[L.table_offset + t]) / prob;
    if (!bf_isfinite(remapped) || remapped < 0.0f || !(remapped < 1.0f)) return false;
    const float root = sqrtf(remapped);
    const float b0 = 1.0f - root, b1 = root * (1.0f - triangle_sample), b2 = root * triangle_sample;
    const bf_dev_geometry G = S.geometries[L.geometry];
    const size_t tri = (size_t)G.triangle_offset + t;
    v3f lv[3], wv[3];
    for (int c = 0; c < 3; ++c) {
        const size_t gv = (size_t)G.vertex_offset + S.indices[3 * tri + c];
        lv[c] = mk(S.positions[3 * gv], S.positions[3 * gv + 1], S.positions[3 * gv + 2]);
        wv[c] = xform_point(I.local_to_world, lv[c]);
    }
    const float4 nrm = S.light_normals[L.table_offset + t];
    const v3f n = mk(nrm.x, nrm.y, nrm.z);
    const float density = prob / area;
    if (!bf_isfinite(density) || !(density > 0.0f)) return false;
    const bf_material_desc& M = S.materials[I.material_index];
    float radiance[4];
    bool black = true;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        radiance[c] = spectrum_at(S, M.emission_spectrum, wl[c]);
        if (!bf_isfinite(radiance[c]) || radiance[c] < 0.0f) return false;
        if (radiance[c] != 0.0f) black = false;
    }
    const v3f traversal = add3(add3(mul3(wv[0], b0), mul3(wv[1], b1)), mul3(wv[2], b2));
    positioned_t obj, world;
    if (!interpolate_position_with_error(lv, mk(b0, b1, b2), obj) ||
        !transform_position_with_error(I.local_to_world, obj, traversal, world))
        return false;
    if (black) return true;
    segment_t seg;
    if (!scaled_segment(sub3(world.position, surf.position), seg)) return false;
    float absolute;
    bool supported;
    if (!light_surface_alignment(n, seg.scaled, absolute, supported)) return false;
    if (!supported) return true;
    const float num[4] = {density, seg.scale, seg.scale, seg.cubed};
    const float den[1] = {absolute};
    float cond;
    if (checked_product_quotient(num, den, cond) != BF_L_OK || !(cond > 0.0f)) return false;
#pragma unroll
    for (int c = 0; c < 4; ++c) inc.radiance[c] = radiance[c];
    inc.direction = seg.direction;
    inc.position = world.position;
    inc.error = world.error;
    inc.normal = n;
    inc.conditional = cond;
    inc.selection = sel;
    inc.kind = BF_INCIDENT_SURFACE;
    inc.present = true;
    return true;
}

/* emissive_hit_mis_weight (:2513-2594) */
BF_D bool emission_weight(const bf_transport_params& P, const bf_dev_scene& S, uint32_t inst, float4 prev,
                          const surface_t& surf, float& w) {
    w = 1.0f;
    if (!(prev.w > 0.0f)) return true;
    const int32_t li = S.instances[inst].light_index;
    if (li < 0) return false;
    float sel;
    const uint32_t slot = S.punctual_count + (uint32_t)li;
    if (S.light_cdf ? !cdf_light_probability(S.light_cdf, S.light_count, slot, sel)
                    : !uniform_light_probability(S.light_count, slot, sel))
        return false;
    const float inv_area = S.lights[li].inverse_area;
    segment_t seg;
    if (!scaled_segment(sub3(surf.position, mk(prev.x, prev.y, prev.z)), seg)) return false;
    float absolute;
    bool supported;
    if (!light_surface_alignment(surf.ng, seg.scaled, absolute, supported) || !supported) return false;
    const float num[4] = {inv_area, seg.scale, seg.scale, seg.cubed};
    const float den[1] = {absolute};
    float cond, joint;
    if (checked_product_quotient(num, den, cond) != BF_L_OK) return false;
    if (!joint_light_pdf(sel, cond, joint)) return false;
    return mis_weight(P.mis, prev.w, joint, w);
}

BF_D bool zero4(const float* v) { return v[0] == 0.0f && v[1] == 0.0f && v[2] == 0.0f && v[3] == 0.0f; }

/* ---- adjoint: texel gradients of one path vertex ---------------------------------------------
 * The path's estimate is a sum of products over its vertices; with sampling, MIS weights and
 * roulette held constant, a parameter p of the closures at vertex j scales every term that passes
 * through that vertex's BSDF value. For the light sample (visible contribution N, closure value
 * f_l) and the continuation (radiance R it carried back, value f_s):
 *   dLoss/dp += sum_lanes a * (N * (df_l/dp) / f_l + R * (df_s/dp) / f_s),
 * a = dLoss/d(radiance lane) of the path's pixel. Texel gradients follow through the bilinear
 * weights and, for weights, the RGB basis spectra. */
/* Adds v to texel-channel e of texture k's gradient, through a lane-chosen replica row when the
 * texture is small. */
BF_D void gradient_add(const bf_transport_params& P, const bf_dev_job& J, uint32_t k, size_t e, float v) {
    const uint32_t offset = J.gradient_replica[k];
    if (offset != BF_INVALID_INDEX) {
        const uint32_t row = (blockIdx.x * blockDim.x + threadIdx.x) % BF_GRADIENT_REPLICAS;
        unsafeAtomicAdd(&P.gradient_replicas[(size_t)row * BF_GRADIENT_ROW + offset + e], v);
    } else {
        unsafeAtomicAdd(&J.texture_gradients[k][e], v);
    }
}

/* Texel gradients of one texture at one footprint, summed over the closures that bind it before
 * they go out as atomics: closures of a material share the hit's texture coordinate, so
 * consecutive closures bound to the same texture add into one set of texels. */
struct texel_pending {
    uint32_t slot, channels;
    bf_texel_footprint f;
    float sc[3];
};

BF_D void texel_flush(const bf_transport_params& P, const bf_dev_job& J, texel_pending& p) {
    if (p.slot == BF_INVALID_INDEX) return;
    for (int q = 0; q < 4; ++q)
        for (uint32_t ch = 0; ch < p.channels; ++ch)
            if (p.f.weight[q] != 0.0f && p.sc[ch] != 0.0f)
                gradient_add(P, J, p.slot, (size_t)p.f.index[q] * p.channels + ch, p.f.weight[q] * p.sc[ch]);
    p.slot = BF_INVALID_INDEX;
}

BF_D void texel_add(const bf_transport_params& P, const bf_dev_job& J, texel_pending& p, uint32_t slot, uint32_t channels,
                    const bf_texel_footprint& f, const float* sc) {
    if (p.slot != slot) {
        texel_flush(P, J, p);
        p.slot = slot;
        p.channels = channels;
        p.f = f;
        p.sc[0] = p.sc[1] = p.sc[2] = 0.0f;
    }
    for (uint32_t ch = 0; ch < channels; ++ch) p.sc[ch] += sc[ch];
}

BF_D void texel_gradients(const bf_transport_params& P, const bf_dev_job& J, const bf_dev_scene& S, const bf_material_desc& M, const mixture_t& mix,
                          const filter_t& filtered, const float* uv, const float* wl, const float* a, bool light, v3f wo,
                          v3f wi_light, const float* f_light, const float* N, bool cont, bool cont_discrete, v3f wi_cont,
                          const float* f_cont, const float* R) {
    texel_pending weights{BF_INVALID_INDEX, 0u, {}, {0.0f, 0.0f, 0.0f}};
    texel_pending widths{BF_INVALID_INDEX, 0u, {}, {0.0f, 0.0f, 0.0f}};
    for (uint32_t i = 0; i < mix.active_count; ++i) {
        const bf_closure_desc D = S.closures[M.closure_offset + filtered.source_indices[i]];
        const bool wgrad = D.weight_texture && J.texture_gradients[D.weight_texture - 1u];
        const bool rgrad = D.roughness_texture && J.texture_gradients[D.roughness_texture - 1u];
        if (!wgrad && !rgrad) continue;
        const closure_t& c = mix.closures[i];
        float gl[4], gs[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            gl[k] = light && f_light[k] > 0.0f ? a[k] * N[k] / f_light[k] : 0.0f;
            gs[k] = cont && f_cont[k] > 0.0f ? a[k] * R[k] / f_cont[k] : 0.0f;
        }
        if (wgrad) {
            float dl[4] = {0.0f, 0.0f, 0.0f, 0.0f}, ds[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            const bool lw = light && closure_weight_derivative(c, filtered.allowed[i], wo, wi_light, dl) == BF_L_OK;
            /* a delta continuation carries the sum of the matching delta atoms: only their weights
             * move it, and each by its unit-weight atom value */
            const bool cw = cont && (cont_discrete ? delta_weight_derivative(c, wo, wi_cont, ds)
                                                   : closure_weight_derivative(c, filtered.allowed[i], wo, wi_cont, ds) == BF_L_OK);
            const uint32_t slot = D.weight_texture - 1u;
            const uint32_t channels = S.textures[slot].channels;
            float t[3];
            bf_texel_footprint f;
            if (texture_value(S, slot, uv, t, f)) {
                float sc[3] = {0.0f, 0.0f, 0.0f};
                for (int k = 0; k < 4; ++k) {
                    const float g = (lw ? gl[k] * dl[k] : 0.0f) + (cw ? gs[k] * ds[k] : 0.0f);
                    if (g == 0.0f) continue;
                    const float w = g * spectrum_at(S, D.weight_spectrum, wl[k]);
                    if (channels == 1u) {
                        if (t[0] > 0.0f && t[0] < 1.0f) sc[0] += w;
                    } else {
                        float b[3];
                        bf_rgb_basis_at(c_rgb_basis, wl[k], b);
                        const float rho = t[0] * b[0] + t[1] * b[1] + t[2] * b[2];
                        if (!(rho > 0.0f) || !(rho < 1.0f)) continue; /* clamped: no gradient */
                        for (int ch = 0; ch < 3; ++ch) sc[ch] += w * b[ch];
                    }
                }
                texel_add(P, J, weights, slot, channels, f, sc);
            }
        }
        if (rgrad && (c.kind == BF_CLOSURE_ROUGH_CONDUCTOR || c.kind == BF_CLOSURE_ROUGH_DIELECTRIC)) {
            alpha_gradient ql{0.0f, 0.0f}, qs{0.0f, 0.0f};
            const bool use_l = light && closure_alpha_log_gradient(c, wo, wi_light, ql);
            const bool use_s = cont && !cont_discrete && closure_alpha_log_gradient(c, wo, wi_cont, qs);
            /* the closure's own value at each direction: the mixture value when it is alone */
            float fl[4], fs[4];
            for (int k = 0; k < 4; ++k) {
                fl[k] = f_light[k];
                fs[k] = f_cont[k];
            }
            if (mix.active_count > 1u) {
                if (use_l && record_eval(c, filtered.allowed[i], wo, wi_light, fl) != BF_L_OK) continue;
                if (use_s && record_eval(c, filtered.allowed[i], wo, wi_cont, fs) != BF_L_OK) continue;
            }
            float hx = 0.0f, hy = 0.0f;
            for (int k = 0; k < 4; ++k) {
                const float el = use_l ? gl[k] * fl[k] : 0.0f, es = use_s ? gs[k] * fs[k] : 0.0f;
                hx += el * ql.x + es * qs.x;
                hy += el * ql.y + es * qs.y;
            }
            const uint32_t slot = D.roughness_texture - 1u;
            const uint32_t channels = S.textures[slot].channels;
            float t[3];
            bf_texel_footprint f;
            if (texture_value(S, slot, uv, t, f)) {
                const float h[3] = {channels == 2u ? hx : hx + hy, hy, 0.0f};
                texel_add(P, J, widths, slot, channels, f, h);
            }
        }
    }
    texel_flush(P, J, weights);
    texel_flush(P, J, widths);
}

/* route codes written per queue item: 0 terminated, 1 shadow, 2 continuation. MODE is
 * BF_SHADE_PLAIN (no texture lookups compiled in), BF_SHADE_TEXTURED, BF_SHADE_MAPPED (textures and
 * normal / bump maps), or BF_SHADE_ADJOINT: the
 * adjoint replay resolves the light sample from the tape, never routes to the shadow queue and
 * accumulates texel gradients. */
template <bool BATCH, uint32_t MODE>
BF_D void shade_item(const bf_transport_params& P, const uint32_t* queue, uint32_t i, uint8_t* route,
                     uint32_t* sample_counts) {
    constexpr bool ADJ = MODE == BF_SHADE_ADJOINT;
    const uint32_t slot = queue[i];
    route[i] = 0;
    const bf_dev_job& J = job_at<BATCH>(P, slot);
    const bf_dev_scene& S = J.scene;
    const sample_key key = key_for_slot(J, slot);
    const float4 ro4 = P.paths.ray_origin[slot];
    const float4 rd4 = P.paths.ray_direction[slot];
    const v3f ro = mk(ro4.x, ro4.y, ro4.z), rd = mk(rd4.x, rd4.y, rd4.z);
    const float4 hit = P.paths.hit[slot];
    const uint2 ids = P.paths.hit_ids[slot];
    const float4 wl4 = P.paths.wavelengths[slot];
    const float wl[4] = {wl4.x, wl4.y, wl4.z, wl4.w};
    float4 beta4 = P.paths.beta[slot];
    float beta[4] = {beta4.x, beta4.y, beta4.z, beta4.w};
    const float4 L4 = P.paths.radiance[slot];
    float L[4] = {L4.x, L4.y, L4.z, L4.w};
    uint4 dstate = P.paths.depth[slot];
    uint32_t diffuse = dstate.x & 0xFFu, glossy = (dstate.x >> 8) & 0xFFu, specular = (dstate.x >> 16) & 0xFFu,
             transmission = (dstate.x >> 24) & 0xFFu;
    uint32_t depth = dstate.y, delta_flags = dstate.z & 3u;
    const uint32_t vertex = depth;
    float eta_scale = __uint_as_float(dstate.w);
    const float4 prev = P.paths.previous[slot];
    const uint32_t visibility = P.paths.visibility[slot];

    surface_t surf;
    if (!surface_data(S, ids.x, ids.y, hit, ro, rd, surf)) return fail_path(P, J, slot, 10);
    if (MODE == BF_SHADE_MAPPED && S.maps &&
        (S.maps[surf.material].normal_texture != BF_INVALID_INDEX || S.maps[surf.material].bump_texture != BF_INVALID_INDEX)) {
        /* the shade stage's surface maps (:4170-4193): the cone advanced to the hit sets the
         * footprint; 46 the cone, 47-49 the maps (invalid, numerical, texel budget) */
        float2 at_hit;
        if (!(J.in_cones || J.frame_cones) || !advance_cone(P.paths.cone[slot], rd, hit.x, at_hit))
            return fail_path(P, J, slot, 46);
        const int st = apply_surface_maps(S, surf.material, at_hit, rd, surf.uv[0], surf.uv[1], surf.ng, surf.dpdu,
                                          surf.dpdv, surf.ns);
        if (st != BF_MAP_OK) return fail_path(P, J, slot, 46u + (uint32_t)st);
    }
    v3f wo;
    if (!normalize3(mk(-rd.x, -rd.y, -rd.z), wo)) return fail_path(P, J, slot, 11);
    bool front = false;
    if (!one_sided_emission_support(surf.ng, wo, front)) return fail_path(P, J, slot, 12);
    const bf_material_desc& M = S.materials[surf.material];
    float emitted[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        emitted[c] = spectrum_at(S, M.emission_spectrum, wl[c]);
        if (!bf_isfinite(emitted[c]) || emitted[c] < 0.0f) return fail_path(P, J, slot, 13);
    }
    mixture_t mix;
    if (!load_mixture<MODE != BF_SHADE_PLAIN>(S, M, wl, surf.uv, mix)) return fail_path(P, J, slot, 42);
    frame3 frame, unrotated;
    const bool created = M.frame_mode == 0u ? make_local_frame(surf.ns, unrotated) : frame_from_tangent(surf.ns, surf.dpdu, unrotated);
    if (!created || !rotate_frame(unrotated, M.tangent_rotation, frame)) return fail_path(P, J, slot, 14);
    if (front && !zero4(emitted)) {
        float weight;
        if (!emission_weight(P, S, ids.x, prev, surf, weight)) return fail_path(P, J, slot, 15);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            if (beta[c] == 0.0f || emitted[c] == 0.0f || weight == 0.0f) continue;
            const float num[3] = {beta[c], emitted[c], weight};
            const float den[1] = {1.0f};
            float contribution;
            if (checked_product_quotient(num, den, contribution) != BF_L_OK) return fail_path(P, J, slot, 16);
            L[c] += contribution;
            if (!bf_isfinite(L[c])) return fail_path(P, J, slot, 16);
        }
        P.paths.radiance[slot] = make_float4(L[0], L[1], L[2], L[3]);
    }
    if (zero4(beta)) {
        P.paths.status[slot] = BF_TERM_ZERO_THROUGHPUT;
        return;
    }
    const uint32_t limits[5] = {P.limits[0], P.limits[1], P.limits[2], P.limits[3], P.limits[4]};
    const uint32_t counters[5] = {diffuse, glossy, specular, transmission, 0u};
    if (diffuse > limits[0] || glossy > limits[1] || specular > limits[2] || transmission > limits[3])
        return fail_path(P, J, slot, 17);
    filter_t filtered;
    filter_by_depth(mix, limits, counters, depth, filtered);
    if (filtered.status != BF_L_OK) return fail_path(P, J, slot, 18);
    if (mix.active_count == 0 && filtered.source_count != 0) {
        /* the depth categories that blocked the path ride above the termination byte */
        P.paths.status[slot] = BF_TERM_DEPTH_LIMIT | (filtered.blocked << 8);
        return;
    }
    const v3f wo_local = to_local(frame, wo);
    const uint64_t first_dim = 4ull + (uint64_t)depth * 10ull;
    const float bsdf_component = sample_1d(P.sampler, key, first_dim + 3);
    const float bsdf_u = sample_1d(P.sampler, key, first_dim + 4);
    const float bsdf_v = sample_1d(P.sampler, key, first_dim + 5);
    wave_count(&sample_counts[0]);
    mixture_sample_t bs;
    sample_mixture(mix, filtered, wo_local, bsdf_component, bsdf_u, bsdf_v, bs);
    const bool has_continuation = bs.status == BF_L_OK;
    if (!has_continuation && bs.status != BF_L_OUTSIDE) return fail_path(P, J, slot, 19);
    if (has_continuation) {
        const bool delta = valid_surface_event(bs.lobes) && (bs.lobes & BF_LOBE_SPECULAR);
        if (!valid_surface_event(bs.lobes) || (bs.discrete != 0u) != delta || !bf_isfinite(bs.pdf) || !(bs.pdf > 0.0f) ||
            (delta && bs.pdf > 1.0f))
            return fail_path(P, J, slot, 20);
    }
    incident_t inc;
    inc.present = false;
    if (S.light_count != 0u && mixture_nonblack(mix)) {
        wave_count(&sample_counts[1]);
        if (!sample_light(P, S, key, first_dim, wl, visibility, surf, inc)) return fail_path(P, J, slot, 21);
    }
    bool has_shadow = false;
    float4 shadow_o = make_float4(0.0f, 0.0f, 0.0f, 0.0f), shadow_d = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float contribution[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    float nee_f[4] = {0.0f, 0.0f, 0.0f, 0.0f}; /* closure value toward the light (adjoint) */
    v3f nee_wi = mk(0.0f, 0.0f, 0.0f);
    if (inc.present) {
        const v3f wi_local = to_local(frame, inc.direction);
        nee_wi = wi_local;
        if (wo_local.z != 0.0f && wi_local.z != 0.0f) {
            const bool singleton_lambert = mix.active_count == 1 && mix.closures[0].kind == BF_CLOSURE_LAMBERTIAN;
            float bsdf[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            if (singleton_lambert) {
                if (wo_local.z > 0.0f && wi_local.z > 0.0f)
                    for (int c = 0; c < 4; ++c) bsdf[c] = mix.closures[0].weight[c];
            } else if (mixture_eval(mix, filtered, wo_local, wi_local, bsdf) != BF_L_OK) {
                return fail_path(P, J, slot, 22);
            }
            for (int c = 0; c < 4; ++c) nee_f[c] = singleton_lambert ? bsdf[c] * BF_INV_PI : bsdf[c];
            if (!zero4(bsdf)) {
                float corr;
                if (!shading_correction(surf.ng, surf.ns, wo, inc.direction, corr)) return fail_path(P, J, slot, 23);
                if (corr > 0.0f) {
                    float w = 1.0f; /* punctual lights are discrete: no MIS (:4437-4470) */
                    if (inc.kind == BF_INCIDENT_SURFACE) {
                        float joint, bsdf_pdf;
                        if (!joint_light_pdf(inc.selection, inc.conditional, joint)) return fail_path(P, J, slot, 24);
                        if (mixture_pdf(mix, filtered, wo_local, wi_local, bsdf_pdf) != BF_L_OK) return fail_path(P, J, slot, 24);
                        if (!mis_weight(P.mis, joint, bsdf_pdf, w)) return fail_path(P, J, slot, 24);
                    }
                    /* the visible-shadow contribution of process_shadow_kernel (:4772-4908),
                     * evaluated now from the same frozen factors */
                    const float coefficient = singleton_lambert ? BF_INV_PI : 1.0f;
                    const float cosine = fabsf(wi_local.z);
                    for (int c = 0; c < 4; ++c) {
                        const float den[2] = {inc.selection, inc.conditional};
                        int st;
                        if (w == 1.0f) {
                            const float num[6] = {beta[c], bsdf[c], coefficient, inc.radiance[c], cosine, corr};
                            st = checked_product_quotient(num, den, contribution[c]);
                        } else {
                            const float num[7] = {beta[c], bsdf[c], coefficient, inc.radiance[c], cosine, corr, w};
                            st = checked_product_quotient(num, den, contribution[c]);
                        }
                        if (st != BF_L_OK) return fail_path(P, J, slot, 41);
                    }
                    /* prepare_shadow_ray (:4042-4105) */
                    v3f src, endpoint;
                    if (!offset_point(surf.position, surf.error, surf.ng, inc.direction, src)) return fail_path(P, J, slot, 25);
                    if (inc.kind == BF_INCIDENT_INFINITE) {
                        shadow_o = make_float4(src.x, src.y, src.z, BF_INFINITY_F);
                        shadow_d = make_float4(inc.direction.x, inc.direction.y, inc.direction.z, 0.0f);
                    } else {
                        segment_t seg;
                        const bool ends = inc.kind == BF_INCIDENT_POINT
                                              ? contract_point_endpoint(inc.position, inc.error, inc.direction, endpoint)
                                              : offset_point(inc.position, inc.error, inc.normal, mul3(inc.direction, -1.0f), endpoint);
                        if (!ends || !scaled_segment(sub3(endpoint, src), seg) || !(dot3(seg.direction, inc.direction) > 0.0f))
                            return fail_path(P, J, slot, 25);
                        const float tmax = bf_next_toward_zero_f(seg.length);
                        if (!(tmax > 0.0f)) return fail_path(P, J, slot, 25);
                        shadow_o = make_float4(src.x, src.y, src.z, tmax);
                        shadow_d = make_float4(seg.direction.x, seg.direction.y, seg.direction.z, 0.0f);
                    }
                    has_shadow = true;
                }
            }
        }
    }
    v3f wi = mk(0.0f, 0.0f, 0.0f);
    bool continuation = has_continuation;
    float corr = 0.0f;
    if (continuation) {
        if (!to_world(frame, bs.incoming, wi)) return fail_path(P, J, slot, 26);
        bool geo = false;
        if (!shading_correction(surf.ng, surf.ns, wo, wi, corr) || !geometric_event_support(surf.ng, wo, wi, bs.lobes, geo))
            return fail_path(P, J, slot, 27);
        continuation = corr > 0.0f && geo;
    }
    uint32_t pending_term = BF_TERM_OUTSIDE_SUPPORT;
    bool cont_pending = false;
    if (continuation) {
        /* update_closure_throughput (:3894-3938) */
        float nb[4];
        const float cosine = fabsf(bs.incoming.z);
        if (!bf_isfinite(cosine) || !(cosine > 0.0f)) return fail_path(P, J, slot, 28);
        if (mix.active_count == 1 && mix.closures[0].kind == BF_CLOSURE_LAMBERTIAN && corr == 1.0f) {
            for (int c = 0; c < 4; ++c)
                if (!checked_product(beta[c], mix.closures[0].weight[c], nb[c])) return fail_path(P, J, slot, 28);
        } else {
            for (int c = 0; c < 4; ++c) {
                const float num[4] = {beta[c], bs.value[c], cosine, corr};
                const float den[1] = {bs.pdf};
                if (checked_product_quotient(num, den, nb[c]) != BF_L_OK) return fail_path(P, J, slot, 28);
            }
        }
        const float next_eta = eta_scale * bs.eta_scale;
        if (!bf_isfinite(bs.eta_scale) || !(bs.eta_scale > 0.0f) || !bf_isfinite(next_eta) || !(next_eta > 0.0f))
            return fail_path(P, J, slot, 29);
        v3f origin;
        if (!offset_point(surf.position, surf.error, surf.ng, wi, origin)) return fail_path(P, J, slot, 30);
        const bool delta_event = (bs.lobes & BF_LOBE_SPECULAR) != 0u;
        delta_flags = ((delta_flags & 2u) ? 2u : 0u) | (delta_event ? 1u : 2u);
        /* advance_depth_event (:3851-3880) */
        if (bs.lobes & BF_LOBE_DIFFUSE) ++diffuse;
        else if (bs.lobes & BF_LOBE_GLOSSY) ++glossy;
        else ++specular;
        if (bs.lobes & BF_LOBE_TRANSMISSION) ++transmission;
        ++depth;
        eta_scale = next_eta;
        beta4 = make_float4(nb[0], nb[1], nb[2], nb[3]);
        P.paths.previous[slot] = delta_event ? make_float4(0.0f, 0.0f, 0.0f, 0.0f)
                                             : make_float4(surf.position.x, surf.position.y, surf.position.z, bs.pdf);
        if (J.in_cones || J.frame_cones) {
            /* the sampled closure's compact index, then the cone advanced to the hit and scattered */
            uint32_t compact = mix.active_count, matches = 0u;
            for (uint32_t k = 0; k < mix.active_count; ++k)
                if (filtered.source_indices[k] == bs.selected) {
                    compact = k;
                    ++matches;
                }
            float2 advanced, next;
            if (matches != 1u || compact >= mix.active_count) return fail_path(P, J, slot, 43);
            if (!advance_cone(P.paths.cone[slot], rd, hit.x, advanced)) return fail_path(P, J, slot, 44);
            if (!propagate_cone(advanced, mix.closures[compact], bs.lobes, wo_local, bs.incoming, next))
                return fail_path(P, J, slot, 45);
            P.paths.cone[slot] = next;
        }
        P.paths.ray_origin[slot] = make_float4(origin.x, origin.y, origin.z, ro4.w);
        P.paths.ray_direction[slot] = make_float4(wi.x, wi.y, wi.z, 0.0f);
        cont_pending = true;
        pending_term = BF_TERM_NONE;
        if (nb[0] == 0.0f && nb[1] == 0.0f && nb[2] == 0.0f && nb[3] == 0.0f) {
            cont_pending = false;
            pending_term = BF_TERM_ZERO_THROUGHPUT;
        } else if (P.rr_enabled && depth >= P.rr_first) {
            const float u = sample_1d(P.sampler, key, first_dim + 9);
            bool survived;
            float* frozen = J.tape_survival ? J.tape_survival + path_of(J, slot) * J.survival_stride + vertex : nullptr;
            if (frozen && J.survival_mode == BF_SURVIVAL_REPLAY) {
                if (!russian_roulette_at(beta4, *frozen, u, survived)) return fail_path(P, J, slot, 32);
            } else {
                float p;
                if (!russian_roulette(beta4, eta_scale, u, P.rr_min, P.rr_max, survived, p)) return fail_path(P, J, slot, 32);
                if (frozen && !ADJ) *frozen = p;
            }
            if (!survived) {
                cont_pending = false;
                pending_term = BF_TERM_RUSSIAN_ROULETTE;
            }
        }
        P.paths.beta[slot] = beta4;
        P.paths.depth[slot] = make_uint4(diffuse | (glossy << 8) | (specular << 16) | (transmission << 24), depth,
                                         delta_flags, __float_as_uint(eta_scale));
    }
    if constexpr (ADJ) {
        /* the light sample's visibility from the tape, then this vertex's gradients */
        const float4 vis4 = P.paths.shadow_direction[slot];
        const unsigned long long vis =
            (unsigned long long)__float_as_uint(vis4.x) | ((unsigned long long)__float_as_uint(vis4.y) << 32);
        const bool visible = has_shadow && vertex < 64u && ((vis >> vertex) & 1ull) != 0ull;
        if (visible) {
            for (int c = 0; c < 4; ++c) {
                L[c] += contribution[c];
                if (!bf_isfinite(L[c]) || L[c] < 0.0f) return fail_path(P, J, slot, 41);
            }
            P.paths.radiance[slot] = make_float4(L[0], L[1], L[2], L[3]);
        }
        if (S.has_textures && J.film_gradient) {
            const float4 total = P.paths.shadow_origin[slot];
            const float4 a4 = P.paths.shadow_contribution[slot];
            const float a[4] = {a4.x, a4.y, a4.z, a4.w};
            const float R[4] = {total.x - L[0], total.y - L[1], total.z - L[2], total.w - L[3]};
            texel_gradients(P, J, S, M, mix, filtered, surf.uv, wl, a, visible, wo_local, nee_wi, nee_f, contribution,
                            cont_pending, bs.discrete != 0u, bs.incoming, bs.value, R);
        }
        if (!cont_pending) {
            P.paths.status[slot] = pending_term;
            return;
        }
        route[i] = 2;
        return;
    }
    if (has_shadow) {
        shadow_d.w = __uint_as_float((cont_pending ? 1u : 0u) | (pending_term << 8) | (vertex << 16));
        P.paths.shadow_origin[slot] = shadow_o;
        P.paths.shadow_direction[slot] = shadow_d;
        P.paths.shadow_contribution[slot] = make_float4(contribution[0], contribution[1], contribution[2], contribution[3]);
        route[i] = 1;
        return;
    }
    if (!cont_pending) {
        P.paths.status[slot] = pending_term;
        return;
    }
    route[i] = 2;
}

template <bool BATCH, uint32_t MODE>
__global__ __launch_bounds__(BF_SHADE_BLOCK, MODE == BF_SHADE_ADJOINT ? BF_ADJOINT_SHADE_WAVES_PER_SIMD : BF_SHADE_WAVES_PER_SIMD) void k_shade(bf_transport_params P, const uint32_t* queue,
                                                                             const uint32_t* count_ptr, const uint32_t* order,
                                                                             uint8_t* route, uint32_t* sample_counts) {
    const uint32_t n = *count_ptr;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        shade_item<BATCH, MODE>(P, queue, order ? order[i] : i, route, sample_counts);
}

/* World-space shadow ray of a slot, re-read from the pending-shadow columns. */
struct shadow_world {
    const bf_dev_paths* paths;
    uint32_t slot;
    BF_D void operator()(v3f& o, v3f& d) const {
        const float4 a = paths->shadow_origin[slot];
        const float4 b = paths->shadow_direction[slot];
        o = mk(a.x, a.y, a.z);
        d = mk(b.x, b.y, b.z);
    }
};

template <bool COUNT, bool BATCH, uint32_t BLOCK>
BF_D uint32_t shadow_start(const bf_transport_params& P, const uint32_t* queue, uint32_t i,
                           trav_stack<BF_SHORT_STACK, BLOCK>& stack, lane_ray& L, trav_counters& ctr) {
    L.item = i;
    L.slot = queue[i];
    const bf_dev_job& J = job_at<BATCH>(P, L.slot);
    const float4 so = P.paths.shadow_origin[L.slot];
    const float4 sd = P.paths.shadow_direction[L.slot];
    trav_ray ray;
    ray.o = mk(so.x, so.y, so.z);
    ray.d = mk(sd.x, sd.y, sd.z);
    ray.t_min = 0.0f;
    ray.t_max = so.w;
    ray.mask = P.paths.visibility[L.slot];
    return traverse_begin<COUNT>(J.scene, ray, stack, L.ts, L.best, L.occluded, ctr);
}

template <bool COUNT, bool BATCH, uint32_t BLOCK>
BF_D uint32_t shadow_step(const bf_transport_params& P, trav_stack<BF_SHORT_STACK, BLOCK>& stack, lane_ray& L,
                          trav_counters& ctr) {
    const bf_dev_job& J = job_at<BATCH>(P, L.slot);
    return traverse_step<true, COUNT>(J.scene, shadow_world{&P.paths, L.slot}, stack, L.ts, L.best, L.occluded, ctr);
}

/* process_shadow (:4772-4908): the frozen contribution is added when the light is visible (and the
 * vertex's tape bit set), then the path continues or ends with its pending termination. */
template <bool BATCH>
BF_D void shadow_finish(const bf_transport_params& P, uint8_t* route, lane_ray& L, uint32_t st) {
    const uint32_t slot = L.slot;
    const bf_dev_job& J = job_at<BATCH>(P, slot);
    const uint32_t packed = __float_as_uint(P.paths.shadow_direction[slot].w);
    uint8_t r = 0;
    if (st != BF_HIT_MISS) {
        fail_path(P, J, slot, 40);
    } else {
        bool ok = true;
        if (!L.occluded) {
            if (J.tape_visibility) J.tape_visibility[path_of(J, slot)] |= 1ull << ((packed >> 16) & 63u);
            const float4 c = P.paths.shadow_contribution[slot];
            float4 Lr = P.paths.radiance[slot];
            Lr.x += c.x;
            Lr.y += c.y;
            Lr.z += c.z;
            Lr.w += c.w;
            if (!bf_isfinite(Lr.x) || !bf_isfinite(Lr.y) || !bf_isfinite(Lr.z) || !bf_isfinite(Lr.w) || Lr.x < 0.0f ||
                Lr.y < 0.0f || Lr.z < 0.0f || Lr.w < 0.0f) {
                ok = false;
                fail_path(P, J, slot, 41);
            } else {
                P.paths.radiance[slot] = Lr;
            }
        }
        if (ok) {
            if (packed & 1u) r = 2;
            else P.paths.status[slot] = (packed >> 8) & 0xFFu;
        }
    }
    route[L.item] = r;
}

/* ---- shadow stage: gather + any-hit + process_shadow (:4729-4908) ---------------------------- */
template <bool COUNT, bool BATCH, uint32_t BLOCK>
__global__ __launch_bounds__(BLOCK, BF_SHADOW_WAVES_PER_SIMD) void k_trace_shadow(bf_transport_params P, const uint32_t* queue,
                                                                          const uint32_t* count_ptr, const uint32_t* order,
                                                                          uint8_t* route, unsigned long long* counters) {
    __shared__ uint2 s_stack[BF_SHORT_STACK * BLOCK];
    const uint32_t n = *count_ptr;
    trav_stack<BF_SHORT_STACK, BLOCK> stack{s_stack + threadIdx.x, P.spill + blockIdx.x * BLOCK + threadIdx.x,
                                            P.spill_stride, 0u, 0u, BF_STACK_DEPTH};
    trav_counters ctr{0u, 0u, 0u};
    lane_ray L;
    struct stepper {
        const bf_transport_params* P;
        trav_stack<BF_SHORT_STACK, BLOCK>* stack;
        lane_ray* L;
        trav_counters* ctr;
        BF_D uint32_t operator()() const { return shadow_step<COUNT, BATCH>(*P, *stack, *L, *ctr); }
        BF_D bool pending() const { return stack_depth(*stack) != 0u; }
        BF_D bool leaf_next() const { return ::leaf_next(*stack); }
    };
    lane_loop<BF_SHADOW_REFILL, BF_SHADOW_LEAF_PHASE>(
        n, P.fetch + 1, [&](uint32_t i) { return shadow_start<COUNT, BATCH>(P, queue, order ? order[i] : i, stack, L, ctr); },
        stepper{&P, &stack, &L, &ctr}, [&](uint32_t st) { shadow_finish<BATCH>(P, route, L, st); });
    if (COUNT) reduce_counters(ctr, counters + 3);
}

/* ---- stable compaction: wave64 ballot/popcount ranks + tile-offset scan -----------------------
 * For route values r = 1, 2: dst_r[base_r + exclusive_prefix(route == r)] = src[i], which is the
 * stable exclusive-scan append of CudaWavefrontQueueCompaction.cu:55-315. The item count is read
 * on the device and the kernels stride over 256-item tiles, so launches can be issued (and
 * captured) before the count is known. */
__global__ __launch_bounds__(BF_BLOCK) void k_compact_count(const uint8_t* route, const uint32_t* count_ptr,
                                                            uint32_t* tile_counts) {
    __shared__ uint32_t wave_counts[2][BF_BLOCK / 64];
    const uint32_t n = *count_ptr;
    const uint32_t tiles = (n + BF_BLOCK - 1u) / BF_BLOCK;
    const uint32_t wave = threadIdx.x >> 6;
    for (uint32_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const uint32_t i = tile * BF_BLOCK + threadIdx.x;
        const uint8_t r = i < n ? route[i] : 0;
        const uint64_t m1 = __ballot(r == 1);
        const uint64_t m2 = __ballot(r == 2);
        if ((threadIdx.x & 63u) == 0u) {
            wave_counts[0][wave] = (uint32_t)__popcll(m1);
            wave_counts[1][wave] = (uint32_t)__popcll(m2);
        }
        __syncthreads();
        if (threadIdx.x < 2u) {
            uint32_t total = 0;
            for (uint32_t w = 0; w < BF_BLOCK / 64; ++w) total += wave_counts[threadIdx.x][w];
            tile_counts[2 * tile + threadIdx.x] = total;
        }
        __syncthreads();
    }
}

/* single workgroup: exclusive scan of the per-tile counts for both routes, plus the queue bases */
__global__ __launch_bounds__(1024) void k_compact_scan(const uint32_t* count_ptr, const uint32_t* tile_counts,
                                                       uint32_t* tile_offsets, uint32_t* size1, uint32_t* size2,
                                                       uint32_t append1, uint32_t append2) {
    __shared__ uint32_t partial[2][1024];
    const uint32_t t = threadIdx.x;
    const uint32_t tiles = (*count_ptr + BF_BLOCK - 1u) / BF_BLOCK;
    const uint32_t per = (tiles + 1023u) / 1024u;
    const uint32_t begin = t * per;
    const uint32_t end = begin + per < tiles ? begin + per : tiles;
    uint32_t s1 = 0, s2 = 0;
    for (uint32_t b = begin; b < end; ++b) {
        s1 += tile_counts[2 * b];
        s2 += tile_counts[2 * b + 1];
    }
    partial[0][t] = s1;
    partial[1][t] = s2;
    __syncthreads();
    for (uint32_t off = 1; off < 1024u; off <<= 1) {
        const uint32_t a = t >= off ? partial[0][t - off] : 0u;
        const uint32_t c = t >= off ? partial[1][t - off] : 0u;
        __syncthreads();
        partial[0][t] += a;
        partial[1][t] += c;
        __syncthreads();
    }
    const uint32_t base1 = append1 ? *size1 : 0u;
    const uint32_t base2 = append2 ? *size2 : 0u;
    uint32_t run1 = base1 + partial[0][t] - s1;
    uint32_t run2 = base2 + partial[1][t] - s2;
    for (uint32_t b = begin; b < end; ++b) {
        tile_offsets[2 * b] = run1;
        tile_offsets[2 * b + 1] = run2;
        run1 += tile_counts[2 * b];
        run2 += tile_counts[2 * b + 1];
    }
    __syncthreads();
    if (t == 1023u) {
        *size1 = base1 + partial[0][1023];
        *size2 = base2 + partial[1][1023];
    }
}

__global__ __launch_bounds__(BF_BLOCK) void k_compact_scatter(const uint8_t* route, const uint32_t* count_ptr,
                                                              const uint32_t* src, const uint32_t* tile_offsets,
                                                              uint32_t* dst1, uint32_t* dst2) {
    __shared__ uint32_t wave_counts[2][BF_BLOCK / 64];
    const uint32_t n = *count_ptr;
    const uint32_t tiles = (n + BF_BLOCK - 1u) / BF_BLOCK;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = threadIdx.x >> 6;
    const uint64_t below = lane == 0u ? 0ull : (~0ull >> (64u - lane));
    for (uint32_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const uint32_t i = tile * BF_BLOCK + threadIdx.x;
        const uint8_t r = i < n ? route[i] : 0;
        const uint64_t m1 = __ballot(r == 1);
        const uint64_t m2 = __ballot(r == 2);
        if (lane == 0u) {
            wave_counts[0][wave] = (uint32_t)__popcll(m1);
            wave_counts[1][wave] = (uint32_t)__popcll(m2);
        }
        __syncthreads();
        if (r != 0) {
            const uint32_t k = r - 1u;
            uint32_t rank = (uint32_t)__popcll((k == 0 ? m1 : m2) & below);
            for (uint32_t w = 0; w < wave; ++w) rank += wave_counts[k][w];
            const uint32_t position = tile_offsets[2 * tile + k] + rank;
            (k == 0 ? dst1 : dst2)[position] = src[i];
        }
        __syncthreads();
    }
}

/* ---- shade order: hit-queue items grouped by the closure kind of their material ---------------
 * Shading a wave of mixed closure kinds issues every kind's code under a partial mask. The shade
 * stage can instead visit the hit queue through a permutation that groups items by kind (stable
 * within a kind); each item still writes its route at its own queue position, so the compactions
 * that follow see the queue order unchanged. Buckets: 0 multi-closure, 1..4 the closure kind. */
#define BF_SHADE_BUCKETS 8u

template <bool BATCH>
BF_D uint32_t shade_bucket(const bf_transport_params& P, const uint32_t* queue, uint32_t i) {
    const uint32_t slot = queue[i];
    const bf_dev_scene& S = job_at<BATCH>(P, slot).scene;
    const bf_material_desc& M = S.materials[S.instances[P.paths.hit_ids[slot].x].material_index];
    if (M.closure_count != 1u) return 0u;
    const uint32_t kind = S.closures[M.closure_offset].kind;
    return kind < BF_SHADE_BUCKETS ? kind : 0u;
}

/* Traversal order: rays grouped by direction octant, so that a wave's rays tend to take the same
 * near/far child order. */
BF_D uint32_t octant(float4 d) { return (d.x < 0.0f ? 1u : 0u) | (d.y < 0.0f ? 2u : 0u) | (d.z < 0.0f ? 4u : 0u); }

#define BF_KEY_SHADE 0u
#define BF_KEY_RAY 1u
#define BF_KEY_SHADOW 2u

template <bool BATCH, uint32_t KEY>
BF_D uint32_t bucket_key(const bf_transport_params& P, const uint32_t* queue, uint32_t i) {
    if constexpr (KEY == BF_KEY_SHADE) return shade_bucket<BATCH>(P, queue, i);
    else if constexpr (KEY == BF_KEY_RAY) return octant(P.paths.ray_direction[queue[i]]);
    else return octant(P.paths.shadow_direction[queue[i]]);
}

template <bool BATCH, uint32_t KEY>
__global__ __launch_bounds__(BF_BLOCK) void k_bucket_count(bf_transport_params P, const uint32_t* queue,
                                                           const uint32_t* count_ptr, uint8_t* keys, uint32_t* tile_counts) {
    __shared__ uint32_t wave_counts[BF_SHADE_BUCKETS][BF_BLOCK / 64];
    const uint32_t n = *count_ptr;
    const uint32_t tiles = (n + BF_BLOCK - 1u) / BF_BLOCK;
    const uint32_t wave = threadIdx.x >> 6;
    for (uint32_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const uint32_t i = tile * BF_BLOCK + threadIdx.x;
        const uint32_t key = i < n ? bucket_key<BATCH, KEY>(P, queue, i) : BF_SHADE_BUCKETS;
        if (i < n) keys[i] = (uint8_t)key;
#pragma unroll
        for (uint32_t b = 0; b < BF_SHADE_BUCKETS; ++b) {
            const uint64_t m = __ballot(key == b);
            if ((threadIdx.x & 63u) == 0u) wave_counts[b][wave] = (uint32_t)__popcll(m);
        }
        __syncthreads();
        if (threadIdx.x < BF_SHADE_BUCKETS) {
            uint32_t total = 0;
            for (uint32_t w = 0; w < BF_BLOCK / 64; ++w) total += wave_counts[threadIdx.x][w];
            tile_counts[BF_SHADE_BUCKETS * tile + threadIdx.x] = total;
        }
        __syncthreads();
    }
}

/* single workgroup: tile offsets = bucket base (items of lower buckets) + earlier tiles' counts */
__global__ __launch_bounds__(1024) void k_bucket_scan(const uint32_t* count_ptr, const uint32_t* tile_counts,
                                                      uint32_t* tile_offsets) {
    __shared__ uint32_t partial[BF_SHADE_BUCKETS][1024];
    __shared__ uint32_t base[BF_SHADE_BUCKETS];
    const uint32_t t = threadIdx.x;
    const uint32_t tiles = (*count_ptr + BF_BLOCK - 1u) / BF_BLOCK;
    const uint32_t per = (tiles + 1023u) / 1024u;
    const uint32_t begin = t * per;
    const uint32_t end = begin + per < tiles ? begin + per : tiles;
    uint32_t own[BF_SHADE_BUCKETS];
#pragma unroll
    for (uint32_t b = 0; b < BF_SHADE_BUCKETS; ++b) own[b] = 0u;
    for (uint32_t tile = begin; tile < end; ++tile)
#pragma unroll
        for (uint32_t b = 0; b < BF_SHADE_BUCKETS; ++b) own[b] += tile_counts[BF_SHADE_BUCKETS * tile + b];
#pragma unroll
    for (uint32_t b = 0; b < BF_SHADE_BUCKETS; ++b) partial[b][t] = own[b];
    __syncthreads();
    for (uint32_t off = 1; off < 1024u; off <<= 1) {
        uint32_t add[BF_SHADE_BUCKETS];
#pragma unroll
        for (uint32_t b = 0; b < BF_SHADE_BUCKETS; ++b) add[b] = t >= off ? partial[b][t - off] : 0u;
        __syncthreads();
#pragma unroll
        for (uint32_t b = 0; b < BF_SHADE_BUCKETS; ++b) partial[b][t] += add[b];
        __syncthreads();
    }
    if (t == 0u) {
        uint32_t run = 0;
        for (uint32_t b = 0; b < BF_SHADE_BUCKETS; ++b) {
            base[b] = run;
            run += partial[b][1023];
        }
    }
    __syncthreads();
    uint32_t run[BF_SHADE_BUCKETS];
#pragma unroll
    for (uint32_t b = 0; b < BF_SHADE_BUCKETS; ++b) run[b] = base[b] + partial[b][t] - own[b];
    for (uint32_t tile = begin; tile < end; ++tile)
#pragma unroll
        for (uint32_t b = 0; b < BF_SHADE_BUCKETS; ++b) {
            tile_offsets[BF_SHADE_BUCKETS * tile + b] = run[b];
            run[b] += tile_counts[BF_SHADE_BUCKETS * tile + b];
        }
}

__global__ __launch_bounds__(BF_BLOCK) void k_bucket_scatter(const uint8_t* keys, const uint32_t* count_ptr,
                                                             const uint32_t* tile_offsets, uint32_t* order) {
    __shared__ uint32_t wave_counts[BF_SHADE_BUCKETS][BF_BLOCK / 64];
    const uint32_t n = *count_ptr;
    const uint32_t tiles = (n + BF_BLOCK - 1u) / BF_BLOCK;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = threadIdx.x >> 6;
    const uint64_t below = lane == 0u ? 0ull : (~0ull >> (64u - lane));
    for (uint32_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const uint32_t i = tile * BF_BLOCK + threadIdx.x;
        const uint32_t key = i < n ? keys[i] : BF_SHADE_BUCKETS;
        uint64_t mine = 0ull;
#pragma unroll
        for (uint32_t b = 0; b < BF_SHADE_BUCKETS; ++b) {
            const uint64_t m = __ballot(key == b);
            if (key == b) mine = m;
            if (lane == 0u) wave_counts[b][wave] = (uint32_t)__popcll(m);
        }
        __syncthreads();
        if (key < BF_SHADE_BUCKETS) {
            uint32_t rank = (uint32_t)__popcll(mine & below);
            for (uint32_t w = 0; w < wave; ++w) rank += wave_counts[key][w];
            order[tile_offsets[BF_SHADE_BUCKETS * tile + key] + rank] = i;
        }
        __syncthreads();
    }
}

/* End of a bounce: record the queue sizes of the bounce, make the continuation queue the next ray
 * queue, clear the scratch sizes, and advance the bounce counter (one thread). */
__global__ void k_bounce_end(uint32_t* sizes, uint32_t* history, uint32_t history_bounces) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const uint32_t b = sizes[7];
    if (b < history_bounces) {
        history[4 * b + 0] = sizes[0];
        history[4 * b + 1] = sizes[1];
        history[4 * b + 2] = sizes[2];
        history[4 * b + 3] = sizes[3];
    }
    sizes[0] = sizes[3];
    sizes[1] = 0u;
    sizes[2] = 0u;
    sizes[3] = 0u;
    sizes[4] = 0u;
    sizes[5] = 0u;
    sizes[7] = b + 1u;
}

/* ---- film: per-pixel accumulation in sample order (Film.hpp:150-248, Cie1931Sensor.cpp) ----- */
BF_D void cie_xyz(const float* radiance, const float* wavelengths, float pdf, float normalization, float* xyz) {
    float sum[3] = {0.0f, 0.0f, 0.0f};
    for (int lane = 0; lane < 4; ++lane) {
        float response[3];
        cie_cmf(wavelengths[lane], response);
        for (int c = 0; c < 3; ++c) {
            /* scaled_sensor_term (Cie1931Sensor.cpp:104-124) */
            float term = 0.0f;
            if (radiance[lane] != 0.0f && response[c] != 0.0f) {
                int ce, re, ne, pe;
                const float cm = bf_frexpf(radiance[lane], &ce);
                const float rm = bf_frexpf(response[c], &re);
                const float nm = bf_frexpf(normalization, &ne);
                const float pm = bf_frexpf(pdf, &pe);
                const float mantissa = cm * rm * nm / pm;
                term = bf_ldexpf(mantissa, ce + re + ne - pe);
            }
            sum[c] += term;
        }
    }
    xyz[0] = sum[0];
    xyz[1] = sum[1];
    xyz[2] = sum[2];
}

__global__ __launch_bounds__(BF_BLOCK) void k_film(bf_dev_paths paths, bf_dev_job J, float pdf, float normalization) {
    const uint32_t pixels = J.pixels;
    const uint32_t passes = J.slot_count / pixels;
    float* film = J.film;
    double* film_compensated = J.film_compensated;
    uint32_t* film_counts = J.film_counts;
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= pixels) return;
    float f[3] = {film[3 * (size_t)p], film[3 * (size_t)p + 1], film[3 * (size_t)p + 2]};
    double s[3], cmp[3];
    for (int c = 0; c < 3; ++c) {
        s[c] = film_compensated[6 * (size_t)p + 2 * c];
        cmp[c] = film_compensated[6 * (size_t)p + 2 * c + 1];
    }
    uint32_t count = film_counts[p];
    for (uint32_t pass = 0; pass < passes; ++pass) {
        const size_t slot = (size_t)J.slot_begin + (size_t)pass * pixels + p;
        const float4 L4 = paths.radiance[slot];
        if (J.tape_radiance) J.tape_radiance[path_of(J, (uint32_t)slot)] = L4;
        const float4 W4 = paths.wavelengths[slot];
        const float L[4] = {L4.x, L4.y, L4.z, L4.w};
        const float W[4] = {W4.x, W4.y, W4.z, W4.w};
        float xyz[3];
        cie_xyz(L, W, pdf, normalization, xyz);
        /* Color.hpp:77-87 */
        const float rgb[3] = {3.2404542f * xyz[0] - 1.5371385f * xyz[1] - 0.4985314f * xyz[2],
                              -0.9692660f * xyz[0] + 1.8760108f * xyz[1] + 0.0415560f * xyz[2],
                              0.0556434f * xyz[0] - 0.2040259f * xyz[1] + 1.0572252f * xyz[2]};
        for (int c = 0; c < 3; ++c) {
            f[c] += rgb[c] * 1.0f;
            const double value = (double)rgb[c] * 1.0;
            const double updated = s[c] + value;
            if (fabs(s[c]) >= fabs(value)) cmp[c] += (s[c] - updated) + value;
            else cmp[c] += (value - updated) + s[c];
            s[c] = updated;
        }
        ++count;
    }
    for (int c = 0; c < 3; ++c) {
        film[3 * (size_t)p + c] = f[c];
        film_compensated[6 * (size_t)p + 2 * c] = s[c];
        film_compensated[6 * (size_t)p + 2 * c + 1] = cmp[c];
    }
    film_counts[p] = count;
}

/* FNV-1a-64 of a queue's slot bytes (diagnostic; one thread) */
__global__ void k_queue_hash(const uint32_t* queue, const uint32_t* count_ptr, unsigned long long* out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    uint64_t h = 14695981039346656037ull;
    const uint32_t n = *count_ptr;
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t v = queue[i];
        for (int b = 0; b < 4; ++b) {
            h ^= (v >> (8 * b)) & 0xFFu;
            h *= 1099511628211ull;
        }
    }
    out[0] = h;
    out[1] = n;
}

/* ---- stand-alone traversal (blackframe_hip_launch_scene_closest_hit / _occlusion) ----------- */
/* One ray per lane with a 64-entry LDS stack and no spill area: a lane whose TLAS and BLAS entries
 * together exceed 64 reports stack overflow (the reference allows 64 per structure; a median-split
 * BVH never holds more than its depth + 1 entries per structure). */
__global__ __launch_bounds__(BF_API_BLOCK) void k_scene_closest_hit(bf_dev_scene S, const bf_transport_ray* rays,
                                                                    uint32_t count, bf_closest_hit_result* results,
                                                                    uint32_t* counters) {
    __shared__ uint2 s_stack[BF_API_STACK * BF_API_BLOCK];
    trav_stack<BF_API_STACK, BF_API_BLOCK> stack{s_stack + threadIdx.x, nullptr, 0u, 0u, 0u, BF_API_STACK};
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const bf_transport_ray r = rays[i];
    bf_closest_hit_result out;
    memset(&out, 0, sizeof out);
    trav_counters ctr{0u, 0u, 0u};
    if (!valid_traversal_ray(r)) {
        out.status = BF_HIT_INVALID_RAY;
    } else {
        trav_ray ray;
        ray.o = mk(r.origin_x, r.origin_y, r.origin_z);
        ray.d = mk(r.direction_x, r.direction_y, r.direction_z);
        ray.t_min = r.t_min;
        ray.t_max = r.t_max;
        ray.mask = r.visibility_mask;
        trav_hit best;
        bool occ;
        const auto world = [&](v3f& o, v3f& d) {
            o = mk(rays[i].origin_x, rays[i].origin_y, rays[i].origin_z);
            d = mk(rays[i].direction_x, rays[i].direction_y, rays[i].direction_z);
        };
        uint32_t st = traverse<false, true>(S, ray, world, stack, best, occ, ctr);
        if (st == BF_HIT_MISS && best.present)
            st = closest_hit_barycentrics(S, ray.o, ray.d, ray.t_min, ray.t_max, best);
        if (st != BF_HIT_MISS) {
            out.status = st;
        } else if (best.present) {
            const bf_dev_instance& I = S.instances[best.scene_instance];
            const float4 q0 = S.leaf_triangles[3 * (size_t)best.leaf + 0];
            const float4 q1 = S.leaf_triangles[3 * (size_t)best.leaf + 1];
            const float4 q2 = S.leaf_triangles[3 * (size_t)best.leaf + 2];
            v3f n;
            if (!world_normal(I.local_to_world, mk(q0.x, q0.y, q0.z), mk(q1.x, q1.y, q1.z), mk(q2.x, q2.y, q2.z), n)) {
                out.status = BF_HIT_NUMERICAL_FAILURE;
            } else {
                out.status = BF_HIT_HIT;
                out.hit.parameter = best.t;
                out.hit.barycentric0 = best.b0;
                out.hit.barycentric1 = best.b1;
                out.hit.barycentric2 = best.b2;
                out.hit.normal_x = n.x;
                out.hit.normal_y = n.y;
                out.hit.normal_z = n.z;
                out.hit.ids.object = I.object_id;
                out.hit.ids.instance = I.id;
                out.hit.ids.geometry = S.geometries[I.geometry_index].id;
                out.hit.ids.primitive = best.primitive;
                out.hit.ids.material = S.materials[I.material_index].id;
            }
        }
    }
    results[i] = out;
    if (counters) {
        counters[3 * (size_t)i] = ctr.nodes;
        counters[3 * (size_t)i + 1] = ctr.triangles;
        counters[3 * (size_t)i + 2] = ctr.instances;
    }
}

__global__ __launch_bounds__(BF_API_BLOCK) void k_scene_occlusion(bf_dev_scene S, const bf_transport_ray* rays,
                                                                  uint32_t count, bf_occlusion_result* results) {
    __shared__ uint2 s_stack[BF_API_STACK * BF_API_BLOCK];
    trav_stack<BF_API_STACK, BF_API_BLOCK> stack{s_stack + threadIdx.x, nullptr, 0u, 0u, 0u, BF_API_STACK};
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const bf_transport_ray r = rays[i];
    bf_occlusion_result out;
    memset(&out, 0, sizeof out);
    if (!valid_traversal_ray(r)) {
        out.status = BF_HIT_INVALID_RAY;
    } else {
        trav_ray ray;
        ray.o = mk(r.origin_x, r.origin_y, r.origin_z);
        ray.d = mk(r.direction_x, r.direction_y, r.direction_z);
        ray.t_min = r.t_min;
        ray.t_max = r.t_max;
        ray.mask = r.visibility_mask;
        trav_hit best;
        bool occ;
        trav_counters ctr{0u, 0u, 0u};
        const auto world = [&](v3f& o, v3f& d) {
            o = mk(rays[i].origin_x, rays[i].origin_y, rays[i].origin_z);
            d = mk(rays[i].direction_x, rays[i].direction_y, rays[i].direction_z);
        };
        const uint32_t st = traverse<true, false>(S, ray, world, stack, best, occ, ctr);
        out.status = st != BF_HIT_MISS ? st : (occ ? BF_OCCLUDED : BF_VISIBLE);
    }
    results[i] = out;
}

/* ---- stand-alone traversal over the reference's device blobs ----------------------------------
 * blackframe_hip_launch_scene_closest_hit_blob / _occlusion_blob take the reference launcher's
 * arguments (SceneClosestHit.hpp, CudaSceneClosestHit.cu:188-1330): "BLKFSOA4" scene bytes and a
 * "BLKFBVH1" BVH in device memory. The first lane of every workgroup validates both layouts
 * (validate_scene_layout / validate_bvh_layout, :199-358) and a bad one answers every ray of the
 * workgroup with invalid_scene / invalid_bvh. Each lane then traverses one ray over the canonical
 * 48-byte nodes in the reference's order: pop, re-test the node's own box against the current
 * t_max, push both intersecting children far then near (ties to the first), enter an instance leaf
 * in object space after checking it against the scene's columns, and test a BLAS leaf's triangle
 * from the scene's index and position columns. The box and triangle decisions are this library's
 * binary64 slab and watertight test, so the results equal those of
 * blackframe_hip_launch_scene_closest_hit on the scene the blobs describe. The TLAS and each BLAS
 * get their own 64-entry LDS stack, as in the reference. */
struct blob_view {
    const uint8_t* scene;
    const bf_scene_soa_header* sh;
    const uint8_t* bvh;
    const bf_bvh_header* bh;
};

template <typename T>
BF_D const T* soa_col(const blob_view& b, uint32_t column) {
    return (const T*)(b.scene + b.sh->columns[column].offset_bytes);
}

template <typename T>
BF_D const T* bvh_arr(const blob_view& b, uint32_t array) {
    return (const T*)(b.bvh + b.bh->arrays[array].offset_bytes);
}

BF_D uint64_t blob_align(uint64_t v, uint64_t a) { return (v + a - 1u) / a * a; }

/* validate_scene_layout (:199-250) */
BF_D uint32_t blob_scene_layout(const uint8_t* bytes, size_t n) {
    if (n < sizeof(bf_scene_soa_header)) return BF_HIT_INVALID_SCENE;
    const bf_scene_soa_header& h = *(const bf_scene_soa_header*)bytes;
    if (h.magic != BF_SCENE_SOA_MAGIC || h.abi_major != BF_SCENE_SOA_ABI_MAJOR || h.abi_minor != BF_SCENE_SOA_ABI_MINOR ||
        h.header_size != sizeof(bf_scene_soa_header) || h.column_count != BF_SCENE_SOA_COLUMN_COUNT ||
        h.hash_algorithm != BF_SCENE_SOA_HASH_FNV1A64 || h.environment_count > 1u || h.total_size_bytes != n ||
        h.total_size_bytes < sizeof(bf_scene_soa_header) || h.texture_count > 0xFFFFFFFFull ||
        h.image_texture_count > 0xFFFFFFFFull || h.image_mip_count > 0xFFFFFFFFull)
        return BF_HIT_INVALID_SCENE;
    for (int k = 0; k < 5; ++k)
        if (h.reserved[k] != 0u) return BF_HIT_INVALID_SCENE;
    const uint64_t max = ~0ull;
    uint64_t cursor = blob_align(sizeof(bf_scene_soa_header), BF_SCENE_SOA_COLUMN_ALIGNMENT);
    for (uint32_t c = 0; c < BF_SCENE_SOA_COLUMN_COUNT; ++c) {
        const bf_scene_soa_column& d = h.columns[c];
        const uint64_t count = bf_soa::column_count(h, c);
        const uint64_t size = bf_soa::column_element_size(c);
        if (d.element_count != count || d.element_size != size || d.reserved != 0u) return BF_HIT_INVALID_SCENE;
        if (count == 0u) {
            if (d.offset_bytes != 0u) return BF_HIT_INVALID_SCENE;
            continue;
        }
        if (cursor > max - (BF_SCENE_SOA_COLUMN_ALIGNMENT - 1u)) return BF_HIT_INVALID_SCENE;
        cursor = blob_align(cursor, BF_SCENE_SOA_COLUMN_ALIGNMENT);
        if (d.offset_bytes != cursor || count > max / size) return BF_HIT_INVALID_SCENE;
        const uint64_t bytes_in_column = count * size;
        if (cursor > max - bytes_in_column) return BF_HIT_INVALID_SCENE;
        cursor += bytes_in_column;
    }
    return cursor == h.total_size_bytes ? BF_HIT_MISS : BF_HIT_INVALID_SCENE;
}

/* validate_bvh_layout (:291-358) */
BF_D uint32_t blob_bvh_layout(const uint8_t* bytes, size_t n, const bf_scene_soa_header& scene) {
    if (n < sizeof(bf_bvh_header)) return BF_HIT_INVALID_BVH;
    const bf_bvh_header& h = *(const bf_bvh_header*)bytes;
    const uint64_t invalid = BF_INVALID_INDEX;
    if (h.magic != BF_BVH_MAGIC || h.abi_major != BF_BVH_ABI_MAJOR || h.abi_minor != BF_BVH_ABI_MINOR ||
        h.header_size != sizeof(bf_bvh_header) || h.array_count != BF_BVH_ARRAY_COUNT ||
        h.hash_algorithm != BF_BVH_HASH_FNV1A64 || h.leaf_capacity != BF_BVH_LEAF_CAPACITY || h.total_size_bytes != n ||
        h.total_size_bytes < sizeof(bf_bvh_header) || h.source_scene_hash != scene.content_hash ||
        h.blas_count != scene.geometry_count || h.primitive_reference_count != scene.triangle_count ||
        h.instance_reference_count != scene.instance_count || h.blas_count > invalid || h.blas_node_count > invalid ||
        h.primitive_reference_count > invalid || h.tlas_node_count > invalid || h.instance_reference_count > invalid)
        return BF_HIT_INVALID_BVH;
    if ((h.blas_count == 0u && (h.blas_node_count != 0u || h.primitive_reference_count != 0u)) ||
        (h.blas_count != 0u && (h.primitive_reference_count < h.blas_count ||
                                h.blas_node_count != h.primitive_reference_count * 2u - h.blas_count)) ||
        (h.instance_reference_count == 0u && h.tlas_node_count != 0u) ||
        (h.instance_reference_count != 0u && h.tlas_node_count != h.instance_reference_count * 2u - 1u) ||
        (h.tlas_node_count == 0u && h.tlas_root_node != BF_INVALID_INDEX) ||
        (h.tlas_node_count != 0u && h.tlas_root_node != 0u))
        return BF_HIT_INVALID_BVH;
    for (int k = 0; k < 5; ++k)
        if (h.reserved[k] != 0u) return BF_HIT_INVALID_BVH;
    const uint64_t counts[BF_BVH_ARRAY_COUNT] = {h.blas_count, h.blas_node_count, h.primitive_reference_count,
                                                 h.tlas_node_count, h.instance_reference_count};
    const uint64_t sizes[BF_BVH_ARRAY_COUNT] = {sizeof(bf_bvh_blas), sizeof(bf_bvh_node), sizeof(uint32_t),
                                                sizeof(bf_bvh_node), sizeof(bf_bvh_instance_ref)};
    const uint64_t max = ~0ull;
    uint64_t cursor = blob_align(sizeof(bf_bvh_header), BF_BVH_ARRAY_ALIGNMENT);
    for (uint32_t a = 0; a < BF_BVH_ARRAY_COUNT; ++a) {
        const bf_bvh_array_descriptor& d = h.arrays[a];
        if (d.element_count != counts[a] || d.element_size != sizes[a] || d.reserved != 0u) return BF_HIT_INVALID_BVH;
        if (counts[a] == 0u) {
            if (d.offset_bytes != 0u) return BF_HIT_INVALID_BVH;
            continue;
        }
        if (cursor > max - (BF_BVH_ARRAY_ALIGNMENT - 1u)) return BF_HIT_INVALID_BVH;
        cursor = blob_align(cursor, BF_BVH_ARRAY_ALIGNMENT);
        if (d.offset_bytes != cursor || counts[a] > max / sizes[a]) return BF_HIT_INVALID_BVH;
        const uint64_t bytes_in_array = counts[a] * sizes[a];
        if (cursor > max - bytes_in_array) return BF_HIT_INVALID_BVH;
        cursor += bytes_in_array;
    }
    return cursor == h.total_size_bytes ? BF_HIT_MISS : BF_HIT_INVALID_BVH;
}

/* intersect_bounds (:407-455): the node's box, validated, in the reference's binary64 slab */
BF_D uint32_t blob_box(const bf_bvh_node& n, const trav_space& s, float t_min, float t_max, bool& hit, float& nearv) {
    if (!bf_isfinite(n.min_x) || !bf_isfinite(n.min_y) || !bf_isfinite(n.min_z) || !bf_isfinite(n.max_x) ||
        !bf_isfinite(n.max_y) || !bf_isfinite(n.max_z) || n.min_x > n.max_x || n.min_y > n.max_y || n.min_z > n.max_z)
        return BF_HIT_INVALID_TOPOLOGY;
    double v = 0.0;
    const uint32_t st = slab_exact(n.min_x, n.min_y, n.min_z, n.max_x, n.max_y, n.max_z, s, t_min, t_max, hit, v);
    nearv = (float)v;
    return st;
}

/* push_intersecting_children (:466-512) */
BF_D uint32_t blob_push_children(const bf_bvh_node* nodes, const bf_bvh_node& n, uint32_t begin, uint32_t count,
                                 const trav_space& s, float t_min, float t_max, uint32_t* stack, uint32_t& size) {
    const uint64_t end = (uint64_t)begin + count;
    const uint32_t a = n.first_child, b = n.second_child;
    if (a == b || a < begin || b < begin || a >= end || b >= end) return BF_HIT_INVALID_TOPOLOGY;
    bool ha, hb;
    float na = 0.0f, nb = 0.0f;
    uint32_t st = blob_box(nodes[a], s, t_min, t_max, ha, na);
    if (st != BF_HIT_MISS) return st;
    st = blob_box(nodes[b], s, t_min, t_max, hb, nb);
    if (st != BF_HIT_MISS) return st;
    if (!ha && !hb) return BF_HIT_MISS;
    if (ha != hb) {
        if (size >= BF_STRUCTURE_STACK) return BF_HIT_STACK_OVERFLOW;
        stack[(size++) * BF_API_BLOCK] = ha ? a : b;
        return BF_HIT_MISS;
    }
    const bool a_near = na < nb || (na == nb && a < b);
    if (size >= BF_STRUCTURE_STACK) return BF_HIT_STACK_OVERFLOW;
    stack[(size++) * BF_API_BLOCK] = a_near ? b : a;
    if (size >= BF_STRUCTURE_STACK) return BF_HIT_STACK_OVERFLOW;
    stack[(size++) * BF_API_BLOCK] = a_near ? a : b;
    return BF_HIT_MISS;
}

/* a 3x4 affine matrix from one of the scene's 16-element matrix column groups (:513-537) */
BF_D uint32_t blob_matrix(const blob_view& b, uint32_t first_column, uint32_t instance, float* m) {
    float full[16];
    for (uint32_t e = 0; e < 16u; ++e) {
        full[e] = soa_col<float>(b, first_column + e)[instance];
        if (!bf_isfinite(full[e])) return BF_HIT_NUMERICAL_FAILURE;
    }
    if (full[12] != 0.0f || full[13] != 0.0f || full[14] != 0.0f || full[15] != 1.0f) return BF_HIT_INVALID_TOPOLOGY;
    for (uint32_t e = 0; e < 12u; ++e) m[e] = full[e];
    return BF_HIT_MISS;
}

struct blob_candidate {
    bool present;
    float t;
    uint32_t instance_id, primitive;
    bf_closest_hit hit;
};

/* traverse_blas (:895-1085) */
template <bool ANY>
BF_D uint32_t blob_blas(const blob_view& b, const bf_bvh_instance_ref& ref, const bf_transport_ray& ray, v3f lo, v3f ld,
                        uint32_t* stack, blob_candidate& best, bool& occluded) {
    const bf_bvh_blas* blases = bvh_arr<bf_bvh_blas>(b, 0);
    const bf_bvh_node* nodes = bvh_arr<bf_bvh_node>(b, 1);
    const uint32_t* prims = bvh_arr<uint32_t>(b, 2);
    if (ref.blas_index >= b.bh->blas_count || ref.scene_geometry_index >= b.sh->geometry_count)
        return BF_HIT_INVALID_TOPOLOGY;
    const bf_bvh_blas B = blases[ref.blas_index];
    const uint64_t node_end = (uint64_t)B.node_offset + B.node_count;
    const uint64_t ref_end = (uint64_t)B.primitive_reference_offset + B.primitive_reference_count;
    if (B.reserved[0] != 0u || B.reserved[1] != 0u || B.geometry_id != ref.geometry_id || B.root_node != B.node_offset ||
        B.node_count == 0u || node_end > b.bh->blas_node_count ||
        B.node_count != (uint64_t)B.primitive_reference_count * 2u - 1u || ref_end > b.bh->primitive_reference_count)
        return BF_HIT_INVALID_TOPOLOGY;
    const uint32_t g = ref.scene_geometry_index;
    const uint64_t vertex_offset = soa_col<uint64_t>(b, bf_soa::C_GEOMETRY_VERTEX_OFFSET)[g];
    const uint64_t vertex_count = soa_col<uint64_t>(b, bf_soa::C_GEOMETRY_VERTEX_COUNT)[g];
    const uint64_t triangle_offset = soa_col<uint64_t>(b, bf_soa::C_GEOMETRY_TRIANGLE_OFFSET)[g];
    const uint64_t triangle_count = soa_col<uint64_t>(b, bf_soa::C_GEOMETRY_TRIANGLE_COUNT)[g];
    if (soa_col<uint32_t>(b, bf_soa::C_GEOMETRY_ID)[g] != ref.geometry_id || triangle_count != B.primitive_reference_count ||
        vertex_count == 0u || vertex_offset > b.sh->vertex_count || vertex_count > b.sh->vertex_count - vertex_offset ||
        triangle_offset > b.sh->triangle_count || triangle_count > b.sh->triangle_count - triangle_offset)
        return BF_HIT_INVALID_TOPOLOGY;
    trav_space s;
    make_space(lo, ld, s);
    trav_shear sh;
    make_shear(ld, sh);
    uint32_t size = 0, visited = 0;
    stack[(size++) * BF_API_BLOCK] = B.root_node;
    while (size != 0u) {
        if (visited++ >= B.node_count) return BF_HIT_INVALID_TOPOLOGY;
        const uint32_t index = stack[(--size) * BF_API_BLOCK];
        if (index < B.node_offset || index >= node_end) return BF_HIT_INVALID_TOPOLOGY;
        const bf_bvh_node n = nodes[index];
        const float t_max = ANY ? ray.t_max : best.t;
        bool h;
        float nearv;
        uint32_t st = blob_box(n, s, ray.t_min, t_max, h, nearv);
        if (st != BF_HIT_MISS) return st;
        if (!h) continue;
        if (n.kind == 0u) {
            if (n.split_axis > 2u || n.reference_offset != 0u || n.reference_count != 0u) return BF_HIT_INVALID_TOPOLOGY;
            st = blob_push_children(nodes, n, B.node_offset, B.node_count, s, ray.t_min, t_max, stack, size);
            if (st != BF_HIT_MISS) return st;
            continue;
        }
        if (n.kind != 1u || n.split_axis != BF_INVALID_INDEX || n.first_child != BF_INVALID_INDEX ||
            n.second_child != BF_INVALID_INDEX || n.reference_count != 1u || n.reference_offset < B.primitive_reference_offset ||
            n.reference_offset >= ref_end)
            return BF_HIT_INVALID_TOPOLOGY;
        const uint32_t prim = prims[n.reference_offset];
        if (prim >= triangle_count) return BF_HIT_INVALID_TOPOLOGY;
        const uint64_t tri = triangle_offset + prim;
        v3f v[3];
        for (uint32_t c = 0; c < 3u; ++c) {
            const uint32_t local = soa_col<uint32_t>(b, bf_soa::C_TRIANGLE_VERTEX_0 + c)[tri];
            if (local >= vertex_count) return BF_HIT_INVALID_TOPOLOGY;
            const uint64_t gv = vertex_offset + local;
            v[c] = mk(soa_col<float>(b, bf_soa::C_POSITION_X)[gv], soa_col<float>(b, bf_soa::C_POSITION_X + 1)[gv],
                      soa_col<float>(b, bf_soa::C_POSITION_X + 2)[gv]);
            if (!fin3(v[c])) return BF_HIT_NUMERICAL_FAILURE;
        }
        tri_result tr;
        tr.hit = false;
        st = watertight<!ANY>(v[0], v[1], v[2], lo, ld, sh, ray.t_min, t_max, tr);
        if (st != BF_HIT_MISS) return st;
        if (!tr.hit) continue;
        if (ANY) {
            occluded = true;
            return BF_HIT_MISS;
        }
        if (best.present && !(tr.t < best.t ||
                              (tr.t == best.t && (ref.instance_id < best.instance_id ||
                                                  (ref.instance_id == best.instance_id && prim < best.primitive)))))
            continue;
        float m[12];
        st = blob_matrix(b, bf_soa::C_INSTANCE_LOCAL_TO_WORLD, ref.scene_instance_index, m);
        if (st != BF_HIT_MISS) return st;
        v3f nrm;
        if (!world_normal(m, v[0], v[1], v[2], nrm)) return BF_HIT_NUMERICAL_FAILURE;
        const uint32_t inst = ref.scene_instance_index;
        if (soa_col<uint32_t>(b, bf_soa::C_INSTANCE_ID)[inst] != ref.instance_id ||
            soa_col<uint32_t>(b, bf_soa::C_INSTANCE_GEOMETRY_ID)[inst] != ref.geometry_id)
            return BF_HIT_INVALID_TOPOLOGY;
        best.present = true;
        best.t = tr.t;
        best.instance_id = ref.instance_id;
        best.primitive = prim;
        bf_closest_hit& hr = best.hit;
        memset(&hr, 0, sizeof hr);
        hr.parameter = tr.t;
        hr.barycentric0 = tr.b0;
        hr.barycentric1 = tr.b1;
        hr.barycentric2 = tr.b2;
        hr.normal_x = nrm.x;
        hr.normal_y = nrm.y;
        hr.normal_z = nrm.z;
        hr.ids.object = soa_col<uint32_t>(b, bf_soa::C_INSTANCE_OBJECT_ID)[inst];
        hr.ids.instance = ref.instance_id;
        hr.ids.geometry = ref.geometry_id;
        hr.ids.primitive = prim;
        hr.ids.material = soa_col<uint32_t>(b, bf_soa::C_INSTANCE_MATERIAL_ID)[inst];
    }
    return BF_HIT_MISS;
}

/* trace_scene (:1087-1186) */
template <bool ANY>
BF_D uint32_t blob_trace(const blob_view& b, const bf_transport_ray& ray, uint32_t* tlas_stack, uint32_t* blas_stack,
                         blob_candidate& best, bool& occluded) {
    if (b.bh->tlas_node_count == 0u) return BF_HIT_MISS;
    const bf_bvh_node* nodes = bvh_arr<bf_bvh_node>(b, 3);
    const bf_bvh_instance_ref* refs = bvh_arr<bf_bvh_instance_ref>(b, 4);
    const uint32_t tlas_count = (uint32_t)b.bh->tlas_node_count;
    trav_space world;
    const v3f o = mk(ray.origin_x, ray.origin_y, ray.origin_z), d = mk(ray.direction_x, ray.direction_y, ray.direction_z);
    make_space(o, d, world);
    uint32_t size = 0, visited = 0;
    tlas_stack[(size++) * BF_API_BLOCK] = b.bh->tlas_root_node;
    while (size != 0u) {
        if (visited++ >= tlas_count) return BF_HIT_INVALID_TOPOLOGY;
        const uint32_t index = tlas_stack[(--size) * BF_API_BLOCK];
        if (index >= tlas_count) return BF_HIT_INVALID_TOPOLOGY;
        const bf_bvh_node n = nodes[index];
        const float t_max = ANY ? ray.t_max : best.t;
        bool h;
        float nearv;
        uint32_t st = blob_box(n, world, ray.t_min, t_max, h, nearv);
        if (st != BF_HIT_MISS) return st;
        if (!h) continue;
        if (n.kind == 0u) {
            if (n.split_axis > 2u || n.reference_offset != 0u || n.reference_count != 0u) return BF_HIT_INVALID_TOPOLOGY;
            st = blob_push_children(nodes, n, 0u, tlas_count, world, ray.t_min, t_max, tlas_stack, size);
            if (st != BF_HIT_MISS) return st;
            continue;
        }
        if (n.kind != 1u || n.split_axis != BF_INVALID_INDEX || n.first_child != BF_INVALID_INDEX ||
            n.second_child != BF_INVALID_INDEX || n.reference_count != 1u ||
            n.reference_offset >= b.bh->instance_reference_count)
            return BF_HIT_INVALID_TOPOLOGY;
        const bf_bvh_instance_ref ref = refs[n.reference_offset];
        if (ref.reserved[0] != 0u || ref.reserved[1] != 0u || ref.scene_instance_index >= b.sh->instance_count ||
            ref.scene_geometry_index >= b.sh->geometry_count || ref.blas_index >= b.bh->blas_count)
            return BF_HIT_INVALID_TOPOLOGY;
        const uint32_t inst = ref.scene_instance_index;
        if (soa_col<uint32_t>(b, bf_soa::C_INSTANCE_ID)[inst] != ref.instance_id ||
            soa_col<uint32_t>(b, bf_soa::C_INSTANCE_GEOMETRY_ID)[inst] != ref.geometry_id ||
            soa_col<uint32_t>(b, bf_soa::C_INSTANCE_VISIBILITY_MASK)[inst] != ref.visibility_mask)
            return BF_HIT_INVALID_TOPOLOGY;
        if ((ray.visibility_mask & ref.visibility_mask) == 0u) continue;
        float m[12];
        st = blob_matrix(b, bf_soa::C_INSTANCE_WORLD_TO_LOCAL, inst, m);
        if (st != BF_HIT_MISS) return st;
        const v3f lo = xform_point(m, o), ld = xform_vector(m, d);
        if (!fin3(lo) || !fin3(ld) || (ld.x == 0.0f && ld.y == 0.0f && ld.z == 0.0f)) return BF_HIT_NUMERICAL_FAILURE;
        st = blob_blas<ANY>(b, ref, ray, lo, ld, blas_stack, best, occluded);
        if (st != BF_HIT_MISS) return st;
        if (ANY && occluded) return BF_HIT_MISS;
    }
    return BF_HIT_MISS;
}

template <bool ANY>
__global__ __launch_bounds__(BF_API_BLOCK) void k_blob_traverse(const uint8_t* scene_bytes, size_t scene_size,
                                                                const uint8_t* bvh_bytes, size_t bvh_size,
                                                                const bf_transport_ray* rays, uint32_t count,
                                                                void* results) {
    __shared__ uint32_t s_layout;
    __shared__ uint32_t s_tlas[BF_STRUCTURE_STACK * BF_API_BLOCK];
    __shared__ uint32_t s_blas[BF_STRUCTURE_STACK * BF_API_BLOCK];
    if (threadIdx.x == 0u) {
        uint32_t st = blob_scene_layout(scene_bytes, scene_size);
        if (st == BF_HIT_MISS) st = blob_bvh_layout(bvh_bytes, bvh_size, *(const bf_scene_soa_header*)scene_bytes);
        s_layout = st;
    }
    __syncthreads();
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    uint32_t status = s_layout;
    blob_candidate best;
    best.present = false;
    bool occluded = false;
    if (status == BF_HIT_MISS) {
        const bf_transport_ray r = rays[i];
        if (!valid_traversal_ray(r)) {
            status = BF_HIT_INVALID_RAY;
        } else {
            const blob_view b{scene_bytes, (const bf_scene_soa_header*)scene_bytes, bvh_bytes,
                              (const bf_bvh_header*)bvh_bytes};
            best.t = r.t_max;
            status = blob_trace<ANY>(b, r, s_tlas + threadIdx.x, s_blas + threadIdx.x, best, occluded);
        }
    }
    if (ANY) {
        bf_occlusion_result out;
        memset(&out, 0, sizeof out);
        out.status = status != BF_HIT_MISS ? status : (occluded ? BF_OCCLUDED : BF_VISIBLE);
        ((bf_occlusion_result*)results)[i] = out;
    } else {
        bf_closest_hit_result out;
        memset(&out, 0, sizeof out);
        if (status != BF_HIT_MISS) out.status = status;
        else if (best.present) {
            out.status = BF_HIT_HIT;
            out.hit = best.hit;
        }
        ((bf_closest_hit_result*)results)[i] = out;
    }
}

/* ---- sampler probe (SampleStreamKernel.hpp:170-176 counterpart) ------------------------------ */
__global__ void k_sample_dump(sampler_ctx sc, const bf_sample_stream_index* indices, const uint64_t* dims, uint32_t count,
                              float* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    sample_key k;
    k.px = indices[i].pixel_x;
    k.py = indices[i].pixel_y;
    k.s = indices[i].sample_index;
    k.seed = indices[i].seed;
    out[i] = sample_1d(sc, k, dims[i]);
}

/* The reference's dimension map (SampleDimensionMap.cpp:13-69, SampleStreamKernel.hpp:60-91):
 * bounce b uses dimensions 4 + 10 b + {light selection, light u/v, bsdf component/u/v, medium x3,
 * roulette} for b <= 2^32 - 1. */
__global__ void k_sample_stream_dump(const bf_sample_dump_request* requests, uint32_t count,
                                     bf_sample_dimension_manifest* manifest, bf_sample_dump_result* results) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0u) {
        bf_sample_dimension_manifest m;
        m.schema_version = BF_SAMPLE_DIMENSION_SCHEMA;
        m.reserved = 0u;
        m.primary_dimensions = bf_sample_primary_dimensions{BF_DIM_RASTER_X, BF_DIM_RASTER_Y, 0u, 1u, 2u, 3u};
        m.first_bounce_dimension = BF_SAMPLE_FIRST_BOUNCE_DIMENSION;
        m.dimensions_per_bounce = BF_SAMPLE_DIMENSIONS_PER_BOUNCE;
        m.maximum_bounce_index = BF_SAMPLE_MAXIMUM_BOUNCE;
        m.bounce_offsets = bf_sample_bounce_dimensions{0u, 1u, 2u, 3u, 4u, 5u, 6u, 7u, 8u, 9u};
        *manifest = m;
    }
    if (i >= count) return;
    const bf_sample_dump_request r = requests[i];
    sample_key k;
    k.px = r.sample_stream.pixel_x;
    k.py = r.sample_stream.pixel_y;
    k.s = r.sample_stream.sample_index;
    k.seed = r.sample_stream.seed;
    bf_sample_dump_result out;
    out.reserved = 0u;
    out.reserved_tail = 0u;
    out.indexed_bits = indexed_bits(k, r.dimension);
    out.sample = (float)(out.indexed_bits >> 40) * 0x1p-24f;
    uint64_t d[10];
    if (r.bounce_index > BF_SAMPLE_MAXIMUM_BOUNCE) {
        out.status = BF_SAMPLE_DUMP_BOUNCE_OUT_OF_RANGE;
        for (int k2 = 0; k2 < 10; ++k2) d[k2] = 0u;
    } else {
        out.status = BF_SAMPLE_DUMP_SUCCESS;
        const uint64_t first = BF_SAMPLE_FIRST_BOUNCE_DIMENSION + r.bounce_index * BF_SAMPLE_DIMENSIONS_PER_BOUNCE;
        for (int k2 = 0; k2 < 10; ++k2) d[k2] = first + (uint64_t)k2;
    }
    out.bounce_dimensions = bf_sample_bounce_dimensions{d[0], d[1], d[2], d[3], d[4], d[5], d[6], d[7], d[8], d[9]};
    results[i] = out;
}

/* ---- small utility kernels (used instead of hipMemsetAsync so the sequence stays capturable) - */
/* one thread per gradient element: the replica rows in row order, then one add to the caller's buffer */
__global__ __launch_bounds__(BF_BLOCK) void k_reduce_replicas(const float* replicas, uint32_t offset, uint32_t count,
                                                              float* dst) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= count) return;
    float sum = 0.0f;
    for (uint32_t row = 0; row < BF_GRADIENT_REPLICAS; ++row) sum += replicas[(size_t)row * BF_GRADIENT_ROW + offset + e];
    dst[e] += sum;
}

/* ---- texture fitting: L2 film loss and Adam ------------------------------------------------- */

__device__ inline void adam_element(float& p, float g, float& m, float& v, const bfk_adam_coefficients& c) {
    m = c.beta1 * m + c.one_minus_beta1 * g;
    v = c.beta2 * v + c.one_minus_beta2 * (g * g);
    const float den = sqrtf(v) / c.bias2_sqrt + c.epsilon;
    p = p - c.step_size * (m / den);
    p = fminf(fmaxf(p, c.lower), c.upper);
}

/* VEC: the four arrays are 16-byte aligned and the body runs on float4 (the tail of count % 4
 * elements is done by the first threads of the grid). */
template <bool VEC>
__global__ __launch_bounds__(BF_BLOCK) void k_adam(float* value, const float* gradient, float* first, float* second,
                                                   uint64_t count, bfk_adam_coefficients c, const float2* schedule,
                                                   const uint32_t* step) {
    if (schedule) {
        /* a replayed step: the bias-corrected step size and sqrt(bias2) of the current step number,
         * computed on the host in the same binary64 arithmetic as a direct call */
        const float2 row = schedule[*step];
        c.step_size = row.x;
        c.bias2_sqrt = row.y;
    }
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t scalar_begin = 0;
    if (VEC) {
        const uint64_t quads = count / 4u;
        for (uint64_t i = t; i < quads; i += stride) {
            float4 p = reinterpret_cast<float4*>(value)[i];
            const float4 g = reinterpret_cast<const float4*>(gradient)[i];
            float4 m = reinterpret_cast<float4*>(first)[i];
            float4 v = reinterpret_cast<float4*>(second)[i];
            adam_element(p.x, g.x, m.x, v.x, c);
            adam_element(p.y, g.y, m.y, v.y, c);
            adam_element(p.z, g.z, m.z, v.z, c);
            adam_element(p.w, g.w, m.w, v.w, c);
            reinterpret_cast<float4*>(value)[i] = p;
            reinterpret_cast<float4*>(first)[i] = m;
            reinterpret_cast<float4*>(second)[i] = v;
        }
        scalar_begin = quads * 4u;
    }
    for (uint64_t i = scalar_begin + t; i < count; i += stride) {
        float p = value[i], m = first[i], v = second[i];
        adam_element(p, gradient[i], m, v, c);
        value[i] = p;
        first[i] = m;
        second[i] = v;
    }
}

__global__ __launch_bounds__(BF_BLOCK) void k_film_l2(const float* film, const uint32_t* counts, const float* target,
                                                      float* gradient, uint32_t pixels, float two_scale, double scale,
                                                      double* loss) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double sum = 0.0;
    if (i < pixels) {
        const float c = (float)(counts[i] > 1u ? counts[i] : 1u);
        for (uint32_t ch = 0; ch < 3u; ++ch) {
            const float d = film[(size_t)i * 3u + ch] / c - target[(size_t)i * 3u + ch];
            gradient[(size_t)i * 3u + ch] = (two_scale * d) / c;
            sum += (double)d * (double)d;
        }
    }
    for (int o = 32; o > 0; o >>= 1) sum += __shfl_down(sum, o, 64);
    __shared__ double partial[BF_BLOCK / 64];
    if ((threadIdx.x & 63u) == 0) partial[threadIdx.x >> 6] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        double block = 0.0;
        for (uint32_t w = 0; w < BF_BLOCK / 64; ++w) block += partial[w];
        atomicAdd(loss, scale * block);
    }
}

__global__ void k_zero_doubles(double* dst, uint32_t count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) dst[i] = 0.0;
}

/* ---- scene blob hash: FNV-1a-64 with the content-hash field read as zero ---------------------- */
/* The byte chain is sequential by definition, so one lane walks it; it reads 16-byte words where
 * the blob is 16-byte aligned and single bytes otherwise. */
/* validate_scene_soa_header on the device copy of a blob (the check every reference stage makes
 * of the uploaded header, CudaWavefrontStages.cu:818-880), one thread; plus the byte count. */
/* validate_scene_soa_header (SceneSoaAbi.hpp:283-341), then the stage check's further conditions
 * (CudaWavefrontStages.cu:818-862): the header's total size equals the device byte count, every
 * record count fits 32 bits, and the mesh-area lights fit the uniform light registry (2^24). */
__global__ void k_scene_soa_validate(const bf_scene_soa_header* header, uint64_t byte_count, uint32_t* status) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const bf_scene_soa_header& h = *header;
    uint32_t st = bf_soa::validate_header(h);
    if (st == BF_SCENE_SOA_VALID && h.total_size_bytes > byte_count) st = BF_SCENE_SOA_TRUNCATED;
    if (st == BF_SCENE_SOA_VALID && h.total_size_bytes != byte_count) st = BF_SCENE_SOA_SIZE_MISMATCH;
    const uint64_t counts[12] = {h.object_count, h.geometry_count, h.vertex_count, h.triangle_count,
                                 h.material_count, h.closure_count, h.instance_count, h.mesh_area_light_count,
                                 h.punctual_light_count, h.texture_count, h.image_texture_count, h.image_mip_count};
    for (int k = 0; k < 12 && st == BF_SCENE_SOA_VALID; ++k)
        if (counts[k] > 0xFFFFFFFFull) st = BF_SCENE_SOA_COUNT_OVERFLOW;
    if (st == BF_SCENE_SOA_VALID && h.mesh_area_light_count > (1ull << 24)) st = BF_SCENE_SOA_TOO_MANY_MESH_LIGHTS;
    *status = st;
}

__global__ void k_scene_hash(const uint8_t* bytes, size_t count, uint64_t* out) {
    if (blockIdx.x != 0u || threadIdx.x != 0u) return;
    uint64_t h = BF_FNV1A64_OFFSET_BASIS;
    const auto step = [&](size_t index, uint32_t byte) {
        const uint32_t v = index >= BF_SCENE_SOA_HASH_OFFSET && index < BF_SCENE_SOA_HASH_OFFSET + 8u ? 0u : byte;
        h ^= v;
        h *= BF_FNV1A64_PRIME;
    };
    size_t i = 0;
    if ((reinterpret_cast<uintptr_t>(bytes) & 15u) == 0u) {
        for (; i + 16u <= count; i += 16u) {
            const uint4 w = *reinterpret_cast<const uint4*>(bytes + i);
            const uint32_t words[4] = {w.x, w.y, w.z, w.w};
            for (int k = 0; k < 4; ++k)
                for (int b = 0; b < 4; ++b) step(i + 4u * k + b, (words[k] >> (8 * b)) & 0xFFu);
        }
    }
    for (; i < count; ++i) step(i, bytes[i]);
    *out = h;
}

/* ---- transport record layout, as device code sees it (CudaTransportAbiProbe.cu:33-160) ------- */
__global__ void k_layout_manifest(bf_layout_manifest* out) {
    if (blockIdx.x != 0u || threadIdx.x != 0u) return;
    bf_layout_manifest m;
    memset(&m, 0, sizeof m);
    m.abi_major = BF_QUEUE_ABI_MAJOR;
    m.abi_minor = BF_QUEUE_ABI_MINOR;
    m.device_cxx_standard = (uint32_t)__cplusplus;
    m.value_count = BF_LAYOUT_VALUE_COUNT;
    uint32_t k = 0;
    const auto put = [&](size_t v) { m.values[k++] = (uint32_t)v; };
    /* the reference's LayoutValue order */
    put(sizeof(bf_transport_spectrum)); put(alignof(bf_transport_spectrum)); put(offsetof(bf_transport_spectrum, values));
    put(sizeof(bf_transport_ray)); put(alignof(bf_transport_ray));
    put(offsetof(bf_transport_ray, origin_x)); put(offsetof(bf_transport_ray, origin_y));
    put(offsetof(bf_transport_ray, origin_z)); put(offsetof(bf_transport_ray, t_min));
    put(offsetof(bf_transport_ray, direction_x)); put(offsetof(bf_transport_ray, direction_y));
    put(offsetof(bf_transport_ray, direction_z)); put(offsetof(bf_transport_ray, t_max));
    put(offsetof(bf_transport_ray, time)); put(offsetof(bf_transport_ray, visibility_mask));
    put(offsetof(bf_transport_ray, current_medium)); put(offsetof(bf_transport_ray, reserved));
    put(sizeof(bf_sample_stream_index)); put(alignof(bf_sample_stream_index));
    put(offsetof(bf_sample_stream_index, pixel_x)); put(offsetof(bf_sample_stream_index, pixel_y));
    put(offsetof(bf_sample_stream_index, sample_index)); put(offsetof(bf_sample_stream_index, seed));
    put(sizeof(bf_path_slot)); put(alignof(bf_path_slot)); put(offsetof(bf_path_slot, value));
    put(sizeof(bf_queue_header)); put(alignof(bf_queue_header));
    put(offsetof(bf_queue_header, abi_major)); put(offsetof(bf_queue_header, abi_minor));
    put(offsetof(bf_queue_header, struct_size)); put(offsetof(bf_queue_header, queue_kind));
    put(offsetof(bf_queue_header, capacity)); put(offsetof(bf_queue_header, size));
    put(offsetof(bf_queue_header, overflow_count)); put(offsetof(bf_queue_header, rejected_count));
    put(offsetof(bf_queue_header, reserved));
    put(sizeof(bf_surface_identifiers)); put(alignof(bf_surface_identifiers));
    put(offsetof(bf_surface_identifiers, object)); put(offsetof(bf_surface_identifiers, instance));
    put(offsetof(bf_surface_identifiers, geometry)); put(offsetof(bf_surface_identifiers, primitive));
    put(offsetof(bf_surface_identifiers, material)); put(offsetof(bf_surface_identifiers, reserved));
    put(sizeof(bf_closest_hit)); put(alignof(bf_closest_hit));
    put(offsetof(bf_closest_hit, parameter)); put(offsetof(bf_closest_hit, barycentric0));
    put(offsetof(bf_closest_hit, barycentric1)); put(offsetof(bf_closest_hit, barycentric2));
    put(offsetof(bf_closest_hit, normal_x)); put(offsetof(bf_closest_hit, normal_y));
    put(offsetof(bf_closest_hit, normal_z)); put(offsetof(bf_closest_hit, reserved)); put(offsetof(bf_closest_hit, ids));
    put(sizeof(bf_path_state_lane)); put(alignof(bf_path_state_lane));
    put(offsetof(bf_path_state_lane, beta)); put(offsetof(bf_path_state_lane, accumulated_radiance));
    put(offsetof(bf_path_state_lane, wavelength_nanometers)); put(offsetof(bf_path_state_lane, wavelength_pdf_values));
    put(offsetof(bf_path_state_lane, diffuse_depth)); put(offsetof(bf_path_state_lane, glossy_depth));
    put(offsetof(bf_path_state_lane, specular_depth)); put(offsetof(bf_path_state_lane, transmission_depth));
    put(offsetof(bf_path_state_lane, volume_depth)); put(offsetof(bf_path_state_lane, depth));
    put(offsetof(bf_path_state_lane, eta_scale)); put(offsetof(bf_path_state_lane, current_medium));
    put(offsetof(bf_path_state_lane, delta_flags)); put(offsetof(bf_path_state_lane, wavelength_pdf_measures));
    put(offsetof(bf_path_state_lane, reserved));
    *out = m;
}

/* ---- queue compaction: CudaWavefrontQueueCompaction.cu:55-250 -------------------------------- */
/* A selected outcome succeeded and names the route; the selected slots keep their outcome order
 * (a stable partition), the append is all or nothing, and a header is used only while its
 * immutable fields and counters satisfy the queue contract (WavefrontQueueDevice.cuh:77-85). */
BF_D bool qc_header_ok(const bf_queue_header& h, uint32_t kind, uint32_t stride) {
    return h.abi_major == BF_QUEUE_ABI_MAJOR && h.abi_minor == BF_QUEUE_ABI_MINOR &&
           h.struct_size == (uint32_t)sizeof(bf_queue_header) && h.queue_kind == kind && h.capacity == stride &&
           h.reserved == 0u;
}

BF_D bool qc_selected(const bf_stage_outcome& o, uint32_t route) { return o.status == 0u && o.route == route; }

BF_D void qc_saturating_add(uint32_t* counter, uint32_t value) {
    uint32_t observed = *counter;
    for (;;) {
        const uint32_t next = value > 0xFFFFFFFFu - observed ? 0xFFFFFFFFu : observed + value;
        const uint32_t previous = atomicCAS(counter, observed, next);
        if (previous == observed) return;
        observed = previous;
    }
}

/* Queue pushes (WavefrontQueueDevice.cuh:118-197): the requests of one wave that name one queue
 * kind reserve a contiguous run with one compare-and-swap on its size, leader lane first; the
 * lanes past the room that was left are rejected and counted. */
__global__ __launch_bounds__(BF_BLOCK) void k_queue_push(bf_queue_soa q, const bf_queue_push* requests, uint32_t n,
                                                         uint32_t* outcomes) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63u;
    const bool live = i < n;
    bf_queue_push r{0xFFFFFFFFu, 0u};
    if (live) r = requests[i];
    uint32_t status = BF_PUSH_INVALID_CONTRACT;
    const bool grouped = live && r.queue_kind < BF_QUEUE_COUNT && qc_header_ok(q.headers[r.queue_kind], r.queue_kind, q.slot_stride);
    uint64_t pending = __ballot(grouped);
    while (pending != 0ull) {
        const uint32_t leader = (uint32_t)__ffsll((long long)pending) - 1u;
        const uint32_t kind = __shfl(r.queue_kind, (int)leader, 64);
        const uint64_t group = __ballot(grouped && r.queue_kind == kind) & pending;
        pending &= ~group;
        const uint32_t size = (uint32_t)__popcll(group);
        uint32_t base = 0u, accepted = 0u, valid = 1u;
        if (lane == leader) {
            bf_queue_header& h = q.headers[kind];
            uint32_t observed = __hip_atomic_load(&h.size, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            for (;;) {
                if (observed > h.capacity) {
                    valid = 0u;
                    break;
                }
                const uint32_t remaining = h.capacity - observed;
                accepted = remaining < size ? remaining : size;
                base = observed;
                if (accepted == 0u) break;
                const uint32_t previous = atomicCAS(&h.size, observed, observed + accepted);
                if (previous == observed) break;
                observed = previous;
            }
            if (valid) {
                qc_saturating_add(&h.overflow_count, size - accepted);
                qc_saturating_add(&h.rejected_count, size - accepted);
            }
        }
        base = __shfl(base, (int)leader, 64);
        accepted = __shfl(accepted, (int)leader, 64);
        valid = __shfl(valid, (int)leader, 64);
        if ((group >> lane) & 1ull) {
            const uint32_t rank = (uint32_t)__popcll(group & ((1ull << lane) - 1ull));
            if (!valid) {
                status = BF_PUSH_INVALID_CONTRACT;
            } else if (rank >= accepted) {
                status = BF_PUSH_CAPACITY_EXHAUSTED;
            } else {
                q.path_slots[(size_t)kind * q.slot_stride + base + rank] = r.slot;
                status = BF_PUSH_PUSHED;
            }
        }
    }
    if (live) outcomes[i] = status;
}

__global__ void k_qc_init(bf_queue_soa q, uint32_t route, uint32_t n, bf_compaction_result* result) {
    if (blockIdx.x != 0u || threadIdx.x != 0u) return;
    bf_compaction_result r;
    memset(&r, 0, sizeof r);
    r.status = BF_COMPACTION_SUCCESS;
    r.queue_kind = route;
    r.route = route;
    r.input_count = n;
    const bf_queue_header h = q.headers[route];
    r.initial_size = h.size;
    if (!qc_header_ok(h, route, q.slot_stride) || h.size > h.capacity || h.overflow_count != 0u || h.rejected_count != 0u)
        r.status = BF_COMPACTION_INVALID_CONTRACT;
    *result = r;
}

__global__ void k_qc_validate(const bf_stage_outcome* outcomes, uint32_t n, uint32_t route, uint32_t capacity,
                              bf_compaction_result* result) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bf_stage_outcome o = outcomes[i];
    if (qc_selected(o, route) && o.path_slot >= capacity)
        atomicCAS(&result->status, (uint32_t)BF_COMPACTION_SUCCESS, (uint32_t)BF_COMPACTION_INVALID_CONTRACT);
}

/* selected outcomes per BF_BLOCK-item tile, from wave ballots */
__global__ __launch_bounds__(BF_BLOCK) void k_qc_count(const bf_stage_outcome* outcomes, uint32_t n, uint32_t route,
                                                       uint32_t* tile_counts) {
    __shared__ uint32_t waves[BF_BLOCK / 64u];
    const uint32_t tiles = (n + BF_BLOCK - 1u) / BF_BLOCK;
    for (uint32_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const uint32_t i = tile * BF_BLOCK + threadIdx.x;
        const bool sel = i < n && qc_selected(outcomes[i], route);
        const uint32_t c = (uint32_t)__popcll(__ballot(sel));
        if ((threadIdx.x & 63u) == 0u) waves[threadIdx.x >> 6] = c;
        __syncthreads();
        if (threadIdx.x == 0u) {
            uint32_t sum = 0u;
            for (uint32_t w = 0; w < BF_BLOCK / 64u; ++w) sum += waves[w];
            tile_counts[tile] = sum;
        }
        __syncthreads();
    }
}

/* one workgroup: exclusive scan of the tile counts, and their total */
__global__ __launch_bounds__(1024) void k_qc_scan(const uint32_t* tile_counts, uint32_t tiles, uint32_t* tile_offsets,
                                                  uint32_t* total) {
    __shared__ uint32_t partial[1024];
    const uint32_t t = threadIdx.x;
    const uint32_t per = (tiles + 1023u) / 1024u;
    const uint32_t begin = t * per < tiles ? t * per : tiles;
    const uint32_t end = begin + per < tiles ? begin + per : tiles;
    uint32_t own = 0u;
    for (uint32_t k = begin; k < end; ++k) own += tile_counts[k];
    partial[t] = own;
    __syncthreads();
    for (uint32_t off = 1u; off < 1024u; off <<= 1) {
        const uint32_t add = t >= off ? partial[t - off] : 0u;
        __syncthreads();
        partial[t] += add;
        __syncthreads();
    }
    uint32_t run = partial[t] - own;
    for (uint32_t k = begin; k < end; ++k) {
        tile_offsets[k] = run;
        run += tile_counts[k];
    }
    if (t == 1023u) *total = partial[1023];
}

/* the transactional capacity decision (CudaWavefrontQueueCompaction.cu:155-192) */
__global__ void k_qc_decide(bf_queue_soa q, const uint32_t* total, bf_compaction_result* result) {
    if (blockIdx.x != 0u || threadIdx.x != 0u) return;
    const uint32_t selected = *total;
    result->selected_count = selected;
    if (result->status != BF_COMPACTION_SUCCESS) return;
    bf_queue_header& h = q.headers[result->queue_kind];
    if (!qc_header_ok(h, result->queue_kind, q.slot_stride) || h.size != result->initial_size || h.overflow_count != 0u ||
        h.rejected_count != 0u) {
        result->status = BF_COMPACTION_INVALID_CONTRACT;
        return;
    }
    if (selected > h.capacity - h.size) {
        qc_saturating_add(&h.overflow_count, selected);
        qc_saturating_add(&h.rejected_count, selected);
        result->status = BF_COMPACTION_CAPACITY_EXHAUSTED;
        result->rejected_count = selected;
    }
}

/* stable scatter: rank = tile offset + earlier waves of the tile + earlier lanes of the wave */
__global__ __launch_bounds__(BF_BLOCK) void k_qc_scatter(bf_queue_soa q, const bf_stage_outcome* outcomes, uint32_t n,
                                                         uint32_t route, const uint32_t* tile_offsets,
                                                         const bf_compaction_result* result) {
    __shared__ uint32_t waves[BF_BLOCK / 64u];
    if (result->status != BF_COMPACTION_SUCCESS) return;
    const uint32_t tiles = (n + BF_BLOCK - 1u) / BF_BLOCK;
    const uint32_t lane = threadIdx.x & 63u;
    const size_t base = (size_t)result->queue_kind * q.slot_stride + result->initial_size;
    for (uint32_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const uint32_t i = tile * BF_BLOCK + threadIdx.x;
        bf_stage_outcome o{0u, 0u, 0u, 0u};
        if (i < n) o = outcomes[i];
        const bool sel = i < n && qc_selected(o, route);
        const uint64_t mask = __ballot(sel);
        if (lane == 0u) waves[threadIdx.x >> 6] = (uint32_t)__popcll(mask);
        __syncthreads();
        uint32_t rank = tile_offsets[tile];
        for (uint32_t w = 0; w < (threadIdx.x >> 6); ++w) rank += waves[w];
        rank += (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
        if (sel) q.path_slots[base + rank] = o.path_slot;
        __syncthreads();
    }
}

__global__ void k_qc_publish(bf_queue_soa q, bf_compaction_result* result) {
    if (blockIdx.x != 0u || threadIdx.x != 0u || result->status != BF_COMPACTION_SUCCESS) return;
    bf_queue_header& h = q.headers[result->queue_kind];
    if (!qc_header_ok(h, result->queue_kind, q.slot_stride) || h.size != result->initial_size || h.overflow_count != 0u ||
        h.rejected_count != 0u || result->selected_count > h.capacity - h.size) {
        result->status = BF_COMPACTION_INVALID_CONTRACT;
        return;
    }
    h.size += result->selected_count;
    result->published_count = result->selected_count;
}

__global__ void k_set_words(uint32_t* dst, uint32_t count, uint32_t value) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) dst[i] = value;
}

__global__ void k_copy_word(const uint32_t* src, uint32_t* dst) {
    if (threadIdx.x == 0 && blockIdx.x == 0) dst[0] = src[0];
}

__global__ void k_copy_words(const uint32_t* src, uint32_t* dst, uint32_t count) {
    for (uint32_t i = threadIdx.x; i < count; i += blockDim.x) dst[i] = src[i];
}

/* advances a replayed step counter (after the kernels that read it) */
__global__ void k_tick(uint32_t* counter) {
    if (threadIdx.x == 0 && blockIdx.x == 0) counter[0] += 1u;
}

/* Adds one launch's per-bounce queue sizes (rows of [ray, hit, shadow, continuation], ended by the
 * first empty ray queue) to totals [closest-hit queries, shade lanes, shadow queries, largest bounce
 * count]: the report of a launch whose host never reads its history. */
__global__ void k_history_totals(const uint32_t* history, uint32_t rows, unsigned long long* totals) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    unsigned long long ray = 0, hit = 0, shadow = 0;
    uint32_t active = 0;
    for (uint32_t b = 0; b < rows; ++b) {
        const uint32_t* h = history + 4 * b;
        if (h[0] == 0u) break;
        ray += h[0];
        hit += h[1];
        shadow += h[2];
        active = b + 1u;
    }
    totals[0] += ray;
    totals[1] += hit;
    totals[2] += shadow;
    if (active > totals[3]) totals[3] = active;
}

__global__ __launch_bounds__(BF_BLOCK) void k_export_paths(bf_dev_paths paths, bf_dev_job J) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= J.slot_count) return;
    const uint32_t i = J.slot_begin + k;
    bf_path_result* out = J.paths + J.path_base;
    bf_path_result r;
    const float4 L = paths.radiance[i], B = paths.beta[i], W = paths.wavelengths[i];
    const uint4 D = paths.depth[i];
    r.radiance[0] = L.x;
    r.radiance[1] = L.y;
    r.radiance[2] = L.z;
    r.radiance[3] = L.w;
    r.beta[0] = B.x;
    r.beta[1] = B.y;
    r.beta[2] = B.z;
    r.beta[3] = B.w;
    r.wavelengths[0] = W.x;
    r.wavelengths[1] = W.y;
    r.wavelengths[2] = W.z;
    r.wavelengths[3] = W.w;
    r.depth = D.y;
    r.termination = paths.status[i] & 0xFFu; /* the failure code stays in the workspace report */
    r.depth_counters = D.x;
    r.eta_scale = __uint_as_float(D.w);
    out[k] = r;
}

/* The terminal records of the path-input mode (WavefrontTransport.cpp:2753-2786): the path state
 * from the slot's columns (wavelength pdfs and measures pass through from the input), the ray the
 * path ended on (its first ray keeps the input's t_max, every later one runs to infinity), the ray
 * cone and a terminated lane control. A failed path returns zero records and termination NONE; its
 * failure is in the report. */
__global__ __launch_bounds__(BF_BLOCK) void k_export_path_records(bf_dev_paths paths, bf_dev_job J) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= J.slot_count) return;
    const uint32_t i = J.slot_begin + k;
    const uint64_t path = J.path_base + k;
    const uint32_t status = paths.status[i];
    const uint32_t termination = status & 0xFFu;
    bf_path_state_lane st;
    bf_transport_ray ray;
    float2 cone = make_float2(0.0f, 0.0f);
    memset(&st, 0, sizeof st);
    memset(&ray, 0, sizeof ray);
    bf_lane_control ctl;
    ctl.phase = BF_PHASE_TERMINATED;
    ctl.termination = 0u;
    ctl.flags = 0u;
    ctl.blocked_depth_limits = 0u;
    if (termination != BF_TERM_FAILED) {
        const bf_path_state_lane& in = J.in_states[path];
        const float4 B = paths.beta[i], L = paths.radiance[i], W = paths.wavelengths[i];
        const uint4 D = paths.depth[i];
        st.beta.values[0] = B.x;
        st.beta.values[1] = B.y;
        st.beta.values[2] = B.z;
        st.beta.values[3] = B.w;
        st.accumulated_radiance.values[0] = L.x;
        st.accumulated_radiance.values[1] = L.y;
        st.accumulated_radiance.values[2] = L.z;
        st.accumulated_radiance.values[3] = L.w;
        st.wavelength_nanometers.values[0] = W.x;
        st.wavelength_nanometers.values[1] = W.y;
        st.wavelength_nanometers.values[2] = W.z;
        st.wavelength_nanometers.values[3] = W.w;
        st.wavelength_pdf_values = in.wavelength_pdf_values;
#pragma unroll
        for (int w = 0; w < 4; ++w) st.wavelength_pdf_measures[w] = in.wavelength_pdf_measures[w];
        st.diffuse_depth = D.x & 0xFFu;
        st.glossy_depth = (D.x >> 8) & 0xFFu;
        st.specular_depth = (D.x >> 16) & 0xFFu;
        st.transmission_depth = D.x >> 24;
        st.depth = D.y;
        st.eta_scale = __uint_as_float(D.w);
        st.delta_flags = D.z & 3u;
        const float4 o = paths.ray_origin[i], d = paths.ray_direction[i];
        ray.origin_x = o.x;
        ray.origin_y = o.y;
        ray.origin_z = o.z;
        ray.t_min = d.w;
        ray.direction_x = d.x;
        ray.direction_y = d.y;
        ray.direction_z = d.z;
        ray.t_max = (D.z & BF_FIRST_RAY) != 0u ? J.in_rays[path].t_max : BF_INFINITY_F;
        ray.time = o.w;
        ray.visibility_mask = paths.visibility[i];
        cone = paths.cone[i];
        ctl.termination = termination;
        ctl.blocked_depth_limits = termination == BF_TERM_DEPTH_LIMIT ? status >> 8 : 0u;
    }
    J.out_states[path] = st;
    J.out_rays[path] = ray;
    J.out_cones[path] = bf_ray_cone{cone.x, cone.y};
    J.out_controls[path] = ctl;
}

/* ---- stage audit (CudaWavefrontStages.cu:4945-5033) -------------------------------------------
 * Each workgroup reduces its outcomes with wave ballots and shuffles into one partial (smallest
 * failing index, closure and light sample counts), written to the audit's reserved words through
 * atomics; a second one-wave launch on the same stream closes the audit (inspected count, first
 * failing outcome) and clears the words it used. */
BF_D bool outcome_failed(const bf_stage_outcome& o, uint32_t allowed_route_mask, uint32_t path_capacity) {
    if (o.status != BF_STAGE_SUCCESS || o.route > BF_ROUTE_TERMINATED) return true;
    return (allowed_route_mask & (1u << o.route)) == 0u || o.path_slot >= path_capacity;
}

__global__ void k_audit_init(uint32_t work_count, uint32_t stage_kind, bf_stage_audit* audit) {
    if (threadIdx.x != 0u) return;
    bf_stage_audit a;
    memset(&a, 0, sizeof a);
    a.abi_major = BF_STAGE_AUDIT_ABI_MAJOR;
    a.abi_minor = BF_STAGE_AUDIT_ABI_MINOR;
    a.struct_size = sizeof(bf_stage_audit);
    a.stage_kind = stage_kind;
    a.expected_work_count = work_count;
    a.first_failure_work_index = 0xFFFFFFFFu;
    *audit = a;
}

__global__ __launch_bounds__(BF_BLOCK) void k_audit_reduce(const bf_stage_outcome* outcomes, uint32_t work_count,
                                                           uint32_t allowed_route_mask, uint32_t path_capacity,
                                                           uint32_t shade, bf_stage_audit* audit) {
    __shared__ uint32_t wave_first[BF_BLOCK / 64], wave_closure[BF_BLOCK / 64], wave_light[BF_BLOCK / 64];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t first = 0xFFFFFFFFu, closure = 0u, light = 0u;
    for (uint32_t i = blockIdx.x * BF_BLOCK + threadIdx.x; i < work_count; i += gridDim.x * BF_BLOCK) {
        const bf_stage_outcome o = outcomes[i];
        if (outcome_failed(o, allowed_route_mask, path_capacity)) first = min(first, i);
        if (shade) {
            closure += (o.detail & BF_SHADE_DETAIL_CLOSURE_SAMPLED) ? 1u : 0u;
            light += (o.detail & BF_SHADE_DETAIL_LIGHT_SAMPLED) ? 1u : 0u;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        first = min(first, (uint32_t)__shfl_xor((int)first, off, 64));
        closure += (uint32_t)__shfl_xor((int)closure, off, 64);
        light += (uint32_t)__shfl_xor((int)light, off, 64);
    }
    if (lane == 0u) {
        wave_first[wave] = first;
        wave_closure[wave] = closure;
        wave_light[wave] = light;
    }
    __syncthreads();
    if (threadIdx.x == 0u) {
        for (uint32_t w = 1; w < BF_BLOCK / 64; ++w) {
            first = min(first, wave_first[w]);
            closure += wave_closure[w];
            light += wave_light[w];
        }
        if (first != 0xFFFFFFFFu) atomicMin(&audit->first_failure_work_index, first);
        if (closure) atomicAdd(&audit->closure_samples, closure);
        if (light) atomicAdd(&audit->light_samples, light);
    }
}

__global__ void k_audit_close(const bf_stage_outcome* outcomes, uint32_t work_count, bf_stage_audit* audit) {
    if (threadIdx.x != 0u) return;
    audit->inspected_work_count = work_count;
    const uint32_t f = audit->first_failure_work_index;
    if (f < work_count) {
        const bf_stage_outcome o = outcomes[f];
        audit->first_failure[0] = o.status;
        audit->first_failure[1] = o.route;
        audit->first_failure[2] = o.path_slot;
        audit->first_failure[3] = o.detail;
    }
}

/* ---- host-side launch helpers (stream-ordered, never synchronising) -------------------------- */
static inline uint32_t blocks_for(uint32_t n, uint32_t b) { return (n + b - 1u) / b; }

extern "C" {

hipError_t bfk_camera(const bf_transport_params* P, uint32_t count, uint32_t* ray_queue, uint8_t* route, bool batch,
                      bool adjoint, hipStream_t s) {
    if (count == 0) return hipSuccess;
    const dim3 grid(blocks_for(count, BF_BLOCK)), block(BF_BLOCK);
    if (batch && adjoint) hipLaunchKernelGGL((k_camera<true, true>), grid, block, 0, s, *P, count, ray_queue, route);
    else if (batch) hipLaunchKernelGGL((k_camera<true, false>), grid, block, 0, s, *P, count, ray_queue, route);
    else if (adjoint) hipLaunchKernelGGL((k_camera<false, true>), grid, block, 0, s, *P, count, ray_queue, route);
    else hipLaunchKernelGGL((k_camera<false, false>), grid, block, 0, s, *P, count, ray_queue, route);
    return hipGetLastError();
}

/* Workgroups of one persistent traversal launch: enough to fill every CU at the kernels'
 * occupancy, never more than the workspace's spill area covers, and no more than the items need. */
static uint32_t trace_blocks(const bf_transport_params* P, uint32_t max_count, uint32_t block) {
    const uint32_t needed = blocks_for((max_count + 63u) / 64u * 64u, block);
    const uint32_t resident = P->spill_stride / block;
    return needed < resident ? needed : resident;
}

} /* extern "C" */

template <bool BATCH, uint32_t KEY>
static void launch_shade_order(const bf_transport_params* P, const uint32_t* queue, const uint32_t* count_ptr,
                               uint32_t max_count, uint8_t* keys, uint32_t* tile_counts, uint32_t* tile_offsets,
                               uint32_t* order, hipStream_t s);

template <bool ANY, bool COUNT, bool BATCH, uint32_t BLOCK>
static void launch_trace(const bf_transport_params* P, uint32_t max_count, const uint32_t* q, const uint32_t* n,
                         uint8_t* route, unsigned long long* counters, hipStream_t s) {
    const dim3 grid(trace_blocks(P, max_count, BLOCK));
    hipLaunchKernelGGL(k_set_words, dim3(1), dim3(64), 0, s, P->fetch + (ANY ? 1 : 0), 1u, 0u);
    /* optional direction-octant order; the route bytes are free until the traversal writes them */
    uint32_t* order = ANY ? P->shadow_order : P->closest_order;
    if (order)
        launch_shade_order<BATCH, ANY ? BF_KEY_SHADOW : BF_KEY_RAY>(P, q, n, max_count, route, P->bucket_counts,
                                                                     P->bucket_offsets, order, s);
    if (ANY)
        hipLaunchKernelGGL((k_trace_shadow<COUNT, BATCH, BLOCK>), grid, dim3(BLOCK), 0, s, *P, q, n, order, route, counters);
    else
        hipLaunchKernelGGL((k_trace_closest<COUNT, BATCH, BLOCK>), grid, dim3(BLOCK), 0, s, *P, q, n, order, route, counters);
}

template <bool ANY, bool COUNT, bool BATCH>
static void launch_trace_block(const bf_transport_params* P, uint32_t max_count, const uint32_t* q, const uint32_t* n,
                               uint8_t* route, unsigned long long* counters, hipStream_t s) {
    if (P->trace_block == 64u) launch_trace<ANY, COUNT, BATCH, 64u>(P, max_count, q, n, route, counters, s);
    else if (P->trace_block == 256u) launch_trace<ANY, COUNT, BATCH, 256u>(P, max_count, q, n, route, counters, s);
    else launch_trace<ANY, COUNT, BATCH, 128u>(P, max_count, q, n, route, counters, s);
}

template <bool ANY>
static hipError_t trace_stage(const bf_transport_params* P, const uint32_t* queue, const uint32_t* count_ptr,
                              uint32_t max_count, uint8_t* route, unsigned long long* counters, bool batch, hipStream_t s) {
    if (max_count == 0) return hipSuccess;
    if (counters) {
        if (batch) launch_trace_block<ANY, true, true>(P, max_count, queue, count_ptr, route, counters, s);
        else launch_trace_block<ANY, true, false>(P, max_count, queue, count_ptr, route, counters, s);
    } else {
        if (batch) launch_trace_block<ANY, false, true>(P, max_count, queue, count_ptr, route, counters, s);
        else launch_trace_block<ANY, false, false>(P, max_count, queue, count_ptr, route, counters, s);
    }
    return hipGetLastError();
}

template <uint32_t BLOCK>
static void trace_kernels(const void** out) {
    out[0] = (const void*)k_trace_closest<false, false, BLOCK>;
    out[1] = (const void*)k_trace_closest<true, false, BLOCK>;
    out[2] = (const void*)k_trace_closest<false, true, BLOCK>;
    out[3] = (const void*)k_trace_closest<true, true, BLOCK>;
    out[4] = (const void*)k_trace_shadow<false, false, BLOCK>;
    out[5] = (const void*)k_trace_shadow<true, false, BLOCK>;
    out[6] = (const void*)k_trace_shadow<false, true, BLOCK>;
    out[7] = (const void*)k_trace_shadow<true, true, BLOCK>;
}

extern "C" {

hipError_t bfk_trace_closest(const bf_transport_params* P, const uint32_t* queue, const uint32_t* count_ptr,
                             uint32_t max_count, uint8_t* route, unsigned long long* counters, bool batch, hipStream_t s) {
    return trace_stage<false>(P, queue, count_ptr, max_count, route, counters, batch, s);
}

hipError_t bfk_trace_shadow(const bf_transport_params* P, const uint32_t* queue, const uint32_t* count_ptr,
                            uint32_t max_count, uint8_t* route, unsigned long long* counters, bool batch, hipStream_t s) {
    return trace_stage<true>(P, queue, count_ptr, max_count, route, counters, batch, s);
}

static int max_resident(const void* const* kernels, int count, uint32_t block, int device) {
    int cus = 0, per_cu = 0, m = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus <= 0) return 0;
    for (int k = 0; k < count; ++k) {
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernels[k], (int)block, 0) != hipSuccess) return 0;
        if (per_cu > m) m = per_cu;
    }
    return cus * m;
}

/* Lanes the device holds at once for the widest-resident traversal kernel variant: the width of
 * the spill area, one column per lane. */
uint32_t bfk_trace_resident_lanes(int device) {
    const void* k64[8];
    const void* k128[8];
    const void* k256[8];
    trace_kernels<64u>(k64);
    trace_kernels<128u>(k128);
    trace_kernels<256u>(k256);
    const uint32_t a = (uint32_t)max_resident(k64, 8, 64u, device) * 64u;
    const uint32_t b = (uint32_t)max_resident(k128, 8, 128u, device) * 128u;
    const uint32_t c = (uint32_t)max_resident(k256, 8, 256u, device) * 256u;
    if (a == 0 || b == 0 || c == 0) return 0;
    return std::max(a, std::max(b, c));
}

/* Resident workgroups of the forward shade variants, or of the adjoint ones (which hold fewer
 * waves per SIMD): the persistent shade grid of each, so that no workgroup waits for a slot. */
uint32_t bfk_shade_resident_blocks(int device, bool adjoint) {
    const void* forward[6] = {(const void*)k_shade<false, BF_SHADE_PLAIN>,    (const void*)k_shade<true, BF_SHADE_PLAIN>,
                              (const void*)k_shade<false, BF_SHADE_TEXTURED>, (const void*)k_shade<true, BF_SHADE_TEXTURED>,
                              (const void*)k_shade<false, BF_SHADE_MAPPED>,   (const void*)k_shade<true, BF_SHADE_MAPPED>};
    const void* backward[2] = {(const void*)k_shade<false, BF_SHADE_ADJOINT>, (const void*)k_shade<true, BF_SHADE_ADJOINT>};
    return adjoint ? (uint32_t)max_resident(backward, 2, BF_SHADE_BLOCK, device)
                   : (uint32_t)max_resident(forward, 6, BF_SHADE_BLOCK, device);
}

} /* extern "C" */

template <bool BATCH>
static void launch_shade(const bf_transport_params* P, dim3 grid, const uint32_t* queue, const uint32_t* count_ptr,
                         const uint32_t* order, uint8_t* route, uint32_t* sample_counts, uint32_t mode, hipStream_t s) {
    const dim3 block(BF_SHADE_BLOCK);
    if (mode == BF_SHADE_ADJOINT)
        hipLaunchKernelGGL((k_shade<BATCH, BF_SHADE_ADJOINT>), grid, block, 0, s, *P, queue, count_ptr, order, route,
                           sample_counts);
    else if (mode == BF_SHADE_TEXTURED)
        hipLaunchKernelGGL((k_shade<BATCH, BF_SHADE_TEXTURED>), grid, block, 0, s, *P, queue, count_ptr, order, route,
                           sample_counts);
    else if (mode == BF_SHADE_MAPPED)
        hipLaunchKernelGGL((k_shade<BATCH, BF_SHADE_MAPPED>), grid, block, 0, s, *P, queue, count_ptr, order, route,
                           sample_counts);
    else
        hipLaunchKernelGGL((k_shade<BATCH, BF_SHADE_PLAIN>), grid, block, 0, s, *P, queue, count_ptr, order, route,
                           sample_counts);
}

template <bool BATCH, uint32_t KEY>
static void launch_shade_order(const bf_transport_params* P, const uint32_t* queue, const uint32_t* count_ptr,
                               uint32_t max_count, uint8_t* keys, uint32_t* tile_counts, uint32_t* tile_offsets,
                               uint32_t* order, hipStream_t s) {
    const uint32_t tiles = bfk_compact_tiles(max_count);
    const dim3 grid(tiles < 4096u ? tiles : 4096u);
    hipLaunchKernelGGL((k_bucket_count<BATCH, KEY>), grid, dim3(BF_BLOCK), 0, s, *P, queue, count_ptr, keys, tile_counts);
    hipLaunchKernelGGL(k_bucket_scan, dim3(1), dim3(1024), 0, s, count_ptr, tile_counts, tile_offsets);
    hipLaunchKernelGGL(k_bucket_scatter, grid, dim3(BF_BLOCK), 0, s, keys, count_ptr, tile_offsets, order);
}

extern "C" {

hipError_t bfk_shade(const bf_transport_params* P, const uint32_t* queue, const uint32_t* count_ptr, uint32_t max_count,
                     uint8_t* route, uint32_t* sample_counts, bool batch, uint32_t mode, hipStream_t s) {
    if (max_count == 0) return hipSuccess;
    const uint32_t* order = nullptr;
    if (P->shade_order) {
        /* the route bytes of the hit queue are free until shade writes them: they hold the keys */
        if (batch)
            launch_shade_order<true, BF_KEY_SHADE>(P, queue, count_ptr, max_count, route, P->bucket_counts,
                                                   P->bucket_offsets, P->shade_order, s);
        else
            launch_shade_order<false, BF_KEY_SHADE>(P, queue, count_ptr, max_count, route, P->bucket_counts,
                                                    P->bucket_offsets, P->shade_order, s);
        order = P->shade_order;
    }
    const uint32_t needed = blocks_for(max_count, BF_SHADE_BLOCK);
    const uint32_t resident = mode == BF_SHADE_ADJOINT ? P->adjoint_shade_blocks : P->shade_blocks;
    const dim3 grid(needed < resident ? needed : resident);
    if (batch) launch_shade<true>(P, grid, queue, count_ptr, order, route, sample_counts, mode, s);
    else launch_shade<false>(P, grid, queue, count_ptr, order, route, sample_counts, mode, s);
    return hipGetLastError();
}

uint32_t bfk_compact_tiles(uint32_t max_count) { return blocks_for(max_count ? max_count : 1u, BF_BLOCK); }

hipError_t bfk_compact(const uint8_t* route, const uint32_t* count_ptr, uint32_t max_count, const uint32_t* src,
                       uint32_t* tile_counts, uint32_t* tile_offsets, uint32_t* dst1, uint32_t* size1, uint32_t append1,
                       uint32_t* dst2, uint32_t* size2, uint32_t append2, hipStream_t s) {
    const uint32_t tiles = bfk_compact_tiles(max_count);
    const dim3 grid(tiles < 4096u ? tiles : 4096u);
    hipLaunchKernelGGL(k_compact_count, grid, dim3(BF_BLOCK), 0, s, route, count_ptr, tile_counts);
    hipLaunchKernelGGL(k_compact_scan, dim3(1), dim3(1024), 0, s, count_ptr, tile_counts, tile_offsets, size1, size2,
                       append1, append2);
    hipLaunchKernelGGL(k_compact_scatter, grid, dim3(BF_BLOCK), 0, s, route, count_ptr, src, tile_offsets, dst1, dst2);
    return hipGetLastError();
}

hipError_t bfk_bounce_end(uint32_t* sizes, uint32_t* history, uint32_t history_bounces, hipStream_t s) {
    hipLaunchKernelGGL(k_bounce_end, dim3(1), dim3(64), 0, s, sizes, history, history_bounces);
    return hipGetLastError();
}

hipError_t bfk_film(const bf_dev_paths* paths, const bf_dev_job* job, float pdf, float normalization, hipStream_t s) {
    hipLaunchKernelGGL(k_film, dim3(blocks_for(job->pixels, BF_BLOCK)), dim3(BF_BLOCK), 0, s, *paths, *job, pdf,
                       normalization);
    return hipGetLastError();
}

hipError_t bfk_queue_hash(const uint32_t* queue, const uint32_t* count_ptr, unsigned long long* out, hipStream_t s) {
    hipLaunchKernelGGL(k_queue_hash, dim3(1), dim3(64), 0, s, queue, count_ptr, out);
    return hipGetLastError();
}

hipError_t bfk_scene_closest_hit(const bf_dev_scene* S, const bf_transport_ray* rays, uint32_t count,
                                 bf_closest_hit_result* results, uint32_t* counters, hipStream_t s) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(k_scene_closest_hit, dim3(blocks_for(count, BF_API_BLOCK)), dim3(BF_API_BLOCK), 0, s, *S, rays,
                       count, results, counters);
    return hipGetLastError();
}

hipError_t bfk_blob_traverse(bool any, const uint8_t* scene_bytes, size_t scene_size, const uint8_t* bvh_bytes,
                             size_t bvh_size, const bf_transport_ray* rays, uint32_t count, void* results, hipStream_t s) {
    if (count == 0) return hipSuccess;
    const dim3 grid(blocks_for(count, BF_API_BLOCK)), block(BF_API_BLOCK);
    if (any)
        hipLaunchKernelGGL(k_blob_traverse<true>, grid, block, 0, s, scene_bytes, scene_size, bvh_bytes, bvh_size, rays,
                           count, results);
    else
        hipLaunchKernelGGL(k_blob_traverse<false>, grid, block, 0, s, scene_bytes, scene_size, bvh_bytes, bvh_size, rays,
                           count, results);
    return hipGetLastError();
}

hipError_t bfk_scene_occlusion(const bf_dev_scene* S, const bf_transport_ray* rays, uint32_t count,
                               bf_occlusion_result* results, hipStream_t s) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(k_scene_occlusion, dim3(blocks_for(count, BF_API_BLOCK)), dim3(BF_API_BLOCK), 0, s, *S, rays,
                       count, results);
    return hipGetLastError();
}

hipError_t bfk_sample_dump(uint32_t kind, const uint64_t* directions, uint32_t dims, const bf_sample_stream_index* indices,
                           const uint64_t* dimensions, uint32_t count, float* out, hipStream_t s) {
    if (count == 0) return hipSuccess;
    sampler_ctx sc;
    sc.kind = kind;
    sc.sobol_dims = dims;
    sc.directions = directions;
    hipLaunchKernelGGL(k_sample_dump, dim3(blocks_for(count, BF_BLOCK)), dim3(BF_BLOCK), 0, s, sc, indices, dimensions, count,
                       out);
    return hipGetLastError();
}

hipError_t bfk_sample_stream_dump(const bf_sample_dump_request* requests, uint32_t count,
                                  bf_sample_dimension_manifest* manifest, bf_sample_dump_result* results) {
    hipLaunchKernelGGL(k_sample_stream_dump, dim3(blocks_for(count, BF_BLOCK)), dim3(BF_BLOCK), 0, nullptr, requests, count,
                       manifest, results);
    return hipGetLastError();
}

hipError_t bfk_reduce_replicas(const float* replicas, uint32_t offset, uint32_t count, float* dst, hipStream_t s) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(k_reduce_replicas, dim3(blocks_for(count, BF_BLOCK)), dim3(BF_BLOCK), 0, s, replicas, offset, count, dst);
    return hipGetLastError();
}

hipError_t bfk_adam(float* value, const float* gradient, float* first, float* second, uint64_t count,
                    const bfk_adam_coefficients* c, const float2* schedule, const uint32_t* step, hipStream_t s) {
    if (count == 0) return hipSuccess;
    const bool vec = ((reinterpret_cast<uintptr_t>(value) | reinterpret_cast<uintptr_t>(gradient) |
                       reinterpret_cast<uintptr_t>(first) | reinterpret_cast<uintptr_t>(second)) & 15u) == 0;
    const uint64_t work = vec ? (count + 3u) / 4u : count;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((work + BF_BLOCK - 1u) / BF_BLOCK, 16384u);
    if (vec)
        hipLaunchKernelGGL(k_adam<true>, dim3(blocks), dim3(BF_BLOCK), 0, s, value, gradient, first, second, count, *c,
                           schedule, step);
    else
        hipLaunchKernelGGL(k_adam<false>, dim3(blocks), dim3(BF_BLOCK), 0, s, value, gradient, first, second, count, *c,
                           schedule, step);
    return hipGetLastError();
}

hipError_t bfk_film_l2(const float* film, const uint32_t* counts, const float* target, float* gradient, uint32_t pixels,
                       float two_scale, double scale, double* loss, hipStream_t s) {
    if (pixels == 0) return hipSuccess;
    hipLaunchKernelGGL(k_film_l2, dim3(blocks_for(pixels, BF_BLOCK)), dim3(BF_BLOCK), 0, s, film, counts, target, gradient,
                       pixels, two_scale, scale, loss);
    return hipGetLastError();
}

hipError_t bfk_zero_doubles(double* dst, uint32_t count, hipStream_t s) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(k_zero_doubles, dim3(blocks_for(count, BF_BLOCK)), dim3(BF_BLOCK), 0, s, dst, count);
    return hipGetLastError();
}

hipError_t bfk_scene_soa_validate(const uint8_t* bytes, size_t count, uint32_t* status, hipStream_t s) {
    hipLaunchKernelGGL(k_scene_soa_validate, dim3(1), dim3(64), 0, s, (const bf_scene_soa_header*)bytes, (uint64_t)count, status);
    return hipGetLastError();
}

hipError_t bfk_scene_hash(const uint8_t* bytes, size_t count, uint64_t* out, hipStream_t s) {
    hipLaunchKernelGGL(k_scene_hash, dim3(1), dim3(64), 0, s, bytes, count, out);
    return hipGetLastError();
}

hipError_t bfk_queue_pushes(const bf_queue_soa* queues, const bf_queue_push* requests, uint32_t n, uint32_t* outcomes,
                            hipStream_t s) {
    if (n == 0u) return hipSuccess;
    hipLaunchKernelGGL(k_queue_push, dim3(blocks_for(n, BF_BLOCK)), dim3(BF_BLOCK), 0, s, *queues, requests, n, outcomes);
    return hipGetLastError();
}

hipError_t bfk_layout_manifest(bf_layout_manifest* out, hipStream_t s) {
    hipLaunchKernelGGL(k_layout_manifest, dim3(1), dim3(64), 0, s, out);
    return hipGetLastError();
}

static size_t align_scratch(size_t bytes) {
    return (bytes + BF_COMPACTION_SCRATCH_ALIGNMENT - 1u) / BF_COMPACTION_SCRATCH_ALIGNMENT * BF_COMPACTION_SCRATCH_ALIGNMENT;
}

size_t bfk_queue_compaction_scratch(uint32_t n) {
    if (n == 0u) return 0u;
    const size_t tiles = (n + BF_BLOCK - 1u) / BF_BLOCK;
    return 2u * align_scratch(tiles * sizeof(uint32_t)) + BF_COMPACTION_SCRATCH_ALIGNMENT;
}

hipError_t bfk_queue_compaction(const bf_queue_soa* queues, const bf_stage_outcome* outcomes, uint32_t n, uint32_t route,
                                void* scratch, bf_compaction_result* result, hipStream_t s) {
    const bf_queue_soa q = *queues;
    hipLaunchKernelGGL(k_qc_init, dim3(1), dim3(64), 0, s, q, route, n, result);
    if (n == 0u) return hipGetLastError();
    const uint32_t tiles = (n + BF_BLOCK - 1u) / BF_BLOCK;
    const size_t part = align_scratch((size_t)tiles * sizeof(uint32_t));
    uint8_t* base = static_cast<uint8_t*>(scratch);
    uint32_t* tile_counts = reinterpret_cast<uint32_t*>(base);
    uint32_t* tile_offsets = reinterpret_cast<uint32_t*>(base + part);
    uint32_t* total = reinterpret_cast<uint32_t*>(base + 2u * part);
    const dim3 grid(tiles < 4096u ? tiles : 4096u);
    hipLaunchKernelGGL(k_qc_validate, dim3(blocks_for(n, BF_BLOCK)), dim3(BF_BLOCK), 0, s, outcomes, n, route, q.slot_stride,
                       result);
    hipLaunchKernelGGL(k_qc_count, grid, dim3(BF_BLOCK), 0, s, outcomes, n, route, tile_counts);
    hipLaunchKernelGGL(k_qc_scan, dim3(1), dim3(1024), 0, s, tile_counts, tiles, tile_offsets, total);
    hipLaunchKernelGGL(k_qc_decide, dim3(1), dim3(64), 0, s, q, total, result);
    hipLaunchKernelGGL(k_qc_scatter, grid, dim3(BF_BLOCK), 0, s, q, outcomes, n, route, tile_offsets, result);
    hipLaunchKernelGGL(k_qc_publish, dim3(1), dim3(64), 0, s, q, result);
    return hipGetLastError();
}

/*
 * Library-owned compaction scratch for callers that cannot allocate: one copy per device, held
 * until the process ends. Its size serves BF_STATIC_COMPACTION_MAX_INPUT outcomes. In host code
 * `scratch` names a host-side stand-in, never the device array: every host use goes through the
 * address hipGetSymbolAddress returned for the current device. Nothing in it must start at zero
 * (the count kernel writes every tile it uses, the scan writes the total), and every call on a
 * device shares it, so calls that overlap on one device race.
 */
#define BF_STATIC_COMPACTION_SCRATCH_BYTES                                                                  \
    (2u * (((BF_STATIC_COMPACTION_MAX_INPUT / BF_BLOCK) * 4u + BF_COMPACTION_SCRATCH_ALIGNMENT - 1u) /     \
           BF_COMPACTION_SCRATCH_ALIGNMENT * BF_COMPACTION_SCRATCH_ALIGNMENT) +                             \
     BF_COMPACTION_SCRATCH_ALIGNMENT)
static __device__ __align__(256) uint8_t scratch[BF_STATIC_COMPACTION_SCRATCH_BYTES];

namespace {
constexpr int kStaticScratchDevices = 64;
std::mutex g_static_scratch_lock;
void* g_static_scratch[kStaticScratchDevices] = {};
} // namespace

hipError_t bfk_static_compaction_scratch(void** out) {
    *out = nullptr;
    int device = 0;
    hipError_t e = hipGetDevice(&device);
    if (e != hipSuccess) return e;
    if (device < 0 || device >= kStaticScratchDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> guard(g_static_scratch_lock);
    if (g_static_scratch[device] == nullptr) {
        void* ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_static_scratch[device] = ptr;
    }
    *out = g_static_scratch[device];
    return hipSuccess;
}

static_assert(BF_STATIC_COMPACTION_MAX_INPUT % BF_BLOCK == 0u, "the static scratch covers whole tiles");

hipError_t bfk_set_words(uint32_t* dst, uint32_t count, uint32_t value, hipStream_t s) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(k_set_words, dim3(blocks_for(count, BF_BLOCK)), dim3(BF_BLOCK), 0, s, dst, count, value);
    return hipGetLastError();
}

hipError_t bfk_copy_words(const uint32_t* src, uint32_t* dst, uint32_t count, hipStream_t s) {
    hipLaunchKernelGGL(k_copy_words, dim3(1), dim3(64), 0, s, src, dst, count);
    return hipGetLastError();
}

hipError_t bfk_tick(uint32_t* counter, hipStream_t s) {
    hipLaunchKernelGGL(k_tick, dim3(1), dim3(64), 0, s, counter);
    return hipGetLastError();
}

hipError_t bfk_history_totals(const uint32_t* history, uint32_t rows, unsigned long long* totals, hipStream_t s) {
    hipLaunchKernelGGL(k_history_totals, dim3(1), dim3(64), 0, s, history, rows, totals);
    return hipGetLastError();
}

hipError_t bfk_copy_word(const uint32_t* src, uint32_t* dst, hipStream_t s) {
    hipLaunchKernelGGL(k_copy_word, dim3(1), dim3(64), 0, s, src, dst);
    return hipGetLastError();
}

hipError_t bfk_export_paths(const bf_dev_paths* paths, const bf_dev_job* job, hipStream_t s) {
    if (job->slot_count == 0 || !job->paths) return hipSuccess;
    hipLaunchKernelGGL(k_export_paths, dim3(blocks_for(job->slot_count, BF_BLOCK)), dim3(BF_BLOCK), 0, s, *paths, *job);
    return hipGetLastError();
}

hipError_t bfk_audit_stage(const bf_stage_outcome* outcomes, uint32_t work_count, uint32_t allowed_route_mask,
                           uint32_t path_capacity, uint32_t stage_kind, bf_stage_audit* audit, hipStream_t s) {
    hipLaunchKernelGGL(k_audit_init, dim3(1), dim3(64), 0, s, work_count, stage_kind, audit);
    if (work_count == 0u) return hipGetLastError();
    const uint32_t blocks = std::min<uint32_t>(blocks_for(work_count, BF_BLOCK), 2048u);
    hipLaunchKernelGGL(k_audit_reduce, dim3(blocks), dim3(BF_BLOCK), 0, s, outcomes, work_count, allowed_route_mask,
                       path_capacity, stage_kind == BF_STAGE_KIND_SHADE ? 1u : 0u, audit);
    hipLaunchKernelGGL(k_audit_close, dim3(1), dim3(64), 0, s, outcomes, work_count, audit);
    return hipGetLastError();
}

hipError_t bfk_export_path_records(const bf_dev_paths* paths, const bf_dev_job* job, hipStream_t s) {
    if (job->slot_count == 0 || !job->out_states) return hipSuccess;
    hipLaunchKernelGGL(k_export_path_records, dim3(blocks_for(job->slot_count, BF_BLOCK)), dim3(BF_BLOCK), 0, s, *paths,
                       *job);
    return hipGetLastError();
}

} /* extern "C" */
