/*
 * Kernel-side parameter block and the stream-ordered launch helpers exported by bf_kernels.hip
 * to the host orchestration (bf_host.cpp). Launch helpers never allocate and never synchronise.
 */
#ifndef BF_KERNELS_H
#define BF_KERNELS_H

#include <hip/hip_runtime.h>

#include "bf_device_math.h"
#include "bf_device_scene.h"

/*
 * One frame of a launch (a batch holds one job per scene). Within a launch the job owns the
 * workspace slots [slot_begin, slot_begin + slot_count): whole sample passes, path slot
 * slot_begin + k being path path_base + k of its frame (path = pass * pixels + pixel).
 */
typedef struct bf_dev_job {
    bf_dev_scene scene;
    bf_camera_desc camera;
    uint64_t seed;
    uint64_t path_base;
    uint32_t slot_begin;
    uint32_t slot_count;
    uint32_t first_sample_index;
    uint32_t pixels;
    uint32_t index;             /* position in the batch (failure reports) */
    uint32_t reserved;
    float* film;                /* per-frame film buffers (see blackframe_hip.h) */
    double* film_compensated;
    uint32_t* film_counts;
    bf_path_result* paths;      /* optional per-path results of the whole frame */
    float4* tape_radiance;      /* optional forward tape / adjoint input, frame path order */
    unsigned long long* tape_visibility;
    const float* film_gradient; /* adjoint: dLoss/d(film RGB sums) */
    float* texture_gradients[BF_MAX_TEXTURES];
    /* offset of texture k's gradient in each replica row, or BF_INVALID_INDEX: accumulate into
     * texture_gradients[k] directly */
    uint32_t gradient_replica[BF_MAX_TEXTURES];
    float* tape_survival;       /* roulette survival probabilities [path * survival_stride + vertex] */
    uint32_t survival_stride;
    uint32_t survival_mode;     /* bf_survival_mode */
    uint32_t frame_cones;       /* camera frames of mapped scenes: the camera stage starts each path's
                                 * cone, the shade stage carries it */
    /* path-input frames (blackframe_hip_trace_wavefront_paths): the inputs replace the camera, the
     * outputs receive the terminal records; all null for camera frames */
    const bf_sample_stream_index* in_samples;
    const bf_transport_ray* in_rays;
    const bf_ray_cone* in_cones;
    const bf_path_state_lane* in_states;
    bf_path_state_lane* out_states;
    bf_transport_ray* out_rays;
    bf_ray_cone* out_cones;
    bf_lane_control* out_controls;
} bf_dev_job;

/* depth.z bit: the slot's ray is still the path's first ray, whose t_max is the camera's or the
 * input ray's (every later ray runs to infinity) */
#define BF_FIRST_RAY 0x80000000u

/* Gradients of small textures (<= BF_GRADIENT_SMALL floats, e.g. 1x1 per-material scalars) are
 * accumulated into BF_GRADIENT_REPLICAS copies, one chosen per lane, and summed into the
 * caller's buffer at the end of the call: same-address atomics from every lane would serialise. */
#define BF_GRADIENT_REPLICAS 256u
#define BF_GRADIENT_SMALL 4096u
#define BF_GRADIENT_ROW 16384u

#define BF_SHADE_PLAIN 0u
#define BF_SHADE_TEXTURED 1u
#define BF_SHADE_ADJOINT 2u
#define BF_SHADE_MAPPED 3u /* texture lookups and normal / bump maps compiled in */

typedef struct bf_transport_params {
    bf_dev_job job0;            /* the job of a single-frame launch (kernel argument, no gather) */
    const bf_dev_job* jobs;     /* the job table of a batched launch (device memory) */
    uint32_t job_count;
    bf_dev_paths paths;
    sampler_ctx sampler;
    uint32_t mis;
    uint32_t wavelength_mode;
    uint32_t limits[5];
    uint32_t rr_enabled;
    uint32_t rr_first;
    float rr_min;
    float rr_max;
    float packet[4];
    uint32_t* failure;          /* [0] failed path count, [2..3] lowest (job << 48 | path << 8 | code) */
    uint2* spill;               /* traversal stack spill: BF_STACK_DEPTH entries per persistent lane */
    uint32_t spill_stride;      /* persistent lanes = resident traversal workgroups * BF_TRACE_BLOCK */
    uint32_t shade_blocks;      /* resident forward shade workgroups (persistent grid) */
    uint32_t adjoint_shade_blocks; /* resident adjoint shade workgroups */
    uint32_t* fetch;            /* traversal work counters [closest, shadow], cleared before each launch */
    uint32_t trace_block;       /* lanes per traversal workgroup: 64, 128 or 256 */
    float film_pdf;             /* wavelength pdf and sensor normalisation of the film (adjoint) */
    float film_normalization;
    uint32_t shade_mode;        /* BF_SHADE_* of this launch (host bookkeeping: part of the graph-cache key) */
    float* gradient_replicas;   /* BF_GRADIENT_REPLICAS rows of BF_GRADIENT_ROW floats */
    uint32_t* shade_order;      /* non-null: shade visits the hit queue grouped by closure kind (a free queue) */
    uint32_t* closest_order;    /* non-null: traversal claims rays grouped by direction octant (free queues) */
    uint32_t* shadow_order;
    uint32_t* bucket_counts;    /* per-tile bucket counts / offsets of these groupings */
    uint32_t* bucket_offsets;
} bf_transport_params;

#ifdef __cplusplus
extern "C" {
#endif
/* adjoint: regenerate the primary rays of a taped frame and load each path's tape and film
 * gradient instead of clearing the tape. Path-input frames: every slot's input is validated, the
 * slot written to ray_queue and route[slot] = 1 for a valid input, 0 for a failed one (the caller
 * compacts the queue with these routes). */
hipError_t bfk_camera(const bf_transport_params* P, uint32_t count, uint32_t* ray_queue, uint8_t* route, bool batch,
                      bool adjoint, hipStream_t s);
hipError_t bfk_trace_closest(const bf_transport_params* P, const uint32_t* queue, const uint32_t* count_ptr,
                             uint32_t max_count, uint8_t* route, unsigned long long* counters, bool batch, hipStream_t s);
/* mode BF_SHADE_ADJOINT: shadow visibility from the tape (route 1 is never produced) and texel
 * gradients; BF_SHADE_TEXTURED: texture lookups compiled in; BF_SHADE_PLAIN: untextured scenes */
hipError_t bfk_shade(const bf_transport_params* P, const uint32_t* queue, const uint32_t* count_ptr, uint32_t max_count,
                     uint8_t* route, uint32_t* sample_counts, bool batch, uint32_t mode, hipStream_t s);
hipError_t bfk_trace_shadow(const bf_transport_params* P, const uint32_t* queue, const uint32_t* count_ptr,
                            uint32_t max_count, uint8_t* route, unsigned long long* counters, bool batch, hipStream_t s);
/* 256-item tiles the compaction of max_count items uses (size of the tile count / offset arrays / 2). */
uint32_t bfk_compact_tiles(uint32_t max_count);
hipError_t bfk_compact(const uint8_t* route, const uint32_t* count_ptr, uint32_t max_count, const uint32_t* src,
                       uint32_t* tile_counts, uint32_t* tile_offsets, uint32_t* dst1, uint32_t* size1, uint32_t append1,
                       uint32_t* dst2, uint32_t* size2, uint32_t append2, hipStream_t s);
/* sizes[0..3] -> history[4 * sizes[7]], sizes[0] = sizes[3], sizes[1..5] = 0, ++sizes[7] */
hipError_t bfk_bounce_end(uint32_t* sizes, uint32_t* history, uint32_t history_bounces, hipStream_t s);
hipError_t bfk_film(const bf_dev_paths* paths, const bf_dev_job* job, float pdf, float normalization, hipStream_t s);
/* Workgroups the device holds at once for the traversal / shade kernels, 0 on error. */
uint32_t bfk_trace_resident_lanes(int device);
uint32_t bfk_shade_resident_blocks(int device, bool adjoint);
hipError_t bfk_queue_hash(const uint32_t* queue, const uint32_t* count_ptr, unsigned long long* out, hipStream_t s);
hipError_t bfk_scene_closest_hit(const bf_dev_scene* S, const bf_transport_ray* rays, uint32_t count,
                                 bf_closest_hit_result* results, uint32_t* counters, hipStream_t s);
/* traversal over the reference's device blobs (scene "BLKFSOA4", BVH "BLKFBVH1"); results are
 * bf_occlusion_result when `any`, bf_closest_hit_result otherwise */
hipError_t bfk_blob_traverse(bool any, const uint8_t* scene_bytes, size_t scene_size, const uint8_t* bvh_bytes,
                             size_t bvh_size, const bf_transport_ray* rays, uint32_t count, void* results, hipStream_t s);
hipError_t bfk_scene_occlusion(const bf_dev_scene* S, const bf_transport_ray* rays, uint32_t count,
                               bf_occlusion_result* results, hipStream_t s);
hipError_t bfk_sample_dump(uint32_t kind, const uint64_t* directions, uint32_t dims, const bf_sample_stream_index* indices,
                           const uint64_t* dimensions, uint32_t count, float* out, hipStream_t s);
hipError_t bfk_sample_stream_dump(const bf_sample_dump_request* requests, uint32_t count,
                                  bf_sample_dimension_manifest* manifest, bf_sample_dump_result* results);
hipError_t bfk_set_words(uint32_t* dst, uint32_t count, uint32_t value, hipStream_t s);
hipError_t bfk_copy_word(const uint32_t* src, uint32_t* dst, hipStream_t s);
hipError_t bfk_export_paths(const bf_dev_paths* paths, const bf_dev_job* job, hipStream_t s);
/* stage outcome audit: init + per-workgroup partials + one finalising workgroup */
hipError_t bfk_audit_stage(const bf_stage_outcome* outcomes, uint32_t work_count, uint32_t allowed_route_mask,
                           uint32_t path_capacity, uint32_t stage_kind, bf_stage_audit* audit, hipStream_t s);
/* the terminal records of a path-input frame's slots: path state, ray, cone and lane control */
hipError_t bfk_export_path_records(const bf_dev_paths* paths, const bf_dev_job* job, hipStream_t s);
/* dst[e] += sum over the replica rows of replicas[row * BF_GRADIENT_ROW + offset + e], e < count */
hipError_t bfk_reduce_replicas(const float* replicas, uint32_t offset, uint32_t count, float* dst, hipStream_t s);
/* Per-array coefficients of one Adam step, rounded once on the host. */
typedef struct bfk_adam_coefficients {
    float beta1, one_minus_beta1, beta2, one_minus_beta2, step_size, bias2_sqrt, epsilon, lower, upper;
} bfk_adam_coefficients;
/* schedule (optional): per step number, (step size, sqrt(bias2)) read at schedule[*step] on the device */
hipError_t bfk_adam(float* value, const float* gradient, float* first, float* second, uint64_t count,
                    const bfk_adam_coefficients* c, const float2* schedule, const uint32_t* step, hipStream_t s);
hipError_t bfk_copy_words(const uint32_t* src, uint32_t* dst, uint32_t count, hipStream_t s);
hipError_t bfk_tick(uint32_t* counter, hipStream_t s);
/* totals[0..2] += closest-hit, shade, shadow lanes of history rows; totals[3] = max(active bounces) */
hipError_t bfk_history_totals(const uint32_t* history, uint32_t rows, unsigned long long* totals, hipStream_t s);
/* loss[0] += scale * sum d^2 and gradient = (two_scale * d) / max(count, 1), d = film / max(count, 1) - target */
hipError_t bfk_film_l2(const float* film, const uint32_t* counts, const float* target, float* gradient, uint32_t pixels,
                       float two_scale, double scale, double* loss, hipStream_t s);
hipError_t bfk_zero_doubles(double* dst, uint32_t count, hipStream_t s);
/* Scratch bytes of bfk_queue_compaction for n outcomes, and the compaction itself (arguments
 * validated by the caller; scratch at least bfk_queue_compaction_scratch(n) bytes, 256-aligned). */
hipError_t bfk_layout_manifest(bf_layout_manifest* out, hipStream_t s);
hipError_t bfk_queue_pushes(const bf_queue_soa* queues, const bf_queue_push* requests, uint32_t n, uint32_t* outcomes,
                            hipStream_t s);
hipError_t bfk_scene_hash(const uint8_t* bytes, size_t count, uint64_t* out, hipStream_t s);
hipError_t bfk_scene_soa_validate(const uint8_t* bytes, size_t count, uint32_t* status, hipStream_t s);
size_t bfk_queue_compaction_scratch(uint32_t n);
hipError_t bfk_queue_compaction(const bf_queue_soa* queues, const bf_stage_outcome* outcomes, uint32_t n, uint32_t route,
                                void* scratch, bf_compaction_result* result, hipStream_t s);
/* Device address, on the current device, of the library-owned compaction scratch (at least
 * bfk_queue_compaction_scratch(BF_STATIC_COMPACTION_MAX_INPUT) bytes, 256-aligned); looked up once
 * per device with hipGetSymbolAddress. */
hipError_t bfk_static_compaction_scratch(void** out);
#ifdef __cplusplus
}
#endif

#endif
