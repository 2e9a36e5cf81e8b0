/*
 * Kernel-side parameter block and the stream-ordered launch helpers exported by bf_kernels.hip
 * to the host orchestration (bf_host.cpp). Launch helpers never allocate and never synchronise.
 */
#ifndef BF_KERNELS_H
#define BF_KERNELS_H

#include <hip/hip_runtime.h>

#include "bf_device_math.h"
#include "bf_device_scene.h"

typedef struct bf_transport_params {
    bf_dev_scene scene;
    bf_dev_paths paths;
    sampler_ctx sampler;
    bf_camera_desc camera;
    uint64_t seed;
    uint64_t slot_base;         /* global path index of slot 0 */
    uint32_t first_sample_index;
    uint32_t pixels;            /* camera width * height */
    uint32_t mis;
    uint32_t wavelength_mode;
    uint32_t limits[5];
    uint32_t rr_enabled;
    uint32_t rr_first;
    float rr_min;
    float rr_max;
    float packet[4];
    uint32_t* failure;          /* [0] failed path count, [2..3] lowest (global slot << 8 | code) */
    uint2* spill;               /* traversal stack spill: BF_STACK_DEPTH entries per persistent lane */
    uint32_t spill_stride;      /* persistent lanes = resident traversal workgroups * BF_TRACE_BLOCK */
} bf_transport_params;

#ifdef __cplusplus
extern "C" {
#endif
hipError_t bfk_camera(const bf_transport_params* P, uint32_t count, uint32_t* ray_queue, hipStream_t s);
hipError_t bfk_trace_closest(const bf_transport_params* P, const uint32_t* queue, const uint32_t* count_ptr,
                             uint32_t max_count, uint8_t* route, unsigned long long* counters, hipStream_t s);
hipError_t bfk_shade(const bf_transport_params* P, const uint32_t* queue, const uint32_t* count_ptr, uint32_t max_count,
                     uint8_t* route, uint32_t* sample_counts, hipStream_t s);
hipError_t bfk_trace_shadow(const bf_transport_params* P, const uint32_t* queue, const uint32_t* count_ptr,
                            uint32_t max_count, uint8_t* route, unsigned long long* counters, hipStream_t s);
uint32_t bfk_compact_blocks(uint32_t max_count);
/* Traversal workgroups the device holds at once (all traversal kernels), 0 on error. */
uint32_t bfk_trace_resident_blocks(int device);
hipError_t bfk_compact(const uint8_t* route, const uint32_t* count_ptr, uint32_t max_count, const uint32_t* src,
                       uint32_t* block_counts, uint32_t* block_offsets, uint32_t* dst1, uint32_t* size1, uint32_t append1,
                       uint32_t* dst2, uint32_t* size2, uint32_t append2, hipStream_t s);
hipError_t bfk_film(const bf_dev_paths* paths, uint32_t pixels, uint32_t passes, float pdf, float normalization,
                    float* film, double* film_compensated, uint32_t* film_counts, hipStream_t s);
hipError_t bfk_queue_hash(const uint32_t* queue, const uint32_t* count_ptr, unsigned long long* out, hipStream_t s);
hipError_t bfk_scene_closest_hit(const bf_dev_scene* S, const bf_transport_ray* rays, uint32_t count,
                                 bf_closest_hit_result* results, uint32_t* counters, hipStream_t s);
hipError_t bfk_scene_occlusion(const bf_dev_scene* S, const bf_transport_ray* rays, uint32_t count,
                               bf_occlusion_result* results, hipStream_t s);
hipError_t bfk_sample_dump(uint32_t kind, const uint64_t* directions, uint32_t dims, const bf_sample_stream_index* indices,
                           const uint64_t* dimensions, uint32_t count, float* out, hipStream_t s);
hipError_t bfk_set_words(uint32_t* dst, uint32_t count, uint32_t value, hipStream_t s);
hipError_t bfk_copy_word(const uint32_t* src, uint32_t* dst, hipStream_t s);
hipError_t bfk_export_paths(const bf_dev_paths* paths, uint32_t count, bf_path_result* out, hipStream_t s);
#ifdef __cplusplus
}
#endif

#endif
