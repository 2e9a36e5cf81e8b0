/*
 * Blackframe XPU/HIP — C ABI of the MI355X wavefront path-tracing backend (libblackframe_hip.so).
 *
 * Conventions (the reference launcher ABI's, Blackframe/XPU/CUDA/Include/Blackframe/XPU/CUDA/
 * WavefrontStageKernel.hpp:324-401 and SceneClosestHit.hpp:1-17):
 *  - every entry point returns an int status (enum bf_status, 0 = success); invalid arguments are
 *    rejected synchronously before any launch;
 *  - `stream` is an opaque hipStream_t (NULL = the legacy default stream);
 *  - launchers never allocate and never synchronise; the caller owns completion;
 *  - pointers documented "device" name storage on the scene's device, "host" name host memory;
 *  - there is no fallback: an unsupported request is an error.
 * No C++ types cross this boundary.
 */
#ifndef BLACKFRAME_HIP_H
#define BLACKFRAME_HIP_H

#include "blackframe_hip_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bf_hip_scene bf_hip_scene;
typedef struct bf_hip_workspace bf_hip_workspace;

/* Human-readable text of the last error raised on this thread (never NULL). */
const char* blackframe_hip_last_error(void);

/* Number of visible HIP devices (0 when the runtime reports none).
 * Replaces: blackframe_xpu_backend_api_v1::get_device_count (XpuBackendAbi.h:45-55). */
int blackframe_hip_device_count(uint32_t* count);

/*
 * Builds the device scene on the current device: the SoA scene upload plus the deterministic
 * BVH (BLAS per geometry + TLAS over instances, median split, leaf size 1) re-laid out as 64-byte
 * wide nodes, and the mesh-area-light tables.
 * Replaces: engine::CudaSceneSoA::upload (Backends/GPU/CUDA/Source/SceneSoA.cpp:1180) followed by
 *           engine::CudaSceneBvh::build (Backends/GPU/CUDA/Source/SceneBvh.cpp:1109-1146).
 */
int blackframe_hip_scene_create(const bf_scene_desc* desc, bf_hip_scene** out_scene);

/* Host-only variant: builds the canonical and wide BVH and the light tables without touching a
 * device (for inspection and export; such a scene cannot be traced). */
int blackframe_hip_scene_prepare(const bf_scene_desc* desc, bf_hip_scene** out_scene);
int blackframe_hip_scene_destroy(bf_hip_scene* scene);

/*
 * Copies the canonical (reference-layout, 48-byte node) BVH arrays to host memory. counts
 * receives {blas, blas_node, primitive_reference, tlas_node, instance_reference}; pass NULL
 * arrays to query the counts only. The serialized form of the reference is SceneBvhAbi.hpp:37-98.
 */
int blackframe_hip_scene_bvh_export(const bf_hip_scene* scene, uint64_t counts[5], bf_bvh_blas* blases,
                                    bf_bvh_node* blas_nodes, uint32_t* primitive_refs, bf_bvh_node* tlas_nodes,
                                    bf_bvh_instance_ref* instance_refs, uint32_t* tlas_root);

/*
 * Writes the scene's canonical BVH as the reference's serialized "BLKFBVH1" blob (host memory):
 * header, the five arrays at 16-byte aligned offsets, content hash. source_scene_hash is the
 * content hash of the scene blob for scenes made by blackframe_hip_scene_create_from_soa, 0 for
 * scenes made from a bf_scene_desc. *size receives the blob's byte count; out == NULL queries it
 * only, and a capacity below it returns BF_RESOURCE_EXHAUSTED. Host only.
 * Replaces: the bytes engine::CudaSceneBvh::build serializes and uploads (Backends/GPU/CUDA/Source/
 *           SceneBvh.cpp:996-1071, 1109-1146; layout SceneBvhAbi.hpp:37-98).
 */
int blackframe_hip_scene_bvh_serialize(const bf_hip_scene* scene, void* out, size_t capacity, size_t* size);

/*
 * Checks a host "BLKFBVH1" blob: validate_scene_bvh_header's checks in its order, then (no
 * reference counterpart) that `byte_count` covers total_size_bytes and, with BF_BVH_VERIFY_HASH, the
 * content hash. *status receives a bf_bvh_status; the call returns BF_OK whenever it could decide.
 * Replaces: validate_scene_bvh_header (XPU/Shared/Include/Blackframe/XPU/Shared/SceneBvhAbi.hpp:146-229).
 */
int blackframe_hip_scene_bvh_validate(const void* blob, size_t byte_count, uint32_t flags, uint32_t* status);

/*
 * The power-weighted light selection CDF of `count` lights from their four-lane spectral powers
 * (host, [count][4] floats, registry order: punctual lights, then mesh-area lights): lanes scaled by
 * the global peak and summed per light, compensated binary32 prefix sums over the compensated
 * total, the last light with power closing at exactly 1, lights without power repeating the
 * previous boundary. `boundaries` (host) receives count + 1 floats; light i is selected with
 * probability boundaries[i + 1] - boundaries[i]. Rejections carry the reference's status classes
 * (no lights, black or non-finite powers: BF_INVALID_ARGUMENT; unrepresentable weights or
 * intervals: BF_RESOURCE_EXHAUSTED). Host only.
 * Replaces: renderer::LightSampler::create_power_weighted (Renderer/Include/Blackframe/Renderer/
 *           LightSampler.hpp:112-113; Renderer/Source/LightSampler.cpp:116-249, 605-624).
 */
int blackframe_hip_light_power_cdf(const float* spectral_powers, uint32_t count, float* boundaries);

/*
 * Binds a power-weighted light sampler to a device scene: the transport then selects the light of
 * each light sample from the CDF above (LightSampler::sample, LightSampler.cpp:1023-1057) and
 * weighs emissive hits with that light's selection probability (::probability, :1133-1151; an
 * emitter without power fails its path there, as the reference's scalar MIS does). count must be
 * the scene's registered light count; spectral_powers == NULL returns the scene to uniform
 * selection. Synchronises the device when it replaces a bound table.
 * Replaces: the CudaWavefrontLightSampler argument of trace_cuda_wavefront_transport
 *           (WavefrontTransport.hpp:160-161, 197-208), whose CUDA path accepts only the uniform
 *           strategy (WavefrontTransport.cpp:575-630).
 */
int blackframe_hip_scene_set_light_powers(bf_hip_scene* scene, const float* spectral_powers, uint32_t count);

/* Copies the device-resident 64-byte wide nodes (host memory): counts = {blas wide, tlas wide}. */
int blackframe_hip_scene_wide_export(const bf_hip_scene* scene, uint64_t counts[2], bf_wide_node* blas_wide,
                                     bf_wide_node* tlas_wide);

/*
 * Closest-hit traversal of `count` device rays into device results; optional device `counters`
 * receives per ray {canonical nodes visited, triangle tests, instances entered}.
 * Replaces: blackframe_cuda_launch_scene_closest_hit (XPU/CUDA/Include/Blackframe/XPU/CUDA/SceneClosestHit.hpp).
 */
int blackframe_hip_launch_scene_closest_hit(const bf_hip_scene* scene, const bf_transport_ray* rays, uint32_t count,
                                            bf_closest_hit_result* results, uint32_t* counters, void* stream);

/* Any-hit traversal. Replaces: blackframe_cuda_launch_scene_occlusion (SceneOcclusion.hpp). */
int blackframe_hip_launch_scene_occlusion(const bf_hip_scene* scene, const bf_transport_ray* rays, uint32_t count,
                                          bf_occlusion_result* results, void* stream);

/*
 * The same traversal in the reference launcher's form: a "BLKFSOA4" scene blob and the "BLKFBVH1"
 * BVH built from it (its source_scene_hash is the scene's content hash), both in device memory,
 * `ray_count` device rays into device results. Every workgroup first checks both layouts as the
 * reference does (validate_scene_layout / validate_bvh_layout) and answers each of its rays with
 * BF_HIT_INVALID_SCENE or BF_HIT_INVALID_BVH when one fails; each ray then traverses the canonical
 * 48-byte nodes in the reference's order (a 64-entry stack per structure), entering instances
 * through the scene's world-to-local columns and testing triangles from its index and position
 * columns. Results equal blackframe_hip_launch_scene_closest_hit's on the scene the blobs
 * describe. Null pointers and blobs shorter than their headers are rejected before any launch;
 * asynchronous on `stream`, nothing allocated.
 * Replaces: blackframe_cuda_launch_scene_closest_hit (XPU/CUDA/Include/Blackframe/XPU/CUDA/
 *           SceneClosestHit.hpp; Source/CudaSceneClosestHit.cu:188-1242,1287-1308).
 */
int blackframe_hip_launch_scene_closest_hit_blob(const uint8_t* scene_bytes, size_t scene_size, const uint8_t* bvh_bytes,
                                                 size_t bvh_size, const bf_transport_ray* rays, uint32_t ray_count,
                                                 bf_closest_hit_result* results, void* stream);

/* Any-hit form of the blob traversal.
 * Replaces: blackframe_cuda_launch_scene_occlusion (SceneOcclusion.hpp; CudaSceneClosestHit.cu:1243-1286,1310-1330). */
int blackframe_hip_launch_scene_occlusion_blob(const uint8_t* scene_bytes, size_t scene_size, const uint8_t* bvh_bytes,
                                               size_t bvh_size, const bf_transport_ray* rays, uint32_t ray_count,
                                               bf_occlusion_result* results, void* stream);

/*
 * For every device request: the indexed-hash bits and binary32 sample of (sample_stream, dimension)
 * and the ten transport dimensions of bounce_index (status bounce_out_of_range with zero dimensions
 * past BF_SAMPLE_MAXIMUM_BOUNCE); thread 0 also writes the dimension manifest. A schema other than
 * BF_SAMPLE_DIMENSION_SCHEMA, a zero count or a null pointer returns BF_INVALID_ARGUMENT (the value
 * of the reference's cudaErrorInvalidValue) before any launch. Default stream, asynchronous.
 * Replaces: blackframe_cuda_launch_sample_stream_dump (XPU/CUDA/Include/Blackframe/XPU/CUDA/
 *           SampleStreamKernel.hpp:170-176, Source/CudaSampleStream.cu:23-63).
 */
int blackframe_hip_launch_sample_stream_dump(uint32_t requested_schema_version, const bf_sample_dump_request* requests,
                                             uint32_t request_count, bf_sample_dimension_manifest* manifest,
                                             bf_sample_dump_result* results);

/*
 * sample_1d of `count` (index, dimension) pairs with either sampler (bf_sampler_kind: the
 * reference's indexed hash, or the Sobol-Owen sequence the transport can use instead) into device
 * floats, on `stream`. No reference counterpart: the reference dumps the indexed hash only.
 */
int blackframe_hip_launch_sample_dump(uint32_t sampler, const bf_sample_stream_index* indices, const uint64_t* dimensions,
                                      uint32_t count, float* samples, void* stream);

/*
 * Fixed-capacity transport workspace: path-state columns, the four queues, compaction scratch.
 * `path_capacity` paths are in flight per launch (a multiple of the frame's pixel count).
 * Replaces: engine::CudaWavefrontTransportWorkspace::create / close (WavefrontTransport.hpp:152-185).
 */
int blackframe_hip_workspace_create(uint64_t path_capacity, bf_hip_workspace** out_workspace);
int blackframe_hip_workspace_destroy(bf_hip_workspace* workspace);

/*
 * Renders `frame` (spp passes over width*height pixels, path slot s*W*H + y*W + x) through the
 * wavefront: device-side camera rays and wavelengths, closest hit, shade (emission + MIS, NEE,
 * BSDF sample, Russian roulette), shadow rays, stable queue compaction, and device film
 * accumulation in sample order. Film buffers (device) accumulate on top of their contents:
 *   film            float   [W*H*3]   running float sums (Film<float32>)
 *   film_compensated double [W*H*6]  (sum, compensation) pairs per channel (Film<float64>)
 *   film_counts     uint32 [W*H]     samples per pixel
 * `paths` (device, optional) receives one bf_path_result per path slot of the frame.
 * `report` (host, optional) receives lane counts, failures and queue fingerprints; this call
 * synchronises `stream` once per bounce to read the queue sizes.
 * `traversal_counters` (host, optional, 6 x uint64): {closest nodes, closest triangles,
 * closest instances, shadow nodes, shadow triangles, shadow instances} over the whole frame.
 * Replaces: engine::trace_cuda_wavefront_transport (Backends/GPU/CUDA/Include/Blackframe/Backends/
 *           GPU/CUDA/WavefrontTransport.hpp:197-208, Source/WavefrontTransport.cpp:1983-2833).
 */
int blackframe_hip_trace_wavefront_transport(bf_hip_workspace* workspace, const bf_hip_scene* scene,
                                             const bf_frame_desc* frame, const bf_transport_options* options,
                                             float* film, double* film_compensated, uint32_t* film_counts,
                                             bf_path_result* paths, bf_transport_report* report,
                                             uint64_t* traversal_counters, void* stream);

/*
 * The path-input form of the transport: per path a sample stream address, a primary ray, its ray
 * cone and an initial path state (throughput, radiance, four wavelengths with their pdfs and
 * measures, depth counters, eta scale, delta flags) come in as device columns; the camera stage
 * validates and loads them, the bounce loop is the frame transport's, and every path returns its
 * terminal path state (wavelength pdfs and measures passed through), the ray it ended on (its
 * sampled continuation when roulette or zero throughput ended it; its last traced ray otherwise),
 * its ray cone advanced to each hit and scattered (spread floored by the sampled closure's
 * roughness, 1 for diffuse; refraction scales by the eta and cosine ratios) and a lane control
 * (phase terminated, termination, the depth categories that blocked a depth-limited path).
 * Invalid inputs (negative t_min, non-unit wavelength measures, media) fail only their paths,
 * which return zero records and termination NONE; failures are counted in `report` under the
 * options' failure policy. inputs->count == outputs->count <= the workspace capacity; the light
 * sampler is the options' strategy. No film is written.
 * Replaces: engine::trace_cuda_wavefront_transport(workspace, scene, bvh, inputs, light_sampler,
 *           options) (Backends/GPU/CUDA/Include/Blackframe/Backends/GPU/CUDA/WavefrontTransport.hpp:
 *           197-208), whose CudaWavefrontPathInput span and CudaWavefrontTransportBatch (paths and
 *           terminal cones) these columns carry.
 */
int blackframe_hip_trace_wavefront_paths(bf_hip_workspace* workspace, const bf_hip_scene* scene,
                                         const bf_path_inputs* inputs, const bf_transport_options* options,
                                         const bf_path_outputs* outputs, bf_transport_report* report, void* stream);

/*
 * One frame of a batch: a scene (on the current device), its frame description and its device
 * film / optional per-path outputs, with the meaning they have in
 * blackframe_hip_trace_wavefront_transport.
 */
typedef struct bf_frame_job {
    const bf_hip_scene* scene;
    bf_frame_desc frame;
    float* film;
    double* film_compensated;
    uint32_t* film_counts;
    bf_path_result* paths;
    /* Optional tape for blackframe_hip_trace_wavefront_adjoint (both or neither), device memory
     * in frame path order (path = pass * pixels + pixel): */
    float* tape_radiance;      /* [paths * 4] final spectral radiance of each path */
    uint64_t* tape_visibility; /* [paths] bit d: the shadow ray of path vertex d reached its light */
    /* Optional Russian-roulette survival probabilities by path vertex, device memory
     * [paths * (diffuse + glossy + specular depth limits + 1)] floats in frame path order (no
     * reference counterpart). survival_mode BF_SURVIVAL_RECORD writes the probability of every
     * roulette decision; BF_SURVIVAL_REPLAY decides with the recorded probability instead of the
     * one the path's throughput gives (u < p survives, the throughput is divided by p). Replay holds
     * the roulette fixed while texels move, so a finite difference of replayed renders is the
     * derivative the adjoint computes. */
    float* tape_survival;
    uint32_t survival_mode;    /* bf_survival_mode; BF_SURVIVAL_NONE ignores tape_survival */
    uint32_t reserved;         /* must be 0 */
} bf_frame_job;


/*
 * Renders many frames of (possibly) different scenes through one set of queues: every launch of
 * the bounce loop carries paths of all frames that fit the workspace (whole sample passes, in
 * job order), so small frames share the device instead of running one after another. Each
 * frame's film, paths and failure semantics are exactly those of a separate
 * blackframe_hip_trace_wavefront_transport call with the same options; the report aggregates
 * the batch (per-bounce queue fingerprints are recorded for the first launch only).
 * The reference has no batched entry point: its callers loop trace_cuda_wavefront_transport
 * (WavefrontTransport.hpp:197-208) over frames, which this call replaces.
 */
int blackframe_hip_trace_wavefront_batch(bf_hip_workspace* workspace, const bf_frame_job* jobs, uint32_t job_count,
                                         const bf_transport_options* options, bf_transport_report* report,
                                         uint64_t* traversal_counters, void* stream);

/*
 * Gradient request of one frame for blackframe_hip_trace_wavefront_adjoint.
 * film_gradient holds dLoss/d(film RGB sums), [pixels * 3], the layout of the float film.
 * texture_gradients[k] receives dLoss/d(texels of scene texture k) in the texture's layout (it must
 * hold width * height * channels floats) and is ACCUMULATED (+=) with float atomics in an
 * unspecified order; null entries are skipped.
 */
typedef struct bf_adjoint_job {
    const float* film_gradient;
    float* texture_gradients[BF_MAX_TEXTURES];
} bf_adjoint_job;

/*
 * Path-replay adjoint of blackframe_hip_trace_wavefront_batch (no counterpart in the reference,
 * which is forward-only). jobs[i] must describe the frame exactly as it was rendered with a tape
 * (same scene and texture values, frame, options); its film pointers are not used. Each path is
 * replayed from its sample stream index through the same BVH decisions; shadow visibility and the
 * final radiance come from the tape, so no shadow ray is traced. Sampling, MIS weights and
 * Russian roulette are treated as constants (detached); derivatives flow through the textured
 * closure weights and GGX widths. Requires diffuse + glossy + specular depth limits <= 63.
 */
int blackframe_hip_trace_wavefront_adjoint(bf_hip_workspace* workspace, const bf_frame_job* jobs,
                                           const bf_adjoint_job* adjoints, uint32_t job_count,
                                           const bf_transport_options* options, bf_transport_report* report,
                                           void* stream);

/* Result of blackframe_hip_autotune_launch. */
typedef struct bf_launch_tuning {
    uint32_t launch;         /* the fastest bf_transport_options.launch word found */
    uint32_t candidates;     /* launch words timed */
    float best_ms;           /* trial time of `launch` */
    float default_ms;        /* trial time of launch = 0 (the library default) */
    uint32_t tried[48];      /* the candidates, in the order timed */
    float tried_ms[48];
} bf_launch_tuning;

/*
 * Launch autotuner (no reference counterpart: the reference launches fixed shapes). Times the
 * transport of `jobs` restricted to their first `trial_passes` sample passes (into scratch films
 * the call allocates; the jobs' own buffers are not touched) for every combination of traversal
 * workgroup size {64, 128, 256}, bounce group size {2, 4, 8} and graph replay {off, on} (graph
 * replay only when `options` allow it), then the fastest of those with the closure-kind shade
 * order, the direction-octant traversal order and both, then the fastest of all with half and a
 * quarter of the workspace's paths in flight per launch (BF_LAUNCH_INFLIGHT, where a sample pass
 * of every frame still fits), with hipEvents on `stream`, and returns the fastest launch word.
 * options->launch is ignored; every other option is used as given. Synchronises.
 */
int blackframe_hip_autotune_launch(bf_hip_workspace* workspace, const bf_frame_job* jobs, uint32_t job_count,
                                   const bf_transport_options* options, uint32_t trial_passes,
                                   bf_launch_tuning* result, void* stream);

/*
 * FNV-1a-64 of `scene_byte_count` device bytes of a scene blob, with the 8 bytes of its
 * content-hash field (offset BF_SCENE_SOA_HASH_OFFSET) read as zero, written to the device word
 * `output_hash` on the default stream. Rejects a blob shorter than its header
 * (BF_SCENE_SOA_HEADER_SIZE) or null pointers before any launch.
 * Replaces: blackframe_cuda_launch_scene_soa_hash (XPU/CUDA/Include/Blackframe/XPU/CUDA/SceneSoaHash.hpp:9,
 *           Source/CudaSceneSoaHash.cu:13-51).
 */
int blackframe_hip_launch_scene_soa_hash(const uint8_t* scene_bytes, size_t scene_byte_count, uint64_t* output_hash);

/*
 * validate_scene_soa_header run by the device on a device copy of a blob (16-byte aligned), then
 * BF_SCENE_SOA_TRUNCATED when total_size_bytes exceeds `scene_byte_count`: the bf_scene_soa_status
 * is written to the device word `output_status` on the default stream. Rejects a blob shorter than
 * its header, a misaligned blob or null pointers before any launch.
 * Replaces: the header check each reference stage makes of the uploaded blob
 *           (CudaWavefrontStages.cu:818-880, 3966-3971; validate_scene_soa_header, SceneSoaAbi.hpp:283-341).
 */
int blackframe_hip_launch_scene_soa_validate(const uint8_t* scene_bytes, size_t scene_byte_count, uint32_t* output_status);

/*
 * Checks a host scene blob: validate_scene_soa_header's checks in its order, then (no reference
 * counterpart) that `byte_count` covers total_size_bytes and, with BF_SCENE_SOA_VERIFY_HASH, that
 * content_hash is the FNV-1a-64 of the bytes with the hash field read as zero. *status receives a
 * bf_scene_soa_status; the call returns BF_OK whenever it could decide. Host only, no GPU call.
 * Replaces: validate_scene_soa_header (XPU/Shared/.../SceneSoaAbi.hpp:283-341), the check every
 *           reference stage makes of the uploaded header (CudaWavefrontStages.cu:818-880).
 */
int blackframe_hip_scene_soa_validate(const void* blob, size_t byte_count, uint32_t flags, uint32_t* status);

/*
 * Builds a device scene from a host "BLKFSOA4" blob, the reference's serialized FrameScene
 * (SceneSoA.cpp:333-1178): the same BVH, tables and upload as blackframe_hip_scene_create.
 * Materials, closures and the environment arrive sampled at one wavelength packet and become
 * point-sampled spectra (BF_SPECTRUM_SAMPLED): render the scene in frame-packet mode with
 * packet_wavelengths equal to the blob's wavelengths, as the reference's stages require
 * (spectra_match, CudaWavefrontStages.cu:1170-1187). Rejects with BF_INVALID_ARGUMENT an invalid
 * header (see blackframe_hip_scene_soa_validate), non-canonical ids, inconsistent ranges,
 * unknown references, a mesh-area-light column that disagrees with the emissive instances, or
 * wavelengths other than one uniform-visible packet (pdf 1/470 per nm, wavelength measure); with
 * BF_UNAVAILABLE normal or bump maps and instanced materials without spectral data, which this
 * transport does not render. Punctual lights keep their registry order, and their spectra apply
 * lane by lane. Constant and image textures are accepted and unused.
 * Replaces: CudaSceneSoA::upload (Backends/GPU/CUDA/Include/Blackframe/Backends/GPU/CUDA/SceneSoA.hpp:33,
 *           Source/SceneSoA.cpp:1180-1210).
 */
int blackframe_hip_scene_create_from_soa(const void* blob, size_t byte_count, uint32_t flags, bf_hip_scene** out);

/*
 * Fills `output` from device code on device 0 with the transport records' sizes, alignments and
 * field offsets (BF_LAYOUT_VALUE_COUNT values in the reference's LayoutValue order) and sets
 * `device_count`. Allocates and frees one manifest of device memory and synchronises.
 * Replaces: blackframe_cuda_query_transport_abi_layout (XPU/CUDA/Include/Blackframe/XPU/CUDA/
 *           TransportAbiProbe.hpp, Source/CudaTransportAbiProbe.cu:168-209).
 */
int blackframe_hip_query_transport_abi_layout(bf_layout_manifest* output, int* device_count);

/*
 * Pushes requests[i].slot onto queue requests[i].queue_kind and writes outcomes[i] (bf_push_status):
 * pushed, capacity_exhausted (the queue's overflow and rejected counters grow, saturating), or
 * invalid_contract (queue kind >= BF_QUEUE_COUNT, a header that breaks the queue contract, or a
 * size beyond capacity). Requests of one wave and kind reserve their slots with one atomic; the
 * order of slots and which requests win under overflow are unspecified, as in the reference.
 * Default stream, asynchronous; invalid launch arguments are rejected before any launch.
 * Replaces: blackframe_cuda_launch_wavefront_queue_pushes (WavefrontQueueKernel.hpp:48-57,
 *           Source/CudaWavefrontQueues.cu:22-68, WavefrontQueueDevice.cuh:118-197).
 */
int blackframe_hip_launch_wavefront_queue_pushes(bf_queue_header* headers, uint32_t* path_slots, uint32_t queue_count,
                                                 uint32_t slot_stride, const bf_queue_push* requests,
                                                 uint32_t request_count, uint32_t* outcomes);

/*
 * Minimum scratch (device bytes, BF_COMPACTION_SCRATCH_ALIGNMENT-aligned) of
 * blackframe_hip_launch_wavefront_queue_compaction for any input_count <= max_input_count (0 for 0).
 * Replaces: blackframe_cuda_query_wavefront_queue_compaction_scratch_bytes
 *           (XPU/CUDA/Include/Blackframe/XPU/CUDA/WavefrontQueueCompaction.hpp:52-59).
 */
int blackframe_hip_query_wavefront_queue_compaction_scratch_bytes(uint32_t max_input_count, size_t* scratch_bytes);

/*
 * Stable append of the successful outcomes whose route equals `route` (1..6) to the queue of that
 * kind: their path slots go, in outcome order, to queue[route][initial size ...], and the header's
 * size grows by their count. Transactional: when they do not all fit, nothing is appended, the
 * header's overflow and rejected counters grow (saturating) by their count and the result says
 * capacity_exhausted. A header that breaks its contract (ABI version, struct size, kind, capacity
 * equal to slot_stride, size within capacity, zero counters) or a selected path slot >=
 * slot_stride gives invalid_contract and leaves the queue unchanged. Invalid launch arguments
 * (queue view, route, input_count > slot_stride, null result, short or misaligned scratch) are
 * rejected synchronously with BF_INVALID_ARGUMENT. Kernels only, on `stream`; no allocation.
 * Replaces: blackframe_cuda_launch_wavefront_queue_compaction (WavefrontQueueCompaction.hpp:61-72,
 *           Source/CudaWavefrontQueueCompaction.cu:253-330).
 */
int blackframe_hip_launch_wavefront_queue_compaction(bf_queue_soa queues, const bf_stage_outcome* outcomes,
                                                     uint32_t input_count, uint32_t route, void* scratch,
                                                     size_t scratch_bytes, bf_compaction_result* device_result,
                                                     void* stream);

/*
 * Audits `work_count` device stage outcomes into one device bf_stage_audit: an outcome fails when
 * its status is not success, its route is unknown or outside allowed_route_mask (bit = route), or
 * its path slot is not below path_capacity; first_failure_work_index is the smallest failing index
 * (UINT32_MAX when none) and first_failure that outcome; for the shade stage the outcomes' closure-
 * and light-sample detail bits are counted. An invalid stage kind or a null pointer returns
 * BF_INVALID_ARGUMENT before any launch. Asynchronous on `stream`.
 * Replaces: blackframe_cuda_launch_wavefront_audit_stage (XPU/CUDA/Include/Blackframe/XPU/CUDA/
 *           WavefrontStageKernel.hpp:398-401; Source/CudaWavefrontStages.cu:4945-5033, 5274-5293).
 */
int blackframe_hip_launch_wavefront_audit_stage(const bf_stage_outcome* outcomes, uint32_t work_count,
                                                uint32_t allowed_route_mask, uint32_t path_capacity,
                                                uint32_t stage_kind, bf_stage_audit* audit, void* stream);

/*
 * blackframe_hip_launch_wavefront_queue_compaction without a scratch argument: the same kernels,
 * the same contract and results, with their scratch in a library-owned device array of fixed size
 * (about 512 KiB, one per device, held until the process ends). It calls no allocator, so callers
 * that cannot allocate device memory (code under stream capture, tight serving loops, harnesses
 * that forbid it) can compact. input_count above BF_STATIC_COMPACTION_MAX_INPUT is rejected with
 * BF_INVALID_ARGUMENT before any launch. Every call on one device shares that array: calls that
 * overlap on one device (concurrent streams, or threads) race and must be serialised by the caller.
 * No counterpart in the reference, whose launcher always takes caller scratch.
 */
int blackframe_hip_launch_wavefront_queue_compaction_static(bf_queue_soa queues, const bf_stage_outcome* outcomes,
                                                            uint32_t input_count, uint32_t route,
                                                            bf_compaction_result* device_result, void* stream);

/* One frame of blackframe_hip_film_l2_loss (device pointers; 40 bytes). */
typedef struct bf_film_loss_job {
    const float* film;            /* [pixels * 3] float RGB sums, the film of a transport call */
    const uint32_t* film_counts;  /* [pixels] samples per pixel */
    const float* target;          /* [pixels * 3] target image (per-pixel means) */
    float* film_gradient;         /* [pixels * 3] out: dLoss/d(film sums), the adjoint's input */
    uint32_t pixel_count;
    uint32_t reserved;            /* must be 0 */
} bf_film_loss_job;

/*
 * L2 image loss of rendered films and its gradient (no counterpart in the reference, which is
 * forward-only; this is the loss the adjoint's texture fits use). Per frame j, with
 * mean = film / max(count, 1) and d = mean - target per channel:
 *   losses[j]        = scale * sum(d * d)            (binary64, summed in an unspecified order)
 *   film_gradient[i] = ((2 * scale) * d) / max(count, 1)   (binary32, in this order)
 * `losses` is device memory of job_count doubles and is overwritten. Kernels only (no memset, no
 * allocation, no synchronisation), so a caller may capture the call in a graph.
 */
int blackframe_hip_film_l2_loss(const bf_film_loss_job* jobs, uint32_t job_count, float scale, double* losses,
                                void* stream);

/* One parameter array of blackframe_hip_adam_step (device pointers; 56 bytes). */
typedef struct bf_adam_param {
    float* value;                 /* [count] parameters, updated in place */
    const float* gradient;        /* [count] */
    float* first_moment;          /* [count] m, zero before step 1 */
    float* second_moment;         /* [count] v, zero before step 1 */
    uint64_t count;
    float learning_rate;
    float lower;                  /* the update is projected onto [lower, upper] */
    float upper;
    uint32_t reserved;            /* must be 0 */
} bf_adam_param;

typedef struct bf_adam_options {
    double beta1;
    double beta2;
    double epsilon;
    uint32_t step;                /* 1 for the first step */
    uint32_t reserved;            /* must be 0 */
} bf_adam_options;

/*
 * One Adam step (Kingma & Ba 2015, the bias-corrected form torch.optim.Adam uses) with a box
 * projection, over every parameter array (no counterpart in the reference). Per element, in
 * binary32 and in this order:
 *   m = beta1 * m + (1 - beta1) * g
 *   v = beta2 * v + (1 - beta2) * (g * g)
 *   p = p - step_size * (m / (sqrt(v) / sqrt(1 - beta2^step) + epsilon)),  p = min(max(p, lower), upper)
 * with step_size = learning_rate / (1 - beta1^step). beta1, beta2, 1 - beta1, 1 - beta2, epsilon,
 * step_size and sqrt(1 - beta2^step) are computed in binary64 on the host and rounded once to
 * binary32. One kernel launch per array; arrays must not overlap.
 */
int blackframe_hip_adam_step(const bf_adam_param* params, uint32_t param_count, const bf_adam_options* options,
                             void* stream);

/*
 * One texture-fit step as one graph (no counterpart in the reference). A step is: clear the films,
 * the taped forward render of every frame (as blackframe_hip_trace_wavefront_batch), each frame's
 * L2 loss and film gradient (blackframe_hip_film_l2_loss with its own scale), clear the fitted
 * gradients, the path-replay adjoint of every frame (blackframe_hip_trace_wavefront_adjoint) and one
 * projected Adam step of every parameter (blackframe_hip_adam_step). _create records all of it,
 * once, into one HIP graph; _launch replays it with no host work and no host read: every launch of
 * a render runs all depth-bound + 1 bounces, an exhausted queue making a bounce's kernels exit at
 * once, and the Adam step number is a device counter advanced by each replay, whose bias-corrected
 * coefficients come from a table computed on the host in the arithmetic of a direct
 * blackframe_hip_adam_step call. A replay's results therefore equal the four direct calls', up to
 * the order of the adjoint's float atomics.
 *
 * The graph holds the descriptor's pointers: every buffer must stay allocated, and the workspace
 * unused by other calls while a replay runs. adam.step is the step number of the first replay;
 * at most max_steps replays. Frames, options and launch word are validated as the direct calls do;
 * stage events, traversal counters and queue records are refused.
 */
typedef struct bf_hip_fit_graph bf_hip_fit_graph;
typedef struct bf_fit_step_desc {
    const bf_frame_job* frames;       /* [frame_count] taped forward jobs: films, tapes, survival */
    const bf_adjoint_job* adjoints;   /* [frame_count] film gradient and texture gradients */
    const bf_film_loss_job* losses;   /* [frame_count] film, counts, target, film gradient */
    const float* loss_scales;         /* [frame_count] host: each frame's loss scale */
    double* loss_values;              /* [frame_count] device: each frame's loss, written by every replay */
    const bf_adam_param* params;      /* [param_count] fitted arrays; gradients are cleared by every replay */
    uint32_t frame_count;
    uint32_t param_count;
    bf_adam_options adam;
    uint32_t max_steps;
    uint32_t reserved;                /* must be 0 */
} bf_fit_step_desc;

int blackframe_hip_fit_graph_create(bf_hip_workspace* workspace, const bf_fit_step_desc* desc,
                                    const bf_transport_options* options, bf_hip_fit_graph** out);
/* Replays one step on `stream` (asynchronous). BF_RESOURCE_EXHAUSTED after max_steps replays. */
int blackframe_hip_fit_graph_launch(bf_hip_fit_graph* graph, void* stream);
/* Waits for `stream`, then reports the last replay: its forward and adjoint lane counts, failures
 * and sample counts (either report may be NULL) and the replays so far. Returns
 * BF_TRANSPORT_FAILURE when a path failed under BF_FAILURES_ABORT. */
int blackframe_hip_fit_graph_report(bf_hip_fit_graph* graph, bf_transport_report* forward, bf_transport_report* adjoint,
                                    uint32_t* steps, void* stream);
int blackframe_hip_fit_graph_destroy(bf_hip_fit_graph* graph);

#ifdef __cplusplus
} /* extern "C" */
#endif

BF_STATIC_ASSERT(sizeof(bf_frame_job) == 184, "frame job is 184 bytes");
BF_STATIC_ASSERT(sizeof(bf_film_loss_job) == 40, "film loss job is 40 bytes");
BF_STATIC_ASSERT(sizeof(bf_adam_param) == 56, "Adam parameter record is 56 bytes");
BF_STATIC_ASSERT(sizeof(bf_adam_options) == 32, "Adam options are 32 bytes");
BF_STATIC_ASSERT(sizeof(bf_fit_step_desc) == 96, "fit step descriptor is 96 bytes");

#endif /* BLACKFRAME_HIP_H */
