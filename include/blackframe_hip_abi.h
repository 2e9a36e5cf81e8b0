/*
 * Blackframe XPU/HIP — host/device record layouts shared by the C ABI (blackframe_hip.h),
 * the HIP kernels and any host that binds the library.
 *
 * Records are plain C: fixed-width scalars, explicit padding, no native pointers. Where a record
 * restates one of the reference's frozen host/device records its size and field offsets are kept
 * identical (static asserts below) so that a host which already fills the reference record can pass
 * the same bytes:
 *   bf_sample_stream_index   <- Blackframe/XPU/Shared/Include/Blackframe/XPU/Shared/TransportAbi.hpp:38-43
 *   bf_transport_ray         <- TransportAbi.hpp:22-35
 *   bf_closest_hit           <- TransportAbi.hpp:100-110 (+ SurfaceIdentifiers :92-99)
 *   bf_closest_hit_result    <- Blackframe/XPU/Shared/Include/Blackframe/XPU/Shared/SceneTraversalAbi.hpp:36-40
 *   bf_occlusion_result      <- SceneTraversalAbi.hpp:44-47
 *   bf_bvh_node              <- Blackframe/XPU/Shared/Include/Blackframe/XPU/Shared/SceneBvhAbi.hpp:37-50
 *   bf_bvh_blas              <- SceneBvhAbi.hpp:52-60
 *   bf_bvh_instance_ref      <- SceneBvhAbi.hpp:62-70
 *   bf_closure_record        <- Blackframe/XPU/CUDA/Include/Blackframe/XPU/CUDA/TransportLobesDevice.cuh:66-71
 *   bf_queue_header          <- TransportAbi.hpp:48-58
 * Everything else (scene description, device-resident wide BVH node, transport options and
 * report) is this backend's own design.
 */
#ifndef BLACKFRAME_HIP_ABI_H
#define BLACKFRAME_HIP_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BF_SPECTRUM_LANES 4u
#define BF_MAX_CLOSURES 8u
#define BF_CLOSURE_PARAMETERS 10u
#define BF_MAX_TEXTURES 8u
#define BF_INVALID_INDEX 0xFFFFFFFFu
#define BF_TRAVERSAL_STACK_CAPACITY 64u

/* ---- status codes returned by every entry point (0 = success) ---------------------------- */
enum bf_status {
    BF_OK = 0,
    BF_INVALID_ARGUMENT = 1,
    BF_INCOMPATIBLE = 2,
    BF_RESOURCE_EXHAUSTED = 3,
    BF_UNAVAILABLE = 4,
    BF_DEVICE_ERROR = 5,
    BF_INTERNAL_ERROR = 6,
    BF_TRANSPORT_FAILURE = 7
};

/* ---- traversal status (SceneTraversalAbi.hpp:11-32) ---------------------------------------- */
enum bf_closest_hit_status {
    BF_HIT_MISS = 0,
    BF_HIT_HIT = 1,
    BF_HIT_INVALID_RAY = 2,
    BF_HIT_INVALID_SCENE = 3,
    BF_HIT_INVALID_BVH = 4,
    BF_HIT_STACK_OVERFLOW = 5,
    BF_HIT_INVALID_TOPOLOGY = 6,
    BF_HIT_NUMERICAL_FAILURE = 7
};
enum bf_occlusion_status { BF_VISIBLE = 0, BF_OCCLUDED = 1 };

/* ---- sampler / wavelength / MIS selectors -------------------------------------------------- */
enum bf_sampler_kind { BF_SAMPLER_INDEXED_HASH = 0, BF_SAMPLER_SOBOL_OWEN = 1 };
enum bf_wavelength_mode { BF_WAVELENGTHS_FRAME_PACKET = 0, BF_WAVELENGTHS_PER_PATH = 1 };
enum bf_mis_heuristic { BF_MIS_BALANCE = 0, BF_MIS_POWER = 1 };
enum bf_jitter_mode { BF_JITTER_CENTER = 0, BF_JITTER_UNIFORM = 1 };

/* ---- scattering lobes and closure kinds (TransportLobesDevice.cuh:25-44) ------------------ */
enum bf_lobe {
    BF_LOBE_DIFFUSE = 0x01u,
    BF_LOBE_GLOSSY = 0x02u,
    BF_LOBE_SPECULAR = 0x04u,
    BF_LOBE_REFLECTION = 0x08u,
    BF_LOBE_TRANSMISSION = 0x10u,
    BF_LOBE_VOLUME = 0x20u
};
enum bf_closure_kind {
    BF_CLOSURE_NONE = 0,
    BF_CLOSURE_LAMBERTIAN = 1,
    BF_CLOSURE_ROUGH_DIFFUSE = 2,
    BF_CLOSURE_ROUGH_CONDUCTOR = 3,
    BF_CLOSURE_ROUGH_DIELECTRIC = 4,
    BF_CLOSURE_SPECULAR_REFLECTION = 5,
    BF_CLOSURE_SPECULAR_TRANSMISSION = 6
};

/* ---- path termination (WavefrontStageKernel.hpp:51-58) ------------------------------------- */
enum bf_termination {
    BF_TERM_NONE = 0,
    BF_TERM_ESCAPED = 1,
    BF_TERM_DEPTH_LIMIT = 2,
    BF_TERM_ZERO_THROUGHPUT = 3,
    BF_TERM_OUTSIDE_SUPPORT = 4,
    BF_TERM_RUSSIAN_ROULETTE = 5,
    BF_TERM_FAILED = 15
};

/* ---- frozen records ------------------------------------------------------------------------ */
/* Records the reference declares alignas(16) (TransportAbi.hpp:16-134) carry the same alignment. */
#define BF_ALIGN16 __attribute__((aligned(16)))

typedef struct BF_ALIGN16 bf_transport_spectrum {
    float values[BF_SPECTRUM_LANES];
} bf_transport_spectrum;

typedef struct bf_path_slot {
    uint32_t value;
} bf_path_slot;

typedef struct bf_sample_stream_index {
    uint32_t pixel_x;
    uint32_t pixel_y;
    uint64_t sample_index;
    uint64_t seed;
} bf_sample_stream_index;

typedef struct BF_ALIGN16 bf_transport_ray {
    float origin_x, origin_y, origin_z, t_min;
    float direction_x, direction_y, direction_z, t_max;
    float time;
    uint32_t visibility_mask;
    uint32_t current_medium;
    uint32_t reserved;
} bf_transport_ray;

typedef struct BF_ALIGN16 bf_surface_identifiers {
    uint32_t object, instance, geometry, primitive, material;
    uint32_t reserved[3];
} bf_surface_identifiers;

typedef struct BF_ALIGN16 bf_closest_hit {
    float parameter;
    float barycentric0, barycentric1, barycentric2;
    float normal_x, normal_y, normal_z;
    uint32_t reserved;
    bf_surface_identifiers ids;
} bf_closest_hit;

typedef struct bf_closest_hit_result {
    uint32_t status;
    uint32_t reserved[3];
    bf_closest_hit hit;
} bf_closest_hit_result;

typedef struct bf_occlusion_result {
    uint32_t status;
    uint32_t reserved[3];
} bf_occlusion_result;

typedef struct BF_ALIGN16 bf_queue_header {
    uint16_t abi_major, abi_minor;
    uint32_t struct_size;
    uint32_t queue_kind;
    uint32_t capacity;
    uint32_t size;
    uint32_t overflow_count;
    uint32_t rejected_count;
    uint32_t reserved;
} bf_queue_header;

/* TransportPathStateLane (TransportAbi.hpp:110-134): one float path lane across the boundary. */
typedef struct BF_ALIGN16 bf_path_state_lane {
    bf_transport_spectrum beta;
    bf_transport_spectrum accumulated_radiance;
    bf_transport_spectrum wavelength_nanometers;
    bf_transport_spectrum wavelength_pdf_values;
    uint32_t diffuse_depth, glossy_depth, specular_depth, transmission_depth, volume_depth, depth;
    float eta_scale;
    uint32_t current_medium;
    uint32_t delta_flags;
    uint8_t wavelength_pdf_measures[BF_SPECTRUM_LANES];
    uint32_t reserved[6];
} bf_path_state_lane;

/* LayoutManifest (TransportAbi.hpp:136-238): 73 sizes, alignments and offsets of the transport
 * records, indexed in the reference's LayoutValue order, filled from device code. */
#define BF_LAYOUT_VALUE_COUNT 73u
typedef struct BF_ALIGN16 bf_layout_manifest {
    uint16_t abi_major, abi_minor;
    uint32_t device_cxx_standard;
    uint32_t value_count;
    uint32_t reserved_header;
    uint32_t values[BF_LAYOUT_VALUE_COUNT];
    uint32_t reserved_tail[3];
} bf_layout_manifest;

/* ---- wavefront stage records (WavefrontStageKernel.hpp:12-203) ------------------------------------
 * The reference's per-stage lane records, layout-identical. The fused kernels of this library keep
 * the same information in their own SoA columns; these records are the form the path-input entry
 * point (blackframe_hip_trace_wavefront_paths) takes and returns and the stage audit reports. */
enum bf_stage_status {
    BF_STAGE_SUCCESS = 0, BF_STAGE_INVALID_CONTRACT = 1, BF_STAGE_INVALID_PATH_SLOT = 2, BF_STAGE_INVALID_LANE_STATE = 3,
    BF_STAGE_INVALID_RAY = 4, BF_STAGE_INVALID_SCENE = 5, BF_STAGE_UNSUPPORTED_TRANSPORT = 6,
    BF_STAGE_NUMERICAL_FAILURE = 7, BF_STAGE_QUEUE_OVERFLOW = 8, BF_STAGE_TRAVERSAL_ERROR = 9
};
enum bf_stage_route {
    BF_ROUTE_NONE = 0, BF_ROUTE_RAY = 1, BF_ROUTE_HIT = 2, BF_ROUTE_MISS = 3, BF_ROUTE_SHADE = 4, BF_ROUTE_SHADOW = 5,
    BF_ROUTE_CONTINUATION = 6, BF_ROUTE_TERMINATED = 7
};
enum bf_lane_phase {
    BF_PHASE_EMPTY = 0, BF_PHASE_CAMERA = 1, BF_PHASE_RAY = 2, BF_PHASE_HIT = 3, BF_PHASE_MISS = 4, BF_PHASE_SHADE = 5,
    BF_PHASE_SHADOW = 6, BF_PHASE_CONTINUATION = 7, BF_PHASE_TERMINATED = 8, BF_PHASE_PROCESSING = 9
};
enum bf_stage_kind {
    BF_STAGE_KIND_CAMERA_SEED = 0, BF_STAGE_KIND_CAMERA = 1, BF_STAGE_KIND_INTERSECTION_GATHER = 2,
    BF_STAGE_KIND_INTERSECTION_CLASSIFY = 3, BF_STAGE_KIND_HIT = 4, BF_STAGE_KIND_MISS = 5, BF_STAGE_KIND_SHADE = 6,
    BF_STAGE_KIND_SHADOW_GATHER = 7, BF_STAGE_KIND_SHADOW_PROCESS = 8, BF_STAGE_KIND_CONTINUATION = 9
};
#define BF_STAGE_AUDIT_ABI_MAJOR 1u
#define BF_STAGE_AUDIT_ABI_MINOR 0u
/* bf_stage_outcome.detail bits of a shade outcome (WavefrontStageKernel.hpp:82-83) */
#define BF_SHADE_DETAIL_CLOSURE_SAMPLED (1u << 30)
#define BF_SHADE_DETAIL_LIGHT_SAMPLED (1u << 31)
#define BF_LANE_CONTINUATION_PENDING (1u << 0)
#define BF_LANE_SHADOW_PENDING (1u << 1)

/* WavefrontTransportConfig (64 B) */
typedef struct BF_ALIGN16 bf_transport_config {
    uint16_t abi_major, abi_minor;
    uint32_t struct_size;
    uint32_t mis_heuristic;
    uint32_t light_sampling_strategy;
    uint32_t light_count;
    uint32_t diffuse_depth_limit, glossy_depth_limit, specular_depth_limit, transmission_depth_limit, volume_depth_limit;
    uint32_t russian_roulette_mode;
    uint32_t russian_roulette_first_depth;
    float russian_roulette_minimum_probability;
    float russian_roulette_maximum_probability;
    uint32_t reserved[2];
} bf_transport_config;

/* WavefrontStageAudit (64 B): the first failure is the failing outcome with the smallest work index
 * (UINT32_MAX when the stage succeeded). */
typedef struct BF_ALIGN16 bf_stage_audit {
    uint16_t abi_major, abi_minor;
    uint32_t struct_size;
    uint32_t stage_kind;
    uint32_t expected_work_count;
    uint32_t inspected_work_count;
    uint32_t first_failure_work_index;
    uint32_t closure_samples;
    uint32_t light_samples;
    uint32_t first_failure[4]; /* bf_stage_outcome */
    uint32_t reserved[4];
} bf_stage_audit;

/* WavefrontLaneControl (16 B): phase, termination (bf_termination), flags, blocked depth limits
 * (the lobes the depth filter removed when the path ended at its depth limit). */
typedef struct BF_ALIGN16 bf_lane_control {
    uint32_t phase;
    uint32_t termination;
    uint32_t flags;
    uint32_t blocked_depth_limits;
} bf_lane_control;

/* WavefrontRayCone (8 B): width at the ray origin and spread angle. */
typedef struct bf_ray_cone {
    float width;
    float spread;
} bf_ray_cone;

/* WavefrontPendingShadow (128 B) */
typedef struct BF_ALIGN16 bf_pending_shadow {
    bf_transport_ray ray;
    bf_transport_spectrum beta, bsdf_factor, incident_radiance;
    float absolute_incoming_cosine, estimator_weight, selection_probability, conditional_probability;
    float shading_normal_correction;
    uint32_t continuation_pending, termination, bsdf_encoding;
} bf_pending_shadow;

/* WavefrontPreviousBsdfSample (48 B) */
typedef struct BF_ALIGN16 bf_previous_bsdf_sample {
    float context_x, context_y, context_z, context_time;
    float incoming_x, incoming_y, incoming_z;
    float probability_value;
    uint32_t probability_measure;
    uint32_t valid;
    uint32_t reserved[2];
} bf_previous_bsdf_sample;

/* WavefrontCameraInputDeviceSoa (48 B): the path inputs as four device columns. */
typedef struct BF_ALIGN16 bf_path_inputs {
    const bf_sample_stream_index* sample_streams;
    const bf_transport_ray* rays;
    const bf_ray_cone* ray_cones;
    const bf_path_state_lane* path_states;
    uint32_t count;
    uint32_t reserved;
    uint32_t reserved_tail[2];
} bf_path_inputs;

/* WavefrontStageDeviceSoa restricted to what a finished path returns (the per-path outputs of the
 * path-input entry point, four device columns of `count` records). */
typedef struct BF_ALIGN16 bf_path_outputs {
    bf_path_state_lane* path_states;
    bf_transport_ray* terminal_rays;
    bf_ray_cone* terminal_cones;
    bf_lane_control* controls;
    uint32_t count;
    uint32_t reserved;
    uint32_t reserved_tail[2];
} bf_path_outputs;

/* ---- scene blob "BLKFSOA4" (SceneSoaAbi.hpp:11-19, 141-189, 350) ------------------------------ */
#define BF_SCENE_SOA_HEADER_SIZE 4848u          /* sizeof(SceneSoaHeader) */
#define BF_SCENE_SOA_HASH_OFFSET 32u            /* offsetof(SceneSoaHeader, content_hash) */
#define BF_FNV1A64_OFFSET_BASIS 14695981039346656037ull
#define BF_FNV1A64_PRIME 1099511628211ull
#define BF_SCENE_SOA_MAGIC 0x34414F53464B4C42ull /* "BLKFSOA4" */
#define BF_SCENE_SOA_ABI_MAJOR 4u
#define BF_SCENE_SOA_ABI_MINOR 0u
#define BF_SCENE_SOA_HASH_FNV1A64 1u
#define BF_SCENE_SOA_COLUMN_COUNT 194u
#define BF_SCENE_SOA_COLUMN_ALIGNMENT 16u

/* SceneSoaHeaderValidationStatus (SceneSoaAbi.hpp:177-189), same values and check order. */
enum bf_scene_soa_status {
    BF_SCENE_SOA_VALID = 0,
    BF_SCENE_SOA_BAD_MAGIC = 1,
    BF_SCENE_SOA_INCOMPATIBLE_VERSION = 2,
    BF_SCENE_SOA_BAD_HEADER_SIZE = 3,
    BF_SCENE_SOA_BAD_COLUMN_COUNT = 4,
    BF_SCENE_SOA_UNSUPPORTED_HASH_ALGORITHM = 5,
    BF_SCENE_SOA_NONZERO_RESERVED = 6,
    BF_SCENE_SOA_INVALID_ENVIRONMENT_COUNT = 7,
    BF_SCENE_SOA_INVALID_COLUMN_DESCRIPTOR = 8,
    BF_SCENE_SOA_SIZE_OVERFLOW = 9,
    BF_SCENE_SOA_BAD_TOTAL_SIZE = 10,
    /* beyond the header (no reference counterpart): checked by the blob entry points */
    BF_SCENE_SOA_TRUNCATED = 11,   /* fewer bytes than the header or than total_size_bytes */
    BF_SCENE_SOA_HASH_MISMATCH = 12, /* content_hash differs from the FNV-1a-64 of the bytes */
    /* the further checks each reference stage makes of the uploaded blob (validate_scene,
     * CudaWavefrontStages.cu:818-862), reported by blackframe_hip_launch_scene_soa_validate: */
    BF_SCENE_SOA_SIZE_MISMATCH = 13,       /* total_size_bytes differs from the blob's byte count (padding) */
    BF_SCENE_SOA_COUNT_OVERFLOW = 14,      /* a record count above 2^32 - 1 */
    BF_SCENE_SOA_TOO_MANY_MESH_LIGHTS = 15 /* mesh_area_light_count above 2^24 (unsupported_transport) */
};

/* Flags of blackframe_hip_scene_soa_validate / blackframe_hip_scene_create_from_soa. */
#define BF_SCENE_SOA_VERIFY_HASH 0x1u

/* SceneSoaColumnDescriptor and SceneSoaHeader (SceneSoaAbi.hpp:141-175): the blob's first
 * 4848 bytes, little-endian; column c holds element_count elements of element_size bytes at
 * offset_bytes, 16-byte aligned, in column order, empty columns at offset 0. */
typedef struct bf_scene_soa_column {
    uint64_t offset_bytes;
    uint64_t element_count;
    uint32_t element_size;
    uint32_t reserved;
} bf_scene_soa_column;

typedef struct BF_ALIGN16 bf_scene_soa_header {
    uint64_t magic;
    uint16_t abi_major, abi_minor;
    uint32_t header_size;
    uint32_t column_count;
    uint32_t hash_algorithm;
    uint64_t total_size_bytes;
    uint64_t content_hash;
    uint64_t object_count, geometry_count, vertex_count, triangle_count, material_count, closure_count;
    uint64_t instance_count, punctual_light_count, mesh_area_light_count, environment_count, texture_count;
    uint64_t image_texture_count, image_mip_count, image_texel_count;
    bf_scene_soa_column columns[BF_SCENE_SOA_COLUMN_COUNT];
    uint64_t reserved[5];
} bf_scene_soa_header;

/* ---- sample-stream dump (the reference's SampleStreamKernel.hpp:10-176 contract) --------------- */
#define BF_SAMPLE_DIMENSION_SCHEMA 1u              /* sample_stream::DimensionMapSchemaVersion */
#define BF_SAMPLE_MAXIMUM_BOUNCE 0xFFFFFFFFull     /* MaximumBounceIndex */
#define BF_SAMPLE_FIRST_BOUNCE_DIMENSION 4ull
#define BF_SAMPLE_DIMENSIONS_PER_BOUNCE 10ull
enum bf_sample_dump_status { BF_SAMPLE_DUMP_SUCCESS = 0, BF_SAMPLE_DUMP_BOUNCE_OUT_OF_RANGE = 1 };

/* SampleStreamPrimaryDimensions (48 B) and SampleStreamBounceDimensions (80 B) */
typedef struct bf_sample_primary_dimensions {
    uint64_t camera_raster_x, camera_raster_y, lens_u, lens_v, time, wavelength;
} bf_sample_primary_dimensions;
typedef struct bf_sample_bounce_dimensions {
    uint64_t light_selection, light_u, light_v, bsdf_component, bsdf_u, bsdf_v;
    uint64_t medium_distance, medium_phase_u, medium_phase_v, russian_roulette;
} bf_sample_bounce_dimensions;

/* SampleStreamDimensionManifest (160 B) */
typedef struct bf_sample_dimension_manifest {
    uint32_t schema_version;
    uint32_t reserved;
    bf_sample_primary_dimensions primary_dimensions;
    uint64_t first_bounce_dimension;
    uint64_t dimensions_per_bounce;
    uint64_t maximum_bounce_index;
    bf_sample_bounce_dimensions bounce_offsets;
} bf_sample_dimension_manifest;

/* SampleStreamDumpRequest (40 B) and SampleStreamDumpResult (104 B) */
typedef struct bf_sample_dump_request {
    bf_sample_stream_index sample_stream;
    uint64_t dimension;
    uint64_t bounce_index;
} bf_sample_dump_request;
typedef struct bf_sample_dump_result {
    uint32_t status; /* bf_sample_dump_status */
    uint32_t reserved;
    uint64_t indexed_bits;
    float sample;
    uint32_t reserved_tail;
    bf_sample_bounce_dimensions bounce_dimensions; /* zero when status is bounce_out_of_range */
} bf_sample_dump_result;

/* ---- queue compaction (the reference's WavefrontQueueCompaction.hpp contract) --------------- */
#define BF_QUEUE_COUNT 7u                    /* camera, ray, hit, miss, shade, shadow, continuation */
#define BF_QUEUE_ABI_MAJOR 1u                /* HostDeviceTransportAbiMajor (TransportAbi.hpp:12) */
#define BF_QUEUE_ABI_MINOR 0u
#define BF_COMPACTION_SCRATCH_ALIGNMENT 256u
/* Largest input_count of blackframe_hip_launch_wavefront_queue_compaction_static (its scratch is
 * library-owned and of fixed size: about 512 KiB per device). */
#define BF_STATIC_COMPACTION_MAX_INPUT (1u << 24)

/* WavefrontStageOutcome (WavefrontStageKernel.hpp:107-112): status 0 = success, route 1..6 names a
 * queue kind (ray, hit, miss, shade, shadow, continuation; 7 = terminated). */
typedef struct bf_stage_outcome {
    uint32_t status;
    uint32_t route;
    uint32_t path_slot;
    uint32_t detail;
} bf_stage_outcome;

/* WavefrontQueueDeviceSoa (WavefrontQueueKernel.hpp:22-27): BF_QUEUE_COUNT headers and queue-major
 * path slots, queue k at path_slots[k * slot_stride]. */
typedef struct bf_queue_soa {
    bf_queue_header* headers;
    uint32_t* path_slots;
    uint32_t queue_count;
    uint32_t slot_stride;
} bf_queue_soa;

/* WavefrontQueueDevicePush (WavefrontQueueKernel.hpp:29-32) and its per-request status. */
typedef struct bf_queue_push {
    uint32_t queue_kind;
    uint32_t slot;
} bf_queue_push;
enum bf_push_status { BF_PUSH_PUSHED = 0, BF_PUSH_CAPACITY_EXHAUSTED = 1, BF_PUSH_INVALID_CONTRACT = 2 };

enum bf_compaction_status {
    BF_COMPACTION_SUCCESS = 0,
    BF_COMPACTION_INVALID_CONTRACT = 1,
    BF_COMPACTION_CAPACITY_EXHAUSTED = 2
};

/* WavefrontQueueCompactionResult (WavefrontQueueCompaction.hpp:21-31), written on the device. */
typedef struct bf_compaction_result {
    uint32_t status;
    uint32_t queue_kind;
    uint32_t route;
    uint32_t input_count;
    uint32_t initial_size;
    uint32_t selected_count;
    uint32_t published_count;
    uint32_t rejected_count;
    uint32_t reserved[4];
} bf_compaction_result;

/* Canonical BVH node, bit-identical to the reference's 48-byte SceneBvhNode. */
typedef struct bf_bvh_node {
    float min_x, min_y, min_z;
    uint32_t kind; /* 0 internal, 1 leaf */
    float max_x, max_y, max_z;
    uint32_t split_axis;
    uint32_t first_child, second_child;
    uint32_t reference_offset, reference_count;
} bf_bvh_node;

typedef struct bf_bvh_blas {
    uint32_t geometry_id, root_node, node_offset, node_count;
    uint32_t primitive_reference_offset, primitive_reference_count;
    uint32_t reserved[2];
} bf_bvh_blas;

typedef struct bf_bvh_instance_ref {
    uint32_t scene_instance_index, instance_id, scene_geometry_index, geometry_id;
    uint32_t blas_index, visibility_mask;
    uint32_t reserved[2];
} bf_bvh_instance_ref;

/* ---- serialized BVH "BLKFBVH1" (SceneBvhAbi.hpp:11-98, 116-229) -------------------------------
 * header, then five 16-byte aligned arrays in this order: BLAS, BLAS nodes, primitive references,
 * TLAS nodes, instance references; FNV-1a-64 of every byte with content_hash read as zero. */
#define BF_BVH_MAGIC 0x31485642464B4C42ull /* "BLKFBVH1" */
#define BF_BVH_ABI_MAJOR 1u
#define BF_BVH_ABI_MINOR 0u
#define BF_BVH_HASH_FNV1A64 1u
#define BF_BVH_ARRAY_ALIGNMENT 16u
#define BF_BVH_LEAF_CAPACITY 1u
#define BF_BVH_ARRAY_COUNT 5u
#define BF_BVH_HEADER_SIZE 256u
#define BF_BVH_HASH_OFFSET 32u

typedef struct bf_bvh_array_descriptor {
    uint64_t offset_bytes;
    uint64_t element_count;
    uint32_t element_size;
    uint32_t reserved;
} bf_bvh_array_descriptor;

typedef struct BF_ALIGN16 bf_bvh_header {
    uint64_t magic;
    uint16_t abi_major, abi_minor;
    uint32_t header_size;
    uint32_t array_count;
    uint32_t hash_algorithm;
    uint64_t total_size_bytes;
    uint64_t content_hash;
    uint64_t source_scene_hash; /* content hash of the scene blob the BVH was built from */
    uint64_t blas_count, blas_node_count, primitive_reference_count, tlas_node_count, instance_reference_count;
    uint32_t tlas_root_node;
    uint32_t leaf_capacity;
    bf_bvh_array_descriptor arrays[BF_BVH_ARRAY_COUNT];
    uint64_t reserved[5];
} bf_bvh_header;

/* SceneBvhHeaderValidationStatus (SceneBvhAbi.hpp:104-118), same values and check order, then the
 * blob checks of blackframe_hip_scene_bvh_validate (no reference counterpart). */
enum bf_bvh_status {
    BF_BVH_VALID = 0,
    BF_BVH_BAD_MAGIC = 1,
    BF_BVH_INCOMPATIBLE_VERSION = 2,
    BF_BVH_BAD_HEADER_SIZE = 3,
    BF_BVH_BAD_ARRAY_COUNT = 4,
    BF_BVH_UNSUPPORTED_HASH_ALGORITHM = 5,
    BF_BVH_NONZERO_RESERVED = 6,
    BF_BVH_BAD_LEAF_CAPACITY = 7,
    BF_BVH_INVALID_TOPOLOGY_COUNTS = 8,
    BF_BVH_INVALID_TLAS_ROOT = 9,
    BF_BVH_INVALID_ARRAY_DESCRIPTOR = 10,
    BF_BVH_SIZE_OVERFLOW = 11,
    BF_BVH_BAD_TOTAL_SIZE = 12,
    BF_BVH_TRUNCATED = 13,    /* fewer bytes than the header or than total_size_bytes */
    BF_BVH_HASH_MISMATCH = 14 /* content_hash differs from the FNV-1a-64 of the bytes */
};
#define BF_BVH_VERIFY_HASH 0x1u

/*
 * Device-resident wide node (this backend's layout). One 64-byte record per INTERNAL canonical
 * node, in canonical pre-order. It stores both children's boxes so one cache line decides both
 * child tests. child[i] is either (internal 64-byte record index) or (BF_WIDE_LEAF | reference
 * offset) for a leaf child; leaves are never fetched as nodes.
 */
#define BF_WIDE_LEAF 0x80000000u
typedef struct bf_wide_node {
    float lo_x[2], lo_y[2], lo_z[2];
    float hi_x[2], hi_y[2], hi_z[2];
    uint32_t child[2];
    uint32_t canonical_index; /* canonical pre-order index of this internal node */
    uint32_t split_axis;
} bf_wide_node;

typedef struct bf_closure_record {
    uint32_t kind;
    uint32_t lobes;
    float weight[BF_SPECTRUM_LANES];
    float parameters[BF_CLOSURE_PARAMETERS];
} bf_closure_record;

/* ---- scene description (input to the host scene builder) ----------------------------------- */

/* A spectrum is `count` samples on a uniform grid spanning [360, 830] nm; count == 1 is constant. */

/* A float32 image: texel (x, y) channel c at data[(y * width + x) * channels + c], row 0 at v = 0.
 * Lookups repeat in u and v and filter bilinearly between texel centres (x + 0.5) / width. */
typedef struct bf_texture_desc {
    uint32_t width, height, channels; /* channels 1..3 */
    uint32_t reserved;
    const float* data;   /* host texels, copied into the scene at creation ... */
    float* device_data;  /* ... or caller-owned device texels, read in place by every render */
} bf_texture_desc;
/* An image texture with a mip chain, read by the EWA filter of the surface maps (the scene blob's
 * image_texture / image_mip / image_texel columns, SceneSoaAbi.hpp; sampled as
 * CudaWavefrontStages.cu:2876-3330). Level k is max(1, floor(level k-1 / 2)) texels in each axis;
 * texels holds every level in order, each as rows of pixels of channel_count values, row 0 at v = 0. */
#define BF_TEXTURE_WRAP_REPEAT 0u
#define BF_TEXTURE_WRAP_CLAMP 1u
#define BF_TEXTURE_WRAP_MIRROR 2u
#define BF_TEXTURE_WRAP_BORDER 3u /* reads outside the image are 0 */
#define BF_MAX_EWA_ANISOTROPY 64u
#define BF_MAX_IMAGE_MIPS 32u
typedef struct bf_image_texture_desc {
    uint32_t width, height;
    uint32_t channel_count;
    uint32_t mip_count;
    const float* texels;
    uint64_t texel_count; /* sum over the levels of width * height * channel_count */
} bf_image_texture_desc;

/* A material's normal and bump maps (the material_normal_map_* / material_bump_map_* columns of the
 * scene blob; applied as CudaWavefrontStages.cu:3395-3673). A normal map's three channels encode a
 * tangent-space normal in [0, 1] (2 c - 1); y_convention 1 flips its y. A bump map's channel is a
 * height, scaled by bump_scale. Both are EWA-filtered over the hit's ray-cone footprint with at most
 * max_anisotropy (1..64) and max_texel_visits texel reads. */
typedef struct bf_surface_maps_desc {
    uint32_t normal_texture; /* image texture index, BF_INVALID_INDEX: none */
    uint32_t normal_channels[3];
    uint32_t normal_u_wrap, normal_v_wrap;
    uint32_t normal_max_anisotropy, normal_max_texel_visits;
    uint32_t normal_y_convention;
    uint32_t bump_texture; /* image texture index, BF_INVALID_INDEX: none */
    uint32_t bump_channel;
    float bump_scale;
    uint32_t bump_u_wrap, bump_v_wrap;
    uint32_t bump_max_anisotropy, bump_max_texel_visits;
} bf_surface_maps_desc;

/* A point-sampled spectrum: count == (BF_SPECTRUM_SAMPLED | BF_SPECTRUM_LANES) and its eight
 * values are four wavelengths (nm) then the value at each. It is defined only at those wavelengths,
 * so a scene holding one renders in frame-packet mode with packet_wavelengths equal to them (every
 * sampled spectrum of a scene shares one wavelength packet). This is how the reference's scene
 * blob stores materials, closures and the environment (SceneSoA.cpp:606-663, 880-906). */
#define BF_SPECTRUM_SAMPLED 0x80000000u
typedef struct bf_spectrum_desc {
    uint32_t offset; /* into bf_scene_desc.spectrum_values */
    uint32_t count;
} bf_spectrum_desc;

/*
 * A closure's spectral inputs are spectrum ids (index into bf_scene_desc.spectra or
 * BF_INVALID_INDEX). They are resolved at the path's wavelengths into the reference's
 * bf_closure_record layout: weight <- weight_spectrum; for rough conductors parameters[0..3] <-
 * eta_spectrum and parameters[4..7] <- k_spectrum; every other parameter is copied verbatim.
 */
typedef struct bf_closure_desc {
    uint32_t kind;
    uint32_t lobes;
    uint32_t weight_spectrum;
    uint32_t eta_spectrum;
    uint32_t k_spectrum;
    /* Texture bindings, as texture slot + 1 (0: none, so zero-filled descriptors are untextured):
     *  weight_texture     reflectance = weight_spectrum(lambda) * clamp01(sum_c t_c * B_c(lambda)),
     *                     t = bilinear texel (3 channels: RGB; 1 channel: grey, t_r = t_g = t_b),
     *                     B_c the RGB basis spectra of blackframe_texture.h;
     *  roughness_texture  GGX widths of rough conductors / dielectrics (alpha_x, alpha_y) =
     *                     (t_0, t_1) for 2 channels, (t_0, t_0) for 1 channel. */
    uint32_t weight_texture;
    uint32_t roughness_texture;
    uint32_t reserved;
    float parameters[BF_CLOSURE_PARAMETERS];
    uint32_t reserved_tail[2];
} bf_closure_desc;

typedef struct bf_material_desc {
    uint32_t id;
    uint32_t closure_offset;
    uint32_t closure_count;
    uint32_t emission_spectrum; /* BF_INVALID_INDEX when not emissive */
    uint32_t frame_mode;        /* 0 shading-normal frame, 1 dpdu-tangent frame */
    float tangent_rotation;
    uint32_t reserved[2];
    float probabilities[BF_MAX_CLOSURES];
    float cdf[BF_MAX_CLOSURES + 1];
    uint32_t reserved_tail[3];
} bf_material_desc;

typedef struct bf_geometry_desc {
    uint32_t id;
    uint32_t reserved;
    uint64_t vertex_offset, vertex_count;
    uint64_t triangle_offset, triangle_count;
} bf_geometry_desc;

typedef struct bf_instance_desc {
    uint32_t id;
    uint32_t object_id;
    uint32_t geometry_index;
    uint32_t material_index;
    uint32_t visibility_mask;
    uint32_t reserved[3];
    float local_to_world[12]; /* row-major 3x4 affine */
    float world_to_local[12];
} bf_instance_desc;

/* A punctual light (the reference's ScenePointLight / SceneSpotLight / SceneDirectionalLight,
 * serialized as SceneSoA.cpp:818-878 and sampled by sample_punctual_light,
 * CudaWavefrontStages.cu:2240-2376). Lights are registered punctual first, then mesh-area lights
 * (sample_registered_light, :2378-2403), and one is selected uniformly. */
enum bf_punctual_kind { BF_PUNCTUAL_POINT = 0, BF_PUNCTUAL_DIRECTIONAL = 1, BF_PUNCTUAL_SPOT = 2 };
typedef struct bf_punctual_desc {
    uint32_t kind;              /* bf_punctual_kind */
    uint32_t spectrum;          /* radiant intensity (point; spot on its axis) or irradiance (directional) */
    float position[3];          /* point and spot */
    float position_error[3];    /* absolute error bound of position, every component >= 0 */
    float direction[3];         /* unit: propagation (directional) or emission axis (spot) */
    float inner_half_angle;     /* spot, radians: 0 <= inner <= outer, 0 < outer < pi */
    float outer_half_angle;
    uint32_t reserved;
} bf_punctual_desc;

typedef struct bf_scene_desc {
    uint32_t geometry_count, instance_count, material_count, closure_count, spectrum_count;
    uint32_t has_environment;
    uint64_t vertex_count, triangle_count, spectrum_value_count;
    const float* positions; /* [vertex_count*3] */
    const float* normals;   /* [vertex_count*3] */
    const float* uvs;       /* [vertex_count*2] */
    const uint32_t* indices; /* [triangle_count*3], geometry-local vertex indices */
    const bf_geometry_desc* geometries;
    const bf_instance_desc* instances;
    const bf_material_desc* materials;
    const bf_closure_desc* closures;
    const bf_spectrum_desc* spectra;
    const float* spectrum_values;
    uint32_t environment_spectrum;
    uint32_t texture_count;                 /* <= BF_MAX_TEXTURES */
    const bf_texture_desc* textures;
    uint32_t punctual_count;
    uint32_t reserved_punctual;
    const bf_punctual_desc* punctuals;
    /* surface maps: image textures and, per material, its maps (NULL: no material has one). A
     * scene with maps renders through the path-input entry only, whose rays carry ray cones. */
    uint32_t image_texture_count;
    uint32_t reserved_maps;
    const bf_image_texture_desc* image_textures;
    const bf_surface_maps_desc* surface_maps; /* [material_count] or NULL */
} bf_scene_desc;

/* ---- camera / render description ------------------------------------------------------------ */
typedef struct bf_camera_desc {
    float origin[3];
    float tangent[3];
    float bitangent[3];
    float normal[3];
    float half_width, half_height; /* tan(fov/2)*aspect and tan(fov/2), precomputed by the host */
    float t_min, t_max, time;
    uint32_t width, height;
    uint32_t visibility_mask;
    uint32_t jitter_mode;
    uint32_t reserved[3];
} bf_camera_desc;

typedef struct bf_transport_options {
    uint32_t sampler;         /* bf_sampler_kind */
    uint32_t wavelength_mode; /* bf_wavelength_mode */
    uint32_t mis_heuristic;   /* bf_mis_heuristic */
    uint32_t diffuse_depth_limit, glossy_depth_limit, specular_depth_limit;
    uint32_t transmission_depth_limit, volume_depth_limit;
    uint32_t russian_roulette_enabled;
    uint32_t russian_roulette_first_depth;
    float russian_roulette_minimum;
    float russian_roulette_maximum;
    float packet_wavelengths[BF_SPECTRUM_LANES]; /* used in frame-packet mode */
    uint32_t record_queues; /* nonzero: fill bf_transport_report queue fingerprints */
    uint32_t instrumentation; /* bf_instrumentation_mode: BF_INSTRUMENT_STAGE_EVENTS fills stage_metrics */
    uint32_t failure_policy;  /* bf_failure_policy */
    /* Launch shape, three fields:
     *   bits 0-15  bounces issued between two host checks of the live-path count (0: 4);
     *   bits 16-19 log2 of the traversal workgroup size: 6, 7 or 8 (0: 7, 128 lanes);
     *   bit 20     BF_LAUNCH_SHADE_ORDER: shade the hit queue grouped by closure kind;
     *   bit 21     BF_LAUNCH_TRACE_ORDER: traverse rays grouped by direction octant;
     *   bits 22-24 k: at most capacity >> k paths in flight per launch (0: the whole workspace);
     *   bit 31     BF_LAUNCH_GRAPH: each bounce group is one replayed hipGraph (size rounded up
     *              to even). */
    uint32_t launch;
} bf_transport_options;

#define BF_LAUNCH_GRAPH 0x80000000u
#define BF_LAUNCH_GROUP_MASK 0x0000FFFFu
#define BF_LAUNCH_TRACE_BLOCK_MASK 0x000F0000u
#define BF_LAUNCH_TRACE_BLOCK(log2_lanes) ((uint32_t)(log2_lanes) << 16)
#define BF_LAUNCH_SHADE_ORDER 0x00100000u
#define BF_LAUNCH_TRACE_ORDER 0x00200000u
#define BF_LAUNCH_INFLIGHT_MASK 0x01C00000u
#define BF_LAUNCH_INFLIGHT(k) (((uint32_t)(k) & 7u) << 22)

/* What a numerically failed path (termination BF_TERM_FAILED) does to the call. The reference
 * aborts the batch on the first failing work item (WavefrontTransport.cpp:880-925), which is
 * BF_FAILURES_ABORT: the frame still completes, then the call returns BF_TRANSPORT_FAILURE.
 * BF_FAILURES_COUNT returns success and leaves the count and first failing slot in the report;
 * a failed path keeps the radiance it gathered before failing, on both the device and the CPU
 * restatement. */
enum bf_failure_policy { BF_FAILURES_ABORT = 0, BF_FAILURES_COUNT = 1 };

/* What a frame does with its Russian-roulette survival tape (bf_frame_job.tape_survival): nothing,
 * record every decision's probability, or decide with the recorded probabilities. */
enum bf_survival_mode { BF_SURVIVAL_NONE = 0, BF_SURVIVAL_RECORD = 1, BF_SURVIVAL_REPLAY = 2 };

/* One frame of work: spp passes over width*height pixels; path slot p = s*W*H + y*W + x. */
typedef struct bf_frame_desc {
    bf_camera_desc camera;
    uint64_t seed;
    uint32_t samples_per_pixel;
    uint32_t first_sample_index;
} bf_frame_desc;

/* Per-stage GPU metrics, mirroring CudaWavefrontStageMetric (WavefrontTransport.hpp:72-96): the
 * dispatch and lane counts are deterministic, the elapsed time (hipEvent pairs recorded on the
 * transport stream around every dispatch of the stage) is observational. */
enum bf_instrumentation_mode { BF_INSTRUMENT_DISABLED = 0, BF_INSTRUMENT_STAGE_EVENTS = 1 };
enum bf_stage {
    BF_STAGE_CAMERA = 0,
    BF_STAGE_CLOSEST_HIT = 1,
    BF_STAGE_SHADE = 2,
    BF_STAGE_SHADOW = 3,
    BF_STAGE_COMPACTION = 4,
    BF_STAGE_FILM = 5,
    BF_STAGE_AUXILIARY = 6, /* queue fingerprints, path export, counter resets */
    BF_STAGE_COUNT = 8
};
typedef struct bf_stage_metric {
    uint64_t kernel_dispatches;
    uint64_t kernel_lanes;            /* lanes launched (queue sizes; grid-stride upper bounds excluded) */
    uint64_t gpu_elapsed_nanoseconds; /* 0 unless instrumentation is BF_INSTRUMENT_STAGE_EVENTS */
} bf_stage_metric;

#define BF_REPORT_MAX_BOUNCES 64u
typedef struct bf_transport_report {
    uint64_t path_count;
    uint64_t closest_hit_queries;
    uint64_t shadow_queries;
    uint64_t shade_lanes;
    uint64_t miss_lanes;
    uint64_t closure_samples;
    uint64_t light_samples;
    uint64_t failed_paths;
    uint32_t bounces;
    uint32_t first_failure_slot;
    uint32_t first_failure_code;
    uint32_t first_failure_job; /* frame of the first failure in a batched call */
    /* per bounce: FNV-1a-64 of the ray queue contents (slot order) and its size */
    uint64_t ray_queue_hash[BF_REPORT_MAX_BOUNCES];
    uint32_t ray_queue_size[BF_REPORT_MAX_BOUNCES];
    uint64_t shadow_queue_hash[BF_REPORT_MAX_BOUNCES];
    uint32_t shadow_queue_size[BF_REPORT_MAX_BOUNCES];
    bf_stage_metric stage_metrics[BF_STAGE_COUNT];
} bf_transport_report;

/* Per-path terminal record written by the transport (optional output). */
typedef struct bf_path_result {
    float radiance[BF_SPECTRUM_LANES];
    float beta[BF_SPECTRUM_LANES];
    float wavelengths[BF_SPECTRUM_LANES];
    uint32_t depth;
    uint32_t termination;
    uint32_t depth_counters; /* diffuse | glossy<<8 | specular<<16 | transmission<<24 */
    float eta_scale;
} bf_path_result;

#ifdef __cplusplus
} /* extern "C" */
#endif

#ifdef __cplusplus
#define BF_STATIC_ASSERT(c, m) static_assert(c, m)
#else
#define BF_STATIC_ASSERT(c, m) _Static_assert(c, m)
#endif
BF_STATIC_ASSERT(sizeof(bf_sample_stream_index) == 24, "SampleStreamIndex is 24 bytes");
BF_STATIC_ASSERT(sizeof(bf_transport_ray) == 48, "TransportRay is 48 bytes");
BF_STATIC_ASSERT(sizeof(bf_closest_hit) == 64, "ClosestHit is 64 bytes");
BF_STATIC_ASSERT(sizeof(bf_closest_hit_result) == 80, "SceneClosestHitResult is 80 bytes");
BF_STATIC_ASSERT(sizeof(bf_occlusion_result) == 16, "SceneOcclusionResult is 16 bytes");
BF_STATIC_ASSERT(sizeof(bf_queue_header) == 32, "QueueHeader is 32 bytes");
BF_STATIC_ASSERT(sizeof(bf_stage_outcome) == 16, "WavefrontStageOutcome is 16 bytes");
BF_STATIC_ASSERT(sizeof(bf_queue_push) == 8, "WavefrontQueueDevicePush is 8 bytes");
BF_STATIC_ASSERT(sizeof(bf_transport_spectrum) == 16, "TransportSpectrum is 16 bytes");
BF_STATIC_ASSERT(sizeof(bf_path_slot) == 4, "PathSlot is 4 bytes");
BF_STATIC_ASSERT(sizeof(bf_path_state_lane) == 128, "TransportPathStateLane is 128 bytes");
BF_STATIC_ASSERT(offsetof(bf_path_state_lane, wavelength_pdf_measures) == 100, "path state measures at 100");
BF_STATIC_ASSERT(sizeof(bf_layout_manifest) == 320, "LayoutManifest is 320 bytes");
BF_STATIC_ASSERT(sizeof(bf_punctual_desc) == 56, "punctual light record is 56 bytes");
BF_STATIC_ASSERT(sizeof(bf_scene_soa_column) == 24, "SceneSoaColumnDescriptor is 24 bytes");
BF_STATIC_ASSERT(sizeof(bf_scene_soa_header) == BF_SCENE_SOA_HEADER_SIZE, "SceneSoaHeader is 4848 bytes");
BF_STATIC_ASSERT(offsetof(bf_scene_soa_header, content_hash) == BF_SCENE_SOA_HASH_OFFSET, "content hash at 32");
BF_STATIC_ASSERT(offsetof(bf_scene_soa_header, object_count) == 40, "counts at 40");
BF_STATIC_ASSERT(offsetof(bf_scene_soa_header, columns) == 152, "columns at 152");
BF_STATIC_ASSERT(offsetof(bf_scene_soa_header, reserved) == 4808, "reserved at 4808");
BF_STATIC_ASSERT(sizeof(bf_queue_soa) == 24, "queue view is 24 bytes");
BF_STATIC_ASSERT(sizeof(bf_sample_primary_dimensions) == 48, "SampleStreamPrimaryDimensions is 48 bytes");
BF_STATIC_ASSERT(sizeof(bf_sample_bounce_dimensions) == 80, "SampleStreamBounceDimensions is 80 bytes");
BF_STATIC_ASSERT(sizeof(bf_sample_dimension_manifest) == 160, "SampleStreamDimensionManifest is 160 bytes");
BF_STATIC_ASSERT(offsetof(bf_sample_dimension_manifest, bounce_offsets) == 80, "manifest bounce offsets at 80");
BF_STATIC_ASSERT(sizeof(bf_sample_dump_request) == 40, "SampleStreamDumpRequest is 40 bytes");
BF_STATIC_ASSERT(sizeof(bf_sample_dump_result) == 104, "SampleStreamDumpResult is 104 bytes");
BF_STATIC_ASSERT(offsetof(bf_sample_dump_result, bounce_dimensions) == 24, "dump result dimensions at 24");
BF_STATIC_ASSERT(sizeof(bf_compaction_result) == 48, "WavefrontQueueCompactionResult is 48 bytes");
BF_STATIC_ASSERT(sizeof(bf_transport_config) == 64, "WavefrontTransportConfig is 64 bytes");
BF_STATIC_ASSERT(offsetof(bf_transport_config, russian_roulette_mode) == 40, "config roulette mode at 40");
BF_STATIC_ASSERT(offsetof(bf_transport_config, reserved) == 56, "config reserved at 56");
BF_STATIC_ASSERT(sizeof(bf_stage_audit) == 64, "WavefrontStageAudit is 64 bytes");
BF_STATIC_ASSERT(offsetof(bf_stage_audit, first_failure) == 32, "audit first failure at 32");
BF_STATIC_ASSERT(sizeof(bf_lane_control) == 16, "WavefrontLaneControl is 16 bytes");
BF_STATIC_ASSERT(sizeof(bf_ray_cone) == 8, "WavefrontRayCone is 8 bytes");
BF_STATIC_ASSERT(sizeof(bf_pending_shadow) == 128, "WavefrontPendingShadow is 128 bytes");
BF_STATIC_ASSERT(offsetof(bf_pending_shadow, absolute_incoming_cosine) == 96, "pending shadow cosine at 96");
BF_STATIC_ASSERT(offsetof(bf_pending_shadow, bsdf_encoding) == 124, "pending shadow encoding at 124");
BF_STATIC_ASSERT(sizeof(bf_previous_bsdf_sample) == 48, "WavefrontPreviousBsdfSample is 48 bytes");
BF_STATIC_ASSERT(offsetof(bf_previous_bsdf_sample, valid) == 36, "previous sample valid at 36");
BF_STATIC_ASSERT(sizeof(bf_path_inputs) == 48, "WavefrontCameraInputDeviceSoa is 48 bytes");
BF_STATIC_ASSERT(offsetof(bf_path_inputs, count) == 32, "path inputs count at 32");
BF_STATIC_ASSERT(sizeof(bf_path_outputs) == 48, "path outputs are 48 bytes");
BF_STATIC_ASSERT(sizeof(bf_bvh_node) == 48, "SceneBvhNode is 48 bytes");
BF_STATIC_ASSERT(sizeof(bf_bvh_array_descriptor) == 24, "SceneBvhArrayDescriptor is 24 bytes");
BF_STATIC_ASSERT(sizeof(bf_bvh_header) == BF_BVH_HEADER_SIZE, "SceneBvhHeader is 256 bytes");
BF_STATIC_ASSERT(offsetof(bf_bvh_header, total_size_bytes) == 24, "BVH total size at 24");
BF_STATIC_ASSERT(offsetof(bf_bvh_header, content_hash) == BF_BVH_HASH_OFFSET, "BVH content hash at 32");
BF_STATIC_ASSERT(offsetof(bf_bvh_header, source_scene_hash) == 40, "BVH source hash at 40");
BF_STATIC_ASSERT(offsetof(bf_bvh_header, tlas_root_node) == 88, "BVH TLAS root at 88");
BF_STATIC_ASSERT(offsetof(bf_bvh_header, arrays) == 96, "BVH arrays at 96");
BF_STATIC_ASSERT(offsetof(bf_bvh_header, reserved) == 216, "BVH reserved at 216");
BF_STATIC_ASSERT(sizeof(bf_bvh_blas) == 32, "SceneBvhBlas is 32 bytes");
BF_STATIC_ASSERT(sizeof(bf_bvh_instance_ref) == 32, "SceneBvhInstanceReference is 32 bytes");
BF_STATIC_ASSERT(sizeof(bf_wide_node) == 64, "wide node is one 64-byte line");
BF_STATIC_ASSERT(sizeof(bf_closure_record) == 64, "ClosureRecord is 64 bytes");
BF_STATIC_ASSERT(sizeof(bf_closure_desc) == 80, "closure descriptor is 80 bytes");
BF_STATIC_ASSERT(sizeof(bf_texture_desc) == 32, "texture descriptor is 32 bytes");
BF_STATIC_ASSERT(sizeof(bf_image_texture_desc) == 32, "image texture descriptor is 32 bytes");
BF_STATIC_ASSERT(sizeof(bf_surface_maps_desc) == 64, "surface maps descriptor is 64 bytes");
BF_STATIC_ASSERT(sizeof(bf_scene_desc) == 184, "scene descriptor is 184 bytes");
BF_STATIC_ASSERT(sizeof(bf_transport_options) == 80, "transport options are 80 bytes");
BF_STATIC_ASSERT(sizeof(bf_stage_metric) == 24, "stage metric is 24 bytes");
BF_STATIC_ASSERT(sizeof(bf_transport_report) == 1808, "transport report is 1808 bytes");
BF_STATIC_ASSERT(offsetof(bf_closest_hit, ids) == 32, "ClosestHit identifiers at 32");
BF_STATIC_ASSERT(offsetof(bf_bvh_node, first_child) == 32, "SceneBvhNode first_child at 32");
BF_STATIC_ASSERT(offsetof(bf_closure_record, parameters) == 24, "ClosureRecord parameters at 24");

#endif /* BLACKFRAME_HIP_ABI_H */
