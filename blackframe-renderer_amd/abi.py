"""ctypes mirrors of include/blackframe_hip_abi.h (record layouts shared by every binding)."""

import ctypes as C

LANES = 4
MAX_CLOSURES = 8
INVALID = 0xFFFFFFFF

# closure kinds / lobes (TransportLobesDevice.cuh:25-44)
LOBE_DIFFUSE, LOBE_GLOSSY, LOBE_SPECULAR, LOBE_REFLECTION, LOBE_TRANSMISSION = 0x1, 0x2, 0x4, 0x8, 0x10
# bf_status
OK, INVALID_ARGUMENT, INCOMPATIBLE, RESOURCE_EXHAUSTED, UNAVAILABLE, DEVICE_ERROR, INTERNAL_ERROR, TRANSPORT_FAILURE = range(8)
# point-sampled spectra: count = SPECTRUM_SAMPLED | 4, values = 4 wavelengths then 4 values
SPECTRUM_SAMPLED = 0x80000000
# scene blob "BLKFSOA4" (bf_scene_soa_status, BF_SCENE_SOA_VERIFY_HASH)
SCENE_SOA_VALID, SCENE_SOA_BAD_MAGIC, SCENE_SOA_INCOMPATIBLE_VERSION, SCENE_SOA_BAD_HEADER_SIZE = 0, 1, 2, 3
SCENE_SOA_BAD_COLUMN_COUNT, SCENE_SOA_UNSUPPORTED_HASH_ALGORITHM, SCENE_SOA_NONZERO_RESERVED = 4, 5, 6
SCENE_SOA_INVALID_ENVIRONMENT_COUNT, SCENE_SOA_INVALID_COLUMN_DESCRIPTOR, SCENE_SOA_SIZE_OVERFLOW = 7, 8, 9
SCENE_SOA_BAD_TOTAL_SIZE, SCENE_SOA_TRUNCATED, SCENE_SOA_HASH_MISMATCH = 10, 11, 12
SCENE_SOA_SIZE_MISMATCH, SCENE_SOA_COUNT_OVERFLOW, SCENE_SOA_TOO_MANY_MESH_LIGHTS = 13, 14, 15
# serialized BVH "BLKFBVH1" (bf_bvh_status, BF_BVH_VERIFY_HASH)
(BVH_VALID, BVH_BAD_MAGIC, BVH_INCOMPATIBLE_VERSION, BVH_BAD_HEADER_SIZE, BVH_BAD_ARRAY_COUNT,
 BVH_UNSUPPORTED_HASH_ALGORITHM, BVH_NONZERO_RESERVED, BVH_BAD_LEAF_CAPACITY, BVH_INVALID_TOPOLOGY_COUNTS,
 BVH_INVALID_TLAS_ROOT, BVH_INVALID_ARRAY_DESCRIPTOR, BVH_SIZE_OVERFLOW, BVH_BAD_TOTAL_SIZE, BVH_TRUNCATED,
 BVH_HASH_MISMATCH) = range(15)
BVH_VERIFY_HASH = 0x1
SCENE_SOA_VERIFY_HASH = 0x1
CLOSURE_LAMBERTIAN, CLOSURE_ROUGH_DIFFUSE, CLOSURE_ROUGH_CONDUCTOR, CLOSURE_ROUGH_DIELECTRIC = 1, 2, 3, 4
CLOSURE_SPECULAR_REFLECTION, CLOSURE_SPECULAR_TRANSMISSION = 5, 6

SAMPLER_INDEXED_HASH, SAMPLER_SOBOL_OWEN = 0, 1
WAVELENGTHS_FRAME_PACKET, WAVELENGTHS_PER_PATH = 0, 1
MIS_BALANCE, MIS_POWER = 0, 1
JITTER_CENTER, JITTER_UNIFORM = 0, 1

TERM_NONE, TERM_ESCAPED, TERM_DEPTH_LIMIT, TERM_ZERO_THROUGHPUT, TERM_OUTSIDE_SUPPORT, TERM_RUSSIAN_ROULETTE = range(6)
TERM_FAILED = 15


class SampleStreamIndex(C.Structure):
    _fields_ = [("pixel_x", C.c_uint32), ("pixel_y", C.c_uint32), ("sample_index", C.c_uint64), ("seed", C.c_uint64)]


class TransportRay(C.Structure):
    _fields_ = [
        ("origin_x", C.c_float), ("origin_y", C.c_float), ("origin_z", C.c_float), ("t_min", C.c_float),
        ("direction_x", C.c_float), ("direction_y", C.c_float), ("direction_z", C.c_float), ("t_max", C.c_float),
        ("time", C.c_float), ("visibility_mask", C.c_uint32), ("current_medium", C.c_uint32), ("reserved", C.c_uint32),
    ]


class PathStateLane(C.Structure):
    """bf_path_state_lane: TransportPathStateLane (TransportAbi.hpp:110-134), 128 bytes."""
    _fields_ = [("beta", C.c_float * 4), ("accumulated_radiance", C.c_float * 4), ("wavelength_nanometers", C.c_float * 4),
                ("wavelength_pdf_values", C.c_float * 4), ("diffuse_depth", C.c_uint32), ("glossy_depth", C.c_uint32),
                ("specular_depth", C.c_uint32), ("transmission_depth", C.c_uint32), ("volume_depth", C.c_uint32),
                ("depth", C.c_uint32), ("eta_scale", C.c_float), ("current_medium", C.c_uint32), ("delta_flags", C.c_uint32),
                ("wavelength_pdf_measures", C.c_uint8 * 4), ("reserved", C.c_uint32 * 6)]


class RayCone(C.Structure):
    """bf_ray_cone: WavefrontRayCone (WavefrontStageKernel.hpp:146-150)."""
    _fields_ = [("width", C.c_float), ("spread", C.c_float)]


class LaneControl(C.Structure):
    """bf_lane_control: WavefrontLaneControl (WavefrontStageKernel.hpp:139-144)."""
    _fields_ = [("phase", C.c_uint32), ("termination", C.c_uint32), ("flags", C.c_uint32),
                ("blocked_depth_limits", C.c_uint32)]


class PathInputs(C.Structure):
    """bf_path_inputs: WavefrontCameraInputDeviceSoa (device column pointers)."""
    _fields_ = [("sample_streams", C.c_void_p), ("rays", C.c_void_p), ("ray_cones", C.c_void_p), ("path_states", C.c_void_p),
                ("count", C.c_uint32), ("reserved", C.c_uint32), ("reserved_tail", C.c_uint32 * 2)]


class PathOutputs(C.Structure):
    """bf_path_outputs: the terminal path states, rays, cones and lane controls (device columns)."""
    _fields_ = [("path_states", C.c_void_p), ("terminal_rays", C.c_void_p), ("terminal_cones", C.c_void_p),
                ("controls", C.c_void_p), ("count", C.c_uint32), ("reserved", C.c_uint32), ("reserved_tail", C.c_uint32 * 2)]


PHASE_TERMINATED = 8
WAVELENGTH_MEASURE = 5


class ClosestHit(C.Structure):
    _fields_ = [
        ("parameter", C.c_float), ("barycentric0", C.c_float), ("barycentric1", C.c_float), ("barycentric2", C.c_float),
        ("normal_x", C.c_float), ("normal_y", C.c_float), ("normal_z", C.c_float), ("reserved", C.c_uint32),
        ("object", C.c_uint32), ("instance", C.c_uint32), ("geometry", C.c_uint32), ("primitive", C.c_uint32),
        ("material", C.c_uint32), ("ids_reserved", C.c_uint32 * 3),
    ]


class ClosestHitResult(C.Structure):
    _fields_ = [("status", C.c_uint32), ("reserved", C.c_uint32 * 3), ("hit", ClosestHit)]


class OcclusionResult(C.Structure):
    _fields_ = [("status", C.c_uint32), ("reserved", C.c_uint32 * 3)]


class BvhNode(C.Structure):
    _fields_ = [
        ("min_x", C.c_float), ("min_y", C.c_float), ("min_z", C.c_float), ("kind", C.c_uint32),
        ("max_x", C.c_float), ("max_y", C.c_float), ("max_z", C.c_float), ("split_axis", C.c_uint32),
        ("first_child", C.c_uint32), ("second_child", C.c_uint32),
        ("reference_offset", C.c_uint32), ("reference_count", C.c_uint32),
    ]


class BvhBlas(C.Structure):
    _fields_ = [
        ("geometry_id", C.c_uint32), ("root_node", C.c_uint32), ("node_offset", C.c_uint32), ("node_count", C.c_uint32),
        ("primitive_reference_offset", C.c_uint32), ("primitive_reference_count", C.c_uint32),
        ("reserved", C.c_uint32 * 2),
    ]


class BvhInstanceRef(C.Structure):
    _fields_ = [
        ("scene_instance_index", C.c_uint32), ("instance_id", C.c_uint32), ("scene_geometry_index", C.c_uint32),
        ("geometry_id", C.c_uint32), ("blas_index", C.c_uint32), ("visibility_mask", C.c_uint32),
        ("reserved", C.c_uint32 * 2),
    ]


class ClosureRecord(C.Structure):
    """A closure resolved at the path's four wavelengths (ClosureRecord, TransportLobesDevice.cuh:66-71)."""
    _fields_ = [("kind", C.c_uint32), ("lobes", C.c_uint32), ("weight", C.c_float * 4), ("parameters", C.c_float * 10)]


class SpectrumDesc(C.Structure):
    _fields_ = [("offset", C.c_uint32), ("count", C.c_uint32)]


class ClosureDesc(C.Structure):
    _fields_ = [
        ("kind", C.c_uint32), ("lobes", C.c_uint32), ("weight_spectrum", C.c_uint32), ("eta_spectrum", C.c_uint32),
        ("k_spectrum", C.c_uint32), ("weight_texture", C.c_uint32), ("roughness_texture", C.c_uint32),
        ("reserved", C.c_uint32), ("parameters", C.c_float * 10),
        ("reserved_tail", C.c_uint32 * 2),
    ]


class MaterialDesc(C.Structure):
    _fields_ = [
        ("id", C.c_uint32), ("closure_offset", C.c_uint32), ("closure_count", C.c_uint32),
        ("emission_spectrum", C.c_uint32), ("frame_mode", C.c_uint32), ("tangent_rotation", C.c_float),
        ("reserved", C.c_uint32 * 2), ("probabilities", C.c_float * 8), ("cdf", C.c_float * 9),
        ("reserved_tail", C.c_uint32 * 3),
    ]


class GeometryDesc(C.Structure):
    _fields_ = [
        ("id", C.c_uint32), ("reserved", C.c_uint32), ("vertex_offset", C.c_uint64), ("vertex_count", C.c_uint64),
        ("triangle_offset", C.c_uint64), ("triangle_count", C.c_uint64),
    ]


class InstanceDesc(C.Structure):
    _fields_ = [
        ("id", C.c_uint32), ("object_id", C.c_uint32), ("geometry_index", C.c_uint32), ("material_index", C.c_uint32),
        ("visibility_mask", C.c_uint32), ("reserved", C.c_uint32 * 3),
        ("local_to_world", C.c_float * 12), ("world_to_local", C.c_float * 12),
    ]


class SceneDesc(C.Structure):
    _fields_ = [
        ("geometry_count", C.c_uint32), ("instance_count", C.c_uint32), ("material_count", C.c_uint32),
        ("closure_count", C.c_uint32), ("spectrum_count", C.c_uint32), ("has_environment", C.c_uint32),
        ("vertex_count", C.c_uint64), ("triangle_count", C.c_uint64), ("spectrum_value_count", C.c_uint64),
        ("positions", C.c_void_p), ("normals", C.c_void_p), ("uvs", C.c_void_p), ("indices", C.c_void_p),
        ("geometries", C.c_void_p), ("instances", C.c_void_p), ("materials", C.c_void_p), ("closures", C.c_void_p),
        ("spectra", C.c_void_p), ("spectrum_values", C.c_void_p),
        ("environment_spectrum", C.c_uint32), ("texture_count", C.c_uint32), ("textures", C.c_void_p),
        ("punctual_count", C.c_uint32), ("reserved_punctual", C.c_uint32), ("punctuals", C.c_void_p),
        ("image_texture_count", C.c_uint32), ("reserved_maps", C.c_uint32), ("image_textures", C.c_void_p),
        ("surface_maps", C.c_void_p),
    ]


TEXTURE_WRAP_REPEAT, TEXTURE_WRAP_CLAMP, TEXTURE_WRAP_MIRROR, TEXTURE_WRAP_BORDER = 0, 1, 2, 3
MAX_EWA_ANISOTROPY = 64


class ImageTextureDesc(C.Structure):
    """bf_image_texture_desc: a mip chain of float texels for the surface maps' EWA filter."""
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("channel_count", C.c_uint32), ("mip_count", C.c_uint32),
                ("texels", C.c_void_p), ("texel_count", C.c_uint64)]


class SurfaceMapsDesc(C.Structure):
    """bf_surface_maps_desc: a material's normal map and bump map (INVALID_INDEX textures: none)."""
    _fields_ = [("normal_texture", C.c_uint32), ("normal_channels", C.c_uint32 * 3), ("normal_u_wrap", C.c_uint32),
                ("normal_v_wrap", C.c_uint32), ("normal_max_anisotropy", C.c_uint32),
                ("normal_max_texel_visits", C.c_uint32), ("normal_y_convention", C.c_uint32),
                ("bump_texture", C.c_uint32), ("bump_channel", C.c_uint32), ("bump_scale", C.c_float),
                ("bump_u_wrap", C.c_uint32), ("bump_v_wrap", C.c_uint32), ("bump_max_anisotropy", C.c_uint32),
                ("bump_max_texel_visits", C.c_uint32)]


PUNCTUAL_POINT, PUNCTUAL_DIRECTIONAL, PUNCTUAL_SPOT = 0, 1, 2


class PunctualDesc(C.Structure):
    """bf_punctual_desc: a point, directional or spot light (radiant intensity / irradiance spectrum)."""
    _fields_ = [("kind", C.c_uint32), ("spectrum", C.c_uint32), ("position", C.c_float * 3),
                ("position_error", C.c_float * 3), ("direction", C.c_float * 3), ("inner_half_angle", C.c_float),
                ("outer_half_angle", C.c_float), ("reserved", C.c_uint32)]


class TextureDesc(C.Structure):
    """bf_texture_desc: float32 texels, host (copied) or device (borrowed); see blackframe_texture.h."""
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("channels", C.c_uint32), ("reserved", C.c_uint32),
                ("data", C.c_void_p), ("device_data", C.c_void_p)]


class CameraDesc(C.Structure):
    _fields_ = [
        ("origin", C.c_float * 3), ("tangent", C.c_float * 3), ("bitangent", C.c_float * 3), ("normal", C.c_float * 3),
        ("half_width", C.c_float), ("half_height", C.c_float), ("t_min", C.c_float), ("t_max", C.c_float),
        ("time", C.c_float), ("width", C.c_uint32), ("height", C.c_uint32), ("visibility_mask", C.c_uint32),
        ("jitter_mode", C.c_uint32), ("reserved", C.c_uint32 * 3),
    ]


class TransportOptions(C.Structure):
    _fields_ = [
        ("sampler", C.c_uint32), ("wavelength_mode", C.c_uint32), ("mis_heuristic", C.c_uint32),
        ("diffuse_depth_limit", C.c_uint32), ("glossy_depth_limit", C.c_uint32), ("specular_depth_limit", C.c_uint32),
        ("transmission_depth_limit", C.c_uint32), ("volume_depth_limit", C.c_uint32),
        ("russian_roulette_enabled", C.c_uint32), ("russian_roulette_first_depth", C.c_uint32),
        ("russian_roulette_minimum", C.c_float), ("russian_roulette_maximum", C.c_float),
        ("packet_wavelengths", C.c_float * 4), ("record_queues", C.c_uint32), ("instrumentation", C.c_uint32),
        ("failure_policy", C.c_uint32), ("launch", C.c_uint32),
    ]


class FrameDesc(C.Structure):
    _fields_ = [("camera", CameraDesc), ("seed", C.c_uint64), ("samples_per_pixel", C.c_uint32),
                ("first_sample_index", C.c_uint32)]


MAX_BOUNCES = 64
INSTRUMENT_DISABLED, INSTRUMENT_STAGE_EVENTS = 0, 1
FAILURES_ABORT, FAILURES_COUNT = 0, 1
HIT_MISS, HIT_HIT, HIT_INVALID_RAY, HIT_INVALID_SCENE, HIT_INVALID_BVH = 0, 1, 2, 3, 4
HIT_STACK_OVERFLOW, HIT_INVALID_TOPOLOGY, HIT_NUMERICAL_FAILURE = 5, 6, 7
LAUNCH_GRAPH = 0x80000000
LAUNCH_GROUP_MASK = 0x0000FFFF
LAUNCH_SHADE_ORDER = 0x00100000
LAUNCH_TRACE_ORDER = 0x00200000
LAUNCH_INFLIGHT_MASK = 0x01C00000


def launch_in_flight(k):
    """The launch-word field limiting the paths in flight per launch to capacity >> k."""
    return (k & 7) << 22



def launch_trace_block(lanes):
    """The launch-word field selecting the traversal workgroup size (64, 128 or 256 lanes)."""
    return {64: 6, 128: 7, 256: 8}[lanes] << 16
STAGES = ("camera", "closest_hit", "shade", "shadow", "compaction", "film", "auxiliary", "reserved")


class StageMetric(C.Structure):
    _fields_ = [("kernel_dispatches", C.c_uint64), ("kernel_lanes", C.c_uint64),
                ("gpu_elapsed_nanoseconds", C.c_uint64)]


class TransportReport(C.Structure):
    _fields_ = [
        ("path_count", C.c_uint64), ("closest_hit_queries", C.c_uint64), ("shadow_queries", C.c_uint64),
        ("shade_lanes", C.c_uint64), ("miss_lanes", C.c_uint64), ("closure_samples", C.c_uint64),
        ("light_samples", C.c_uint64), ("failed_paths", C.c_uint64), ("bounces", C.c_uint32),
        ("first_failure_slot", C.c_uint32), ("first_failure_code", C.c_uint32), ("first_failure_job", C.c_uint32),
        ("ray_queue_hash", C.c_uint64 * MAX_BOUNCES), ("ray_queue_size", C.c_uint32 * MAX_BOUNCES),
        ("shadow_queue_hash", C.c_uint64 * MAX_BOUNCES), ("shadow_queue_size", C.c_uint32 * MAX_BOUNCES),
        ("stage_metrics", StageMetric * 8),
    ]


class FrameJob(C.Structure):
    """One frame of a batched call (bf_frame_job): scene handle, frame, device film and path buffers."""
    _fields_ = [("scene", C.c_void_p), ("frame", FrameDesc), ("film", C.c_void_p), ("film_compensated", C.c_void_p),
                ("film_counts", C.c_void_p), ("paths", C.c_void_p), ("tape_radiance", C.c_void_p),
                ("tape_visibility", C.c_void_p), ("tape_survival", C.c_void_p), ("survival_mode", C.c_uint32),
                ("reserved", C.c_uint32)]


SURVIVAL_NONE, SURVIVAL_RECORD, SURVIVAL_REPLAY = 0, 1, 2


MAX_TEXTURES = 8


class LaunchTuning(C.Structure):
    """bf_launch_tuning: the autotuner's choice and the timings behind it."""
    _fields_ = [("launch", C.c_uint32), ("candidates", C.c_uint32), ("best_ms", C.c_float), ("default_ms", C.c_float),
                ("tried", C.c_uint32 * 48), ("tried_ms", C.c_float * 48)]


class AdjointJob(C.Structure):
    """bf_adjoint_job: dLoss/d(film RGB sums) in, accumulated texel gradients out (device pointers)."""
    _fields_ = [("film_gradient", C.c_void_p), ("texture_gradients", C.c_void_p * MAX_TEXTURES)]


SAMPLE_DIMENSION_SCHEMA = 1
SAMPLE_MAXIMUM_BOUNCE = 0xFFFFFFFF
SAMPLE_DUMP_SUCCESS, SAMPLE_DUMP_BOUNCE_OUT_OF_RANGE = 0, 1


class SamplePrimaryDimensions(C.Structure):
    """SampleStreamPrimaryDimensions (SampleStreamKernel.hpp:15-22)."""
    _fields_ = [(n, C.c_uint64) for n in ("camera_raster_x", "camera_raster_y", "lens_u", "lens_v", "time", "wavelength")]


class SampleBounceDimensions(C.Structure):
    """SampleStreamBounceDimensions (SampleStreamKernel.hpp:24-35)."""
    _fields_ = [(n, C.c_uint64) for n in ("light_selection", "light_u", "light_v", "bsdf_component", "bsdf_u", "bsdf_v",
                                          "medium_distance", "medium_phase_u", "medium_phase_v", "russian_roulette")]


class SampleDimensionManifest(C.Structure):
    """SampleStreamDimensionManifest (SampleStreamKernel.hpp:37-47)."""
    _fields_ = [("schema_version", C.c_uint32), ("reserved", C.c_uint32), ("primary_dimensions", SamplePrimaryDimensions),
                ("first_bounce_dimension", C.c_uint64), ("dimensions_per_bounce", C.c_uint64),
                ("maximum_bounce_index", C.c_uint64), ("bounce_offsets", SampleBounceDimensions)]


class SampleDumpRequest(C.Structure):
    """SampleStreamDumpRequest (SampleStreamKernel.hpp:49-53)."""
    _fields_ = [("sample_stream", SampleStreamIndex), ("dimension", C.c_uint64), ("bounce_index", C.c_uint64)]


class SampleDumpResult(C.Structure):
    """SampleStreamDumpResult (SampleStreamKernel.hpp:55-62)."""
    _fields_ = [("status", C.c_uint32), ("reserved", C.c_uint32), ("indexed_bits", C.c_uint64), ("sample", C.c_float),
                ("reserved_tail", C.c_uint32), ("bounce_dimensions", SampleBounceDimensions)]


QUEUE_COUNT = 7
QUEUE_ABI_MAJOR, QUEUE_ABI_MINOR = 1, 0
COMPACTION_SCRATCH_ALIGNMENT = 256
STATIC_COMPACTION_MAX_INPUT = 1 << 24  # BF_STATIC_COMPACTION_MAX_INPUT
COMPACTION_SUCCESS, COMPACTION_INVALID_CONTRACT, COMPACTION_CAPACITY_EXHAUSTED = 0, 1, 2
ROUTE_RAY, ROUTE_HIT, ROUTE_MISS, ROUTE_SHADE, ROUTE_SHADOW, ROUTE_CONTINUATION, ROUTE_TERMINATED = 1, 2, 3, 4, 5, 6, 7
STAGE_KIND_SHADE, STAGE_KIND_CONTINUATION = 6, 9
SHADE_DETAIL_CLOSURE_SAMPLED, SHADE_DETAIL_LIGHT_SAMPLED = 1 << 30, 1 << 31


class StageAudit(C.Structure):
    """bf_stage_audit: WavefrontStageAudit (64 B)."""
    _fields_ = [("abi_major", C.c_uint16), ("abi_minor", C.c_uint16), ("struct_size", C.c_uint32),
                ("stage_kind", C.c_uint32), ("expected_work_count", C.c_uint32), ("inspected_work_count", C.c_uint32),
                ("first_failure_work_index", C.c_uint32), ("closure_samples", C.c_uint32),
                ("light_samples", C.c_uint32), ("first_failure", C.c_uint32 * 4), ("reserved", C.c_uint32 * 4)]


LAYOUT_VALUE_COUNT = 73


class LayoutManifest(C.Structure):
    """bf_layout_manifest: LayoutManifest (TransportAbi.hpp:136-238)."""
    _fields_ = [("abi_major", C.c_uint16), ("abi_minor", C.c_uint16), ("device_cxx_standard", C.c_uint32),
                ("value_count", C.c_uint32), ("reserved_header", C.c_uint32),
                ("values", C.c_uint32 * LAYOUT_VALUE_COUNT), ("reserved_tail", C.c_uint32 * 3)]


class QueueSoa(C.Structure):
    """bf_queue_soa: WavefrontQueueDeviceSoa (device pointers)."""
    _fields_ = [("headers", C.c_void_p), ("path_slots", C.c_void_p), ("queue_count", C.c_uint32),
                ("slot_stride", C.c_uint32)]


class CompactionResult(C.Structure):
    _fields_ = [("status", C.c_uint32), ("queue_kind", C.c_uint32), ("route", C.c_uint32), ("input_count", C.c_uint32),
                ("initial_size", C.c_uint32), ("selected_count", C.c_uint32), ("published_count", C.c_uint32),
                ("rejected_count", C.c_uint32), ("reserved", C.c_uint32 * 4)]


class FilmLossJob(C.Structure):
    """bf_film_loss_job: one film of blackframe_hip_film_l2_loss (device pointers)."""
    _fields_ = [("film", C.c_void_p), ("film_counts", C.c_void_p), ("target", C.c_void_p), ("film_gradient", C.c_void_p),
                ("pixel_count", C.c_uint32), ("reserved", C.c_uint32)]


class AdamParam(C.Structure):
    """bf_adam_param: one parameter array of blackframe_hip_adam_step (device pointers)."""
    _fields_ = [("value", C.c_void_p), ("gradient", C.c_void_p), ("first_moment", C.c_void_p),
                ("second_moment", C.c_void_p), ("count", C.c_uint64), ("learning_rate", C.c_float),
                ("lower", C.c_float), ("upper", C.c_float), ("reserved", C.c_uint32)]


class AdamOptions(C.Structure):
    _fields_ = [("beta1", C.c_double), ("beta2", C.c_double), ("epsilon", C.c_double), ("step", C.c_uint32),
                ("reserved", C.c_uint32)]


class FitStepDesc(C.Structure):
    """bf_fit_step_desc: the buffers of one texture-fit step recorded as one graph."""
    _fields_ = [("frames", C.POINTER(FrameJob)), ("adjoints", C.POINTER(AdjointJob)), ("losses", C.POINTER(FilmLossJob)),
                ("loss_scales", C.POINTER(C.c_float)), ("loss_values", C.c_void_p), ("params", C.POINTER(AdamParam)),
                ("frame_count", C.c_uint32), ("param_count", C.c_uint32), ("adam", AdamOptions),
                ("max_steps", C.c_uint32), ("reserved", C.c_uint32)]


class PathResult(C.Structure):
    _fields_ = [
        ("radiance", C.c_float * 4), ("beta", C.c_float * 4), ("wavelengths", C.c_float * 4),
        ("depth", C.c_uint32), ("termination", C.c_uint32), ("depth_counters", C.c_uint32), ("eta_scale", C.c_float),
    ]


for _record, _size in ((SampleStreamIndex, 24), (TransportRay, 48), (ClosestHitResult, 80), (OcclusionResult, 16),
                       (BvhNode, 48), (BvhBlas, 32), (ClosureRecord, 64), (BvhInstanceRef, 32), (ClosureDesc, 80), (MaterialDesc, 112),
                       (InstanceDesc, 128), (PathResult, 64), (SceneDesc, 184), (ImageTextureDesc, 32), (SurfaceMapsDesc, 64), (PunctualDesc, 56), (CameraDesc, 96), (TextureDesc, 32), (AdjointJob, 72), (LaunchTuning, 400),
                       (FilmLossJob, 40), (AdamParam, 56), (AdamOptions, 32), (QueueSoa, 24), (CompactionResult, 48), (LayoutManifest, 320),
                       (TransportOptions, 80), (FrameDesc, 112), (SampleDimensionManifest, 160), (SampleDumpRequest, 40),
                       (SampleDumpResult, 104), (PathStateLane, 128), (RayCone, 8), (LaneControl, 16),
                       (PathInputs, 48), (PathOutputs, 48), (FrameJob, 184), (TransportReport, 1808), (StageMetric, 24),
                       (FitStepDesc, 96)):
    assert C.sizeof(_record) == _size, (_record.__name__, C.sizeof(_record))


# ---- XPU extension (plugin) ABI: ExtensionAbi.h / XpuBackendAbi.h ------------------------------
XPU_BACKEND_INTERFACE_ID = (0x96C6E70DD06B4B93, 0xA9514AB777F49D37)
EXT_STATUS = {0: "success", 1: "invalid_argument", 2: "unsupported_abi", 3: "interface_not_found",
              4: "invalid_state", 5: "out_of_range", 6: "extension_failure"}


class InterfaceId(C.Structure):
    _fields_ = [("high", C.c_uint64), ("low", C.c_uint64)]


class Utf8View(C.Structure):
    _fields_ = [("data", C.c_char_p), ("size", C.c_uint32), ("reserved", C.c_uint32)]

    def text(self):
        return C.string_at(self.data, self.size).decode() if self.size else ""


class ExtensionQuery(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("abi_major", C.c_uint32), ("abi_minor", C.c_uint32),
                ("reserved", C.c_uint32)]


GetInterfaceFn = C.CFUNCTYPE(C.c_uint32, C.c_void_p, InterfaceId, C.c_uint32, C.c_uint32, C.c_void_p, C.c_uint32)
ShutdownFn = C.CFUNCTYPE(None, C.c_void_p)


class ExtensionApi(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("abi_major", C.c_uint32), ("abi_minor", C.c_uint32),
                ("reserved", C.c_uint32), ("extension_context", C.c_void_p), ("get_interface", GetInterfaceFn),
                ("shutdown", ShutdownFn)]


class DeviceDescriptor(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("kind", C.c_uint32), ("identifier_high", C.c_uint64),
                ("identifier_low", C.c_uint64), ("name", Utf8View), ("vendor", Utf8View),
                ("reserved", C.c_uint64 * 4)]


GetDeviceCountFn = C.CFUNCTYPE(C.c_uint32, C.c_void_p, C.POINTER(C.c_uint32))
GetDeviceDescriptorFn = C.CFUNCTYPE(C.c_uint32, C.c_void_p, C.c_uint32, C.POINTER(DeviceDescriptor))


class XpuBackendApi(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("abi_major", C.c_uint32), ("abi_minor", C.c_uint32),
                ("reserved", C.c_uint32), ("backend_context", C.c_void_p), ("backend_name", Utf8View),
                ("get_device_count", GetDeviceCountFn), ("get_device_descriptor", GetDeviceDescriptorFn),
                ("reserved_slots", C.c_uint64 * 4)]


for _record, _size in ((ExtensionQuery, 16), (ExtensionApi, 40), (DeviceDescriptor, 88), (XpuBackendApi, 88)):
    assert C.sizeof(_record) == _size, (_record.__name__, C.sizeof(_record))
