/*
 * The "BLKFSOA4" scene blob's column schema and header validation, shared by the host ingest
 * (bf_scene_soa.cpp) and the device validation kernel (bf_kernels.hip).
 *
 * Column schema: XPU/Shared/Include/Blackframe/XPU/Shared/SceneSoaAbi.hpp:22-276.
 */
#ifndef BF_SCENE_SOA_H
#define BF_SCENE_SOA_H

#include <stdint.h>

#include "../../include/blackframe_hip_abi.h"

#if defined(__HIPCC__)
#define BF_SOA_HD __host__ __device__ static inline
#else
#define BF_SOA_HD static inline
#endif

namespace bf_soa {

/* First column of each group (SceneSoaAbi.hpp:24-132). */
enum : uint32_t {
    C_OBJECT_ID = 0,
    C_GEOMETRY_ID = 1, C_GEOMETRY_VERTEX_OFFSET = 2, C_GEOMETRY_VERTEX_COUNT = 3,
    C_GEOMETRY_TRIANGLE_OFFSET = 4, C_GEOMETRY_TRIANGLE_COUNT = 5,
    C_POSITION_X = 6, C_NORMAL_X = 9, C_TEXCOORD_X = 12, C_TRIANGLE_VERTEX_0 = 14,
    C_MATERIAL_ID = 17, C_MATERIAL_SPECTRAL_PRESENT = 18, C_MATERIAL_WAVELENGTH_NM = 19,
    C_MATERIAL_WAVELENGTH_PDF = 23, C_MATERIAL_WAVELENGTH_MEASURE = 27, C_MATERIAL_CLOSURE_OFFSET = 31,
    C_MATERIAL_CLOSURE_COUNT = 32, C_MATERIAL_FRAME_MODE = 33, C_MATERIAL_TANGENT_ROTATION = 34,
    C_MATERIAL_EMITTED_RADIANCE = 35,
    C_CLOSURE_KIND = 39, C_CLOSURE_LOBES = 40, C_CLOSURE_WEIGHT = 41, C_CLOSURE_PARAMETERS = 45,
    C_CLOSURE_PROBABILITY = 55,
    C_INSTANCE_ID = 56, C_INSTANCE_PARENT_PRESENT = 57, C_INSTANCE_PARENT_ID = 58, C_INSTANCE_OBJECT_ID = 59,
    C_INSTANCE_GEOMETRY_ID = 60, C_INSTANCE_MATERIAL_ID = 61, C_INSTANCE_VISIBILITY_MASK = 62,
    C_INSTANCE_LOCAL_TO_PARENT = 63, C_INSTANCE_PARENT_TO_LOCAL = 79, C_INSTANCE_LOCAL_TO_WORLD = 95,
    C_INSTANCE_WORLD_TO_LOCAL = 111,
    C_PUNCTUAL_KIND = 127, C_PUNCTUAL_POSITION_X = 128, C_PUNCTUAL_POSITION_ERROR_X = 131, C_PUNCTUAL_DIRECTION_X = 134,
    C_PUNCTUAL_INNER = 137, C_PUNCTUAL_OUTER = 138, C_PUNCTUAL_SPECTRUM = 139, C_MESH_AREA_LIGHT_INSTANCE_ID = 143,
    C_ENVIRONMENT_WAVELENGTH_NM = 144, C_ENVIRONMENT_WAVELENGTH_PDF = 148, C_ENVIRONMENT_WAVELENGTH_MEASURE = 152,
    C_ENVIRONMENT_RADIANCE = 156,
    C_TEXTURE_ID = 160, C_TEXTURE_KIND = 161, C_TEXTURE_VALUE = 162,
    C_NORMAL_MAP_PRESENT = 166, C_BUMP_MAP_PRESENT = 176, C_BUMP_MAP_SCALE = 179,
    C_IMAGE_TEXTURE_ID = 184, C_IMAGE_TEXTURE_MIP_OFFSET = 185, C_IMAGE_TEXTURE_MIP_COUNT = 186,
    C_IMAGE_MIP_WIDTH = 189, C_IMAGE_MIP_TEXEL_OFFSET = 191, C_IMAGE_MIP_TEXEL_COUNT = 192,
    C_IMAGE_TEXEL_VALUE = 193,
};

static constexpr uint8_t kMeasureWavelength = 5u; /* ProbabilityMeasure::wavelength (TransportConventions.hpp:11-18) */

/* scene_soa_column_count (SceneSoaAbi.hpp:193-240): the header count that sizes column c. */
BF_SOA_HD uint64_t column_count(const bf_scene_soa_header& h, uint32_t c) {
    if (c == C_OBJECT_ID) return h.object_count;
    if (c >= C_GEOMETRY_ID && c <= C_GEOMETRY_TRIANGLE_COUNT) return h.geometry_count;
    if (c >= C_POSITION_X && c < C_TRIANGLE_VERTEX_0) return h.vertex_count;
    if (c >= C_TRIANGLE_VERTEX_0 && c < C_MATERIAL_ID) return h.triangle_count;
    if (c >= C_MATERIAL_ID && c < C_CLOSURE_KIND) return h.material_count;
    if (c >= C_CLOSURE_KIND && c < C_INSTANCE_ID) return h.closure_count;
    if (c >= C_INSTANCE_ID && c < C_PUNCTUAL_KIND) return h.instance_count;
    if (c >= C_PUNCTUAL_KIND && c < C_MESH_AREA_LIGHT_INSTANCE_ID) return h.punctual_light_count;
    if (c == C_MESH_AREA_LIGHT_INSTANCE_ID) return h.mesh_area_light_count;
    if (c >= C_ENVIRONMENT_WAVELENGTH_NM && c < C_TEXTURE_ID) return h.environment_count;
    if (c >= C_TEXTURE_ID && c < C_NORMAL_MAP_PRESENT) return h.texture_count;
    if (c >= C_NORMAL_MAP_PRESENT && c < C_IMAGE_TEXTURE_ID) return h.material_count;
    if (c >= C_IMAGE_TEXTURE_ID && c < C_IMAGE_MIP_WIDTH) return h.image_texture_count;
    if (c >= C_IMAGE_MIP_WIDTH && c < C_IMAGE_TEXEL_VALUE) return h.image_mip_count;
    if (c == C_IMAGE_TEXEL_VALUE) return h.image_texel_count;
    return 0;
}

/* scene_soa_column_element_size (SceneSoaAbi.hpp:242-276). */
BF_SOA_HD uint32_t column_element_size(uint32_t c) {
    if (c >= BF_SCENE_SOA_COLUMN_COUNT) return 0;
    if ((c >= C_GEOMETRY_VERTEX_OFFSET && c <= C_GEOMETRY_TRIANGLE_COUNT) || c == C_MATERIAL_CLOSURE_OFFSET ||
        c == C_MATERIAL_CLOSURE_COUNT)
        return 8;
    if (c == C_MATERIAL_SPECTRAL_PRESENT || (c >= C_MATERIAL_WAVELENGTH_MEASURE && c < C_MATERIAL_CLOSURE_OFFSET) ||
        c == C_MATERIAL_FRAME_MODE || c == C_INSTANCE_PARENT_PRESENT || c == C_NORMAL_MAP_PRESENT ||
        c == C_BUMP_MAP_PRESENT || (c >= C_ENVIRONMENT_WAVELENGTH_MEASURE && c < C_ENVIRONMENT_RADIANCE))
        return 1;
    if ((c >= C_POSITION_X && c < C_TRIANGLE_VERTEX_0) ||
        (c >= C_MATERIAL_WAVELENGTH_NM && c < C_MATERIAL_WAVELENGTH_MEASURE) || c == C_MATERIAL_TANGENT_ROTATION ||
        (c >= C_MATERIAL_EMITTED_RADIANCE && c < C_CLOSURE_KIND) || (c >= C_CLOSURE_WEIGHT && c < C_INSTANCE_ID) ||
        (c >= C_INSTANCE_LOCAL_TO_PARENT && c < C_PUNCTUAL_KIND) ||
        (c > C_PUNCTUAL_KIND && c < C_MESH_AREA_LIGHT_INSTANCE_ID) ||
        (c >= C_ENVIRONMENT_WAVELENGTH_NM && c < C_ENVIRONMENT_WAVELENGTH_MEASURE) ||
        (c >= C_ENVIRONMENT_RADIANCE && c < C_TEXTURE_ID) || (c >= C_TEXTURE_VALUE && c < C_NORMAL_MAP_PRESENT) ||
        c == C_BUMP_MAP_SCALE || c == C_IMAGE_TEXEL_VALUE)
        return 4;
    if (c == C_IMAGE_TEXTURE_MIP_OFFSET || c == C_IMAGE_TEXTURE_MIP_COUNT || c == C_IMAGE_MIP_TEXEL_OFFSET ||
        c == C_IMAGE_MIP_TEXEL_COUNT)
        return 8;
    return 4;
}

BF_SOA_HD uint64_t align_up(uint64_t v) {
    const uint64_t r = v % BF_SCENE_SOA_COLUMN_ALIGNMENT;
    return r == 0 ? v : v + (BF_SCENE_SOA_COLUMN_ALIGNMENT - r);
}

/* validate_scene_soa_header (SceneSoaAbi.hpp:283-341), check for check. */
BF_SOA_HD uint32_t validate_header(const bf_scene_soa_header& h) {
    if (h.magic != BF_SCENE_SOA_MAGIC) return BF_SCENE_SOA_BAD_MAGIC;
    if (h.abi_major != BF_SCENE_SOA_ABI_MAJOR || h.abi_minor != BF_SCENE_SOA_ABI_MINOR)
        return BF_SCENE_SOA_INCOMPATIBLE_VERSION;
    if (h.header_size != BF_SCENE_SOA_HEADER_SIZE) return BF_SCENE_SOA_BAD_HEADER_SIZE;
    if (h.column_count != BF_SCENE_SOA_COLUMN_COUNT) return BF_SCENE_SOA_BAD_COLUMN_COUNT;
    if (h.hash_algorithm != BF_SCENE_SOA_HASH_FNV1A64) return BF_SCENE_SOA_UNSUPPORTED_HASH_ALGORITHM;
    for (int r = 0; r < 5; ++r)
        if (h.reserved[r] != 0) return BF_SCENE_SOA_NONZERO_RESERVED;
    if (h.environment_count > 1) return BF_SCENE_SOA_INVALID_ENVIRONMENT_COUNT;
    if (h.total_size_bytes < BF_SCENE_SOA_HEADER_SIZE) return BF_SCENE_SOA_BAD_TOTAL_SIZE;
    const uint64_t top = ~0ull;
    uint64_t cursor = align_up(BF_SCENE_SOA_HEADER_SIZE);
    for (uint32_t c = 0; c < BF_SCENE_SOA_COLUMN_COUNT; ++c) {
        const bf_scene_soa_column& d = h.columns[c];
        const uint64_t count = column_count(h, c);
        const uint32_t size = column_element_size(c);
        if (d.element_count != count || d.element_size != size || d.reserved != 0)
            return BF_SCENE_SOA_INVALID_COLUMN_DESCRIPTOR;
        if (count == 0) {
            if (d.offset_bytes != 0) return BF_SCENE_SOA_INVALID_COLUMN_DESCRIPTOR;
            continue;
        }
        if (cursor > top - (BF_SCENE_SOA_COLUMN_ALIGNMENT - 1)) return BF_SCENE_SOA_SIZE_OVERFLOW;
        cursor = align_up(cursor);
        if (d.offset_bytes != cursor) return BF_SCENE_SOA_INVALID_COLUMN_DESCRIPTOR;
        if (count > top / size) return BF_SCENE_SOA_SIZE_OVERFLOW;
        const uint64_t bytes = count * size;
        if (cursor > top - bytes) return BF_SCENE_SOA_SIZE_OVERFLOW;
        cursor += bytes;
    }
    if (h.total_size_bytes != cursor) return BF_SCENE_SOA_BAD_TOTAL_SIZE;
    return BF_SCENE_SOA_VALID;
}

} // namespace bf_soa

#endif
