/* ORACLE — test infrastructure only (see bf_oracle.h). Internal types of the C restatement. */
#ifndef BF_ORACLE_INTERNAL_H
#define BF_ORACLE_INTERNAL_H

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/blackframe_math.h"
#include "bf_oracle.h"

#define BFO_DIM_RASTER_X 0xA24BAED4963EE407ull
#define BFO_DIM_RASTER_Y 0x9FB21C651E98DF25ull
#define BFO_DIM_WAVELENGTH 3ull
#define BFO_FIRST_BOUNCE_DIM 4ull
#define BFO_DIMS_PER_BOUNCE 10ull

#define BFO_FLOAT_EPSILON 0x1p-23f
#define BFO_DENORM_MIN 0x1p-149f
#define BFO_PI 3.14159265358979323846f
#define BFO_HALF_PI (BFO_PI / 2.0f)
#define BFO_INV_PI 0.31830988618379067154f
#define BFO_QUARTER_PI 0.78539816339744830962f

typedef struct {
    float x, y, z;
} v3;

typedef struct {
    float bounds[6]; /* min xyz, max xyz */
    double centroid[3];
    uint32_t stable_key;
    bf_bvh_instance_ref reference;
} bfo_build_item;

typedef struct {
    uint32_t instance;       /* scene instance index */
    uint32_t geometry;       /* scene geometry index */
    uint32_t triangle_count;
    float inverse_area;
    float* areas;
    float* probabilities;
    float* cumulative_before;
    float* next_probabilities;
} bfo_light;

struct bfo_scene {
    bf_scene_desc desc;
    /* owned copies of every description array */
    float* positions;
    float* normals;
    float* uvs;
    uint32_t* indices;
    bf_geometry_desc* geometries;
    bf_instance_desc* instances;
    bf_material_desc* materials;
    bf_closure_desc* closures;
    bf_spectrum_desc* spectra;
    float* spectrum_values;
    /* textures (host copies; desc.textures points at the owned descriptors) */
    bf_texture_desc textures[BF_MAX_TEXTURES];
    float* texels[BF_MAX_TEXTURES];
    /* canonical BVH */
    bf_bvh_blas* blases;
    bf_bvh_node* blas_nodes;
    uint32_t* primitive_refs;
    bf_bvh_node* tlas_nodes;
    bf_bvh_instance_ref* instance_refs;
    uint64_t blas_node_count, primitive_ref_count, tlas_node_count;
    uint32_t tlas_root;
    /* mesh area lights (instances with emissive materials, instance order) */
    uint32_t light_count;
    bfo_light* lights;
    int32_t* instance_light; /* instance -> light index or -1 */
    /* punctual lights, registered before the mesh-area lights; spot-cone cosines (inner, outer) */
    uint32_t punctual_count;
    bf_punctual_desc* punctuals;
    float* punctual_cosines;
    /* power-weighted light selection CDF (punctual_count + light_count + 1 boundaries), or NULL */
    float* light_cdf;
    /* surface maps: image textures (owned texels) and per material its maps, or NULL when no
     * material has one */
    uint32_t image_texture_count;
    bf_image_texture_desc* image_textures;
    bf_surface_maps_desc* surface_maps;
};

/* vector helpers restated from Blackframe/XPU/CUDA/Source/CudaWavefrontStages.cu:298-349 */
static inline v3 v3_make(float x, float y, float z) {
    v3 r = {x, y, z};
    return r;
}
static inline int v3_finite(v3 v) { return bf_isfinite(v.x) && bf_isfinite(v.y) && bf_isfinite(v.z); }
static inline v3 v3_add(v3 a, v3 b) { return v3_make(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 v3_sub(v3 a, v3 b) { return v3_make(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 v3_mul(v3 a, float s) { return v3_make(a.x * s, a.y * s, a.z * s); }
static inline float v3_dot(v3 a, v3 b) { return fmaf(a.x, b.x, fmaf(a.y, b.y, a.z * b.z)); }
static inline v3 v3_cross(v3 a, v3 b) {
    return v3_make(fmaf(a.y, b.z, -a.z * b.y), fmaf(a.z, b.x, -a.x * b.z), fmaf(a.x, b.y, -a.y * b.x));
}
static inline float v3_comp(v3 v, uint32_t axis) { return axis == 0 ? v.x : (axis == 1 ? v.y : v.z); }
static inline void v3_set(v3* v, uint32_t axis, float value) {
    if (axis == 0) v->x = value;
    else if (axis == 1) v->y = value;
    else v->z = value;
}

/* CudaSceneClosestHit.cu:514-536 / CudaWavefrontStages.cu:1024-1034 */
static inline v3 bfo_transform_point(const float* m, v3 p) {
    return v3_make(fmaf(m[0], p.x, fmaf(m[1], p.y, fmaf(m[2], p.z, m[3]))),
                   fmaf(m[4], p.x, fmaf(m[5], p.y, fmaf(m[6], p.z, m[7]))),
                   fmaf(m[8], p.x, fmaf(m[9], p.y, fmaf(m[10], p.z, m[11]))));
}
static inline v3 bfo_transform_vector(const float* m, v3 v) {
    return v3_make(fmaf(m[0], v.x, fmaf(m[1], v.y, m[2] * v.z)), fmaf(m[4], v.x, fmaf(m[5], v.y, m[6] * v.z)),
                   fmaf(m[8], v.x, fmaf(m[9], v.y, m[10] * v.z)));
}

v3 bfo_vertex(const bfo_scene* scene, uint32_t geometry, uint32_t local_vertex);
int bfo_triangle_vertices(const bfo_scene* scene, uint32_t geometry, uint32_t primitive, uint32_t idx[3]);
int bfo_triangle_area_v(const v3* vertices, float* area, v3* normal);
int bfo_world_normal(const bfo_scene* scene, uint32_t instance, const v3* local_vertices, v3* normal);

#endif
