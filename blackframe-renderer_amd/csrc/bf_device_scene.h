/*
 * Device-resident scene and path-state layout of the HIP wavefront backend.
 *
 * Everything the bounce kernels touch is a flat array in HBM, laid out for 64-lane coalescing:
 *  - wide BVH nodes: one 64-byte line per internal node holding both child boxes (bf_wide_node);
 *  - leaf triangles: 48-byte records in BLAS leaf order (3 x float4: vertex xyz + payload word),
 *    so a leaf costs one 48-byte fetch and never a node fetch;
 *  - TLAS leaves: one 64-byte record per instance reference (world->local 3x4 + ids + mask);
 *  - path state: structure-of-arrays columns indexed by path slot, float4-wide where possible.
 */
#ifndef BF_DEVICE_SCENE_H
#define BF_DEVICE_SCENE_H

#include <stdint.h>

#include "../../include/blackframe_hip_abi.h"

/* The reference keeps one 64-entry stack for the TLAS and one for each BLAS it enters
 * (CudaSceneClosestHit.cu:44, 941, 1092) and reports stack_overflow for a lane that needs more.
 * Here one stack per lane holds both parts: BF_STRUCTURE_STACK entries per structure, checked per
 * structure, BF_STACK_DEPTH entries in all. */
#define BF_STRUCTURE_STACK 64u
#define BF_STACK_DEPTH 128u  /* traversal stack entries per lane (TLAS + BLAS) */
#define BF_API_STACK 64u     /* LDS entries per lane of the stand-alone traversal launchers (no spill) */
#define BF_API_BLOCK 64u     /* lanes per workgroup of the stand-alone traversal launchers */
/* A traversal stack entry is (reference, near word). Reference bits: 31 BF_WIDE_LEAF, 30
 * BF_NEAR_EXACT (the near word is the binary64 near value rounded up to binary32; otherwise it
 * holds bounds of it), 0-29 the wide-node or leaf index. */
#define BF_NEAR_EXACT 0x40000000u
#define BF_REF_MASK 0x3FFFFFFFu
#ifndef BF_SHORT_STACK
#define BF_SHORT_STACK 8u    /* of which the transport kernels keep this many in LDS */
#endif
#define BF_TRACE_BLOCK 128u  /* lanes per traversal workgroup */
#ifndef BF_TRACE_WAVES_PER_SIMD
#define BF_TRACE_WAVES_PER_SIMD 6 /* register budget of the closest-hit traversal (waves per SIMD) */
#endif
#ifndef BF_LEAF_PHASE
/* Traversal waves step interior nodes and leaves in separate iterations: leaves when at least this
 * many lanes (or every active lane) have one next, interior nodes otherwise (0: every kind in
 * every iteration, each under its own lane mask). 16 measured best of {1, 4, 8, 12, 16, 32, 64}. */
#define BF_LEAF_PHASE 16u
#endif
#ifndef BF_SHADOW_LEAF_PHASE
#define BF_SHADOW_LEAF_PHASE 16u /* the same for any-hit waves (8, 24 and 32 measured slower) */
#endif
#ifndef BF_WORLD_INV_CACHE
#define BF_WORLD_INV_CACHE 1 /* closest-hit lanes keep their world-space reciprocals in LDS */
#endif
#ifndef BF_SHADOW_WAVES_PER_SIMD
#define BF_SHADOW_WAVES_PER_SIMD 7 /* the any-hit traversal keeps less state than closest hit; 7 waves (72 VGPRs) measured fastest */
#endif
#ifndef BF_TRACE_REFILL
#define BF_TRACE_REFILL 24u /* idle lanes of a closest-hit wave that trigger a fetch of new rays (16, 20, 28, 32, 40 measured slower) */
#endif
#ifndef BF_SHADOW_REFILL
#define BF_SHADOW_REFILL 16u /* the same for an any-hit wave (8 and 24 measured slower) */
#endif
#ifndef BF_SHADE_WAVES_PER_SIMD
#define BF_SHADE_WAVES_PER_SIMD 3 /* register budget of the forward shade kernels (waves per SIMD; 1 and 2 measured slower) */
#endif
#ifndef BF_SHADE_BLOCK
#define BF_SHADE_BLOCK 256u /* lanes per shade workgroup (64 measured slower, 128 equal) */
#endif
#ifndef BF_ADJOINT_SHADE_WAVES_PER_SIMD
#define BF_ADJOINT_SHADE_WAVES_PER_SIMD 3 /* the adjoint shade kernel (1 and 2 measured slower) */
#endif

/* A punctual light with its spot-cone cosines formed once on the host. */
typedef struct bf_dev_punctual {
    uint32_t kind, spectrum;
    float position[3], position_error[3], direction[3];
    float cosine_inner, cosine_outer;
    uint32_t smooth; /* spot with inner != outer: smoothstep falloff between the cones */
} bf_dev_punctual;

typedef struct bf_dev_blas {
    uint32_t root_ref;         /* wide-node index, or BF_WIDE_LEAF | leaf-triangle index */
    uint32_t geometry_index;
    uint32_t canonical_node_offset;
    uint32_t depth;
    float root_box[6];         /* min xyz, max xyz of the canonical root node */
    uint32_t reserved[2];
} bf_dev_blas;

typedef struct bf_dev_tlas_leaf {
    float world_to_local[12];
    uint32_t scene_instance;
    uint32_t instance_id;
    uint32_t blas_index;
    uint32_t visibility_mask;
} bf_dev_tlas_leaf;

typedef struct bf_dev_instance {
    float local_to_world[12];
    float world_to_local[12];
    uint32_t geometry_index;
    uint32_t material_index;
    uint32_t id;
    uint32_t object_id;
    uint32_t visibility_mask;
    int32_t light_index;       /* mesh-area light registry slot or -1 */
    uint32_t reserved[2];
} bf_dev_instance;

typedef struct bf_dev_geometry {
    uint32_t id;
    uint32_t vertex_offset;
    uint32_t vertex_count;
    uint32_t triangle_offset;
    uint32_t triangle_count;
    uint32_t reserved[3];
} bf_dev_geometry;

typedef struct bf_dev_light {
    uint32_t instance;
    uint32_t geometry;
    uint32_t triangle_count;
    uint32_t table_offset;     /* into area / probability / cumulative / next tables */
    float inverse_area;
    uint32_t reserved[3];
} bf_dev_light;

typedef struct bf_dev_texture {
    const float* data;         /* texels, (y * width + x) * channels + c */
    uint32_t width, height, channels, reserved;
} bf_dev_texture;

/* An image texture of the surface maps: level 0 extent, channels, levels, and the index of its
 * first value in the scene's texel array (levels follow each other, halving). */
typedef struct bf_dev_image {
    uint32_t width, height, channels, mips;
    uint64_t texel_base;
} bf_dev_image;

typedef struct bf_dev_scene {
    /* traversal */
    const bf_wide_node* tlas_nodes;
    const bf_wide_node* blas_nodes;
    const bf_dev_blas* blases;
    const float4* leaf_triangles;      /* 3 float4 per BLAS leaf; [0].w = primitive (bits) */
    const bf_dev_tlas_leaf* tlas_leaves;
    /* every wide node's and leaf's own box (6 floats: lo xyz, hi xyz), by wide index and by leaf
     * reference: read only when the near bounds a stack entry keeps cannot decide its pop test */
    const float* blas_node_boxes;
    const float* blas_leaf_boxes;
    const float* tlas_node_boxes;
    const float* tlas_leaf_boxes;
    uint32_t tlas_root_ref;            /* BF_INVALID_INDEX when the scene has no instances */
    uint32_t instance_count;
    float tlas_root_box[6];
    /* shading */
    const bf_dev_instance* instances;
    const bf_dev_geometry* geometries;
    const float* positions;
    const float* normals;
    const float* uvs;
    const uint32_t* indices;
    const bf_material_desc* materials;
    const bf_closure_desc* closures;
    const bf_spectrum_desc* spectra;
    const float* spectrum_values;
    const bf_dev_light* lights;
    const float* light_areas;
    const float* light_probabilities;
    const float* light_cumulative;
    const float* light_next;
    const float4* light_normals;       /* exact world normal of every emitter triangle */
    const bf_dev_punctual* punctuals;
    uint32_t punctual_count;           /* registered lights 0 .. punctual_count-1 */
    uint32_t light_count;              /* punctual + mesh-area lights; lights[i - punctual_count] */
    const float* light_cdf;            /* power-weighted selection CDF (light_count + 1), or null: uniform */
    uint32_t has_environment;
    uint32_t environment_spectrum;
    uint32_t material_count;
    const bf_dev_texture* textures;
    uint32_t texture_count;
    uint32_t has_textures;             /* some closure binds a texture */
    /* surface maps: images, their texels and per material its maps; maps is null when no material
     * has one */
    const bf_dev_image* images;
    const float* image_texels;
    const bf_surface_maps_desc* maps;
} bf_dev_scene;

/* Path-state columns (one entry per path slot). */
typedef struct bf_dev_paths {
    float4* ray_origin;      /* xyz origin, w = ray time */
    float4* ray_direction;   /* xyz direction, w = t_min */
    float4* beta;
    float4* radiance;
    float4* wavelengths;
    uint4* depth;            /* x: packed counters, y: depth, z: delta flags, w: eta scale bits */
    float4* previous;        /* xyz context, w = pdf (0 = no previous continuous sample) */
    float4* hit;             /* t, b0, b1, b2 */
    uint2* hit_ids;          /* scene instance, primitive */
    float4* shadow_origin;   /* xyz, w = t_max */
    float4* shadow_direction;/* xyz, w = bits(continuation | termination << 8) */
    float4* shadow_contribution;
    uint32_t* status;        /* termination | failure code << 8 */
    uint32_t* visibility;    /* ray visibility mask */
    uint16_t* job;           /* job (frame) of the slot in a batched launch */
    float2* cone;            /* ray cone (width, spread) of path-input frames */
} bf_dev_paths;

#endif
