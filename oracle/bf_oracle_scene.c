/*
 * ORACLE — test infrastructure only (see bf_oracle.h).
 * Scene ingestion, the deterministic BVH builder and BVH traversal restated from
 *   Blackframe/Backends/GPU/CUDA/Source/SceneBvh.cpp            (builder)
 *   Blackframe/XPU/CUDA/Source/CudaSceneClosestHit.cu           (closest-hit / any-hit traversal)
 *   Blackframe/XPU/CUDA/Source/CudaWavefrontStages.cu:1729-2042 (exact triangle area)
 */
#include "bf_oracle_internal.h"

#include <stdio.h>

#define DUP(dst, src, n, type)                                            \
    do {                                                                  \
        (dst) = (type*)malloc((((n) != 0) ? (n) : 1) * sizeof(type));            \
        if (((n) != 0) && (src)) memcpy((dst), (src), (n) * sizeof(type));       \
    } while (0)

v3 bfo_vertex(const bfo_scene* scene, uint32_t geometry, uint32_t local_vertex) {
    const uint64_t v = scene->geometries[geometry].vertex_offset + local_vertex;
    return v3_make(scene->positions[3 * v], scene->positions[3 * v + 1], scene->positions[3 * v + 2]);
}

int bfo_triangle_vertices(const bfo_scene* scene, uint32_t geometry, uint32_t primitive, uint32_t idx[3]) {
    const bf_geometry_desc* g = &scene->geometries[geometry];
    if (primitive >= g->triangle_count) return 0;
    const uint64_t t = g->triangle_offset + primitive;
    for (int c = 0; c < 3; ++c) {
        idx[c] = scene->indices[3 * t + c];
        if (idx[c] >= g->vertex_count) return 0;
    }
    return 1;
}

/* ---- BVH builder (SceneBvh.cpp:86-140, 355-427, 477-736) ------------------------------------ */
static float canonical_zero(float v) { return v == 0.0f ? 0.0f : v; }

static void centroid_of(const float* b, double* c) {
    for (int a = 0; a < 3; ++a) c[a] = ((double)b[a] + (double)b[3 + a]) * 0.5;
}

/* (centroid on the split axis, stable key): the axis comes in as the comparator's argument, so
 * builds on several threads do not share state */
static int compare_items(const void* lp, const void* rp, void* axis_arg) {
    const uint32_t axis = *(const uint32_t*)axis_arg;
    const bfo_build_item* l = (const bfo_build_item*)lp;
    const bfo_build_item* r = (const bfo_build_item*)rp;
    if (l->centroid[axis] < r->centroid[axis]) return -1;
    if (r->centroid[axis] < l->centroid[axis]) return 1;
    return l->stable_key < r->stable_key ? -1 : (l->stable_key > r->stable_key ? 1 : 0);
}

typedef struct {
    bf_bvh_node* nodes;
    uint64_t count;
    /* emit: blas -> primitive refs, tlas -> instance refs */
    uint32_t* primitive_refs;
    uint64_t* primitive_ref_count;
    bf_bvh_instance_ref* instance_refs;
    uint64_t* instance_ref_count;
} build_ctx;

static void make_node(const float* b, bf_bvh_node* n) {
    memset(n, 0, sizeof *n);
    n->min_x = b[0];
    n->min_y = b[1];
    n->min_z = b[2];
    n->max_x = b[3];
    n->max_y = b[4];
    n->max_z = b[5];
    n->kind = 0;
    n->split_axis = BF_INVALID_INDEX;
    n->first_child = BF_INVALID_INDEX;
    n->second_child = BF_INVALID_INDEX;
}

static uint32_t build_node(bfo_build_item* items, size_t begin, size_t end, build_ctx* ctx) {
    const uint32_t node_index = (uint32_t)ctx->count++;
    if (end - begin == 1) {
        bf_bvh_node node;
        make_node(items[begin].bounds, &node);
        node.kind = 1;
        if (ctx->primitive_refs) {
            node.reference_offset = (uint32_t)(*ctx->primitive_ref_count);
            ctx->primitive_refs[(*ctx->primitive_ref_count)++] = items[begin].stable_key;
        } else {
            node.reference_offset = (uint32_t)(*ctx->instance_ref_count);
            ctx->instance_refs[(*ctx->instance_ref_count)++] = items[begin].reference;
        }
        node.reference_count = 1;
        ctx->nodes[node_index] = node;
        return node_index;
    }
    double cmin[3], cmax[3];
    for (int a = 0; a < 3; ++a) cmin[a] = cmax[a] = items[begin].centroid[a];
    for (size_t i = begin + 1; i < end; ++i) {
        for (int a = 0; a < 3; ++a) {
            cmin[a] = fmin(cmin[a], items[i].centroid[a]);
            cmax[a] = fmax(cmax[a], items[i].centroid[a]);
        }
    }
    uint32_t axis = 0;
    double largest = cmax[0] - cmin[0];
    for (uint32_t a = 1; a < 3; ++a) {
        const double extent = cmax[a] - cmin[a];
        if (extent > largest) {
            largest = extent;
            axis = a;
        }
    }
    qsort_r(items + begin, end - begin, sizeof(bfo_build_item), compare_items, &axis);
    const size_t middle = begin + (end - begin) / 2;
    const uint32_t first = build_node(items, begin, middle, ctx);
    const uint32_t second = build_node(items, middle, end, ctx);
    const bf_bvh_node* f = &ctx->nodes[first];
    const bf_bvh_node* s = &ctx->nodes[second];
    float b[6] = {
        canonical_zero(fminf(f->min_x, s->min_x)), canonical_zero(fminf(f->min_y, s->min_y)),
        canonical_zero(fminf(f->min_z, s->min_z)), canonical_zero(fmaxf(f->max_x, s->max_x)),
        canonical_zero(fmaxf(f->max_y, s->max_y)), canonical_zero(fmaxf(f->max_z, s->max_z)),
    };
    bf_bvh_node node;
    make_node(b, &node);
    node.split_axis = axis;
    node.first_child = first;
    node.second_child = second;
    ctx->nodes[node_index] = node;
    return node_index;
}

/* SceneBvh.cpp:316-352 */
static int primitive_bounds(const bfo_scene* scene, uint32_t geometry, uint32_t primitive, float* b) {
    uint32_t idx[3];
    if (!bfo_triangle_vertices(scene, geometry, primitive, idx)) return 0;
    const v3 first = bfo_vertex(scene, geometry, idx[0]);
    b[0] = b[3] = canonical_zero(first.x);
    b[1] = b[4] = canonical_zero(first.y);
    b[2] = b[5] = canonical_zero(first.z);
    for (int c = 1; c < 3; ++c) {
        const v3 p = bfo_vertex(scene, geometry, idx[c]);
        const float comp[3] = {p.x, p.y, p.z};
        for (int a = 0; a < 3; ++a) {
            if (!bf_isfinite(comp[a])) return 0;
            b[a] = canonical_zero(fminf(b[a], comp[a]));
            b[3 + a] = canonical_zero(fmaxf(b[3 + a], comp[a]));
        }
    }
    return 1;
}

/* SceneBvh.cpp:555-601: conservative binary64-bounded world bounds of one transformed point */
static int transform_point_bounds(const float* m, v3 point, float* b) {
    const double eps = (double)BFO_FLOAT_EPSILON;
    const double gamma7 = (7.0 * eps) / (1.0 - 7.0 * eps);
    const double underflow_allowance = 7.0 * (double)BFO_DENORM_MIN;
    const double maximum = (double)3.40282346638528859812e+38f;
    const float src[3] = {point.x, point.y, point.z};
    for (int row = 0; row < 3; ++row) {
        double exact = (double)m[row * 4 + 3];
        double magnitude = fabs(exact);
        for (int col = 0; col < 3; ++col) {
            const double product = (double)m[row * 4 + col] * (double)src[col];
            exact += product;
            magnitude += fabs(product);
        }
        const double error = fma(gamma7, magnitude, underflow_allowance);
        const double lower = exact - error;
        const double upper = exact + error;
        if (!bf_isfinited(exact) || !bf_isfinited(magnitude) || !bf_isfinited(error) || lower < -maximum ||
            upper > maximum)
            return 0;
        float rl = (float)lower;
        float ru = (float)upper;
        if ((double)rl > lower) rl = bf_next_down_f(rl);
        if ((double)ru < upper) ru = bf_next_up_f(ru);
        b[row] = canonical_zero(rl);
        b[3 + row] = canonical_zero(ru);
    }
    return 1;
}

static int build_bvh(bfo_scene* scene) {
    const bf_scene_desc* d = &scene->desc;
    uint64_t blas_nodes = 0;
    for (uint32_t g = 0; g < d->geometry_count; ++g) blas_nodes += 2 * scene->geometries[g].triangle_count - 1;
    scene->blases = (bf_bvh_blas*)calloc(d->geometry_count ? d->geometry_count : 1, sizeof(bf_bvh_blas));
    scene->blas_nodes = (bf_bvh_node*)calloc(blas_nodes ? blas_nodes : 1, sizeof(bf_bvh_node));
    scene->primitive_refs = (uint32_t*)calloc(d->triangle_count ? d->triangle_count : 1, sizeof(uint32_t));
    scene->tlas_nodes = (bf_bvh_node*)calloc(d->instance_count ? 2 * d->instance_count : 1, sizeof(bf_bvh_node));
    scene->instance_refs = (bf_bvh_instance_ref*)calloc(d->instance_count ? d->instance_count : 1, sizeof(bf_bvh_instance_ref));
    uint64_t prim_count = 0, inst_count = 0;
    build_ctx ctx = {scene->blas_nodes, 0, scene->primitive_refs, &prim_count, NULL, NULL};
    /* SceneBvh.cpp:477-521 */
    for (uint32_t g = 0; g < d->geometry_count; ++g) {
        const bf_geometry_desc* geo = &scene->geometries[g];
        /* SceneBvh.cpp:264-268: strictly ordered geometry identifiers */
        if (g > 0 && geo->id <= scene->geometries[g - 1].id) return BF_INVALID_ARGUMENT;
        if (geo->triangle_count == 0 || geo->vertex_count == 0) return BF_INVALID_ARGUMENT;
        bfo_build_item* items = (bfo_build_item*)calloc(geo->triangle_count, sizeof(bfo_build_item));
        for (uint32_t p = 0; p < geo->triangle_count; ++p) {
            if (!primitive_bounds(scene, g, p, items[p].bounds)) {
                free(items);
                return BF_INVALID_ARGUMENT;
            }
            centroid_of(items[p].bounds, items[p].centroid);
            items[p].stable_key = p;
        }
        const uint32_t node_offset = (uint32_t)ctx.count;
        const uint32_t ref_offset = (uint32_t)prim_count;
        const uint32_t root = build_node(items, 0, geo->triangle_count, &ctx);
        free(items);
        bf_bvh_blas* blas = &scene->blases[g];
        memset(blas, 0, sizeof *blas);
        blas->geometry_id = geo->id;
        blas->root_node = root;
        blas->node_offset = node_offset;
        blas->node_count = (uint32_t)(ctx.count - node_offset);
        blas->primitive_reference_offset = ref_offset;
        blas->primitive_reference_count = (uint32_t)geo->triangle_count;
    }
    scene->blas_node_count = ctx.count;
    scene->primitive_ref_count = prim_count;
    /* SceneBvh.cpp:603-736 */
    scene->tlas_root = BF_INVALID_INDEX;
    scene->tlas_node_count = 0;
    if (d->instance_count == 0) return BF_OK;
    bfo_build_item* items = (bfo_build_item*)calloc(d->instance_count, sizeof(bfo_build_item));
    for (uint32_t i = 0; i < d->instance_count; ++i) {
        const bf_instance_desc* inst = &scene->instances[i];
        /* SceneBvh.cpp:677-683: strictly ordered instance identifiers */
        if (i > 0 && inst->id <= scene->instances[i - 1].id) {
            free(items);
            return BF_INVALID_ARGUMENT;
        }
        const uint32_t g = inst->geometry_index;
        const bf_geometry_desc* geo = &scene->geometries[g];
        int have = 0;
        float b[6];
        for (uint32_t p = 0; p < geo->triangle_count; ++p) {
            uint32_t idx[3];
            if (!bfo_triangle_vertices(scene, g, p, idx)) {
                free(items);
                return BF_INVALID_ARGUMENT;
            }
            for (int c = 0; c < 3; ++c) {
                float pb[6];
                if (!transform_point_bounds(inst->local_to_world, bfo_vertex(scene, g, idx[c]), pb)) {
                    free(items);
                    return BF_INVALID_ARGUMENT;
                }
                if (!have) {
                    memcpy(b, pb, sizeof b);
                    have = 1;
                } else {
                    for (int a = 0; a < 3; ++a) {
                        b[a] = canonical_zero(fminf(b[a], pb[a]));
                        b[3 + a] = canonical_zero(fmaxf(b[3 + a], pb[3 + a]));
                    }
                }
            }
        }
        memcpy(items[i].bounds, b, sizeof b);
        centroid_of(b, items[i].centroid);
        items[i].stable_key = inst->id;
        bf_bvh_instance_ref* r = &items[i].reference;
        memset(r, 0, sizeof *r);
        r->scene_instance_index = i;
        r->instance_id = inst->id;
        r->scene_geometry_index = g;
        r->geometry_id = geo->id;
        r->blas_index = g;
        r->visibility_mask = inst->visibility_mask;
    }
    build_ctx tctx = {scene->tlas_nodes, 0, NULL, NULL, scene->instance_refs, &inst_count};
    scene->tlas_root = build_node(items, 0, d->instance_count, &tctx);
    scene->tlas_node_count = tctx.count;
    free(items);
    return BF_OK;
}

/* ---- exact triangle area (CudaWavefrontStages.cu:1724-2042) ---------------------------------- */
typedef struct {
    float c[32];
    uint32_t size;
    uint32_t capacity;
} expansion;

static int normal_float(float v) { return bf_isfinite(v) && fabsf(v) >= 0x1p-126f; }

static int add_component(expansion* e, float component) {
    if (!bf_isfinite(component)) return 0;
    float accumulated = component;
    uint32_t out = 0;
    const uint32_t in_size = e->size;
    for (uint32_t i = 0; i < in_size; ++i) {
        const float existing = e->c[i];
        const float rounded_sum = accumulated + existing;
        const float existing_from_sum = rounded_sum - accumulated;
        const float accumulated_from_sum = rounded_sum - existing_from_sum;
        const float existing_remainder = existing - existing_from_sum;
        const float accumulated_remainder = accumulated - accumulated_from_sum;
        const float error = existing_remainder + accumulated_remainder;
        if (!bf_isfinite(rounded_sum) || !bf_isfinite(error)) return 0;
        if (error != 0.0f) {
            if (out == e->capacity) return 0;
            e->c[out++] = error;
        }
        accumulated = rounded_sum;
    }
    if (accumulated != 0.0f || out == 0) {
        if (out == e->capacity) return 0;
        e->c[out++] = accumulated;
    }
    e->size = out;
    return 1;
}

static int add_product(expansion* e, float l, float r, float sign) {
    const float rounded = l * r;
    const float err = fmaf(l, r, -rounded);
    if (!bf_isfinite(rounded) || !bf_isfinite(err) || (l != 0.0f && r != 0.0f && !normal_float(rounded)) ||
        (err != 0.0f && !normal_float(err)))
        return 0;
    return add_component(e, sign * err) && add_component(e, sign * rounded);
}

static int multiply_expansions(const expansion* l, const expansion* r, expansion* p) {
    p->size = 0;
    for (uint32_t i = 0; i < l->size; ++i)
        for (uint32_t j = 0; j < r->size; ++j)
            if (!add_product(p, l->c[i], r->c[j], 1.0f)) return 0;
    return 1;
}

static int expansion_sign(const expansion* e) {
    for (uint32_t i = e->size; i > 0; --i) {
        const float c = e->c[i - 1];
        if (c != 0.0f) return c > 0.0f ? 1 : -1;
    }
    return 0;
}

static int resolve(const expansion* e, int* sign, float* value) {
    const int s = expansion_sign(e);
    if (s == 0) {
        *sign = 0;
        *value = 0.0f;
        return 1;
    }
    float resolved = 0.0f;
    for (uint32_t i = 0; i < e->size; ++i) resolved += e->c[i];
    if (!bf_isfinite(resolved) || resolved == 0.0f || (resolved > 0.0f ? 1 : -1) != s) return 0;
    *sign = s;
    *value = resolved;
    return 1;
}

static int exact_difference(float l, float r, expansion* d) {
    d->size = 0;
    d->capacity = 2;
    return add_component(d, l) && add_component(d, -r);
}

static int resolved_difference_of_products(const expansion* a, const expansion* b, const expansion* c,
                                           const expansion* d, int* sign, float* value) {
    expansion p1 = {.size = 0, .capacity = 16}, p2 = {.size = 0, .capacity = 16}, diff = {.size = 0, .capacity = 32};
    if (!multiply_expansions(a, b, &p1) || !multiply_expansions(c, d, &p2)) return 0;
    for (uint32_t i = 0; i < p1.size; ++i)
        if (!add_component(&diff, p1.c[i])) return 0;
    for (uint32_t i = 0; i < p2.size; ++i)
        if (!add_component(&diff, -1.0f * p2.c[i])) return 0;
    return resolve(&diff, sign, value);
}

static int scaling_exponent(const v3* v, uint32_t axis) {
    float m = 0.0f;
    for (int i = 0; i < 3; ++i) m = fmaxf(m, fabsf(v3_comp(v[i], axis)));
    int e = 0;
    (void)bf_frexpf(m, &e);
    return e;
}

static int preserves(float original, float scaled) { return original == 0.0f || normal_float(scaled); }

int bfo_triangle_area_v(const v3* vertices, float* area, v3* normal) {
    const int ex = scaling_exponent(vertices, 0), ey = scaling_exponent(vertices, 1), ez = scaling_exponent(vertices, 2);
    v3 s[3];
    for (int i = 0; i < 3; ++i) {
        s[i] = v3_make(bf_ldexpf(vertices[i].x, -ex), bf_ldexpf(vertices[i].y, -ey), bf_ldexpf(vertices[i].z, -ez));
        if (!preserves(vertices[i].x, s[i].x) || !preserves(vertices[i].y, s[i].y) || !preserves(vertices[i].z, s[i].z))
            return 0;
    }
    expansion f[3], g[3];
    for (int a = 0; a < 3; ++a) {
        if (!exact_difference(v3_comp(s[1], a), v3_comp(s[0], a), &f[a]) ||
            !exact_difference(v3_comp(s[2], a), v3_comp(s[0], a), &g[a]))
            return 0;
    }
    int xs, ys, zs;
    float xv, yv, zv;
    if (!resolved_difference_of_products(&f[1], &g[2], &f[2], &g[1], &xs, &xv) ||
        !resolved_difference_of_products(&f[2], &g[0], &f[0], &g[2], &ys, &yv) ||
        !resolved_difference_of_products(&f[0], &g[1], &f[1], &g[0], &zs, &zv))
        return 0;
    const int cx = ey + ez, cy = ez + ex, cz = ex + ey;
    int mag = (-2147483647 - 1);
    int e;
    if (xs != 0) {
        (void)bf_frexpf(fabsf(xv), &e);
        if (e + cx > mag) mag = e + cx;
    }
    if (ys != 0) {
        (void)bf_frexpf(fabsf(yv), &e);
        if (e + cy > mag) mag = e + cy;
    }
    if (zs != 0) {
        (void)bf_frexpf(fabsf(zv), &e);
        if (e + cz > mag) mag = e + cz;
    }
    if (mag == (-2147483647 - 1)) return 0;
    const float ax = xs == 0 ? 0.0f : bf_ldexpf(xv, cx - mag);
    const float ay = ys == 0 ? 0.0f : bf_ldexpf(yv, cy - mag);
    const float az = zs == 0 ? 0.0f : bf_ldexpf(zv, cz - mag);
    if ((xs != 0 && ax == 0.0f) || (ys != 0 && ay == 0.0f) || (zs != 0 && az == 0.0f) || !bf_isfinite(ax) ||
        !bf_isfinite(ay) || !bf_isfinite(az))
        return 0;
    const float maxc = fmaxf(fabsf(ax), fmaxf(fabsf(ay), fabsf(az)));
    if (!bf_isfinite(maxc) || !(maxc > 0.0f)) return 0;
    const float sx = ax / maxc, sy = ay / maxc, sz = az / maxc;
    const float sq = fmaf(sx, sx, fmaf(sy, sy, sz * sz));
    const float sl = sqrtf(sq);
    const float al = maxc * sl;
    *area = bf_ldexpf(al * 0.5f, mag);
    if (!bf_isfinite(sl) || !(sl > 0.0f) || !bf_isfinite(al) || !(al > 0.0f) || !bf_isfinite(*area) || !(*area > 0.0f))
        return 0;
    if (normal) {
        if (!preserves(xv, ax) || !preserves(yv, ay) || !preserves(zv, az)) return 0;
        *normal = v3_make(sx / sl, sy / sl, sz / sl);
        if (!v3_finite(*normal)) return 0;
    }
    return 1;
}

int bfo_triangle_area(const float vertices[9], float* area, float normal[3]) {
    v3 v[3] = {v3_make(vertices[0], vertices[1], vertices[2]), v3_make(vertices[3], vertices[4], vertices[5]),
               v3_make(vertices[6], vertices[7], vertices[8])};
    v3 n;
    const int ok = bfo_triangle_area_v(v, area, normal ? &n : NULL);
    if (ok && normal) {
        normal[0] = n.x;
        normal[1] = n.y;
        normal[2] = n.z;
    }
    return ok;
}

/* Mesh-area-light tables: the per-sample loops of CudaWavefrontStages.cu:2085-2147 (area sum with
 * compensation, per-triangle probability and cumulative bound) and :2487-2505 (inverse area)
 * depend only on the scene, so they are evaluated once here with the identical float sequence. */
static int build_lights(bfo_scene* scene) {
    const bf_scene_desc* d = &scene->desc;
    scene->instance_light = (int32_t*)malloc((d->instance_count ? d->instance_count : 1) * sizeof(int32_t));
    scene->lights = (bfo_light*)calloc(d->instance_count ? d->instance_count : 1, sizeof(bfo_light));
    scene->light_count = 0;
    for (uint32_t i = 0; i < d->instance_count; ++i) {
        scene->instance_light[i] = -1;
        const bf_material_desc* m = &scene->materials[scene->instances[i].material_index];
        if (m->emission_spectrum == BF_INVALID_INDEX) continue;
        const bf_spectrum_desc* sp = &scene->spectra[m->emission_spectrum];
        int black = 1;
        const int sampled = (sp->count & BF_SPECTRUM_SAMPLED) != 0;
        const uint32_t first = sampled ? 4u : 0u, count = sampled ? 4u : sp->count;
        for (uint32_t k = 0; k < count; ++k)
            if (scene->spectrum_values[sp->offset + first + k] != 0.0f) black = 0;
        if (black) continue;
        bfo_light* light = &scene->lights[scene->light_count];
        light->instance = i;
        light->geometry = scene->instances[i].geometry_index;
        const uint32_t n = (uint32_t)scene->geometries[light->geometry].triangle_count;
        light->triangle_count = n;
        light->areas = (float*)malloc(n * sizeof(float));
        light->probabilities = (float*)malloc(n * sizeof(float));
        light->cumulative_before = (float*)malloc(n * sizeof(float));
        light->next_probabilities = (float*)malloc(n * sizeof(float));
        float sum = 0.0f, comp = 0.0f;
        for (uint32_t t = 0; t < n; ++t) {
            uint32_t idx[3];
            v3 w[3];
            if (!bfo_triangle_vertices(scene, light->geometry, t, idx)) return BF_INVALID_ARGUMENT;
            for (int c = 0; c < 3; ++c)
                w[c] = bfo_transform_point(scene->instances[i].local_to_world, bfo_vertex(scene, light->geometry, idx[c]));
            float area;
            if (!bfo_triangle_area_v(w, &area, NULL)) return BF_INVALID_ARGUMENT;
            light->areas[t] = area;
            const float next = sum + area;
            comp += fabsf(sum) >= fabsf(area) ? (sum - next) + area : (area - next) + sum;
            sum = next;
        }
        const float total = sum + comp;
        light->inverse_area = 1.0f / total;
        float cumulative = 0.0f;
        for (uint32_t t = 0; t < n; ++t) {
            float probability = 1.0f - cumulative;
            float next = 1.0f;
            if (t + 1 != n) {
                probability = light->areas[t] / total;
                next = cumulative + probability;
                if (!(probability > 0.0f) || !(next > cumulative) || !(next < 1.0f)) return BF_INVALID_ARGUMENT;
            } else if (!(probability > 0.0f) || probability > 1.0f) {
                return BF_INVALID_ARGUMENT;
            }
            light->probabilities[t] = probability;
            light->cumulative_before[t] = cumulative;
            light->next_probabilities[t] = next;
            cumulative = next;
        }
        scene->instance_light[i] = (int32_t)scene->light_count;
        scene->light_count++;
    }
    return BF_OK;
}

int bfo_scene_build(const bf_scene_desc* desc, bfo_scene** out) {
    bfo_scene* s = (bfo_scene*)calloc(1, sizeof(bfo_scene));
    s->desc = *desc;
    DUP(s->positions, desc->positions, desc->vertex_count * 3, float);
    DUP(s->normals, desc->normals, desc->vertex_count * 3, float);
    DUP(s->uvs, desc->uvs, desc->vertex_count * 2, float);
    DUP(s->indices, desc->indices, desc->triangle_count * 3, uint32_t);
    DUP(s->geometries, desc->geometries, desc->geometry_count, bf_geometry_desc);
    DUP(s->instances, desc->instances, desc->instance_count, bf_instance_desc);
    DUP(s->materials, desc->materials, desc->material_count, bf_material_desc);
    DUP(s->closures, desc->closures, desc->closure_count, bf_closure_desc);
    DUP(s->spectra, desc->spectra, desc->spectrum_count, bf_spectrum_desc);
    DUP(s->spectrum_values, desc->spectrum_values, desc->spectrum_value_count, float);
    /* textures: the rules of the product's scene builder; the oracle reads host texels only */
    int status = BF_OK;
    if (desc->texture_count > BF_MAX_TEXTURES || (desc->texture_count && !desc->textures)) status = BF_INVALID_ARGUMENT;
    for (uint32_t t = 0; t < desc->texture_count && status == BF_OK; ++t) {
        const bf_texture_desc* T = &desc->textures[t];
        if (T->width == 0 || T->height == 0 || T->channels == 0 || T->channels > 3 || !T->data ||
            (uint64_t)T->width * T->height * T->channels > 0xFFFFFFFFull) {
            status = BF_INVALID_ARGUMENT;
            break;
        }
        const size_t count = (size_t)T->width * T->height * T->channels;
        DUP(s->texels[t], T->data, count, float);
        s->textures[t] = *T;
        s->textures[t].data = s->texels[t];
        s->textures[t].device_data = NULL;
    }
    s->desc.textures = s->textures;
    for (uint32_t c = 0; c < desc->closure_count && status == BF_OK; ++c) {
        const bf_closure_desc* C = &desc->closures[c];
        if (C->weight_texture) {
            if (C->weight_texture > desc->texture_count) status = BF_INVALID_ARGUMENT;
            else if (desc->textures[C->weight_texture - 1].channels == 2) status = BF_INVALID_ARGUMENT;
        }
        if (C->roughness_texture) {
            if (C->roughness_texture > desc->texture_count ||
                (C->kind != BF_CLOSURE_ROUGH_CONDUCTOR && C->kind != BF_CLOSURE_ROUGH_DIELECTRIC) ||
                desc->textures[C->roughness_texture - 1].channels > 2)
                status = BF_INVALID_ARGUMENT;
        }
    }
    /* surface maps: mip chains whose sizes add up, and maps that name existing channels */
    s->desc.image_textures = NULL;
    s->desc.surface_maps = NULL;
    if (desc->image_texture_count && !desc->image_textures) status = BF_INVALID_ARGUMENT;
    if (status == BF_OK && desc->image_texture_count) {
        s->image_texture_count = desc->image_texture_count;
        s->image_textures = (bf_image_texture_desc*)calloc(desc->image_texture_count, sizeof(bf_image_texture_desc));
        for (uint32_t t = 0; t < desc->image_texture_count && status == BF_OK; ++t) {
            const bf_image_texture_desc* T = &desc->image_textures[t];
            uint64_t w = T->width, h = T->height, total = 0;
            for (uint32_t k = 0; k < T->mip_count; ++k) {
                total += w * h * T->channel_count;
                w = w > 1 ? w / 2 : 1;
                h = h > 1 ? h / 2 : 1;
            }
            if (!T->width || !T->height || !T->channel_count || !T->mip_count || T->mip_count > BF_MAX_IMAGE_MIPS ||
                !T->texels || total != T->texel_count) {
                status = BF_INVALID_ARGUMENT;
                break;
            }
            s->image_textures[t] = *T;
            float* copy = NULL;
            DUP(copy, T->texels, total, float);
            s->image_textures[t].texels = copy;
        }
    }
    int mapped = 0;
    for (uint32_t m = 0; desc->surface_maps && m < desc->material_count && status == BF_OK; ++m) {
        const bf_surface_maps_desc* M = &desc->surface_maps[m];
        if (M->normal_texture != BF_INVALID_INDEX) {
            if (M->normal_texture >= desc->image_texture_count) {
                status = BF_INVALID_ARGUMENT;
                break;
            }
            const uint32_t cc = desc->image_textures[M->normal_texture].channel_count;
            const uint32_t* ch = M->normal_channels;
            if (ch[0] >= cc || ch[1] >= cc || ch[2] >= cc || ch[0] == ch[1] || ch[0] == ch[2] || ch[1] == ch[2] ||
                M->normal_u_wrap > 3 || M->normal_v_wrap > 3 || !M->normal_max_anisotropy ||
                M->normal_max_anisotropy > BF_MAX_EWA_ANISOTROPY || !M->normal_max_texel_visits || M->normal_y_convention > 1)
                status = BF_INVALID_ARGUMENT;
            mapped = 1;
        }
        if (M->bump_texture != BF_INVALID_INDEX) {
            if (M->bump_texture >= desc->image_texture_count ||
                M->bump_channel >= desc->image_textures[M->bump_texture].channel_count || M->bump_u_wrap > 3 ||
                M->bump_v_wrap > 3 || !M->bump_max_anisotropy || M->bump_max_anisotropy > BF_MAX_EWA_ANISOTROPY ||
                !M->bump_max_texel_visits || !bf_isfinite(M->bump_scale))
                status = BF_INVALID_ARGUMENT;
            mapped = 1;
        }
    }
    if (status == BF_OK && mapped) {
        DUP(s->surface_maps, desc->surface_maps, desc->material_count, bf_surface_maps_desc);
        s->desc.surface_maps = s->surface_maps;
        s->desc.image_textures = s->image_textures;
    }
    /* punctual lights: the exact_special_cosine of CudaWavefrontStages.cu:515-526 for spot cones */
    s->punctual_count = desc->punctual_count;
    s->desc.punctuals = NULL;
    if (desc->punctual_count) {
        DUP(s->punctuals, desc->punctuals, desc->punctual_count, bf_punctual_desc);
        s->punctual_cosines = (float*)calloc(2 * (size_t)desc->punctual_count, sizeof(float));
        s->desc.punctuals = s->punctuals;
        const float pi = 3.14159265358979323846f;
        for (uint32_t i = 0; i < desc->punctual_count; ++i) {
            const float angles[2] = {s->punctuals[i].inner_half_angle, s->punctuals[i].outer_half_angle};
            for (int k = 0; k < 2; ++k)
                s->punctual_cosines[2 * i + k] =
                    angles[k] == 0.0f ? 1.0f : (angles[k] == pi / 2.0f ? 0.0f : bf_cosf(angles[k]));
        }
    }
    if (status == BF_OK) status = build_bvh(s);
    if (status == BF_OK) status = build_lights(s);
    if (status != BF_OK) {
        bfo_scene_free(s);
        return status;
    }
    *out = s;
    return BF_OK;
}

void bfo_scene_free(bfo_scene* s) {
    if (!s) return;
    free(s->positions);
    free(s->normals);
    free(s->uvs);
    free(s->indices);
    free(s->geometries);
    free(s->instances);
    free(s->materials);
    free(s->closures);
    free(s->spectra);
    free(s->spectrum_values);
    free(s->punctuals);
    free(s->punctual_cosines);
    free(s->light_cdf);
    for (uint32_t t = 0; t < BF_MAX_TEXTURES; ++t) free(s->texels[t]);
    for (uint32_t t = 0; s->image_textures && t < s->image_texture_count; ++t) free((void*)s->image_textures[t].texels);
    free(s->image_textures);
    free(s->surface_maps);
    free(s->blases);
    free(s->blas_nodes);
    free(s->primitive_refs);
    free(s->tlas_nodes);
    free(s->instance_refs);
    if (s->lights) {
        for (uint32_t i = 0; i < s->light_count; ++i) {
            free(s->lights[i].areas);
            free(s->lights[i].probabilities);
            free(s->lights[i].cumulative_before);
            free(s->lights[i].next_probabilities);
        }
    }
    free(s->lights);
    free(s->instance_light);
    free(s);
}

int bfo_scene_bvh(const bfo_scene* s, uint64_t counts[5], bf_bvh_blas* blases, bf_bvh_node* blas_nodes,
                  uint32_t* primitive_refs, bf_bvh_node* tlas_nodes, bf_bvh_instance_ref* instance_refs) {
    counts[0] = s->desc.geometry_count;
    counts[1] = s->blas_node_count;
    counts[2] = s->primitive_ref_count;
    counts[3] = s->tlas_node_count;
    counts[4] = s->desc.instance_count;
    if (blases) memcpy(blases, s->blases, counts[0] * sizeof(bf_bvh_blas));
    if (blas_nodes) memcpy(blas_nodes, s->blas_nodes, counts[1] * sizeof(bf_bvh_node));
    if (primitive_refs) memcpy(primitive_refs, s->primitive_refs, counts[2] * sizeof(uint32_t));
    if (tlas_nodes) memcpy(tlas_nodes, s->tlas_nodes, counts[3] * sizeof(bf_bvh_node));
    if (instance_refs) memcpy(instance_refs, s->instance_refs, counts[4] * sizeof(bf_bvh_instance_ref));
    return BF_OK;
}

uint32_t bfo_scene_tlas_root(const bfo_scene* s) { return s->tlas_root; }
uint32_t bfo_scene_light_count(const bfo_scene* s) { return s->light_count; }

int bfo_scene_light(const bfo_scene* s, uint32_t light, uint32_t* instance, uint32_t* triangle_count,
                    float* inverse_area, float* areas, float* probabilities, float* next_probabilities) {
    if (light >= s->light_count) return BF_INVALID_ARGUMENT;
    const bfo_light* l = &s->lights[light];
    *instance = l->instance;
    *triangle_count = l->triangle_count;
    *inverse_area = l->inverse_area;
    if (areas) memcpy(areas, l->areas, l->triangle_count * sizeof(float));
    if (probabilities) memcpy(probabilities, l->probabilities, l->triangle_count * sizeof(float));
    if (next_probabilities) memcpy(next_probabilities, l->next_probabilities, l->triangle_count * sizeof(float));
    return BF_OK;
}

/* ---- traversal (CudaSceneClosestHit.cu:395-1186) ------------------------------------------- */
typedef struct {
    int present;
    float parameter;
    uint32_t instance_id;
    uint32_t primitive_id;
    uint32_t scene_instance;
    uint32_t geometry;
    float b0, b1, b2;
} candidate;

typedef struct {
    int hit;
    float t, b0, b1, b2;
} tri_hit;

/* CudaSceneClosestHit.cu:397-405 */
static int valid_traversal_ray(const bf_transport_ray* r) {
    const v3 o = v3_make(r->origin_x, r->origin_y, r->origin_z);
    const v3 d = v3_make(r->direction_x, r->direction_y, r->direction_z);
    return v3_finite(o) && v3_finite(d) && (d.x != 0.0f || d.y != 0.0f || d.z != 0.0f) && bf_isfinite(r->t_min) &&
           r->t_min >= 0.0f && !bf_isnan(r->t_max) && r->t_max >= r->t_min && bf_isfinite(r->time) &&
           r->time >= 0.0f && r->time <= 1.0f && r->reserved == 0u;
}

/* CudaSceneClosestHit.cu:407-455 */
static uint32_t intersect_bounds(const bf_bvh_node* node, v3 origin, v3 direction, float t_min, float t_max,
                                 int* intersects, float* near_parameter) {
    const v3 mn = v3_make(node->min_x, node->min_y, node->min_z);
    const v3 mx = v3_make(node->max_x, node->max_y, node->max_z);
    if (!v3_finite(mn) || !v3_finite(mx) || mn.x > mx.x || mn.y > mx.y || mn.z > mx.z) return BF_HIT_INVALID_TOPOLOGY;
    double near_value = (double)t_min;
    double far_value = (double)t_max;
    for (uint32_t axis = 0; axis < 3; ++axis) {
        const float o = v3_comp(origin, axis), d = v3_comp(direction, axis);
        const float lo = v3_comp(mn, axis), hi = v3_comp(mx, axis);
        if (d == 0.0f) {
            if (o < lo || o > hi) {
                *intersects = 0;
                return BF_HIT_MISS;
            }
            continue;
        }
        const double first = ((double)lo - (double)o) / (double)d;
        const double second = ((double)hi - (double)o) / (double)d;
        if (bf_isnand(first) || bf_isnand(second)) return BF_HIT_NUMERICAL_FAILURE;
        const double axis_near = bf_next_down(fmin(first, second));
        const double axis_far = bf_next_up(fmax(first, second));
        near_value = fmax(near_value, axis_near);
        far_value = fmin(far_value, axis_far);
        if (near_value > far_value) {
            *intersects = 0;
            return BF_HIT_MISS;
        }
    }
    *intersects = 1;
    *near_parameter = (float)near_value;
    return BF_HIT_MISS;
}

/* CudaSceneClosestHit.cu:466-512 */
static uint32_t push_children(const bf_bvh_node* nodes, uint32_t first, uint32_t second, uint32_t begin, uint32_t count,
                              v3 o, v3 d, float t_min, float t_max, uint32_t* stack, uint32_t* size) {
    const uint64_t end = (uint64_t)begin + count;
    if (first == second || first < begin || second < begin || first >= end || second >= end) return BF_HIT_INVALID_TOPOLOGY;
    int fh = 0, sh = 0;
    float fn = 0.0f, sn = 0.0f;
    uint32_t st = intersect_bounds(&nodes[first], o, d, t_min, t_max, &fh, &fn);
    if (st != BF_HIT_MISS) return st;
    st = intersect_bounds(&nodes[second], o, d, t_min, t_max, &sh, &sn);
    if (st != BF_HIT_MISS) return st;
    if (!fh && !sh) return BF_HIT_MISS;
    if (fh && !sh) {
        if (*size >= BF_TRAVERSAL_STACK_CAPACITY) return BF_HIT_STACK_OVERFLOW;
        stack[(*size)++] = first;
        return BF_HIT_MISS;
    }
    if (!fh && sh) {
        if (*size >= BF_TRAVERSAL_STACK_CAPACITY) return BF_HIT_STACK_OVERFLOW;
        stack[(*size)++] = second;
        return BF_HIT_MISS;
    }
    const int first_is_near = fn < sn || (fn == sn && first < second);
    const uint32_t near_child = first_is_near ? first : second;
    const uint32_t far_child = first_is_near ? second : first;
    if (*size + 2 > BF_TRAVERSAL_STACK_CAPACITY) return BF_HIT_STACK_OVERFLOW;
    stack[(*size)++] = far_child;
    stack[(*size)++] = near_child;
    return BF_HIT_MISS;
}

/* CudaSceneClosestHit.cu:583-692 (double refinement of ambiguous edges) */
static uint32_t refine_edges(int with_bary, v3 v0, v3 v1, v3 v2, v3 o, v3 d, uint32_t kx, uint32_t ky, uint32_t kz,
                             float t_min, float t_max, tri_hit* r) {
    const double dx = (double)v3_comp(d, kx), dy = (double)v3_comp(d, ky), dz = (double)v3_comp(d, kz);
    if (dz == 0.0) return BF_HIT_NUMERICAL_FAILURE;
    const double p0x = (double)v3_comp(v0, kx) - (double)v3_comp(o, kx);
    const double p0y = (double)v3_comp(v0, ky) - (double)v3_comp(o, ky);
    const double p0z = (double)v3_comp(v0, kz) - (double)v3_comp(o, kz);
    const double p1x = (double)v3_comp(v1, kx) - (double)v3_comp(o, kx);
    const double p1y = (double)v3_comp(v1, ky) - (double)v3_comp(o, ky);
    const double p1z = (double)v3_comp(v1, kz) - (double)v3_comp(o, kz);
    const double p2x = (double)v3_comp(v2, kx) - (double)v3_comp(o, kx);
    const double p2y = (double)v3_comp(v2, ky) - (double)v3_comp(o, ky);
    const double p2z = (double)v3_comp(v2, kz) - (double)v3_comp(o, kz);
    const double ax = p0x * dz - dx * p0z;
    const double ay = p0y * dz - dy * p0z;
    const double bx = p1x * dz - dx * p1z;
    const double by = p1y * dz - dy * p1z;
    const double cx = p2x * dz - dx * p2z;
    const double cy = p2y * dz - dy * p2z;
    const double e0 = bx * cy - by * cx;
    const double e1 = cx * ay - cy * ax;
    const double e2 = ax * by - ay * bx;
    if (!bf_isfinited(e0) || !bf_isfinited(e1) || !bf_isfinited(e2)) return BF_HIT_NUMERICAL_FAILURE;
    const int neg = e0 < 0.0 || e1 < 0.0 || e2 < 0.0;
    const int pos = e0 > 0.0 || e1 > 0.0 || e2 > 0.0;
    if (neg && pos) return BF_HIT_MISS;
    const double det = e0 + e1 + e2;
    if (!bf_isfinited(det)) return BF_HIT_NUMERICAL_FAILURE;
    if (det == 0.0) return BF_HIT_MISS;
    const double num = fma(e0, p0z, fma(e1, p1z, e2 * p2z));
    const double den = det * dz;
    if (!bf_isfinited(num) || !bf_isfinited(den) || den == 0.0) return BF_HIT_NUMERICAL_FAILURE;
    const double t = num / den;
    if (!with_bary) {
        if (!bf_isfinited(t)) return BF_HIT_NUMERICAL_FAILURE;
        if (t < (double)t_min || t > (double)t_max) return BF_HIT_MISS;
        const float ft = (float)t;
        if (!bf_isfinite(ft)) return BF_HIT_NUMERICAL_FAILURE;
        r->hit = 1;
        r->t = ft == 0.0f ? 0.0f : ft;
        return BF_HIT_MISS;
    }
    const double b0 = e0 / det, b1 = e1 / det, b2 = e2 / det;
    if (!bf_isfinited(t) || !bf_isfinited(b0) || !bf_isfinited(b1) || !bf_isfinited(b2)) return BF_HIT_NUMERICAL_FAILURE;
    if (t < (double)t_min || t > (double)t_max) return BF_HIT_MISS;
    const float ft = (float)t, f0 = (float)b0, f1 = (float)b1, f2 = (float)b2;
    if (!bf_isfinite(ft) || !bf_isfinite(f0) || !bf_isfinite(f1) || !bf_isfinite(f2)) return BF_HIT_NUMERICAL_FAILURE;
    r->hit = 1;
    r->t = ft == 0.0f ? 0.0f : ft;
    r->b0 = f0 == 0.0f ? 0.0f : f0;
    r->b1 = f1 == 0.0f ? 0.0f : f1;
    r->b2 = f2 == 0.0f ? 0.0f : f2;
    return BF_HIT_MISS;
}

/* CudaSceneClosestHit.cu:694-824 */
static uint32_t watertight(int with_bary, v3 v0, v3 v1, v3 v2, v3 o, v3 d, float t_min, float t_max, tri_hit* r) {
    uint32_t kz = 0;
    float maxd = fabsf(d.x);
    if (fabsf(d.y) > maxd) {
        kz = 1;
        maxd = fabsf(d.y);
    }
    if (fabsf(d.z) > maxd) kz = 2;
    uint32_t kx = kz + 1;
    if (kx == 3) kx = 0;
    uint32_t ky = kx + 1;
    if (ky == 3) ky = 0;
    if (v3_comp(d, kz) < 0.0f) {
        const uint32_t t = kx;
        kx = ky;
        ky = t;
    }
    const v3 t0 = v3_sub(v0, o), t1 = v3_sub(v1, o), t2 = v3_sub(v2, o);
    if (!v3_finite(t0) || !v3_finite(t1) || !v3_finite(t2)) return BF_HIT_NUMERICAL_FAILURE;
    const float rz = 1.0f / v3_comp(d, kz);
    const float sx = v3_comp(d, kx) * rz;
    const float sy = v3_comp(d, ky) * rz;
    if (!bf_isfinite(rz) || !bf_isfinite(sx) || !bf_isfinite(sy)) return BF_HIT_NUMERICAL_FAILURE;
    const float ax = fmaf(-sx, v3_comp(t0, kz), v3_comp(t0, kx));
    const float ay = fmaf(-sy, v3_comp(t0, kz), v3_comp(t0, ky));
    const float bx = fmaf(-sx, v3_comp(t1, kz), v3_comp(t1, kx));
    const float by = fmaf(-sy, v3_comp(t1, kz), v3_comp(t1, ky));
    const float cx = fmaf(-sx, v3_comp(t2, kz), v3_comp(t2, kx));
    const float cy = fmaf(-sy, v3_comp(t2, kz), v3_comp(t2, ky));
    const float az = v3_comp(t0, kz) * rz;
    const float bz = v3_comp(t1, kz) * rz;
    const float cz = v3_comp(t2, kz) * rz;
    if (!bf_isfinite(ax) || !bf_isfinite(ay) || !bf_isfinite(bx) || !bf_isfinite(by) || !bf_isfinite(cx) ||
        !bf_isfinite(cy) || !bf_isfinite(az) || !bf_isfinite(bz) || !bf_isfinite(cz))
        return BF_HIT_NUMERICAL_FAILURE;
    const float e0 = bx * cy - by * cx;
    const float e1 = cx * ay - cy * ax;
    const float e2 = ax * by - ay * bx;
    if (!bf_isfinite(e0) || !bf_isfinite(e1) || !bf_isfinite(e2)) return BF_HIT_NUMERICAL_FAILURE;
    const int neg = e0 < 0.0f || e1 < 0.0f || e2 < 0.0f;
    const int pos = e0 > 0.0f || e1 > 0.0f || e2 > 0.0f;
    if (e0 == 0.0f || e1 == 0.0f || e2 == 0.0f || (neg && pos))
        return refine_edges(with_bary, v0, v1, v2, o, d, kx, ky, kz, t_min, t_max, r);
    const float det = e0 + e1 + e2;
    if (!bf_isfinite(det)) return BF_HIT_NUMERICAL_FAILURE;
    if (det == 0.0f) return BF_HIT_MISS;
    const float sp = fmaf(e0, az, fmaf(e1, bz, e2 * cz));
    if (!bf_isfinite(sp)) return BF_HIT_NUMERICAL_FAILURE;
    const float smin = t_min * det, smax = t_max * det;
    if (bf_isnan(smin) || bf_isnan(smax)) return BF_HIT_NUMERICAL_FAILURE;
    if ((det > 0.0f && (sp < smin || sp > smax)) || (det < 0.0f && (sp > smin || sp < smax))) return BF_HIT_MISS;
    const float inv = 1.0f / det;
    const float t = sp * inv;
    if (!with_bary) {
        if (!bf_isfinite(t)) return BF_HIT_NUMERICAL_FAILURE;
        if (t < t_min || t > t_max) return BF_HIT_MISS;
        r->hit = 1;
        r->t = t == 0.0f ? 0.0f : t;
        return BF_HIT_MISS;
    }
    const float b0 = e0 * inv, b1 = e1 * inv, b2 = e2 * inv;
    if (!bf_isfinite(t) || !bf_isfinite(b0) || !bf_isfinite(b1) || !bf_isfinite(b2)) return BF_HIT_NUMERICAL_FAILURE;
    if (t < t_min || t > t_max) return BF_HIT_MISS;
    r->hit = 1;
    r->t = t == 0.0f ? 0.0f : t;
    r->b0 = b0 == 0.0f ? 0.0f : b0;
    r->b1 = b1 == 0.0f ? 0.0f : b1;
    r->b2 = b2 == 0.0f ? 0.0f : b2;
    return BF_HIT_MISS;
}

/* CudaSceneClosestHit.cu:826-871 */
int bfo_world_normal(const bfo_scene* scene, uint32_t instance, const v3* lv, v3* normal) {
    const float* m = scene->instances[instance].local_to_world;
    const v3 w0 = bfo_transform_point(m, lv[0]), w1 = bfo_transform_point(m, lv[1]), w2 = bfo_transform_point(m, lv[2]);
    if (!v3_finite(w0) || !v3_finite(w1) || !v3_finite(w2)) return 0;
    const v3 f = v3_sub(w1, w0), s = v3_sub(w2, w0);
    if (!v3_finite(f) || !v3_finite(s)) return 0;
    const double cx = (double)f.y * (double)s.z - (double)f.z * (double)s.y;
    const double cy = (double)f.z * (double)s.x - (double)f.x * (double)s.z;
    const double cz = (double)f.x * (double)s.y - (double)f.y * (double)s.x;
    const double mx = fmax(fabs(cx), fmax(fabs(cy), fabs(cz)));
    if (!bf_isfinited(cx) || !bf_isfinited(cy) || !bf_isfinited(cz) || !bf_isfinited(mx) || mx == 0.0) return 0;
    const double sx = cx / mx, sy = cy / mx, sz = cz / mx;
    const double len = sqrt(sx * sx + sy * sy + sz * sz);
    if (!bf_isfinited(len) || len == 0.0) return 0;
    *normal = v3_make((float)(sx / len), (float)(sy / len), (float)(sz / len));
    return v3_finite(*normal);
}

/* CudaSceneClosestHit.cu:873-881 */
static int candidate_precedes(float t, uint32_t inst, uint32_t prim, const candidate* c) {
    return !c->present || t < c->parameter ||
           (t == c->parameter && (inst < c->instance_id || (inst == c->instance_id && prim < c->primitive_id)));
}

/* CudaSceneClosestHit.cu:893-1079 */
static uint32_t traverse_blas(const bfo_scene* scene, int stop_at_first, const bf_bvh_instance_ref* ref,
                              const bf_transport_ray* ray, v3 lo, v3 ld, candidate* closest, int* occluded,
                              uint32_t* counters) {
    const bf_bvh_blas* blas = &scene->blases[ref->blas_index];
    const bf_bvh_node* nodes = scene->blas_nodes;
    const uint64_t node_end = (uint64_t)blas->node_offset + blas->node_count;
    const uint64_t ref_end = (uint64_t)blas->primitive_reference_offset + blas->primitive_reference_count;
    const uint32_t g = ref->scene_geometry_index;
    uint32_t stack[BF_TRAVERSAL_STACK_CAPACITY];
    uint32_t size = 0;
    stack[size++] = blas->root_node;
    uint32_t visited = 0;
    while (size != 0) {
        if (visited++ >= blas->node_count) return BF_HIT_INVALID_TOPOLOGY;
        const uint32_t ni = stack[--size];
        if (ni < blas->node_offset || ni >= node_end) return BF_HIT_INVALID_TOPOLOGY;
        const bf_bvh_node* node = &nodes[ni];
        if (counters) counters[0]++;
        int hit = 0;
        float nearp = 0.0f;
        const float tmax = stop_at_first ? ray->t_max : closest->parameter;
        uint32_t st = intersect_bounds(node, lo, ld, ray->t_min, tmax, &hit, &nearp);
        if (st != BF_HIT_MISS) return st;
        if (!hit) continue;
        if (node->kind == 0) {
            st = push_children(nodes, node->first_child, node->second_child, blas->node_offset, blas->node_count, lo, ld,
                               ray->t_min, tmax, stack, &size);
            if (st != BF_HIT_MISS) return st;
            continue;
        }
        if (node->reference_count != 1 || node->reference_offset < blas->primitive_reference_offset ||
            node->reference_offset >= ref_end)
            return BF_HIT_INVALID_TOPOLOGY;
        const uint32_t prim = scene->primitive_refs[node->reference_offset];
        uint32_t idx[3];
        if (!bfo_triangle_vertices(scene, g, prim, idx)) return BF_HIT_INVALID_TOPOLOGY;
        v3 v[3];
        for (int c = 0; c < 3; ++c) {
            v[c] = bfo_vertex(scene, g, idx[c]);
            if (!v3_finite(v[c])) return BF_HIT_NUMERICAL_FAILURE;
        }
        if (counters) counters[1]++;
        tri_hit th = {0};
        st = watertight(!stop_at_first, v[0], v[1], v[2], lo, ld, ray->t_min, tmax, &th);
        if (st != BF_HIT_MISS) return st;
        if (!th.hit) continue;
        if (stop_at_first) {
            *occluded = 1;
            return BF_HIT_MISS;
        }
        if (!candidate_precedes(th.t, ref->instance_id, prim, closest)) continue;
        closest->present = 1;
        closest->parameter = th.t;
        closest->instance_id = ref->instance_id;
        closest->primitive_id = prim;
        closest->scene_instance = ref->scene_instance_index;
        closest->geometry = g;
        closest->b0 = th.b0;
        closest->b1 = th.b1;
        closest->b2 = th.b2;
    }
    return BF_HIT_MISS;
}

/* CudaSceneClosestHit.cu:1081-1186 */
static uint32_t trace_scene(const bfo_scene* scene, int stop_at_first, const bf_transport_ray* ray, candidate* closest,
                            int* occluded, uint32_t* counters) {
    if (scene->tlas_node_count == 0) return BF_HIT_MISS;
    uint32_t stack[BF_TRAVERSAL_STACK_CAPACITY];
    uint32_t size = 0;
    stack[size++] = scene->tlas_root;
    uint32_t visited = 0;
    const v3 wo = v3_make(ray->origin_x, ray->origin_y, ray->origin_z);
    const v3 wd = v3_make(ray->direction_x, ray->direction_y, ray->direction_z);
    while (size != 0) {
        if (visited++ >= scene->tlas_node_count) return BF_HIT_INVALID_TOPOLOGY;
        const uint32_t ni = stack[--size];
        if (ni >= scene->tlas_node_count) return BF_HIT_INVALID_TOPOLOGY;
        const bf_bvh_node* node = &scene->tlas_nodes[ni];
        if (counters) counters[0]++;
        int hit = 0;
        float nearp = 0.0f;
        const float tmax = stop_at_first ? ray->t_max : closest->parameter;
        uint32_t st = intersect_bounds(node, wo, wd, ray->t_min, tmax, &hit, &nearp);
        if (st != BF_HIT_MISS) return st;
        if (!hit) continue;
        if (node->kind == 0) {
            st = push_children(scene->tlas_nodes, node->first_child, node->second_child, 0, (uint32_t)scene->tlas_node_count,
                               wo, wd, ray->t_min, tmax, stack, &size);
            if (st != BF_HIT_MISS) return st;
            continue;
        }
        if (node->reference_count != 1 || node->reference_offset >= scene->desc.instance_count) return BF_HIT_INVALID_TOPOLOGY;
        const bf_bvh_instance_ref* ref = &scene->instance_refs[node->reference_offset];
        if ((ray->visibility_mask & ref->visibility_mask) == 0u) continue;
        const float* w2l = scene->instances[ref->scene_instance_index].world_to_local;
        const v3 lo = bfo_transform_point(w2l, wo);
        const v3 ld = bfo_transform_vector(w2l, wd);
        if (!v3_finite(lo) || !v3_finite(ld) || (ld.x == 0.0f && ld.y == 0.0f && ld.z == 0.0f)) return BF_HIT_NUMERICAL_FAILURE;
        if (counters) counters[2]++;
        st = traverse_blas(scene, stop_at_first, ref, ray, lo, ld, closest, occluded, counters);
        if (st != BF_HIT_MISS) return st;
        if (*occluded) return BF_HIT_MISS;
    }
    return BF_HIT_MISS;
}

static void fill_hit(const bfo_scene* scene, const candidate* c, bf_closest_hit_result* out, uint32_t* status) {
    uint32_t idx[3];
    v3 lv[3], n;
    bfo_triangle_vertices(scene, c->geometry, c->primitive_id, idx);
    for (int k = 0; k < 3; ++k) lv[k] = bfo_vertex(scene, c->geometry, idx[k]);
    if (!bfo_world_normal(scene, c->scene_instance, lv, &n)) {
        *status = BF_HIT_NUMERICAL_FAILURE;
        return;
    }
    const bf_instance_desc* inst = &scene->instances[c->scene_instance];
    out->hit.parameter = c->parameter;
    out->hit.barycentric0 = c->b0;
    out->hit.barycentric1 = c->b1;
    out->hit.barycentric2 = c->b2;
    out->hit.normal_x = n.x;
    out->hit.normal_y = n.y;
    out->hit.normal_z = n.z;
    out->hit.ids.object = inst->object_id;
    out->hit.ids.instance = inst->id;
    out->hit.ids.geometry = scene->geometries[c->geometry].id;
    out->hit.ids.primitive = c->primitive_id;
    out->hit.ids.material = scene->materials[inst->material_index].id;
    *status = BF_HIT_HIT;
}

void bfo_closest_hit(const bfo_scene* scene, const bf_transport_ray* rays, uint32_t count, bf_closest_hit_result* results,
                     uint32_t* counters) {
    for (uint32_t i = 0; i < count; ++i) {
        bf_closest_hit_result out;
        memset(&out, 0, sizeof out);
        uint32_t* ctr = counters ? counters + 3 * (size_t)i : NULL;
        if (ctr) ctr[0] = ctr[1] = ctr[2] = 0;
        if (!valid_traversal_ray(&rays[i])) {
            out.status = BF_HIT_INVALID_RAY;
            results[i] = out;
            continue;
        }
        candidate c;
        memset(&c, 0, sizeof c);
        c.parameter = rays[i].t_max;
        int occ = 0;
        uint32_t st = trace_scene(scene, 0, &rays[i], &c, &occ, ctr);
        if (st != BF_HIT_MISS) {
            out.status = st;
        } else if (c.present) {
            uint32_t s;
            fill_hit(scene, &c, &out, &s);
            if (s != BF_HIT_HIT) memset(&out, 0, sizeof out);
            out.status = s;
        }
        results[i] = out;
    }
}

void bfo_occlusion(const bfo_scene* scene, const bf_transport_ray* rays, uint32_t count, bf_occlusion_result* results,
                   uint32_t* counters) {
    for (uint32_t i = 0; i < count; ++i) {
        bf_occlusion_result out;
        memset(&out, 0, sizeof out);
        uint32_t* ctr = counters ? counters + 3 * (size_t)i : NULL;
        if (ctr) ctr[0] = ctr[1] = ctr[2] = 0;
        if (!valid_traversal_ray(&rays[i])) {
            out.status = BF_HIT_INVALID_RAY;
            results[i] = out;
            continue;
        }
        int occ = 0;
        uint32_t st = trace_scene(scene, 1, &rays[i], NULL, &occ, ctr);
        out.status = st != BF_HIT_MISS ? st : (occ ? BF_OCCLUDED : BF_VISIBLE);
        results[i] = out;
    }
}

/* Brute force over every triangle of every visible instance with the identical watertight test
 * and candidate order: the BVH-independent closed form the traversal must reproduce. */
void bfo_closest_hit_brute_force(const bfo_scene* scene, const bf_transport_ray* rays, uint32_t count,
                                 bf_closest_hit_result* results) {
    for (uint32_t i = 0; i < count; ++i) {
        const bf_transport_ray* ray = &rays[i];
        bf_closest_hit_result out;
        memset(&out, 0, sizeof out);
        candidate c;
        memset(&c, 0, sizeof c);
        c.parameter = ray->t_max;
        uint32_t status = BF_HIT_MISS;
        for (uint32_t inst = 0; inst < scene->desc.instance_count && status == BF_HIT_MISS; ++inst) {
            const bf_instance_desc* id = &scene->instances[inst];
            if ((ray->visibility_mask & id->visibility_mask) == 0u) continue;
            const v3 wo = v3_make(ray->origin_x, ray->origin_y, ray->origin_z);
            const v3 wd = v3_make(ray->direction_x, ray->direction_y, ray->direction_z);
            const v3 lo = bfo_transform_point(id->world_to_local, wo);
            const v3 ld = bfo_transform_vector(id->world_to_local, wd);
            const uint32_t g = id->geometry_index;
            for (uint32_t p = 0; p < scene->geometries[g].triangle_count; ++p) {
                uint32_t idx[3];
                bfo_triangle_vertices(scene, g, p, idx);
                tri_hit th = {0};
                uint32_t st = watertight(1, bfo_vertex(scene, g, idx[0]), bfo_vertex(scene, g, idx[1]),
                                         bfo_vertex(scene, g, idx[2]), lo, ld, ray->t_min, c.parameter, &th);
                if (st != BF_HIT_MISS) {
                    status = st;
                    break;
                }
                if (!th.hit || !candidate_precedes(th.t, id->id, p, &c)) continue;
                c.present = 1;
                c.parameter = th.t;
                c.instance_id = id->id;
                c.primitive_id = p;
                c.scene_instance = inst;
                c.geometry = g;
                c.b0 = th.b0;
                c.b1 = th.b1;
                c.b2 = th.b2;
            }
        }
        if (status != BF_HIT_MISS) {
            out.status = status;
        } else if (c.present) {
            uint32_t s;
            fill_hit(scene, &c, &out, &s);
            if (s != BF_HIT_HIT) memset(&out, 0, sizeof out);
            out.status = s;
        }
        results[i] = out;
    }
}
