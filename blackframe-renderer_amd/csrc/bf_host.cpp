/*
 * Host side of the MI355X wavefront backend: scene ingestion, the deterministic BVH builder and
 * its 64-byte relayout, mesh-area-light tables, the transport workspace, the bounce loop and the
 * C ABI of include/blackframe_hip.h.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <atomic>
#include <mutex>
#include <new>
#include <thread>
#include <string>
#include <utility>
#include <vector>

#include "../../include/blackframe_hip.h"
#include "../../include/blackframe_math.h"
#include "bf_kernels.h"

namespace {

thread_local std::string g_last_error;

int set_error(int status, const char* fmt, ...) {
    char buffer[512];
    va_list args;
    va_start(args, fmt);
    std::vsnprintf(buffer, sizeof buffer, fmt, args);
    va_end(args);
    g_last_error = buffer;
    return status;
}

int hip_error(hipError_t e, const char* what) {
    return set_error(BF_DEVICE_ERROR, "%s failed: %s", what, hipGetErrorString(e));
}

#define HIP_TRY(call)                                        \
    do {                                                     \
        const hipError_t _e = (call);                        \
        if (_e != hipSuccess) return hip_error(_e, #call);   \
    } while (0)

float canonical_zero(float v) { return v == 0.0f ? 0.0f : v; }
float next_up(float v) { return std::nextafter(v, INFINITY); }
float next_down(float v) { return std::nextafter(v, -INFINITY); }

/* ---- Sobol direction numbers (Blackframe/Renderer/Source/SobolSampling.cpp:53-88) ----------- */
struct JoeKuoRow {
    uint32_t degree;
    uint32_t a;
    uint32_t m[18];
};
const JoeKuoRow kJoeKuo[] = {
#include "data/sobol_joe_kuo.inc"
};
constexpr uint32_t kSobolDimensions = sizeof(kJoeKuo) / sizeof(kJoeKuo[0]) + 1u;

/* CIE 1931 table: the film normalisation (1/4) / integral(ybar) of Cie1931Sensor.cpp:26-40,131 */
const double kCieTable[471][3] = {
#include "data/cie1931_cmf.inc"
};
float film_normalization() {
    double integral = 0.0;
    for (int i = 0; i + 1 < 471; ++i) integral += (kCieTable[i][1] + kCieTable[i + 1][1]) / 2.0;
    return (1.0f / 4.0f) / (float)integral;
}

std::vector<uint64_t> sobol_directions() {
    std::vector<uint64_t> d((size_t)kSobolDimensions * 64u);
    for (uint32_t c = 0; c < 64u; ++c) d[c] = 1ull << (63u - c);
    for (uint32_t dim = 1; dim < kSobolDimensions; ++dim) {
        const JoeKuoRow& row = kJoeKuo[dim - 1];
        uint64_t* v = d.data() + (size_t)dim * 64u;
        const uint32_t s = row.degree;
        for (uint32_t c = 0; c < s; ++c) v[c] = (uint64_t)row.m[c] << (63u - c);
        for (uint32_t c = s; c < 64u; ++c) {
            uint64_t x = v[c - s] ^ (v[c - s] >> s);
            for (uint32_t k = 1; k < s; ++k)
                if ((row.a >> (s - 1u - k)) & 1u) x ^= v[c - k];
            v[c] = x;
        }
    }
    return d;
}

struct SobolTable {
    std::mutex lock;
    int device = -1;
    uint64_t* directions = nullptr;
};
SobolTable g_sobol;

int sobol_table_for_current_device(const uint64_t** out) {
    std::lock_guard<std::mutex> guard(g_sobol.lock);
    int device = 0;
    HIP_TRY(hipGetDevice(&device));
    if (g_sobol.directions == nullptr || g_sobol.device != device) {
        const std::vector<uint64_t> table = sobol_directions();
        uint64_t* ptr = nullptr;
        HIP_TRY(hipMalloc(&ptr, table.size() * sizeof(uint64_t)));
        HIP_TRY(hipMemcpy(ptr, table.data(), table.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
        g_sobol.directions = ptr; /* one table per process and device, kept until exit */
        g_sobol.device = device;
    }
    *out = g_sobol.directions;
    return BF_OK;
}

/* ---- exact triangle area (CudaWavefrontStages.cu:1724-2042, restated on the host) ----------- */
struct Expansion {
    float c[32];
    uint32_t size = 0;
    uint32_t capacity = 32;
};
bool normal_float(float v) { return std::isfinite(v) && std::fabs(v) >= 0x1p-126f; }
bool expansion_add(Expansion& e, float component) {
    if (!std::isfinite(component)) return false;
    float acc = component;
    uint32_t out = 0;
    const uint32_t in = e.size;
    for (uint32_t i = 0; i < in; ++i) {
        const float existing = e.c[i];
        const float rs = acc + existing;
        const float efs = rs - acc;
        const float afs = rs - efs;
        const float error = (existing - efs) + (acc - afs);
        if (!std::isfinite(rs) || !std::isfinite(error)) return false;
        if (error != 0.0f) {
            if (out == e.capacity) return false;
            e.c[out++] = error;
        }
        acc = rs;
    }
    if (acc != 0.0f || out == 0) {
        if (out == e.capacity) return false;
        e.c[out++] = acc;
    }
    e.size = out;
    return true;
}
bool expansion_add_product(Expansion& e, float l, float r) {
    const float rounded = l * r;
    const float err = std::fma(l, r, -rounded);
    if (!std::isfinite(rounded) || !std::isfinite(err) || (l != 0.0f && r != 0.0f && !normal_float(rounded)) ||
        (err != 0.0f && !normal_float(err)))
        return false;
    return expansion_add(e, err) && expansion_add(e, rounded);
}
bool resolved_difference(const Expansion& a, const Expansion& b, const Expansion& c, const Expansion& d, int& sign,
                         float& value) {
    Expansion p1, p2, diff;
    p1.capacity = p2.capacity = 16;
    for (uint32_t i = 0; i < a.size; ++i)
        for (uint32_t j = 0; j < b.size; ++j)
            if (!expansion_add_product(p1, a.c[i], b.c[j])) return false;
    for (uint32_t i = 0; i < c.size; ++i)
        for (uint32_t j = 0; j < d.size; ++j)
            if (!expansion_add_product(p2, c.c[i], d.c[j])) return false;
    for (uint32_t i = 0; i < p1.size; ++i)
        if (!expansion_add(diff, p1.c[i])) return false;
    for (uint32_t i = 0; i < p2.size; ++i)
        if (!expansion_add(diff, -1.0f * p2.c[i])) return false;
    int s = 0;
    for (uint32_t i = diff.size; i > 0; --i)
        if (diff.c[i - 1] != 0.0f) {
            s = diff.c[i - 1] > 0.0f ? 1 : -1;
            break;
        }
    if (s == 0) {
        sign = 0;
        value = 0.0f;
        return true;
    }
    float resolved = 0.0f;
    for (uint32_t i = 0; i < diff.size; ++i) resolved += diff.c[i];
    if (!std::isfinite(resolved) || resolved == 0.0f || (resolved > 0.0f ? 1 : -1) != s) return false;
    sign = s;
    value = resolved;
    return true;
}

bool triangle_area(const float v[3][3], float& area, float normal[3]) {
    int ex[3];
    for (int a = 0; a < 3; ++a) {
        float m = 0.0f;
        for (int i = 0; i < 3; ++i) m = std::fmax(m, std::fabs(v[i][a]));
        std::frexp(m, &ex[a]);
    }
    float s[3][3];
    for (int i = 0; i < 3; ++i)
        for (int a = 0; a < 3; ++a) {
            s[i][a] = std::ldexp(v[i][a], -ex[a]);
            if (v[i][a] != 0.0f && !normal_float(s[i][a])) return false;
        }
    Expansion f[3], g[3];
    for (int a = 0; a < 3; ++a) {
        f[a].capacity = g[a].capacity = 2;
        if (!expansion_add(f[a], s[1][a]) || !expansion_add(f[a], -s[0][a]) || !expansion_add(g[a], s[2][a]) ||
            !expansion_add(g[a], -s[0][a]))
            return false;
    }
    int xs, ys, zs;
    float xv, yv, zv;
    if (!resolved_difference(f[1], g[2], f[2], g[1], xs, xv) || !resolved_difference(f[2], g[0], f[0], g[2], ys, yv) ||
        !resolved_difference(f[0], g[1], f[1], g[0], zs, zv))
        return false;
    const int cx = ex[1] + ex[2], cy = ex[2] + ex[0], cz = ex[0] + ex[1];
    int mag = INT32_MIN, e;
    if (xs != 0) {
        std::frexp(std::fabs(xv), &e);
        mag = std::max(mag, e + cx);
    }
    if (ys != 0) {
        std::frexp(std::fabs(yv), &e);
        mag = std::max(mag, e + cy);
    }
    if (zs != 0) {
        std::frexp(std::fabs(zv), &e);
        mag = std::max(mag, e + cz);
    }
    if (mag == INT32_MIN) return false;
    const float ax = xs == 0 ? 0.0f : std::ldexp(xv, cx - mag);
    const float ay = ys == 0 ? 0.0f : std::ldexp(yv, cy - mag);
    const float az = zs == 0 ? 0.0f : std::ldexp(zv, cz - mag);
    if ((xs != 0 && ax == 0.0f) || (ys != 0 && ay == 0.0f) || (zs != 0 && az == 0.0f)) return false;
    const float maxc = std::fmax(std::fabs(ax), std::fmax(std::fabs(ay), std::fabs(az)));
    if (!(maxc > 0.0f)) return false;
    const float sx = ax / maxc, sy = ay / maxc, sz = az / maxc;
    const float sq = std::fma(sx, sx, std::fma(sy, sy, sz * sz));
    const float sl = std::sqrt(sq);
    const float al = maxc * sl;
    area = std::ldexp(al * 0.5f, mag);
    if (!(sl > 0.0f) || !(al > 0.0f) || !std::isfinite(area) || !(area > 0.0f)) return false;
    if ((xv != 0.0f && !normal_float(ax)) || (yv != 0.0f && !normal_float(ay)) || (zv != 0.0f && !normal_float(az)))
        return false;
    normal[0] = sx / sl;
    normal[1] = sy / sl;
    normal[2] = sz / sl;
    return std::isfinite(normal[0]) && std::isfinite(normal[1]) && std::isfinite(normal[2]);
}

/* ---- deterministic BVH builder (SceneBvh.cpp:355-427, 477-736) ------------------------------ */
struct Item {
    float b[6];
    double c[3];
    uint32_t key;
    bf_bvh_instance_ref ref;
};

/* Where a build writes: the whole node array and the whole primitive- or instance-reference
 * array of its kind (indices are absolute). */
struct Builder {
    bf_bvh_node* nodes;
    uint32_t* prim_refs;            /* BLAS builds */
    bf_bvh_instance_ref* inst_refs; /* the TLAS build */
};

/* Ranges of at least this many items build their two halves on two threads. */
constexpr size_t kParallelBuildItems = 16384;

/*
 * Builds items[0, n) as the subtree whose root is node `index` and whose leaves take references
 * ref .. ref + n - 1. With one item per leaf a subtree of m items holds 2m - 1 nodes in pre-order,
 * so the first child's subtree starts at index + 1 and the second's at index + 2 m_first: every
 * node's position is known before its subtree is built, and the two halves (disjoint item ranges,
 * disjoint node and reference ranges) can be built concurrently with the bytes a sequential
 * pre-order build writes. `spawn` levels below this one may still split onto new threads.
 */
void build_range(Item* items, size_t n, uint32_t index, uint32_t ref, const Builder& b, int spawn) {
    auto init = [](const float* box) {
        bf_bvh_node node{};
        node.min_x = box[0];
        node.min_y = box[1];
        node.min_z = box[2];
        node.max_x = box[3];
        node.max_y = box[4];
        node.max_z = box[5];
        node.split_axis = node.first_child = node.second_child = BF_INVALID_INDEX;
        return node;
    };
    if (n == 1) {
        bf_bvh_node leaf = init(items[0].b);
        leaf.kind = 1;
        leaf.reference_offset = ref;
        leaf.reference_count = 1;
        if (b.prim_refs) b.prim_refs[ref] = items[0].key;
        else b.inst_refs[ref] = items[0].ref;
        b.nodes[index] = leaf;
        return;
    }
    /* split axis: the largest binary64 centroid extent, the first axis on ties */
    double mn[3], mx[3];
    for (int a = 0; a < 3; ++a) mn[a] = mx[a] = items[0].c[a];
    for (size_t i = 1; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            mn[a] = std::min(mn[a], items[i].c[a]);
            mx[a] = std::max(mx[a], items[i].c[a]);
        }
    uint32_t axis = 0;
    double largest = mx[0] - mn[0];
    for (uint32_t a = 1; a < 3; ++a)
        if (mx[a] - mn[a] > largest) {
            largest = mx[a] - mn[a];
            axis = a;
        }
    /* a strict total order (centroid, then key): any correct sort gives the same permutation */
    std::sort(items, items + n, [axis](const Item& l, const Item& r) {
        if (l.c[axis] < r.c[axis]) return true;
        if (r.c[axis] < l.c[axis]) return false;
        return l.key < r.key;
    });
    const size_t half = n / 2;
    const uint32_t first = index + 1, second = index + (uint32_t)(2 * half);
    if (spawn > 0 && n >= kParallelBuildItems) {
        std::thread worker([&] { build_range(items, half, first, ref, b, spawn - 1); });
        build_range(items + half, n - half, second, ref + (uint32_t)half, b, spawn - 1);
        worker.join();
    } else {
        build_range(items, half, first, ref, b, 0);
        build_range(items + half, n - half, second, ref + (uint32_t)half, b, 0);
    }
    const bf_bvh_node f = b.nodes[first], g = b.nodes[second];
    const float box[6] = {canonical_zero(std::min(f.min_x, g.min_x)), canonical_zero(std::min(f.min_y, g.min_y)),
                          canonical_zero(std::min(f.min_z, g.min_z)), canonical_zero(std::max(f.max_x, g.max_x)),
                          canonical_zero(std::max(f.max_y, g.max_y)), canonical_zero(std::max(f.max_z, g.max_z))};
    bf_bvh_node node = init(box);
    node.split_axis = axis;
    node.first_child = first;
    node.second_child = second;
    b.nodes[index] = node;
}

/* Host threads for scene preparation. */
uint32_t build_threads() {
    const unsigned t = std::thread::hardware_concurrency();
    return t == 0 ? 1u : std::min(t, 64u);
}

/* levels of build_range that may split onto new threads so that about `threads` run at once */
int spawn_levels(uint32_t threads) {
    int levels = 0;
    while ((1u << levels) < threads && levels < 6) ++levels;
    return levels;
}

/* Runs task(0 .. count-1) on up to `threads` threads (dynamic assignment; tasks are independent). */
template <typename Task>
void parallel_for(uint32_t count, uint32_t threads, Task task) {
    const uint32_t workers = std::min(threads, count);
    if (workers <= 1) {
        for (uint32_t i = 0; i < count; ++i) task(i);
        return;
    }
    std::atomic<uint32_t> next{0};
    auto run = [&] {
        for (uint32_t i = next.fetch_add(1); i < count; i = next.fetch_add(1)) task(i);
    };
    std::vector<std::thread> pool;
    pool.reserve(workers - 1);
    for (uint32_t w = 1; w < workers; ++w) pool.emplace_back(run);
    run();
    for (std::thread& t : pool) t.join();
}

void item_centroid(Item& it) {
    for (int a = 0; a < 3; ++a) it.c[a] = ((double)it.b[a] + (double)it.b[3 + a]) * 0.5;
}

/* SceneBvh.cpp:555-601 */
bool point_bounds(const float* m, const float* p, float* out) {
    const double eps = (double)0x1p-23f;
    const double gamma7 = (7.0 * eps) / (1.0 - 7.0 * eps);
    const double underflow = 7.0 * (double)0x1p-149f;
    const double maximum = (double)3.40282346638528859812e+38f;
    for (int row = 0; row < 3; ++row) {
        double exact = (double)m[row * 4 + 3];
        double magnitude = std::fabs(exact);
        for (int col = 0; col < 3; ++col) {
            const double product = (double)m[row * 4 + col] * (double)p[col];
            exact += product;
            magnitude += std::fabs(product);
        }
        const double error = std::fma(gamma7, magnitude, underflow);
        const double lower = exact - error, upper = exact + error;
        if (!std::isfinite(exact) || !std::isfinite(magnitude) || lower < -maximum || upper > maximum) return false;
        float rl = (float)lower, ru = (float)upper;
        if ((double)rl > lower) rl = next_down(rl);
        if ((double)ru < upper) ru = next_up(ru);
        out[row] = canonical_zero(rl);
        out[3 + row] = canonical_zero(ru);
    }
    return true;
}

uint32_t tree_depth(const std::vector<bf_bvh_node>& nodes, uint32_t root) {
    uint32_t best = 0;
    std::vector<std::pair<uint32_t, uint32_t>> stack{{root, 1u}};
    while (!stack.empty()) {
        auto [n, d] = stack.back();
        stack.pop_back();
        best = std::max(best, d);
        if (nodes[n].kind == 0) {
            stack.push_back({nodes[n].first_child, d + 1});
            stack.push_back({nodes[n].second_child, d + 1});
        }
    }
    return best;
}

} /* namespace */

/* ---- scene ----------------------------------------------------------------------------------- */
struct bf_hip_scene {
    int device = -1;
    std::vector<bf_bvh_blas> blases;
    std::vector<bf_bvh_node> blas_nodes, tlas_nodes;
    std::vector<uint32_t> prim_refs;
    std::vector<bf_bvh_instance_ref> inst_refs;
    uint32_t tlas_root = BF_INVALID_INDEX;
    std::vector<bf_wide_node> blas_wide, tlas_wide;
    void* device_block = nullptr;
    float* light_cdf = nullptr; /* device copy of the power-weighted selection CDF, if bound */
    bf_dev_scene dev{};
    uint32_t max_blas_depth = 0, tlas_depth = 0;
    uint32_t light_count = 0;
    std::vector<uint32_t> texture_floats; /* width * height * channels per texture slot */
    bool has_maps = false;                /* some material has a normal or bump map */
    bool sampled = false;                 /* holds point-sampled spectra at sampled_wavelengths */
    float sampled_wavelengths[BF_SPECTRUM_LANES] = {};
    uint64_t source_hash = 0;             /* content hash of the scene blob it was made from, if any */
};

void bf_scene_set_source_hash(bf_hip_scene* s, uint64_t hash) { s->source_hash = hash; }

namespace {

/* Canonical pre-order nodes -> wide nodes (one per internal node, same relative order). */
void relayout(const std::vector<bf_bvh_node>& nodes, uint32_t node_offset, uint32_t node_count, uint32_t leaf_base_is_ref,
              std::vector<bf_wide_node>& wide, std::vector<uint32_t>& canonical_to_ref) {
    (void)leaf_base_is_ref;
    for (uint32_t i = node_offset; i < node_offset + node_count; ++i) {
        if (nodes[i].kind == 0) {
            canonical_to_ref[i] = (uint32_t)wide.size();
            wide.emplace_back();
        } else {
            canonical_to_ref[i] = BF_WIDE_LEAF | nodes[i].reference_offset;
        }
    }
    for (uint32_t i = node_offset; i < node_offset + node_count; ++i) {
        if (nodes[i].kind != 0) continue;
        bf_wide_node& w = wide[canonical_to_ref[i]];
        const bf_bvh_node* c[2] = {&nodes[nodes[i].first_child], &nodes[nodes[i].second_child]};
        for (int k = 0; k < 2; ++k) {
            w.lo_x[k] = c[k]->min_x;
            w.lo_y[k] = c[k]->min_y;
            w.lo_z[k] = c[k]->min_z;
            w.hi_x[k] = c[k]->max_x;
            w.hi_y[k] = c[k]->max_y;
            w.hi_z[k] = c[k]->max_z;
        }
        w.child[0] = canonical_to_ref[nodes[i].first_child];
        w.child[1] = canonical_to_ref[nodes[i].second_child];
        w.canonical_index = i;
        w.split_axis = nodes[i].split_axis;
    }
}

bool valid_box(const bf_bvh_node& n) {
    return std::isfinite(n.min_x) && std::isfinite(n.min_y) && std::isfinite(n.min_z) && std::isfinite(n.max_x) &&
           std::isfinite(n.max_y) && std::isfinite(n.max_z) && n.min_x <= n.max_x && n.min_y <= n.max_y &&
           n.min_z <= n.max_z;
}

struct Arena {
    std::vector<std::pair<size_t, const void*>> parts;
    std::vector<size_t> sizes;
    size_t total = 0;
    size_t add(const void* data, size_t bytes) {
        const size_t offset = total;
        parts.push_back({offset, data});
        sizes.push_back(bytes);
        total += (bytes + 255u) & ~(size_t)255u;
        return offset;
    }
};

int build_scene(const bf_scene_desc* d, bf_hip_scene* s, bool upload) {
    if (d->geometry_count == 0 || d->instance_count == 0 || d->material_count == 0 || !d->positions || !d->normals ||
        !d->uvs || !d->indices || !d->geometries || !d->instances || !d->materials || !d->spectra ||
        !d->spectrum_values)
        return set_error(BF_INVALID_ARGUMENT, "scene description is missing arrays or counts");
    if (d->vertex_count > 0xFFFFFFFFull || d->triangle_count > 0x7FFFFFFFull || d->instance_count > 0x7FFFFFFFu)
        return set_error(BF_RESOURCE_EXHAUSTED, "scene exceeds the 31-bit reference domain of the wide BVH");
    /* -- BLAS per geometry -- *
     * Geometries are checked in order first; their BLASes then build concurrently into node and
     * reference ranges fixed in advance (2k - 1 nodes and k references for k triangles), which gives
     * the bytes of a build in geometry order. The error reported is the first geometry's in order. */
    const uint32_t threads = build_threads();
    uint32_t prev_id = 0;
    std::vector<uint64_t> node_base(d->geometry_count), ref_base(d->geometry_count);
    uint64_t total_nodes = 0, total_refs = 0;
    uint32_t checked = d->geometry_count; /* geometries before this one passed the header checks */
    std::string header_error;
    for (uint32_t g = 0; g < d->geometry_count; ++g) {
        const bf_geometry_desc& geo = d->geometries[g];
        if (g > 0 && geo.id <= prev_id) {
            checked = g;
            header_error = "geometry identifiers must be strictly increasing";
            break;
        }
        prev_id = geo.id;
        if (geo.triangle_count == 0 || geo.vertex_count == 0 || geo.vertex_offset + geo.vertex_count > d->vertex_count ||
            geo.triangle_offset + geo.triangle_count > d->triangle_count) {
            checked = g;
            header_error = "geometry " + std::to_string(g) + " has an invalid range";
            break;
        }
        node_base[g] = total_nodes;
        ref_base[g] = total_refs;
        total_nodes += 2 * geo.triangle_count - 1;
        total_refs += geo.triangle_count;
    }
    if (total_nodes > 0x7FFFFFFFull) return set_error(BF_RESOURCE_EXHAUSTED, "scene exceeds the 31-bit BVH node domain");
    s->blas_nodes.assign(total_nodes, bf_bvh_node{});
    s->prim_refs.assign(total_refs, 0u);
    s->blases.assign(checked, bf_bvh_blas{});
    std::vector<std::string> geometry_error(checked);
    const Builder bb{s->blas_nodes.data(), s->prim_refs.data(), nullptr};
    const bool few = checked < threads;
    parallel_for(checked, few ? 1u : threads, [&](uint32_t g) {
        const bf_geometry_desc& geo = d->geometries[g];
        std::vector<Item> items(geo.triangle_count);
        for (uint32_t p = 0; p < geo.triangle_count; ++p) {
            const uint64_t t = geo.triangle_offset + p;
            Item& it = items[p];
            for (int c = 0; c < 3; ++c) {
                const uint32_t vi = d->indices[3 * t + c];
                if (vi >= geo.vertex_count) {
                    geometry_error[g] = "triangle index out of range";
                    return;
                }
                const float* pos = d->positions + 3 * (geo.vertex_offset + vi);
                for (int a = 0; a < 3; ++a) {
                    if (!std::isfinite(pos[a])) {
                        geometry_error[g] = "non-finite vertex position";
                        return;
                    }
                    if (c == 0) it.b[a] = it.b[3 + a] = canonical_zero(pos[a]);
                    else {
                        it.b[a] = canonical_zero(std::min(it.b[a], pos[a]));
                        it.b[3 + a] = canonical_zero(std::max(it.b[3 + a], pos[a]));
                    }
                }
            }
            item_centroid(it);
            it.key = p;
        }
        const uint32_t root = (uint32_t)node_base[g];
        build_range(items.data(), items.size(), root, (uint32_t)ref_base[g], bb, few ? spawn_levels(threads) : 0);
        bf_bvh_blas& blas = s->blases[g];
        blas.geometry_id = geo.id;
        blas.root_node = root;
        blas.node_offset = root;
        blas.node_count = (uint32_t)(2 * geo.triangle_count - 1);
        blas.primitive_reference_offset = (uint32_t)ref_base[g];
        blas.primitive_reference_count = (uint32_t)geo.triangle_count;
    });
    for (uint32_t g = 0; g < checked; ++g)
        if (!geometry_error[g].empty()) return set_error(BF_INVALID_ARGUMENT, "%s", geometry_error[g].c_str());
    if (checked < d->geometry_count) return set_error(BF_INVALID_ARGUMENT, "%s", header_error.c_str());
    for (uint32_t g = 0; g < d->geometry_count; ++g)
        s->max_blas_depth = std::max(s->max_blas_depth, tree_depth(s->blas_nodes, s->blases[g].root_node));
    /* -- TLAS over instances (SceneBvh.cpp:603-736); instance bounds are formed concurrently -- */
    std::vector<Item> items(d->instance_count);
    prev_id = 0;
    uint32_t valid = d->instance_count; /* instances before this one passed their checks */
    std::string instance_error;
    for (uint32_t i = 0; i < d->instance_count && valid == d->instance_count; ++i) {
        const bf_instance_desc& inst = d->instances[i];
        char message[128] = {};
        if (i > 0 && inst.id <= prev_id) std::snprintf(message, sizeof message, "instance identifiers must be strictly increasing");
        else if (inst.geometry_index >= d->geometry_count || inst.material_index >= d->material_count)
            std::snprintf(message, sizeof message, "instance %u references a missing geometry or material", i);
        else
            for (int k = 0; k < 12; ++k)
                if (!std::isfinite(inst.local_to_world[k]) || !std::isfinite(inst.world_to_local[k]))
                    std::snprintf(message, sizeof message, "instance %u has a non-finite transform", i);
        prev_id = inst.id;
        if (message[0]) {
            valid = i;
            instance_error = message;
        }
    }
    std::vector<uint8_t> bounds_ok(valid, 1u);
    parallel_for(valid, threads, [&](uint32_t i) {
        const bf_instance_desc& inst = d->instances[i];
        const bf_geometry_desc& geo = d->geometries[inst.geometry_index];
        bool have = false;
        float box[6];
        for (uint64_t p = 0; p < geo.triangle_count; ++p)
            for (int c = 0; c < 3; ++c) {
                const float* pos = d->positions + 3 * (geo.vertex_offset + d->indices[3 * (geo.triangle_offset + p) + c]);
                float pb[6];
                if (!point_bounds(inst.local_to_world, pos, pb)) {
                    bounds_ok[i] = 0u;
                    return;
                }
                if (!have) {
                    std::memcpy(box, pb, sizeof box);
                    have = true;
                } else
                    for (int a = 0; a < 3; ++a) {
                        box[a] = canonical_zero(std::min(box[a], pb[a]));
                        box[3 + a] = canonical_zero(std::max(box[3 + a], pb[3 + a]));
                    }
            }
        Item& it = items[i];
        std::memcpy(it.b, box, sizeof box);
        item_centroid(it);
        it.key = inst.id;
        it.ref = bf_bvh_instance_ref{};
        it.ref.scene_instance_index = i;
        it.ref.instance_id = inst.id;
        it.ref.scene_geometry_index = inst.geometry_index;
        it.ref.geometry_id = geo.id;
        it.ref.blas_index = inst.geometry_index;
        it.ref.visibility_mask = inst.visibility_mask;
    });
    for (uint32_t i = 0; i < valid; ++i)
        if (!bounds_ok[i]) return set_error(BF_INVALID_ARGUMENT, "instance %u bounds are not representable", i);
    if (valid < d->instance_count) return set_error(BF_INVALID_ARGUMENT, "%s", instance_error.c_str());
    s->tlas_nodes.assign(2 * (size_t)d->instance_count - 1, bf_bvh_node{});
    s->inst_refs.assign(d->instance_count, bf_bvh_instance_ref{});
    const Builder tb{s->tlas_nodes.data(), nullptr, s->inst_refs.data()};
    build_range(items.data(), items.size(), 0u, 0u, tb, spawn_levels(threads));
    s->tlas_root = 0u;
    s->tlas_depth = tree_depth(s->tlas_nodes, s->tlas_root);
    for (const auto& n : s->blas_nodes)
        if (!valid_box(n)) return set_error(BF_INVALID_ARGUMENT, "BLAS node box is not finite and ordered");
    for (const auto& n : s->tlas_nodes)
        if (!valid_box(n)) return set_error(BF_INVALID_ARGUMENT, "TLAS node box is not finite and ordered");

    /* -- wide relayout -- */
    std::vector<uint32_t> blas_ref(s->blas_nodes.size()), tlas_ref(s->tlas_nodes.size());
    std::vector<bf_dev_blas> dev_blases(d->geometry_count);
    for (uint32_t g = 0; g < d->geometry_count; ++g) {
        const bf_bvh_blas& b = s->blases[g];
        relayout(s->blas_nodes, b.node_offset, b.node_count, 0, s->blas_wide, blas_ref);
        bf_dev_blas& db = dev_blases[g];
        db = bf_dev_blas{};
        db.root_ref = blas_ref[b.root_node];
        db.geometry_index = g;
        db.canonical_node_offset = b.node_offset;
        const bf_bvh_node& r = s->blas_nodes[b.root_node];
        const float rb[6] = {r.min_x, r.min_y, r.min_z, r.max_x, r.max_y, r.max_z};
        std::memcpy(db.root_box, rb, sizeof rb);
    }
    relayout(s->tlas_nodes, 0, (uint32_t)s->tlas_nodes.size(), 1, s->tlas_wide, tlas_ref);
    if (s->blas_wide.size() > BF_REF_MASK || s->prim_refs.size() > BF_REF_MASK || s->tlas_wide.size() > BF_REF_MASK ||
        s->inst_refs.size() > BF_REF_MASK)
        return set_error(BF_RESOURCE_EXHAUSTED, "scene exceeds the 30-bit node and leaf references of the traversal");
    /* own boxes of wide nodes and leaves, for the traversal's exact pop-test fallback */
    auto own_boxes = [](const std::vector<bf_bvh_node>& nodes, const std::vector<uint32_t>& refs, size_t wide_count,
                        size_t leaf_count, std::vector<float>& node_boxes, std::vector<float>& leaf_boxes) {
        node_boxes.assign(6 * wide_count, 0.0f);
        leaf_boxes.assign(6 * leaf_count, 0.0f);
        for (size_t i = 0; i < nodes.size(); ++i) {
            const bf_bvh_node& n = nodes[i];
            const float box[6] = {n.min_x, n.min_y, n.min_z, n.max_x, n.max_y, n.max_z};
            const uint32_t ref = refs[i];
            float* dst = (ref & BF_WIDE_LEAF) ? &leaf_boxes[6 * (size_t)(ref & BF_REF_MASK)] : &node_boxes[6 * (size_t)ref];
            std::memcpy(dst, box, sizeof box);
        }
    };
    std::vector<float> blas_node_boxes, blas_leaf_boxes, tlas_node_boxes, tlas_leaf_boxes;
    own_boxes(s->blas_nodes, blas_ref, s->blas_wide.size(), s->prim_refs.size(), blas_node_boxes, blas_leaf_boxes);
    own_boxes(s->tlas_nodes, tlas_ref, s->tlas_wide.size(), s->inst_refs.size(), tlas_node_boxes, tlas_leaf_boxes);
    /* leaf triangles in BLAS leaf order: object-space vertices + primitive index */
    std::vector<float> leaf_tris(s->prim_refs.size() * 12);
    for (uint32_t g = 0; g < d->geometry_count; ++g) {
        const bf_bvh_blas& b = s->blases[g];
        const bf_geometry_desc& geo = d->geometries[g];
        for (uint32_t r = 0; r < b.primitive_reference_count; ++r) {
            const uint32_t ref = b.primitive_reference_offset + r;
            const uint32_t prim = s->prim_refs[ref];
            for (int c = 0; c < 3; ++c) {
                const float* pos = d->positions + 3 * (geo.vertex_offset + d->indices[3 * (geo.triangle_offset + prim) + c]);
                leaf_tris[12 * (size_t)ref + 4 * c + 0] = pos[0];
                leaf_tris[12 * (size_t)ref + 4 * c + 1] = pos[1];
                leaf_tris[12 * (size_t)ref + 4 * c + 2] = pos[2];
                float payload = 0.0f;
                if (c == 0) std::memcpy(&payload, &prim, sizeof payload);
                leaf_tris[12 * (size_t)ref + 4 * c + 3] = payload;
            }
        }
    }
    std::vector<bf_dev_tlas_leaf> tlas_leaves(s->inst_refs.size());
    for (size_t r = 0; r < s->inst_refs.size(); ++r) {
        const bf_bvh_instance_ref& ir = s->inst_refs[r];
        bf_dev_tlas_leaf& L = tlas_leaves[r];
        std::memcpy(L.world_to_local, d->instances[ir.scene_instance_index].world_to_local, sizeof L.world_to_local);
        L.scene_instance = ir.scene_instance_index;
        L.instance_id = ir.instance_id;
        L.blas_index = ir.blas_index;
        L.visibility_mask = ir.visibility_mask;
    }
    /* -- shading tables -- */
    std::vector<bf_dev_instance> instances(d->instance_count);
    std::vector<bf_dev_geometry> geometries(d->geometry_count);
    for (uint32_t g = 0; g < d->geometry_count; ++g) {
        const bf_geometry_desc& geo = d->geometries[g];
        geometries[g] = bf_dev_geometry{geo.id, (uint32_t)geo.vertex_offset, (uint32_t)geo.vertex_count,
                                        (uint32_t)geo.triangle_offset, (uint32_t)geo.triangle_count, {0, 0, 0}};
    }
    /* mesh-area lights: instances with non-black emission, in instance order (FrameScene.cpp:200-264) */
    std::vector<bf_dev_light> lights;
    std::vector<float> areas, probs, cumul, nexts;
    std::vector<float> normals4;
    for (uint32_t i = 0; i < d->instance_count; ++i) {
        const bf_instance_desc& inst = d->instances[i];
        bf_dev_instance& di = instances[i];
        di = bf_dev_instance{};
        std::memcpy(di.local_to_world, inst.local_to_world, sizeof di.local_to_world);
        std::memcpy(di.world_to_local, inst.world_to_local, sizeof di.world_to_local);
        di.geometry_index = inst.geometry_index;
        di.material_index = inst.material_index;
        di.id = inst.id;
        di.object_id = inst.object_id;
        di.visibility_mask = inst.visibility_mask;
        di.light_index = -1;
        const bf_material_desc& m = d->materials[inst.material_index];
        if (m.emission_spectrum == BF_INVALID_INDEX) continue;
        if (m.emission_spectrum >= d->spectrum_count) return set_error(BF_INVALID_ARGUMENT, "emission spectrum out of range");
        const bf_spectrum_desc& sp = d->spectra[m.emission_spectrum];
        bool black = true;
        const bool sampled = (sp.count & BF_SPECTRUM_SAMPLED) != 0;
        const uint32_t first = sampled ? BF_SPECTRUM_LANES : 0u, count = sampled ? BF_SPECTRUM_LANES : sp.count;
        for (uint32_t k = 0; k < count; ++k)
            if (d->spectrum_values[sp.offset + first + k] != 0.0f) black = false;
        if (black) continue;
        const bf_geometry_desc& geo = d->geometries[inst.geometry_index];
        const uint32_t n = (uint32_t)geo.triangle_count;
        bf_dev_light L{};
        L.instance = i;
        L.geometry = inst.geometry_index;
        L.triangle_count = n;
        L.table_offset = (uint32_t)areas.size();
        std::vector<float> tri_area(n);
        float sum = 0.0f, comp = 0.0f;
        for (uint32_t t = 0; t < n; ++t) {
            float w[3][3];
            for (int c = 0; c < 3; ++c) {
                const float* p = d->positions + 3 * (geo.vertex_offset + d->indices[3 * (geo.triangle_offset + t) + c]);
                const float* M = inst.local_to_world;
                w[c][0] = std::fma(M[0], p[0], std::fma(M[1], p[1], std::fma(M[2], p[2], M[3])));
                w[c][1] = std::fma(M[4], p[0], std::fma(M[5], p[1], std::fma(M[6], p[2], M[7])));
                w[c][2] = std::fma(M[8], p[0], std::fma(M[9], p[1], std::fma(M[10], p[2], M[11])));
            }
            float area, nrm[3];
            if (!triangle_area(w, area, nrm)) return set_error(BF_INVALID_ARGUMENT, "emitter triangle area is not representable");
            tri_area[t] = area;
            normals4.insert(normals4.end(), {nrm[0], nrm[1], nrm[2], 0.0f});
            /* compensated running sum (CudaWavefrontStages.cu:2085-2107) */
            const float next = sum + area;
            comp += std::fabs(sum) >= std::fabs(area) ? (sum - next) + area : (area - next) + sum;
            sum = next;
        }
        const float total = sum + comp;
        L.inverse_area = 1.0f / total;
        float cumulative = 0.0f;
        for (uint32_t t = 0; t < n; ++t) {
            float probability = 1.0f - cumulative, next = 1.0f;
            if (t + 1 != n) {
                probability = tri_area[t] / total;
                next = cumulative + probability;
                if (!(probability > 0.0f) || !(next > cumulative) || !(next < 1.0f))
                    return set_error(BF_INVALID_ARGUMENT, "emitter area distribution is not representable");
            } else if (!(probability > 0.0f) || probability > 1.0f) {
                return set_error(BF_INVALID_ARGUMENT, "emitter area distribution is not representable");
            }
            areas.push_back(tri_area[t]);
            probs.push_back(probability);
            cumul.push_back(cumulative);
            nexts.push_back(next);
            cumulative = next;
        }
        di.light_index = (int32_t)lights.size();
        lights.push_back(L);
    }
    if (lights.size() > (1u << 24)) return set_error(BF_RESOURCE_EXHAUSTED, "more than 2^24 lights");

    if (d->has_environment && d->environment_spectrum >= d->spectrum_count)
        return set_error(BF_INVALID_ARGUMENT, "environment spectrum out of range");

    /* -- punctual lights: the scene checks sample_punctual_light makes per sample
     *    (CudaWavefrontStages.cu:2246-2340) are made once here, with the spot-cone cosines -- */
    if (d->punctual_count && !d->punctuals) return set_error(BF_INVALID_ARGUMENT, "punctual lights without an array");
    std::vector<bf_dev_punctual> punctuals(d->punctual_count);
    for (uint32_t i = 0; i < d->punctual_count; ++i) {
        const bf_punctual_desc& L = d->punctuals[i];
        bf_dev_punctual& D = punctuals[i];
        D = bf_dev_punctual{};
        D.kind = L.kind;
        D.spectrum = L.spectrum;
        std::memcpy(D.position, L.position, sizeof D.position);
        std::memcpy(D.position_error, L.position_error, sizeof D.position_error);
        std::memcpy(D.direction, L.direction, sizeof D.direction);
        if (L.kind > BF_PUNCTUAL_SPOT) return set_error(BF_INVALID_ARGUMENT, "punctual light %u: unknown kind %u", i, L.kind);
        if (L.spectrum >= d->spectrum_count) return set_error(BF_INVALID_ARGUMENT, "punctual light %u: spectrum out of range", i);
        if (L.kind != BF_PUNCTUAL_DIRECTIONAL)
            for (int a = 0; a < 3; ++a)
                if (!std::isfinite(L.position[a]) || !std::isfinite(L.position_error[a]) || !(L.position_error[a] >= 0.0f))
                    return set_error(BF_INVALID_ARGUMENT, "punctual light %u: position and its error must be finite, error >= 0", i);
        if (L.kind != BF_PUNCTUAL_POINT) {
            /* transport::unit_vector (TransportDevice.cuh:325-333) */
            const float* v = L.direction;
            const float sq = std::fma(v[0], v[0], std::fma(v[1], v[1], v[2] * v[2]));
            if (!std::isfinite(v[0]) || !std::isfinite(v[1]) || !std::isfinite(v[2]) || !std::isfinite(sq) ||
                !(std::fabs(sq - 1.0f) <= 128.0f * FLT_EPSILON))
                return set_error(BF_INVALID_ARGUMENT, "punctual light %u: direction must be a unit vector", i);
        }
        if (L.kind == BF_PUNCTUAL_SPOT) {
            const float pi = 3.14159265358979323846f, inner = L.inner_half_angle, outer = L.outer_half_angle;
            if (!std::isfinite(inner) || !std::isfinite(outer) || inner < 0.0f || inner > outer || !(outer > 0.0f) ||
                !(outer < pi))
                return set_error(BF_INVALID_ARGUMENT, "punctual light %u: spot angles need 0 <= inner <= outer < pi, outer > 0", i);
            /* exact_special_cosine (:515-526) */
            auto special_cosine = [pi](float angle) {
                if (angle == 0.0f) return 1.0f;
                if (angle == pi / 2.0f) return 0.0f;
                return bf_cosf(angle);
            };
            D.cosine_inner = special_cosine(inner);
            D.cosine_outer = special_cosine(outer);
            if (!std::isfinite(D.cosine_inner) || !std::isfinite(D.cosine_outer) || D.cosine_inner < D.cosine_outer)
                return set_error(BF_INVALID_ARGUMENT, "punctual light %u: spot cone cosines are not representable", i);
            D.smooth = inner != outer ? 1u : 0u;
        }
    }
    if (lights.size() + punctuals.size() > (1u << 24)) return set_error(BF_RESOURCE_EXHAUSTED, "more than 2^24 lights");
    s->light_count = (uint32_t)(lights.size() + punctuals.size());

    /* -- textures: albedo / roughness images bound to closures (no reference counterpart) -- */
    if (d->texture_count > BF_MAX_TEXTURES || (d->texture_count && !d->textures))
        return set_error(BF_INVALID_ARGUMENT, "at most %u textures, and a texture array when any", BF_MAX_TEXTURES);
    for (uint32_t t = 0; t < d->texture_count; ++t) {
        const bf_texture_desc& T = d->textures[t];
        if (T.width == 0 || T.height == 0 || T.channels == 0 || T.channels > 3 ||
            (uint64_t)T.width * T.height * T.channels > 0xFFFFFFFFull)
            return set_error(BF_INVALID_ARGUMENT, "texture %u has an invalid size", t);
        if ((T.data == nullptr) == (T.device_data == nullptr))
            return set_error(BF_INVALID_ARGUMENT, "texture %u needs exactly one of host data and device data", t);
    }
    s->texture_floats.clear();
    for (uint32_t t = 0; t < d->texture_count; ++t)
        s->texture_floats.push_back(d->textures[t].width * d->textures[t].height * d->textures[t].channels);
    bool textured = false;
    for (uint32_t c = 0; c < d->closure_count; ++c) {
        const bf_closure_desc& C = d->closures[c];
        if (C.weight_texture) {
            if (C.weight_texture > d->texture_count) return set_error(BF_INVALID_ARGUMENT, "closure %u: no texture %u", c, C.weight_texture - 1);
            const uint32_t ch = d->textures[C.weight_texture - 1].channels;
            if (ch != 1 && ch != 3) return set_error(BF_INVALID_ARGUMENT, "closure %u: weight textures have 1 or 3 channels", c);
            textured = true;
        }
        if (C.roughness_texture) {
            if (C.roughness_texture > d->texture_count)
                return set_error(BF_INVALID_ARGUMENT, "closure %u: no texture %u", c, C.roughness_texture - 1);
            if (C.kind != BF_CLOSURE_ROUGH_CONDUCTOR && C.kind != BF_CLOSURE_ROUGH_DIELECTRIC)
                return set_error(BF_INVALID_ARGUMENT, "closure %u: roughness textures drive GGX widths (rough conductors and dielectrics)", c);
            const uint32_t ch = d->textures[C.roughness_texture - 1].channels;
            if (ch != 1 && ch != 2) return set_error(BF_INVALID_ARGUMENT, "closure %u: roughness textures have 1 or 2 channels", c);
            textured = true;
        }
    }
    /* -- surface maps: image mip chains and each material's normal / bump map, checked as the
     * reference's stages check them per lookup (load_image_texture, load_image_mip,
     * apply_normal_map, apply_bump_map: CudaWavefrontStages.cu:2876-2990, 3395-3560) -- */
    if (d->image_texture_count && !d->image_textures)
        return set_error(BF_INVALID_ARGUMENT, "image textures are declared without an array");
    if (d->reserved_maps != 0) return set_error(BF_INVALID_ARGUMENT, "reserved word after image_texture_count must be zero");
    std::vector<bf_dev_image> images(d->image_texture_count);
    uint64_t image_texels = 0;
    for (uint32_t t = 0; t < d->image_texture_count; ++t) {
        const bf_image_texture_desc& T = d->image_textures[t];
        if (T.width == 0 || T.height == 0 || T.channel_count == 0 || T.mip_count == 0 || T.mip_count > BF_MAX_IMAGE_MIPS ||
            T.width > (1u << 24) || T.height > (1u << 24) || !T.texels)
            return set_error(BF_INVALID_ARGUMENT, "image texture %u has an invalid size, mip count or no texels", t);
        uint64_t w = T.width, h = T.height, total = 0;
        for (uint32_t k = 0; k < T.mip_count; ++k) {
            total += w * h * T.channel_count;
            w = w > 1 ? w / 2 : 1;
            h = h > 1 ? h / 2 : 1;
        }
        if (total != T.texel_count)
            return set_error(BF_INVALID_ARGUMENT, "image texture %u holds %llu values, its mip chain %llu", t,
                             (unsigned long long)T.texel_count, (unsigned long long)total);
        images[t] = bf_dev_image{T.width, T.height, T.channel_count, T.mip_count, image_texels};
        image_texels += total;
    }
    s->has_maps = false;
    const auto map_ok = [&](uint32_t texture, uint32_t u_wrap, uint32_t v_wrap, uint32_t anisotropy, uint32_t visits) {
        return texture < d->image_texture_count && u_wrap <= BF_TEXTURE_WRAP_BORDER && v_wrap <= BF_TEXTURE_WRAP_BORDER &&
               anisotropy >= 1 && anisotropy <= BF_MAX_EWA_ANISOTROPY && visits >= 1;
    };
    for (uint32_t m = 0; d->surface_maps && m < d->material_count; ++m) {
        const bf_surface_maps_desc& M = d->surface_maps[m];
        if (M.normal_texture != BF_INVALID_INDEX) {
            const uint32_t* ch = M.normal_channels;
            if (!map_ok(M.normal_texture, M.normal_u_wrap, M.normal_v_wrap, M.normal_max_anisotropy, M.normal_max_texel_visits) ||
                M.normal_y_convention > 1u)
                return set_error(BF_INVALID_ARGUMENT, "material %u: invalid normal map", m);
            const uint32_t cc = d->image_textures[M.normal_texture].channel_count;
            if (ch[0] >= cc || ch[1] >= cc || ch[2] >= cc || ch[0] == ch[1] || ch[0] == ch[2] || ch[1] == ch[2])
                return set_error(BF_INVALID_ARGUMENT, "material %u: normal map channels must be distinct channels of its image", m);
            s->has_maps = true;
        }
        if (M.bump_texture != BF_INVALID_INDEX) {
            if (!map_ok(M.bump_texture, M.bump_u_wrap, M.bump_v_wrap, M.bump_max_anisotropy, M.bump_max_texel_visits) ||
                M.bump_channel >= d->image_textures[M.bump_texture].channel_count || !std::isfinite(M.bump_scale))
                return set_error(BF_INVALID_ARGUMENT, "material %u: invalid bump map", m);
            s->has_maps = true;
        }
    }
    /* -- spectra: grid spectra in range; point-sampled ones share one wavelength packet -- */
    s->sampled = false;
    for (uint32_t i = 0; i < d->spectrum_count; ++i) {
        const bf_spectrum_desc& sp = d->spectra[i];
        const bool sampled = (sp.count & BF_SPECTRUM_SAMPLED) != 0;
        const uint64_t span = sampled ? 2u * BF_SPECTRUM_LANES : (sp.count ? sp.count : 1u);
        if ((sampled && sp.count != (BF_SPECTRUM_SAMPLED | BF_SPECTRUM_LANES)) || (uint64_t)sp.offset + span > d->spectrum_value_count)
            return set_error(BF_INVALID_ARGUMENT, "spectrum %u has an invalid value range", i);
        if (!sampled) continue;
        const float* wl = d->spectrum_values + sp.offset;
        for (uint32_t k = 0; k < BF_SPECTRUM_LANES; ++k)
            if (!std::isfinite(wl[k]) || !(wl[k] > 0.0f))
                return set_error(BF_INVALID_ARGUMENT, "spectrum %u: sampled wavelengths must be finite and positive", i);
        if (!s->sampled) {
            std::memcpy(s->sampled_wavelengths, wl, sizeof s->sampled_wavelengths);
            s->sampled = true;
        } else if (std::memcmp(s->sampled_wavelengths, wl, sizeof s->sampled_wavelengths) != 0) {
            return set_error(BF_INVALID_ARGUMENT, "spectrum %u: every point-sampled spectrum of a scene shares one wavelength packet", i);
        }
    }
    if (!upload) return BF_OK;
    /* -- device upload: one allocation, 256-byte aligned parts -- */
    Arena A;
    const size_t o_tlas = A.add(s->tlas_wide.data(), s->tlas_wide.size() * sizeof(bf_wide_node));
    const size_t o_blasw = A.add(s->blas_wide.data(), s->blas_wide.size() * sizeof(bf_wide_node));
    const size_t o_blas = A.add(dev_blases.data(), dev_blases.size() * sizeof(bf_dev_blas));
    const size_t o_tris = A.add(leaf_tris.data(), leaf_tris.size() * sizeof(float));
    const size_t o_bnb = A.add(blas_node_boxes.data(), blas_node_boxes.size() * sizeof(float));
    const size_t o_blb = A.add(blas_leaf_boxes.data(), blas_leaf_boxes.size() * sizeof(float));
    const size_t o_tnb = A.add(tlas_node_boxes.data(), tlas_node_boxes.size() * sizeof(float));
    const size_t o_tlb = A.add(tlas_leaf_boxes.data(), tlas_leaf_boxes.size() * sizeof(float));
    const size_t o_leaves = A.add(tlas_leaves.data(), tlas_leaves.size() * sizeof(bf_dev_tlas_leaf));
    const size_t o_inst = A.add(instances.data(), instances.size() * sizeof(bf_dev_instance));
    const size_t o_geo = A.add(geometries.data(), geometries.size() * sizeof(bf_dev_geometry));
    const size_t o_pos = A.add(d->positions, d->vertex_count * 3 * sizeof(float));
    const size_t o_nrm = A.add(d->normals, d->vertex_count * 3 * sizeof(float));
    const size_t o_uv = A.add(d->uvs, d->vertex_count * 2 * sizeof(float));
    const size_t o_idx = A.add(d->indices, d->triangle_count * 3 * sizeof(uint32_t));
    const size_t o_mat = A.add(d->materials, d->material_count * sizeof(bf_material_desc));
    const size_t o_clo = A.add(d->closures, d->closure_count * sizeof(bf_closure_desc));
    const size_t o_spe = A.add(d->spectra, d->spectrum_count * sizeof(bf_spectrum_desc));
    const size_t o_spv = A.add(d->spectrum_values, d->spectrum_value_count * sizeof(float));
    const size_t o_lig = A.add(lights.data(), lights.size() * sizeof(bf_dev_light));
    const size_t o_are = A.add(areas.data(), areas.size() * sizeof(float));
    const size_t o_pro = A.add(probs.data(), probs.size() * sizeof(float));
    const size_t o_cum = A.add(cumul.data(), cumul.size() * sizeof(float));
    const size_t o_nex = A.add(nexts.data(), nexts.size() * sizeof(float));
    const size_t o_lno = A.add(normals4.data(), normals4.size() * sizeof(float));
    const size_t o_pun = A.add(punctuals.data(), punctuals.size() * sizeof(bf_dev_punctual));
    const size_t o_tex = A.add(nullptr, d->texture_count * sizeof(bf_dev_texture));
    const size_t o_img = A.add(images.data(), s->has_maps ? images.size() * sizeof(bf_dev_image) : 0);
    const size_t o_maps = A.add(d->surface_maps, s->has_maps ? d->material_count * sizeof(bf_surface_maps_desc) : 0);
    std::vector<float> all_texels;
    if (s->has_maps) {
        all_texels.reserve(image_texels);
        for (uint32_t t = 0; t < d->image_texture_count; ++t)
            all_texels.insert(all_texels.end(), d->image_textures[t].texels,
                              d->image_textures[t].texels + d->image_textures[t].texel_count);
    }
    const size_t o_itx = A.add(all_texels.data(), all_texels.size() * sizeof(float));
    size_t o_texels[BF_MAX_TEXTURES] = {};
    for (uint32_t t = 0; t < d->texture_count; ++t) {
        const bf_texture_desc& T = d->textures[t];
        if (T.data) o_texels[t] = A.add(T.data, (size_t)T.width * T.height * T.channels * sizeof(float));
    }
    HIP_TRY(hipGetDevice(&s->device));
    HIP_TRY(hipMalloc(&s->device_block, A.total ? A.total : 256));
    uint8_t* base = (uint8_t*)s->device_block;
    std::vector<uint8_t> staging(A.total, 0);
    for (size_t k = 0; k < A.parts.size(); ++k)
        if (A.sizes[k] && A.parts[k].second) std::memcpy(staging.data() + A.parts[k].first, A.parts[k].second, A.sizes[k]);
    for (uint32_t t = 0; t < d->texture_count; ++t) {
        const bf_texture_desc& T = d->textures[t];
        const bf_dev_texture D{T.data ? (const float*)(base + o_texels[t]) : T.device_data, T.width, T.height, T.channels, 0u};
        std::memcpy(staging.data() + o_tex + t * sizeof(bf_dev_texture), &D, sizeof D);
    }
    HIP_TRY(hipMemcpy(s->device_block, staging.data(), A.total, hipMemcpyHostToDevice));
    bf_dev_scene& S = s->dev;
    S.tlas_nodes = (const bf_wide_node*)(base + o_tlas);
    S.blas_nodes = (const bf_wide_node*)(base + o_blasw);
    S.blases = (const bf_dev_blas*)(base + o_blas);
    S.leaf_triangles = (const float4*)(base + o_tris);
    S.blas_node_boxes = (const float*)(base + o_bnb);
    S.blas_leaf_boxes = (const float*)(base + o_blb);
    S.tlas_node_boxes = (const float*)(base + o_tnb);
    S.tlas_leaf_boxes = (const float*)(base + o_tlb);
    S.tlas_leaves = (const bf_dev_tlas_leaf*)(base + o_leaves);
    S.tlas_root_ref = tlas_ref[s->tlas_root];
    S.instance_count = d->instance_count;
    const bf_bvh_node& tr = s->tlas_nodes[s->tlas_root];
    const float trb[6] = {tr.min_x, tr.min_y, tr.min_z, tr.max_x, tr.max_y, tr.max_z};
    std::memcpy(S.tlas_root_box, trb, sizeof trb);
    S.instances = (const bf_dev_instance*)(base + o_inst);
    S.geometries = (const bf_dev_geometry*)(base + o_geo);
    S.positions = (const float*)(base + o_pos);
    S.normals = (const float*)(base + o_nrm);
    S.uvs = (const float*)(base + o_uv);
    S.indices = (const uint32_t*)(base + o_idx);
    S.materials = (const bf_material_desc*)(base + o_mat);
    S.closures = (const bf_closure_desc*)(base + o_clo);
    S.spectra = (const bf_spectrum_desc*)(base + o_spe);
    S.spectrum_values = (const float*)(base + o_spv);
    S.lights = (const bf_dev_light*)(base + o_lig);
    S.light_areas = (const float*)(base + o_are);
    S.light_probabilities = (const float*)(base + o_pro);
    S.light_cumulative = (const float*)(base + o_cum);
    S.light_next = (const float*)(base + o_nex);
    S.light_normals = (const float4*)(base + o_lno);
    S.punctuals = (const bf_dev_punctual*)(base + o_pun);
    S.punctual_count = (uint32_t)punctuals.size();
    S.light_count = (uint32_t)(lights.size() + punctuals.size());
    S.has_environment = d->has_environment ? 1u : 0u;
    S.environment_spectrum = d->environment_spectrum;
    S.material_count = d->material_count;
    S.textures = (const bf_dev_texture*)(base + o_tex);
    S.texture_count = d->texture_count;
    S.has_textures = textured ? 1u : 0u;
    S.images = nullptr;
    S.image_texels = nullptr;
    S.maps = nullptr;
    if (s->has_maps) {
        S.images = (const bf_dev_image*)(base + o_img);
        S.image_texels = (const float*)(base + o_itx);
        S.maps = (const bf_surface_maps_desc*)(base + o_maps);
    }
    return BF_OK;
}

} /* namespace */

/* ---- workspace --------------------------------------------------------------------------------- */
#define BF_MAX_BATCH_JOBS 4096u /* frames per batched launch (job ids are 16-bit) */
#define BF_HISTORY_BOUNCES 256u  /* per-bounce queue sizes kept on the device per launch */

struct bf_hip_workspace {
    uint64_t capacity = 0;
    void* block = nullptr;
    bf_dev_paths paths{};
    uint32_t* queues[4] = {nullptr, nullptr, nullptr, nullptr}; /* ray, hit, shadow, continuation */
    uint8_t* route = nullptr;
    uint32_t* tile_counts = nullptr;
    uint32_t* tile_offsets = nullptr;
    uint32_t* sizes = nullptr;     /* [0] ray [1] hit [2] shadow [3] cont [4..5] scratch [7] bounce */
    uint32_t* history = nullptr;   /* [bounce][ray, hit, shadow, cont] */
    uint32_t* failure = nullptr;   /* [0] count, [2..3] lowest (job << 48 | path << 8 | failure code) */
    uint32_t* sample_counts = nullptr;
    unsigned long long* traversal = nullptr; /* 6 counters */
    uint32_t* fetch = nullptr;               /* traversal work counters (2 words after the counters) */
    float* replicas = nullptr;               /* BF_GRADIENT_REPLICAS x BF_GRADIENT_ROW adjoint scratch */
    uint32_t* bucket_counts = nullptr;       /* shade-order grouping: per-tile bucket counts / offsets */
    uint32_t* bucket_offsets = nullptr;
    unsigned long long* hashes = nullptr;    /* [bounce][ray, shadow] FNV-1a-64 of the queues */
    uint32_t* host_words = nullptr;          /* pinned readback */
    uint32_t* host_history = nullptr;        /* pinned readback of the history */
    unsigned long long* host_hashes = nullptr;
    bf_dev_job* jobs = nullptr;              /* device job table */
    bf_dev_job* host_jobs = nullptr;         /* pinned staging of the job table */
    uint2* spill = nullptr;                  /* traversal stack spill, BF_STACK_DEPTH x spill_lanes */
    uint32_t spill_lanes = 0;
    uint32_t shade_blocks = 0, adjoint_shade_blocks = 0;
    int device = 0;
    /* a captured group of bounces, replayed while paths remain (keyed by what its arguments hold) */
    hipGraphExec_t graph = nullptr;
    bf_transport_params graph_params{};
    uint32_t graph_n = 0, graph_group = 0;
    const uint32_t* graph_ray_q = nullptr;
    /* bounce groups are captured on this stream (the caller's may be the legacy null stream,
     * which cannot capture) and replayed on the caller's */
    hipStream_t capture_stream = nullptr;
    /* stage instrumentation: a reusable pool of hipEvent pairs, drained at every host sync */
    std::vector<hipEvent_t> events;
    std::vector<std::pair<uint32_t, uint32_t>> pending; /* (stage, first event of the pair) */
    uint32_t events_used = 0;
};

namespace {

/* Records the launch(es) of one stage, bracketed by an event pair when instrumentation is on. */
template <typename Launch>
hipError_t staged(bf_hip_workspace* w, bf_transport_report& rep, bool instrument, uint32_t stage, uint64_t dispatches,
                  hipStream_t stream, Launch&& launch) {
    rep.stage_metrics[stage].kernel_dispatches += dispatches;
    if (!instrument) return launch();
    while (w->events.size() < (size_t)w->events_used + 2) {
        hipEvent_t e = nullptr;
        const hipError_t st = hipEventCreate(&e);
        if (st != hipSuccess) return st;
        w->events.push_back(e);
    }
    const uint32_t first = w->events_used;
    w->events_used += 2;
    hipError_t st = hipEventRecord(w->events[first], stream);
    if (st != hipSuccess) return st;
    st = launch();
    if (st != hipSuccess) return st;
    st = hipEventRecord(w->events[first + 1], stream);
    if (st != hipSuccess) return st;
    w->pending.emplace_back(stage, first);
    return hipSuccess;
}

/* Folds the elapsed times of every recorded pair into the report; the stream must be idle. */
hipError_t drain_stage_events(bf_hip_workspace* w, bf_transport_report& rep) {
    for (const auto& p : w->pending) {
        float ms = 0.0f;
        const hipError_t st = hipEventElapsedTime(&ms, w->events[p.second], w->events[p.second + 1]);
        if (st != hipSuccess) return st;
        rep.stage_metrics[p.first].gpu_elapsed_nanoseconds += (uint64_t)std::llround((double)ms * 1.0e6);
    }
    w->pending.clear();
    w->events_used = 0;
    return hipSuccess;
}

} // namespace

extern "C" {

const char* blackframe_hip_last_error(void) { return g_last_error.c_str(); }

int blackframe_hip_device_count(uint32_t* count) {
    if (!count) return set_error(BF_INVALID_ARGUMENT, "count is NULL");
    int n = 0;
    const hipError_t e = hipGetDeviceCount(&n);
    *count = e == hipSuccess ? (uint32_t)n : 0u;
    return BF_OK;
}

static int create_scene(const bf_scene_desc* desc, bf_hip_scene** out, bool upload) {
    if (!desc || !out) return set_error(BF_INVALID_ARGUMENT, "scene description or output is NULL");
    *out = nullptr;
    bf_hip_scene* s = new (std::nothrow) bf_hip_scene();
    if (!s) return set_error(BF_RESOURCE_EXHAUSTED, "host allocation failed");
    int status;
    try {
        status = build_scene(desc, s, upload);
    } catch (const std::bad_alloc&) {
        status = set_error(BF_RESOURCE_EXHAUSTED, "host allocation failed during scene build");
    }
    if (status != BF_OK) {
        blackframe_hip_scene_destroy(s);
        return status;
    }
    *out = s;
    return BF_OK;
}

int blackframe_hip_scene_create(const bf_scene_desc* desc, bf_hip_scene** out) { return create_scene(desc, out, true); }

int blackframe_hip_scene_prepare(const bf_scene_desc* desc, bf_hip_scene** out) { return create_scene(desc, out, false); }

/* ---- power-weighted light selection (Renderer/Source/LightSampler.cpp) ------------------------ */
namespace {

/* add_compensated (:85-102): a running binary32 sum with its Neumaier correction; false when a
 * term is negative or not finite, or the sum overflows */
bool add_compensated(float value, float& sum, float& correction) {
    if (!std::isfinite(value) || value < 0.0f) return false;
    const float next = sum + value;
    if (!std::isfinite(next)) return false;
    correction += sum >= value ? (sum - next) + value : (value - next) + sum;
    if (!std::isfinite(correction)) return false;
    sum = next;
    return true;
}

} // namespace

extern "C" int blackframe_hip_light_power_cdf(const float* powers, uint32_t count, float* boundaries) {
    if (!powers || !boundaries) return set_error(BF_INVALID_ARGUMENT, "NULL powers or boundaries");
    if (count == 0) return set_error(BF_INVALID_ARGUMENT, "power-weighted light sampling needs at least one light");
    /* power_weights (:201-249): lanes scaled by the global peak, each light's four lanes summed */
    float peak = 0.0f;
    for (uint64_t k = 0; k < 4ull * count; ++k) {
        if (!std::isfinite(powers[k]) || powers[k] < 0.0f)
            return set_error(BF_INVALID_ARGUMENT, "light %llu: spectral power lanes must be finite and non-negative",
                             (unsigned long long)(k / 4));
        peak = std::max(peak, powers[k]);
    }
    if (peak == 0.0f) return set_error(BF_INVALID_ARGUMENT, "power-weighted light sampling needs positive power");
    std::vector<float> weights(count);
    for (uint32_t i = 0; i < count; ++i) {
        float sum = 0.0f, correction = 0.0f;
        bool positive = false;
        for (int lane = 0; lane < 4; ++lane) {
            const float v = powers[4ull * i + lane];
            const float ratio = v / peak;
            if (v > 0.0f) {
                positive = true;
                if (!(ratio > 0.0f))
                    return set_error(BF_RESOURCE_EXHAUSTED, "light %u: a positive lane underflows after scaling", i);
            }
            if (!add_compensated(ratio, sum, correction))
                return set_error(BF_RESOURCE_EXHAUSTED, "light %u: the selection weight is not representable", i);
        }
        const float w = sum + correction;
        if (!std::isfinite(w) || w < 0.0f || (positive && !(w > 0.0f)))
            return set_error(BF_RESOURCE_EXHAUSTED, "light %u: the selection weight is not representable", i);
        weights[i] = w == 0.0f ? 0.0f : w;
    }
    /* build_power_cdf (:116-199): compensated prefix sums over the compensated total; the last
     * light with power closes at exactly 1, lights without power repeat the previous boundary */
    float total_sum = 0.0f, total_correction = 0.0f;
    for (float w : weights)
        if (!add_compensated(w, total_sum, total_correction))
            return set_error(BF_RESOURCE_EXHAUSTED, "the total light power is not representable");
    float total = total_sum + total_correction;
    if (!std::isfinite(total) || total < 0.0f) return set_error(BF_RESOURCE_EXHAUSTED, "the total light power is not representable");
    if (total == 0.0f) return set_error(BF_INVALID_ARGUMENT, "power-weighted light sampling needs positive power");
    uint32_t last_positive = count;
    for (uint32_t i = count; i > 0; --i)
        if (weights[i - 1] > 0.0f) {
            last_positive = i - 1;
            break;
        }
    if (last_positive == count) return set_error(BF_INTERNAL_ERROR, "positive power has no light");
    boundaries[0] = 0.0f;
    float prefix_sum = 0.0f, prefix_correction = 0.0f, previous = 0.0f;
    for (uint32_t i = 0; i < count; ++i) {
        const float w = weights[i];
        if (w == 0.0f) {
            boundaries[i + 1] = previous;
            continue;
        }
        if (!add_compensated(w, prefix_sum, prefix_correction))
            return set_error(BF_RESOURCE_EXHAUSTED, "a light power prefix is not representable");
        const float prefix = prefix_sum + prefix_correction;
        if (!std::isfinite(prefix) || !(prefix > 0.0f))
            return set_error(BF_RESOURCE_EXHAUSTED, "a light power prefix is not representable");
        float boundary = 1.0f;
        if (i != last_positive) {
            boundary = prefix / total;
            if (!std::isfinite(boundary) || !(boundary > previous) || !(boundary < 1.0f))
                return set_error(BF_RESOURCE_EXHAUSTED, "light %u: its selection probability is not representable", i);
        } else {
            const float nominal = w / total;
            const float candidate = previous + nominal;
            if (!std::isfinite(nominal) || !(nominal > 0.0f) || !std::isfinite(candidate) || !(candidate > previous))
                return set_error(BF_RESOURCE_EXHAUSTED, "light %u: its selection probability is not representable", i);
        }
        boundaries[i + 1] = boundary;
        previous = boundary;
    }
    if (boundaries[count] != 1.0f) return set_error(BF_INTERNAL_ERROR, "the light selection CDF does not end at 1");
    for (uint32_t i = 0; i < count; ++i) {
        const float p = boundaries[i + 1] - boundaries[i];
        if (!std::isfinite(p) || p < 0.0f || ((weights[i] > 0.0f) != (p > 0.0f)))
            return set_error(BF_RESOURCE_EXHAUSTED, "light %u: the selection CDF lost or created support", i);
    }
    return BF_OK;
}

extern "C" int blackframe_hip_scene_set_light_powers(bf_hip_scene* s, const float* powers, uint32_t count) {
    if (!s) return set_error(BF_INVALID_ARGUMENT, "NULL scene");
    if (!s->device_block) return set_error(BF_INVALID_ARGUMENT, "scene was prepared without a device upload");
    if (powers && count != s->light_count)
        return set_error(BF_INVALID_ARGUMENT, "%u light powers for a scene with %u registered lights", count, s->light_count);
    float* cdf = nullptr;
    if (powers) {
        std::vector<float> host(count + 1);
        const int st = blackframe_hip_light_power_cdf(powers, count, host.data());
        if (st != BF_OK) return st;
        int device = 0;
        HIP_TRY(hipGetDevice(&device));
        if (device != s->device)
            return set_error(BF_INCOMPATIBLE, "scene lives on device %d, current device is %d", s->device, device);
        HIP_TRY(hipMalloc(&cdf, host.size() * sizeof(float)));
        const hipError_t e = hipMemcpy(cdf, host.data(), host.size() * sizeof(float), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            (void)hipFree(cdf);
            return hip_error(e, "hipMemcpy(light cdf)");
        }
    }
    if (s->light_cdf) {
        /* a launch still reading the old table must finish before it goes */
        HIP_TRY(hipDeviceSynchronize());
        HIP_TRY(hipFree(s->light_cdf));
    }
    s->light_cdf = cdf;
    s->dev.light_cdf = cdf;
    return BF_OK;
}

int blackframe_hip_scene_destroy(bf_hip_scene* s) {
    if (!s) return BF_OK;
    hipError_t e = hipSuccess;
    if (s->light_cdf) e = hipFree(s->light_cdf);
    if (s->device_block) {
        const hipError_t f = hipFree(s->device_block);
        if (e == hipSuccess) e = f;
    }
    delete s;
    return e == hipSuccess ? BF_OK : hip_error(e, "hipFree(scene)");
}

int blackframe_hip_scene_bvh_export(const bf_hip_scene* s, uint64_t counts[5], bf_bvh_blas* blases, bf_bvh_node* blas_nodes,
                                    uint32_t* primitive_refs, bf_bvh_node* tlas_nodes, bf_bvh_instance_ref* instance_refs,
                                    uint32_t* tlas_root) {
    if (!s || !counts) return set_error(BF_INVALID_ARGUMENT, "scene or counts is NULL");
    counts[0] = s->blases.size();
    counts[1] = s->blas_nodes.size();
    counts[2] = s->prim_refs.size();
    counts[3] = s->tlas_nodes.size();
    counts[4] = s->inst_refs.size();
    if (blases) std::memcpy(blases, s->blases.data(), counts[0] * sizeof(bf_bvh_blas));
    if (blas_nodes) std::memcpy(blas_nodes, s->blas_nodes.data(), counts[1] * sizeof(bf_bvh_node));
    if (primitive_refs) std::memcpy(primitive_refs, s->prim_refs.data(), counts[2] * sizeof(uint32_t));
    if (tlas_nodes) std::memcpy(tlas_nodes, s->tlas_nodes.data(), counts[3] * sizeof(bf_bvh_node));
    if (instance_refs) std::memcpy(instance_refs, s->inst_refs.data(), counts[4] * sizeof(bf_bvh_instance_ref));
    if (tlas_root) *tlas_root = s->tlas_root;
    return BF_OK;
}

} /* extern "C" */

namespace {

uint64_t bvh_array_count(const bf_bvh_header& h, uint32_t a) {
    const uint64_t counts[BF_BVH_ARRAY_COUNT] = {h.blas_count, h.blas_node_count, h.primitive_reference_count,
                                                 h.tlas_node_count, h.instance_reference_count};
    return counts[a];
}

uint32_t bvh_array_element_size(uint32_t a) {
    const uint32_t sizes[BF_BVH_ARRAY_COUNT] = {sizeof(bf_bvh_blas), sizeof(bf_bvh_node), sizeof(uint32_t),
                                                sizeof(bf_bvh_node), sizeof(bf_bvh_instance_ref)};
    return sizes[a];
}

uint64_t bvh_align(uint64_t v) { return (v + BF_BVH_ARRAY_ALIGNMENT - 1u) / BF_BVH_ARRAY_ALIGNMENT * BF_BVH_ARRAY_ALIGNMENT; }

/* validate_scene_bvh_header (SceneBvhAbi.hpp:146-229), in its check order */
uint32_t validate_bvh_header(const bf_bvh_header& h) {
    if (h.magic != BF_BVH_MAGIC) return BF_BVH_BAD_MAGIC;
    if (h.abi_major != BF_BVH_ABI_MAJOR || h.abi_minor != BF_BVH_ABI_MINOR) return BF_BVH_INCOMPATIBLE_VERSION;
    if (h.header_size != BF_BVH_HEADER_SIZE) return BF_BVH_BAD_HEADER_SIZE;
    if (h.array_count != BF_BVH_ARRAY_COUNT) return BF_BVH_BAD_ARRAY_COUNT;
    if (h.hash_algorithm != BF_BVH_HASH_FNV1A64) return BF_BVH_UNSUPPORTED_HASH_ALGORITHM;
    for (uint64_t r : h.reserved)
        if (r != 0u) return BF_BVH_NONZERO_RESERVED;
    if (h.leaf_capacity != BF_BVH_LEAF_CAPACITY) return BF_BVH_BAD_LEAF_CAPACITY;
    const uint64_t u32max = 0xFFFFFFFFull;
    if (h.blas_count > u32max || h.blas_node_count > u32max || h.primitive_reference_count > u32max ||
        h.tlas_node_count > u32max || h.instance_reference_count > u32max ||
        (h.blas_count == 0u && (h.blas_node_count != 0u || h.primitive_reference_count != 0u)) ||
        (h.blas_count != 0u && (h.primitive_reference_count < h.blas_count ||
                                h.blas_node_count != h.primitive_reference_count * 2u - h.blas_count)) ||
        (h.instance_reference_count == 0u && h.tlas_node_count != 0u) ||
        (h.instance_reference_count != 0u && h.tlas_node_count != h.instance_reference_count * 2u - 1u))
        return BF_BVH_INVALID_TOPOLOGY_COUNTS;
    if ((h.tlas_node_count == 0u && h.tlas_root_node != BF_INVALID_INDEX) || (h.tlas_node_count != 0u && h.tlas_root_node != 0u))
        return BF_BVH_INVALID_TLAS_ROOT;
    if (h.total_size_bytes < BF_BVH_HEADER_SIZE) return BF_BVH_BAD_TOTAL_SIZE;
    const uint64_t top = ~0ull;
    uint64_t cursor = bvh_align(BF_BVH_HEADER_SIZE);
    for (uint32_t a = 0; a < BF_BVH_ARRAY_COUNT; ++a) {
        const bf_bvh_array_descriptor& d = h.arrays[a];
        const uint64_t count = bvh_array_count(h, a);
        const uint32_t size = bvh_array_element_size(a);
        if (d.element_count != count || d.element_size != size || d.reserved != 0u) return BF_BVH_INVALID_ARRAY_DESCRIPTOR;
        if (count == 0u) {
            if (d.offset_bytes != 0u) return BF_BVH_INVALID_ARRAY_DESCRIPTOR;
            continue;
        }
        if (cursor > top - (BF_BVH_ARRAY_ALIGNMENT - 1u)) return BF_BVH_SIZE_OVERFLOW;
        cursor = bvh_align(cursor);
        if (d.offset_bytes != cursor) return BF_BVH_INVALID_ARRAY_DESCRIPTOR;
        if (count > top / size) return BF_BVH_SIZE_OVERFLOW;
        const uint64_t bytes = count * size;
        if (cursor > top - bytes) return BF_BVH_SIZE_OVERFLOW;
        cursor += bytes;
    }
    return h.total_size_bytes == cursor ? (uint32_t)BF_BVH_VALID : (uint32_t)BF_BVH_BAD_TOTAL_SIZE;
}

/* FNV-1a-64 with the content-hash field read as zero (SceneBvh.cpp:153-163) */
uint64_t bvh_blob_hash(const uint8_t* bytes, size_t n) {
    uint64_t h = BF_FNV1A64_OFFSET_BASIS;
    for (size_t i = 0; i < n; ++i) {
        h ^= (i >= BF_BVH_HASH_OFFSET && i < BF_BVH_HASH_OFFSET + 8u) ? 0u : bytes[i];
        h *= BF_FNV1A64_PRIME;
    }
    return h;
}

} // namespace

extern "C" {

int blackframe_hip_scene_bvh_serialize(const bf_hip_scene* s, void* out, size_t capacity, size_t* size) {
    if (!s || !size) return set_error(BF_INVALID_ARGUMENT, "scene or size is NULL");
    /* SceneBvh.cpp:882-1071: header, then the arrays at aligned offsets in their fixed order */
    bf_bvh_header h{};
    h.magic = BF_BVH_MAGIC;
    h.abi_major = BF_BVH_ABI_MAJOR;
    h.abi_minor = BF_BVH_ABI_MINOR;
    h.header_size = BF_BVH_HEADER_SIZE;
    h.array_count = BF_BVH_ARRAY_COUNT;
    h.hash_algorithm = BF_BVH_HASH_FNV1A64;
    h.source_scene_hash = s->source_hash;
    h.blas_count = s->blases.size();
    h.blas_node_count = s->blas_nodes.size();
    h.primitive_reference_count = s->prim_refs.size();
    h.tlas_node_count = s->tlas_nodes.size();
    h.instance_reference_count = s->inst_refs.size();
    h.tlas_root_node = s->tlas_nodes.empty() ? BF_INVALID_INDEX : s->tlas_root;
    h.leaf_capacity = BF_BVH_LEAF_CAPACITY;
    const void* data[BF_BVH_ARRAY_COUNT] = {s->blases.data(), s->blas_nodes.data(), s->prim_refs.data(),
                                            s->tlas_nodes.data(), s->inst_refs.data()};
    uint64_t cursor = bvh_align(BF_BVH_HEADER_SIZE);
    for (uint32_t a = 0; a < BF_BVH_ARRAY_COUNT; ++a) {
        bf_bvh_array_descriptor& d = h.arrays[a];
        d.element_count = bvh_array_count(h, a);
        d.element_size = bvh_array_element_size(a);
        if (d.element_count == 0u) continue;
        cursor = bvh_align(cursor);
        d.offset_bytes = cursor;
        cursor += d.element_count * d.element_size;
    }
    h.total_size_bytes = cursor;
    *size = (size_t)cursor;
    if (!out) return BF_OK;
    if (capacity < cursor) return set_error(BF_RESOURCE_EXHAUSTED, "BVH blob needs %zu bytes, %zu given", (size_t)cursor, capacity);
    uint8_t* bytes = static_cast<uint8_t*>(out);
    std::memset(bytes, 0, (size_t)cursor);
    for (uint32_t a = 0; a < BF_BVH_ARRAY_COUNT; ++a)
        if (h.arrays[a].element_count)
            std::memcpy(bytes + h.arrays[a].offset_bytes, data[a], (size_t)(h.arrays[a].element_count * h.arrays[a].element_size));
    std::memcpy(bytes, &h, sizeof h);
    h.content_hash = bvh_blob_hash(bytes, (size_t)cursor);
    std::memcpy(bytes, &h, sizeof h);
    if (validate_bvh_header(h) != BF_BVH_VALID) return set_error(BF_INTERNAL_ERROR, "serialized BVH header does not validate");
    return BF_OK;
}

int blackframe_hip_scene_bvh_validate(const void* blob, size_t byte_count, uint32_t flags, uint32_t* status) {
    if (!blob || !status) return set_error(BF_INVALID_ARGUMENT, "BVH blob or status is NULL");
    if (flags & ~BF_BVH_VERIFY_HASH) return set_error(BF_INVALID_ARGUMENT, "unknown BVH blob flags 0x%x", flags);
    if (byte_count < BF_BVH_HEADER_SIZE) {
        *status = BF_BVH_TRUNCATED;
        return BF_OK;
    }
    bf_bvh_header h;
    std::memcpy(&h, blob, sizeof h);
    uint32_t st = validate_bvh_header(h);
    if (st == BF_BVH_VALID && h.total_size_bytes > byte_count) st = BF_BVH_TRUNCATED;
    if (st == BF_BVH_VALID && (flags & BF_BVH_VERIFY_HASH) &&
        bvh_blob_hash(static_cast<const uint8_t*>(blob), (size_t)h.total_size_bytes) != h.content_hash)
        st = BF_BVH_HASH_MISMATCH;
    *status = st;
    return BF_OK;
}

int blackframe_hip_scene_wide_export(const bf_hip_scene* s, uint64_t counts[2], bf_wide_node* blas_wide,
                                     bf_wide_node* tlas_wide) {
    if (!s || !counts) return set_error(BF_INVALID_ARGUMENT, "scene or counts is NULL");
    counts[0] = s->blas_wide.size();
    counts[1] = s->tlas_wide.size();
    if (blas_wide) std::memcpy(blas_wide, s->blas_wide.data(), counts[0] * sizeof(bf_wide_node));
    if (tlas_wide) std::memcpy(tlas_wide, s->tlas_wide.data(), counts[1] * sizeof(bf_wide_node));
    return BF_OK;
}

int blackframe_hip_launch_scene_closest_hit(const bf_hip_scene* s, const bf_transport_ray* rays, uint32_t count,
                                            bf_closest_hit_result* results, uint32_t* counters, void* stream) {
    if (!s || (count != 0 && (!rays || !results))) return set_error(BF_INVALID_ARGUMENT, "NULL scene, rays or results");
    if (!s->device_block) return set_error(BF_INVALID_ARGUMENT, "scene was prepared without a device upload");
    HIP_TRY(bfk_scene_closest_hit(&s->dev, rays, count, results, counters, (hipStream_t)stream));
    return BF_OK;
}

int blackframe_hip_launch_scene_occlusion(const bf_hip_scene* s, const bf_transport_ray* rays, uint32_t count,
                                          bf_occlusion_result* results, void* stream) {
    if (!s || (count != 0 && (!rays || !results))) return set_error(BF_INVALID_ARGUMENT, "NULL scene, rays or results");
    if (!s->device_block) return set_error(BF_INVALID_ARGUMENT, "scene was prepared without a device upload");
    HIP_TRY(bfk_scene_occlusion(&s->dev, rays, count, results, (hipStream_t)stream));
    return BF_OK;
}

int blackframe_hip_launch_sample_stream_dump(uint32_t requested_schema_version, const bf_sample_dump_request* requests,
                                             uint32_t request_count, bf_sample_dimension_manifest* manifest,
                                             bf_sample_dump_result* results) {
    /* CudaSampleStream.cu:52-56: every rejection before any launch */
    if (requested_schema_version != BF_SAMPLE_DIMENSION_SCHEMA || !requests || request_count == 0u || !manifest || !results)
        return set_error(BF_INVALID_ARGUMENT, "sample-stream dump: schema %u, %u requests, or a null pointer",
                         requested_schema_version, request_count);
    HIP_TRY(bfk_sample_stream_dump(requests, request_count, manifest, results));
    return BF_OK;
}

int blackframe_hip_launch_sample_dump(uint32_t sampler, const bf_sample_stream_index* indices, const uint64_t* dimensions,
                                      uint32_t count, float* samples, void* stream) {
    if (sampler > BF_SAMPLER_SOBOL_OWEN) return set_error(BF_INVALID_ARGUMENT, "unknown sampler %u", sampler);
    if (count != 0 && (!indices || !dimensions || !samples)) return set_error(BF_INVALID_ARGUMENT, "NULL argument");
    const uint64_t* directions = nullptr;
    const int st = sobol_table_for_current_device(&directions);
    if (st != BF_OK) return st;
    HIP_TRY(bfk_sample_dump(sampler, directions, kSobolDimensions, indices, dimensions, count, samples, (hipStream_t)stream));
    return BF_OK;
}

int blackframe_hip_workspace_create(uint64_t capacity, bf_hip_workspace** out) {
    if (!out || capacity == 0 || capacity > 0x7FFFFFFFull)
        return set_error(BF_INVALID_ARGUMENT, "workspace capacity must be in [1, 2^31)");
    *out = nullptr;
    bf_hip_workspace* w = new (std::nothrow) bf_hip_workspace();
    if (!w) return set_error(BF_RESOURCE_EXHAUSTED, "host allocation failed");
    w->capacity = capacity;
    const uint64_t n = capacity;
    const uint32_t tiles = bfk_compact_tiles((uint32_t)n);
    hipError_t de = hipGetDevice(&w->device);
    const uint32_t resident = de == hipSuccess ? bfk_trace_resident_lanes(w->device) : 0u;
    w->shade_blocks = de == hipSuccess ? bfk_shade_resident_blocks(w->device, false) : 0u;
    w->adjoint_shade_blocks = de == hipSuccess ? bfk_shade_resident_blocks(w->device, true) : 0u;
    if (resident == 0 || w->shade_blocks == 0 || w->adjoint_shade_blocks == 0) {
        delete w;
        return set_error(BF_DEVICE_ERROR, "cannot query the kernels' occupancy on the current device");
    }
    w->spill_lanes = resident;
    Arena A;
    const size_t o_ro = A.add(nullptr, n * 16), o_rd = A.add(nullptr, n * 16), o_beta = A.add(nullptr, n * 16),
                 o_rad = A.add(nullptr, n * 16), o_wl = A.add(nullptr, n * 16), o_dep = A.add(nullptr, n * 16),
                 o_prev = A.add(nullptr, n * 16), o_hit = A.add(nullptr, n * 16), o_ids = A.add(nullptr, n * 8),
                 o_so = A.add(nullptr, n * 16), o_sd = A.add(nullptr, n * 16), o_sc = A.add(nullptr, n * 16),
                 o_st = A.add(nullptr, n * 4), o_vis = A.add(nullptr, n * 4), o_job = A.add(nullptr, n * 2),
                 o_cone = A.add(nullptr, n * 8);
    size_t o_q[4];
    for (int k = 0; k < 4; ++k) o_q[k] = A.add(nullptr, n * 4);
    const size_t o_bc = A.add(nullptr, (size_t)tiles * 32), o_bo = A.add(nullptr, (size_t)tiles * 32);
    const size_t o_route = A.add(nullptr, n), o_tc = A.add(nullptr, (size_t)tiles * 8), o_to = A.add(nullptr, (size_t)tiles * 8),
                 o_sizes = A.add(nullptr, 64), o_hist = A.add(nullptr, BF_HISTORY_BOUNCES * 16), o_fail = A.add(nullptr, 16),
                 o_samp = A.add(nullptr, 16), o_trav = A.add(nullptr, 64), o_hash = A.add(nullptr, BF_REPORT_MAX_BOUNCES * 16),
                 o_jobs = A.add(nullptr, BF_MAX_BATCH_JOBS * sizeof(bf_dev_job)),
                 o_replicas = A.add(nullptr, (size_t)BF_GRADIENT_REPLICAS * BF_GRADIENT_ROW * sizeof(float)),
                 o_spill = A.add(nullptr, (size_t)w->spill_lanes * BF_STACK_DEPTH * sizeof(uint2));
    hipError_t e = hipMalloc(&w->block, A.total);
    if (e != hipSuccess) {
        delete w;
        return set_error(BF_RESOURCE_EXHAUSTED, "workspace allocation of %zu bytes failed: %s", A.total, hipGetErrorString(e));
    }
    const size_t host_bytes = 64 + BF_HISTORY_BOUNCES * 16 + BF_REPORT_MAX_BOUNCES * 16 + BF_MAX_BATCH_JOBS * sizeof(bf_dev_job);
    void* host = nullptr;
    e = hipHostMalloc(&host, host_bytes, hipHostMallocDefault);
    if (e != hipSuccess) {
        (void)hipFree(w->block);
        delete w;
        return hip_error(e, "hipHostMalloc");
    }
    w->host_words = (uint32_t*)host;
    w->host_history = (uint32_t*)((uint8_t*)host + 64);
    w->host_hashes = (unsigned long long*)((uint8_t*)host + 64 + BF_HISTORY_BOUNCES * 16);
    w->host_jobs = (bf_dev_job*)((uint8_t*)host + 64 + BF_HISTORY_BOUNCES * 16 + BF_REPORT_MAX_BOUNCES * 16);
    uint8_t* b = (uint8_t*)w->block;
    w->paths.ray_origin = (float4*)(b + o_ro);
    w->paths.ray_direction = (float4*)(b + o_rd);
    w->paths.beta = (float4*)(b + o_beta);
    w->paths.radiance = (float4*)(b + o_rad);
    w->paths.wavelengths = (float4*)(b + o_wl);
    w->paths.depth = (uint4*)(b + o_dep);
    w->paths.previous = (float4*)(b + o_prev);
    w->paths.hit = (float4*)(b + o_hit);
    w->paths.hit_ids = (uint2*)(b + o_ids);
    w->paths.shadow_origin = (float4*)(b + o_so);
    w->paths.shadow_direction = (float4*)(b + o_sd);
    w->paths.shadow_contribution = (float4*)(b + o_sc);
    w->paths.status = (uint32_t*)(b + o_st);
    w->paths.visibility = (uint32_t*)(b + o_vis);
    w->paths.job = (uint16_t*)(b + o_job);
    w->paths.cone = (float2*)(b + o_cone);
    for (int k = 0; k < 4; ++k) w->queues[k] = (uint32_t*)(b + o_q[k]);
    w->route = b + o_route;
    w->tile_counts = (uint32_t*)(b + o_tc);
    w->bucket_counts = (uint32_t*)(b + o_bc);
    w->bucket_offsets = (uint32_t*)(b + o_bo);
    w->tile_offsets = (uint32_t*)(b + o_to);
    w->sizes = (uint32_t*)(b + o_sizes);
    w->history = (uint32_t*)(b + o_hist);
    w->failure = (uint32_t*)(b + o_fail);
    w->sample_counts = (uint32_t*)(b + o_samp);
    w->traversal = (unsigned long long*)(b + o_trav);
    w->fetch = (uint32_t*)(b + o_trav + 48);
    w->hashes = (unsigned long long*)(b + o_hash);
    w->jobs = (bf_dev_job*)(b + o_jobs);
    w->replicas = (float*)(b + o_replicas);
    w->spill = (uint2*)(b + o_spill);
    *out = w;
    return BF_OK;
}

int blackframe_hip_workspace_destroy(bf_hip_workspace* w) {
    if (!w) return BF_OK;
    for (hipEvent_t e : w->events) (void)hipEventDestroy(e);
    if (w->graph) (void)hipGraphExecDestroy(w->graph);
    if (w->capture_stream) (void)hipStreamDestroy(w->capture_stream);
    hipError_t e1 = w->block ? hipFree(w->block) : hipSuccess;
    hipError_t e2 = w->host_words ? hipHostFree(w->host_words) : hipSuccess;
    delete w;
    if (e1 != hipSuccess) return hip_error(e1, "hipFree(workspace)");
    if (e2 != hipSuccess) return hip_error(e2, "hipHostFree(workspace)");
    return BF_OK;
}

} /* extern "C" */

namespace {

struct chunk_job {
    uint32_t job;        /* index into the caller's jobs */
    uint32_t first_pass; /* first sample pass of the frame in this launch */
    uint32_t passes;
};

int validate_options(const bf_transport_options* o) {
    if (o->sampler > BF_SAMPLER_SOBOL_OWEN || o->wavelength_mode > BF_WAVELENGTHS_PER_PATH || o->mis_heuristic > BF_MIS_POWER)
        return set_error(BF_INVALID_ARGUMENT, "unknown sampler, wavelength mode or MIS heuristic");
    if (o->volume_depth_limit != 0) return set_error(BF_UNAVAILABLE, "participating media are not supported");
    if (o->russian_roulette_enabled &&
        (o->russian_roulette_first_depth == 0 || !(o->russian_roulette_minimum > 0.0f) || !(o->russian_roulette_minimum < 1.0f) ||
         o->russian_roulette_maximum < o->russian_roulette_minimum || o->russian_roulette_maximum > 1.0f))
        return set_error(BF_INVALID_ARGUMENT, "invalid Russian-roulette policy");
    if (o->diffuse_depth_limit > 255 || o->glossy_depth_limit > 255 || o->specular_depth_limit > 255 ||
        o->transmission_depth_limit > 255)
        return set_error(BF_INVALID_ARGUMENT, "depth limits above 255 are not supported");
    if (o->instrumentation > BF_INSTRUMENT_STAGE_EVENTS) return set_error(BF_INVALID_ARGUMENT, "unknown instrumentation mode");
    if (o->failure_policy > BF_FAILURES_COUNT) return set_error(BF_INVALID_ARGUMENT, "unknown failure policy");
    const uint32_t block_log2 = (o->launch & BF_LAUNCH_TRACE_BLOCK_MASK) >> 16;
    if (block_log2 != 0 && (block_log2 < 6 || block_log2 > 8))
        return set_error(BF_INVALID_ARGUMENT, "traversal workgroups of 2^%u lanes are not available (64, 128 or 256)", block_log2);
    if ((o->launch & ~(BF_LAUNCH_GRAPH | BF_LAUNCH_TRACE_BLOCK_MASK | BF_LAUNCH_GROUP_MASK | BF_LAUNCH_SHADE_ORDER |
                       BF_LAUNCH_TRACE_ORDER | BF_LAUNCH_INFLIGHT_MASK)) != 0)
        return set_error(BF_INVALID_ARGUMENT, "unknown launch bits 0x%08x", o->launch);
    return BF_OK;
}

/* Issues one bounce of the loop on `stream` (no host synchronisation). The adjoint replay stops
 * after the shade stage: its light samples are resolved from the tape, so no shadow ray exists. */
int issue_bounce(bf_hip_workspace* w, bf_transport_report& rep, const bf_transport_params& P, bool batch, bool instrument,
                 uint32_t n_max, uint32_t* ray_q, uint32_t* cont_q, unsigned long long* counters,
                 unsigned long long* hash_slot, hipStream_t stream) {
    const bool adjoint = P.shade_mode == BF_SHADE_ADJOINT;
#define STAGE(stage, dispatches, call) HIP_TRY(staged(w, rep, instrument, (stage), (dispatches), stream, [&] { return (call); }))
    if (hash_slot) STAGE(BF_STAGE_AUXILIARY, 1, bfk_queue_hash(ray_q, &w->sizes[0], hash_slot, stream));
    STAGE(BF_STAGE_CLOSEST_HIT, 1, bfk_trace_closest(&P, ray_q, &w->sizes[0], n_max, w->route, counters, batch, stream));
    STAGE(BF_STAGE_COMPACTION, 3,
          bfk_compact(w->route, &w->sizes[0], n_max, ray_q, w->tile_counts, w->tile_offsets, w->queues[1], &w->sizes[1], 0u,
                      w->queues[2], &w->sizes[4], 0u, stream));
    STAGE(BF_STAGE_SHADE, P.shade_order ? 4 : 1,
          bfk_shade(&P, w->queues[1], &w->sizes[1], n_max, w->route, w->sample_counts, batch, P.shade_mode, stream));
    STAGE(BF_STAGE_COMPACTION, 3,
          bfk_compact(w->route, &w->sizes[1], n_max, w->queues[1], w->tile_counts, w->tile_offsets, w->queues[2],
                      &w->sizes[2], 0u, cont_q, &w->sizes[3], 0u, stream));
    if (adjoint) {
        STAGE(BF_STAGE_AUXILIARY, 1, bfk_bounce_end(w->sizes, w->history, BF_HISTORY_BOUNCES, stream));
        return BF_OK;
    }
    if (hash_slot) STAGE(BF_STAGE_AUXILIARY, 1, bfk_queue_hash(w->queues[2], &w->sizes[2], hash_slot + 1, stream));
    STAGE(BF_STAGE_SHADOW, 1, bfk_trace_shadow(&P, w->queues[2], &w->sizes[2], n_max, w->route, counters, batch, stream));
    STAGE(BF_STAGE_COMPACTION, 3,
          bfk_compact(w->route, &w->sizes[2], n_max, w->queues[2], w->tile_counts, w->tile_offsets, w->queues[1],
                      &w->sizes[5], 0u, cont_q, &w->sizes[3], 1u, stream));
    STAGE(BF_STAGE_AUXILIARY, 1, bfk_bounce_end(w->sizes, w->history, BF_HISTORY_BOUNCES, stream));
#undef STAGE
    return BF_OK;
}

/* Whole sample passes of the frames packed into launches of at most `in_flight` paths (and
 * BF_MAX_BATCH_JOBS frames), in job order: one contiguous slot range per frame and launch. */
std::vector<std::vector<chunk_job>> pack_launches(const bf_frame_job* jobs, uint32_t job_count, uint64_t in_flight) {
    std::vector<std::vector<chunk_job>> launches;
    std::vector<uint32_t> next_pass(job_count, 0u);
    uint32_t cursor = 0;
    for (;;) {
        std::vector<chunk_job> chunk;
        uint64_t used = 0;
        for (uint32_t j = cursor; j < job_count && chunk.size() < BF_MAX_BATCH_JOBS; ++j) {
            const bf_frame_job& J = jobs[j];
            const uint64_t pixels = (uint64_t)J.frame.camera.width * J.frame.camera.height;
            const uint32_t remaining = J.frame.samples_per_pixel - next_pass[j];
            if (remaining == 0) continue;
            const uint64_t fit = (in_flight - used) / pixels;
            if (fit == 0) break;
            const uint32_t take = (uint32_t)std::min<uint64_t>(remaining, fit);
            chunk.push_back({j, next_pass[j], take});
            used += take * pixels;
            if (take < remaining) break;
        }
        if (chunk.empty()) break;
        for (const chunk_job& c : chunk) next_pass[c.job] += c.passes;
        while (cursor < job_count && next_pass[cursor] == jobs[cursor].frame.samples_per_pixel) ++cursor;
        launches.push_back(std::move(chunk));
    }
    return launches;
}

/* Record mode of trace_jobs, for a caller that captures the whole call into one graph: every
 * launch is issued as copies and kernels only, its job table from a pinned staging row of its own,
 * all depth-bound + 1 bounces without a host read (an exhausted queue makes a bounce's kernels
 * exit at once), and the report's lane counts and failure words stay on the device. */
struct device_loop {
    bf_dev_job* staging = nullptr;        /* pinned, BF_MAX_BATCH_JOBS job records per launch */
    uint32_t launches = 0;                /* staging rows */
    uint32_t used = 0;                    /* rows taken by the launches issued so far */
    unsigned long long* totals = nullptr; /* device: bfk_history_totals accumulators */
};

/*
 * The transport driver shared by the single-frame and the batched entry points: frames are packed
 * into launches of whole sample passes (one contiguous slot range per frame and launch), and each
 * launch runs camera -> bounce groups (live-path count checked between groups) -> film.
 */
int trace_jobs(bf_hip_workspace* w, const bf_frame_job* jobs, const bf_adjoint_job* adjoints, uint32_t job_count,
               const bf_transport_options* o, bf_transport_report* report, uint64_t* traversal_counters,
               hipStream_t stream, const bf_path_inputs* inputs = nullptr, const bf_path_outputs* outputs = nullptr,
               device_loop* loop = nullptr) {
    if (!w || !jobs || !o) return set_error(BF_INVALID_ARGUMENT, "NULL workspace, jobs or options");
    if (job_count == 0) return set_error(BF_INVALID_ARGUMENT, "empty batch");
    const int ost = validate_options(o);
    if (ost != BF_OK) return ost;
    const bool adjoint = adjoints != nullptr;
    /* paths in flight per launch: the workspace, or a power-of-two fraction of it (launch bits 22-24) */
    const uint64_t in_flight = w->capacity >> ((o->launch & BF_LAUNCH_INFLIGHT_MASK) >> 22);
    /* tape bits are indexed by path vertex: at most diffuse + glossy + specular + 1 vertices */
    const uint64_t vertices = (uint64_t)o->diffuse_depth_limit + o->glossy_depth_limit + o->specular_depth_limit + 1u;
    int device = 0;
    HIP_TRY(hipGetDevice(&device));
    if (device != w->device)
        return set_error(BF_INCOMPATIBLE, "workspace lives on device %d, current device is %d", w->device, device);
    for (uint32_t j = 0; j < job_count; ++j) {
        const bf_frame_job& J = jobs[j];
        if (!J.scene) return set_error(BF_INVALID_ARGUMENT, "job %u has no scene", j);
        if (!J.scene->device_block) return set_error(BF_INVALID_ARGUMENT, "scene was prepared without a device upload");
        if (J.scene->device != device)
            return set_error(BF_INCOMPATIBLE, "scene lives on device %d, current device is %d", J.scene->device, device);
        if (J.scene->has_maps && adjoint)
            return set_error(BF_UNAVAILABLE,
                             "job %u: the adjoint replay does not differentiate normal and bump maps", j);
        if (J.scene->sampled && (o->wavelength_mode != BF_WAVELENGTHS_FRAME_PACKET ||
                                 std::memcmp(o->packet_wavelengths, J.scene->sampled_wavelengths, sizeof J.scene->sampled_wavelengths) != 0))
            return set_error(BF_INVALID_ARGUMENT,
                             "job %u: the scene's spectra are point-sampled; render in frame-packet mode at their wavelengths", j);
        const uint64_t pixels = (uint64_t)J.frame.camera.width * J.frame.camera.height;
        if (pixels == 0 || J.frame.samples_per_pixel == 0) return set_error(BF_INVALID_ARGUMENT, "job %u: empty frame", j);
        if (pixels > in_flight)
            return set_error(BF_RESOURCE_EXHAUSTED, "%llu paths in flight per launch are below one sample pass (%llu paths)",
                             (unsigned long long)in_flight, (unsigned long long)pixels);
        if (!adjoint && !inputs && (!J.film || !J.film_compensated || !J.film_counts))
            return set_error(BF_INVALID_ARGUMENT, "job %u: film buffers are required", j);
        if ((J.tape_radiance == nullptr) != (J.tape_visibility == nullptr))
            return set_error(BF_INVALID_ARGUMENT, "job %u: a tape needs both its radiance and visibility buffers", j);
        if ((J.tape_radiance || adjoint) && vertices > 64u)
            return set_error(BF_INVALID_ARGUMENT, "taped frames need diffuse + glossy + specular depth limits <= 63");
        if (J.reserved != 0u) return set_error(BF_INVALID_ARGUMENT, "job %u: reserved word must be zero", j);
        if (J.survival_mode > BF_SURVIVAL_REPLAY) return set_error(BF_INVALID_ARGUMENT, "job %u: unknown survival mode", j);
        if (J.survival_mode != BF_SURVIVAL_NONE && (!J.tape_survival || vertices > 64u))
            return set_error(BF_INVALID_ARGUMENT,
                             "job %u: a survival mode needs its buffer and diffuse + glossy + specular depth limits <= 63", j);
        if (adjoint) {
            if (!J.tape_radiance) return set_error(BF_INVALID_ARGUMENT, "job %u: the adjoint needs the frame's tape", j);
            if (!adjoints[j].film_gradient) return set_error(BF_INVALID_ARGUMENT, "job %u: no film gradient", j);
            for (uint32_t k = J.scene->dev.texture_count; k < BF_MAX_TEXTURES; ++k)
                if (adjoints[j].texture_gradients[k])
                    return set_error(BF_INVALID_ARGUMENT, "job %u: gradient for texture %u, scene has %u", j, k,
                                     J.scene->dev.texture_count);
        }
    }

    if (loop && (inputs || traversal_counters || o->record_queues || o->instrumentation == BF_INSTRUMENT_STAGE_EVENTS))
        return set_error(BF_INVALID_ARGUMENT, "a recorded step takes frames without counters, queue records or stage events");
    const bool instrument = o->instrumentation == BF_INSTRUMENT_STAGE_EVENTS;
    uint32_t group = o->launch & BF_LAUNCH_GROUP_MASK;
    if (group == 0) group = 4;
    const bool use_graph =
        (o->launch & BF_LAUNCH_GRAPH) != 0 && !instrument && !o->record_queues && !traversal_counters && !loop;
    if (use_graph) group += group & 1u; /* even: the ray / continuation queue roles repeat */

    bf_transport_params P{};
    P.paths = w->paths;
    const uint64_t* directions = nullptr;
    const int sst = sobol_table_for_current_device(&directions);
    if (sst != BF_OK) return sst;
    P.sampler.kind = o->sampler;
    P.sampler.sobol_dims = kSobolDimensions;
    P.sampler.directions = directions;
    P.mis = o->mis_heuristic;
    P.wavelength_mode = o->wavelength_mode;
    P.limits[0] = o->diffuse_depth_limit;
    P.limits[1] = o->glossy_depth_limit;
    P.limits[2] = o->specular_depth_limit;
    P.limits[3] = o->transmission_depth_limit;
    P.limits[4] = 0;
    P.rr_enabled = o->russian_roulette_enabled;
    P.rr_first = o->russian_roulette_first_depth;
    P.rr_min = o->russian_roulette_minimum;
    P.rr_max = o->russian_roulette_maximum;
    for (int k = 0; k < 4; ++k) P.packet[k] = o->packet_wavelengths[k];
    P.failure = w->failure;
    P.spill = w->spill;
    P.fetch = w->fetch;
    P.gradient_replicas = w->replicas;
    P.bucket_counts = w->bucket_counts;
    P.bucket_offsets = w->bucket_offsets;
    /* orders live in queues that are free while the stage runs: the shadow queue is built after
     * shade (and is empty during closest hit); the hit queue is consumed before the shadow stage */
    if (o->launch & BF_LAUNCH_SHADE_ORDER) P.shade_order = w->queues[2];
    if (o->launch & BF_LAUNCH_TRACE_ORDER) {
        P.closest_order = w->queues[2];
        P.shadow_order = w->queues[1];
    }
    P.spill_stride = w->spill_lanes;
    P.shade_blocks = w->shade_blocks;
    P.adjoint_shade_blocks = w->adjoint_shade_blocks;
    const uint32_t block_log2 = (o->launch & BF_LAUNCH_TRACE_BLOCK_MASK) >> 16;
    P.trace_block = block_log2 == 0 ? 128u : 1u << block_log2;
    P.jobs = w->jobs;
    P.film_pdf = 1.0f / 470.0f;
    P.film_normalization = film_normalization();

    bf_transport_report rep{};
    rep.first_failure_slot = BF_INVALID_INDEX;
    w->pending.clear();
    w->events_used = 0;
#define STAGE(stage, dispatches, call) HIP_TRY(staged(w, rep, instrument, (stage), (dispatches), stream, [&] { return (call); }))
    STAGE(BF_STAGE_AUXILIARY, 4, [&] {
        hipError_t e = bfk_set_words((uint32_t*)w->traversal, 12, 0u, stream);
        if (e == hipSuccess) e = bfk_set_words(w->failure, 2, 0u, stream);
        if (e == hipSuccess) e = bfk_set_words(w->failure + 2, 2, 0xFFFFFFFFu, stream);
        if (e == hipSuccess) e = bfk_set_words(w->sample_counts, 2, 0u, stream);
        return e;
    }());
    unsigned long long* counters = traversal_counters ? w->traversal : nullptr;
    const uint64_t bound = (uint64_t)o->diffuse_depth_limit + o->glossy_depth_limit + o->specular_depth_limit + 2u;
    const float normalization = film_normalization();

    /* small gradient buffers accumulate through replica rows (one offset per distinct buffer) */
    struct replica_slot {
        float* dst;
        uint32_t offset, count;
    };
    std::vector<replica_slot> replica_slots;
    std::vector<std::array<uint32_t, BF_MAX_TEXTURES>> replica_offsets(adjoint ? job_count : 0u);
    uint32_t replica_used = 0;
    for (uint32_t j = 0; adjoint && j < job_count; ++j) {
        for (uint32_t k = 0; k < BF_MAX_TEXTURES; ++k) {
            replica_offsets[j][k] = BF_INVALID_INDEX;
            float* dst = adjoints[j].texture_gradients[k];
            if (!dst) continue;
            const uint32_t count = jobs[j].scene->texture_floats[k];
            for (const replica_slot& r : replica_slots)
                if (r.dst == dst) replica_offsets[j][k] = r.offset;
            if (replica_offsets[j][k] != BF_INVALID_INDEX || count > BF_GRADIENT_SMALL || replica_used + count > BF_GRADIENT_ROW)
                continue;
            replica_slots.push_back({dst, replica_used, count});
            replica_offsets[j][k] = replica_used;
            replica_used += count;
        }
    }
    if (replica_used)
        STAGE(BF_STAGE_AUXILIARY, 1,
              bfk_set_words((uint32_t*)w->replicas, BF_GRADIENT_REPLICAS * BF_GRADIENT_ROW, 0u, stream));

    const std::vector<std::vector<chunk_job>> launches = pack_launches(jobs, job_count, in_flight);
    if (loop && launches.size() > loop->launches - loop->used)
        return set_error(BF_RESOURCE_EXHAUSTED, "the recorded step has job tables for %u more launches, the frames need %zu",
                         loop->launches - loop->used, launches.size());
    bool first_launch = true;
    for (const std::vector<chunk_job>& chunk : launches) {
        const bool batch = chunk.size() > 1 || use_graph || loop;
        uint32_t slot = 0;
        for (size_t c = 0; c < chunk.size(); ++c) {
            const bf_frame_job& J = jobs[chunk[c].job];
            bf_dev_job& D = w->host_jobs[c];
            D = bf_dev_job{};
            D.scene = J.scene->dev;
            D.camera = J.frame.camera;
            D.seed = J.frame.seed;
            D.pixels = J.frame.camera.width * J.frame.camera.height;
            D.first_sample_index = J.frame.first_sample_index;
            D.slot_begin = slot;
            D.slot_count = chunk[c].passes * D.pixels;
            D.path_base = (uint64_t)chunk[c].first_pass * D.pixels;
            D.index = chunk[c].job;
            D.film = J.film;
            D.film_compensated = J.film_compensated;
            D.film_counts = J.film_counts;
            D.paths = J.paths;
            D.tape_radiance = (float4*)J.tape_radiance;
            D.tape_visibility = (unsigned long long*)J.tape_visibility;
            if (inputs) {
                D.in_samples = inputs->sample_streams;
                D.in_rays = inputs->rays;
                D.in_cones = inputs->ray_cones;
                D.in_states = inputs->path_states;
                D.out_states = outputs->path_states;
                D.out_rays = outputs->terminal_rays;
                D.out_cones = outputs->terminal_cones;
                D.out_controls = outputs->controls;
            }
            /* camera paths of a mapped scene start with the camera's cone (path inputs bring theirs) */
            D.frame_cones = !inputs && J.scene->has_maps ? 1u : 0u;
            if (J.survival_mode != BF_SURVIVAL_NONE) {
                D.tape_survival = J.tape_survival;
                D.survival_stride = (uint32_t)vertices;
                D.survival_mode = J.survival_mode;
            }
            if (adjoint) {
                const bf_adjoint_job& A = adjoints[chunk[c].job];
                D.film = nullptr;
                D.film_compensated = nullptr;
                D.film_counts = nullptr;
                D.paths = nullptr;
                D.film_gradient = A.film_gradient;
                for (uint32_t k = 0; k < BF_MAX_TEXTURES; ++k) {
                    D.texture_gradients[k] = A.texture_gradients[k];
                    D.gradient_replica[k] = replica_offsets[chunk[c].job][k];
                }
            }
            slot += D.slot_count;
        }
        const uint32_t n = slot;
        P.job_count = (uint32_t)chunk.size();
        bool textured = false;
        bool mapped = false;
        for (size_t c = 0; c < chunk.size(); ++c) {
            textured = textured || jobs[chunk[c].job].scene->dev.has_textures != 0u;
            mapped = mapped || jobs[chunk[c].job].scene->has_maps;
        }
        P.shade_mode = adjoint ? BF_SHADE_ADJOINT
                               : (mapped ? BF_SHADE_MAPPED : (textured ? BF_SHADE_TEXTURED : BF_SHADE_PLAIN));
        if (loop) {
            /* the graph's copy node reads this launch's own staging row at every replay */
            bf_dev_job* row = loop->staging + (size_t)loop->used++ * BF_MAX_BATCH_JOBS;
            std::memcpy(row, w->host_jobs, chunk.size() * sizeof(bf_dev_job));
            HIP_TRY(hipMemcpyAsync(w->jobs, row, chunk.size() * sizeof(bf_dev_job), hipMemcpyHostToDevice, stream));
        } else if (batch) {
            HIP_TRY(hipMemcpyAsync(w->jobs, w->host_jobs, chunk.size() * sizeof(bf_dev_job), hipMemcpyHostToDevice, stream));
        } else {
            P.job0 = w->host_jobs[0];
        }
        uint32_t* ray_q = w->queues[0];
        uint32_t* cont_q = w->queues[3];
        STAGE(BF_STAGE_CAMERA, 1, bfk_camera(&P, n, inputs ? cont_q : ray_q, w->route, batch, adjoint, stream));
        rep.stage_metrics[BF_STAGE_CAMERA].kernel_lanes += n;
        STAGE(BF_STAGE_AUXILIARY, 2, [&] {
            hipError_t e = bfk_set_words(w->sizes, 1, n, stream);
            if (e == hipSuccess) e = bfk_set_words(w->sizes + 1, 7, 0u, stream);
            return e;
        }());
        if (inputs) {
            /* path inputs: only the paths whose inputs validated enter the ray queue (sizes[6] holds
             * the slot count the compaction reads) */
            STAGE(BF_STAGE_AUXILIARY, 1, bfk_set_words(w->sizes + 6, 1, n, stream));
            STAGE(BF_STAGE_COMPACTION, 3,
                  bfk_compact(w->route, &w->sizes[6], n, cont_q, w->tile_counts, w->tile_offsets, ray_q, &w->sizes[0], 0u,
                              w->queues[2], &w->sizes[4], 0u, stream));
        }
        const bool record = o->record_queues && first_launch;
        uint32_t bounce = 0;
        bool live = true;
        if (loop) {
            for (; bounce <= bound; ++bounce) {
                const int ist = issue_bounce(w, rep, P, batch, false, n, ray_q, cont_q, nullptr, nullptr, stream);
                if (ist != BF_OK) return ist;
                std::swap(ray_q, cont_q);
            }
            HIP_TRY(bfk_history_totals(w->history, std::min<uint32_t>(bounce, BF_HISTORY_BOUNCES), loop->totals, stream));
            live = false;
        }
        while (live && bounce <= bound) {
            const uint32_t steps = (uint32_t)std::min<uint64_t>(group, bound + 1 - bounce);
            if (use_graph && steps == group) {
                /* a group is replayed while its launch arguments are unchanged: the parameter block (the
                 * frames themselves are in the device job table), the item bound and the queue roles */
                const bool same = w->graph != nullptr && w->graph_n == n && w->graph_group == group &&
                                  w->graph_ray_q == ray_q && std::memcmp(&w->graph_params, &P, sizeof P) == 0;
                if (!same) {
                    if (w->graph) {
                        (void)hipGraphExecDestroy(w->graph);
                        w->graph = nullptr;
                    }
                    bf_transport_report unused{};
                    if (!w->capture_stream) HIP_TRY(hipStreamCreateWithFlags(&w->capture_stream, hipStreamNonBlocking));
                    HIP_TRY(hipStreamBeginCapture(w->capture_stream, hipStreamCaptureModeThreadLocal));
                    int ist = BF_OK;
                    uint32_t* rq = ray_q;
                    uint32_t* cq = cont_q;
                    for (uint32_t k = 0; k < group && ist == BF_OK; ++k) {
                        ist = issue_bounce(w, unused, P, batch, false, n, rq, cq, nullptr, nullptr, w->capture_stream);
                        std::swap(rq, cq);
                    }
                    hipGraph_t g = nullptr;
                    const hipError_t ce = hipStreamEndCapture(w->capture_stream, &g);
                    if (ist != BF_OK) {
                        if (g) (void)hipGraphDestroy(g);
                        return ist;
                    }
                    if (ce != hipSuccess) return hip_error(ce, "hipStreamEndCapture");
                    const hipError_t ie = hipGraphInstantiate(&w->graph, g, nullptr, nullptr, 0);
                    (void)hipGraphDestroy(g);
                    if (ie != hipSuccess) {
                        w->graph = nullptr;
                        return hip_error(ie, "hipGraphInstantiate");
                    }
                    w->graph_params = P;
                    w->graph_n = n;
                    w->graph_group = group;
                    w->graph_ray_q = ray_q;
                }
                HIP_TRY(hipGraphLaunch(w->graph, stream));
                rep.stage_metrics[BF_STAGE_CLOSEST_HIT].kernel_dispatches += group;
                rep.stage_metrics[BF_STAGE_SHADE].kernel_dispatches += group;
                if (!adjoint) rep.stage_metrics[BF_STAGE_SHADOW].kernel_dispatches += group;
                rep.stage_metrics[BF_STAGE_COMPACTION].kernel_dispatches += (adjoint ? 6u : 9u) * group;
                rep.stage_metrics[BF_STAGE_AUXILIARY].kernel_dispatches += group;
            } else {
                for (uint32_t k = 0; k < steps; ++k) {
                    const bool rec = record && bounce + k < BF_REPORT_MAX_BOUNCES;
                    const int ist = issue_bounce(w, rep, P, batch, instrument, n, ray_q, cont_q, counters,
                                                 rec ? w->hashes + 2 * (bounce + k) : nullptr, stream);
                    if (ist != BF_OK) return ist;
                    std::swap(ray_q, cont_q);
                }
            }
            bounce += steps;
            HIP_TRY(hipMemcpyAsync(w->host_words, w->sizes, 4, hipMemcpyDeviceToHost, stream));
            HIP_TRY(hipStreamSynchronize(stream));
            HIP_TRY(drain_stage_events(w, rep));
            live = w->host_words[0] != 0u;
        }
        /* per-bounce sizes of this launch */
        const uint32_t recorded = loop ? 0u : std::min<uint32_t>(bounce, BF_HISTORY_BOUNCES);
        if (!loop) {
            HIP_TRY(hipMemcpyAsync(w->host_history, w->history, (size_t)recorded * 16, hipMemcpyDeviceToHost, stream));
            if (record)
                HIP_TRY(hipMemcpyAsync(w->host_hashes, w->hashes,
                                       (size_t)std::min<uint32_t>(recorded, BF_REPORT_MAX_BOUNCES) * 16,
                                       hipMemcpyDeviceToHost, stream));
            HIP_TRY(hipStreamSynchronize(stream));
        }
        uint32_t active_bounces = 0;
        for (uint32_t b = 0; b < recorded; ++b) {
            const uint32_t* h = w->host_history + 4 * b;
            if (h[0] == 0u) break;
            active_bounces = b + 1;
            rep.closest_hit_queries += h[0];
            rep.shade_lanes += h[1];
            rep.miss_lanes += h[0] - h[1];
            rep.shadow_queries += h[2];
            rep.stage_metrics[BF_STAGE_CLOSEST_HIT].kernel_lanes += h[0];
            rep.stage_metrics[BF_STAGE_SHADE].kernel_lanes += h[1];
            rep.stage_metrics[BF_STAGE_SHADOW].kernel_lanes += h[2];
            rep.stage_metrics[BF_STAGE_COMPACTION].kernel_lanes += (uint64_t)h[0] + h[1] + h[2];
            if (record && b < BF_REPORT_MAX_BOUNCES) {
                rep.ray_queue_size[b] = h[0];
                rep.shadow_queue_size[b] = h[2];
                rep.ray_queue_hash[b] = w->host_hashes[2 * b];
                rep.shadow_queue_hash[b] = w->host_hashes[2 * b + 1];
            }
        }
        rep.bounces = std::max<uint32_t>(rep.bounces, active_bounces);
        for (size_t c = 0; c < chunk.size() && !adjoint; ++c) {
            const bf_dev_job& D = w->host_jobs[c];
            if (inputs) {
                STAGE(BF_STAGE_AUXILIARY, 1, bfk_export_path_records(&w->paths, &D, stream));
                continue;
            }
            STAGE(BF_STAGE_FILM, 1, bfk_film(&w->paths, &D, P.film_pdf, normalization, stream));
            rep.stage_metrics[BF_STAGE_FILM].kernel_lanes += D.slot_count;
            if (D.paths) STAGE(BF_STAGE_AUXILIARY, 1, bfk_export_paths(&w->paths, &D, stream));
        }
        first_launch = false;
    }
    for (const replica_slot& r : replica_slots)
        STAGE(BF_STAGE_AUXILIARY, 1, bfk_reduce_replicas(w->replicas, r.offset, r.count, r.dst, stream));
    for (uint32_t j = 0; j < job_count; ++j)
        rep.path_count += (uint64_t)jobs[j].frame.camera.width * jobs[j].frame.camera.height * jobs[j].frame.samples_per_pixel;
    if (loop) {
        /* the failure words and sample counts stay on the device for the caller to copy */
        if (report) *report = rep;
        return BF_OK;
    }
    HIP_TRY(hipMemcpyAsync(w->host_words, w->failure, 16, hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipMemcpyAsync(w->host_words + 4, w->sample_counts, 8, hipMemcpyDeviceToHost, stream));
    unsigned long long trav[6] = {0, 0, 0, 0, 0, 0};
    if (traversal_counters) HIP_TRY(hipMemcpyAsync(trav, w->traversal, sizeof trav, hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    HIP_TRY(drain_stage_events(w, rep));
#undef STAGE
    rep.failed_paths = w->host_words[0];
    if (rep.failed_paths != 0) {
        const uint64_t key = (uint64_t)w->host_words[2] | ((uint64_t)w->host_words[3] << 32);
        rep.first_failure_job = (uint32_t)(key >> 48);
        rep.first_failure_slot = (uint32_t)((key >> 8) & 0xFFFFFFFFFFull);
        rep.first_failure_code = (uint32_t)(key & 0xFFu);
    }
    rep.closure_samples = w->host_words[4];
    rep.light_samples = w->host_words[5];
    if (traversal_counters)
        for (int k = 0; k < 6; ++k) traversal_counters[k] = trav[k];
    if (report) *report = rep;
    if (rep.failed_paths != 0 && o->failure_policy == BF_FAILURES_ABORT)
        return set_error(BF_TRANSPORT_FAILURE, "%llu paths failed; first failing path %u of frame %u",
                         (unsigned long long)rep.failed_paths, rep.first_failure_slot, rep.first_failure_job);
    return BF_OK;
}

} // namespace

extern "C" {

int blackframe_hip_trace_wavefront_transport(bf_hip_workspace* w, const bf_hip_scene* s, const bf_frame_desc* frame,
                                             const bf_transport_options* o, float* film, double* film_compensated,
                                             uint32_t* film_counts, bf_path_result* paths, bf_transport_report* report,
                                             uint64_t* traversal_counters, void* stream_handle) {
    if (!w || !s || !frame || !o) return set_error(BF_INVALID_ARGUMENT, "NULL workspace, scene, frame or options");
    bf_frame_job job{};
    job.scene = s;
    job.frame = *frame;
    job.film = film;
    job.film_compensated = film_compensated;
    job.film_counts = film_counts;
    job.paths = paths;
    return trace_jobs(w, &job, nullptr, 1, o, report, traversal_counters, (hipStream_t)stream_handle);
}

int blackframe_hip_trace_wavefront_paths(bf_hip_workspace* w, const bf_hip_scene* s, const bf_path_inputs* inputs,
                                         const bf_transport_options* o, const bf_path_outputs* outputs,
                                         bf_transport_report* report, void* stream_handle) {
    if (!w || !s || !inputs || !o || !outputs)
        return set_error(BF_INVALID_ARGUMENT, "NULL workspace, scene, inputs, options or outputs");
    if (!inputs->sample_streams || !inputs->rays || !inputs->ray_cones || !inputs->path_states)
        return set_error(BF_INVALID_ARGUMENT, "path inputs need all four columns");
    if (!outputs->path_states || !outputs->terminal_rays || !outputs->terminal_cones || !outputs->controls)
        return set_error(BF_INVALID_ARGUMENT, "path outputs need all four columns");
    if (inputs->reserved || inputs->reserved_tail[0] || inputs->reserved_tail[1] || outputs->reserved ||
        outputs->reserved_tail[0] || outputs->reserved_tail[1])
        return set_error(BF_INVALID_ARGUMENT, "reserved words of the path inputs and outputs must be zero");
    if (inputs->count == 0) return set_error(BF_INVALID_ARGUMENT, "no paths");
    if (outputs->count != inputs->count)
        return set_error(BF_INVALID_ARGUMENT, "outputs hold %u paths, inputs %u", outputs->count, inputs->count);
    if (inputs->count > w->capacity)
        return set_error(BF_RESOURCE_EXHAUSTED, "%u paths exceed the workspace capacity %llu", inputs->count,
                         (unsigned long long)w->capacity);
    /* a one-row frame of `count` pixels, one pass: only its extent is read (slot packing) */
    bf_frame_job job{};
    job.scene = s;
    job.frame.camera.width = inputs->count;
    job.frame.camera.height = 1;
    job.frame.samples_per_pixel = 1;
    return trace_jobs(w, &job, nullptr, 1, o, report, nullptr, (hipStream_t)stream_handle, inputs, outputs);
}

int blackframe_hip_trace_wavefront_batch(bf_hip_workspace* w, const bf_frame_job* jobs, uint32_t job_count,
                                         const bf_transport_options* o, bf_transport_report* report,
                                         uint64_t* traversal_counters, void* stream_handle) {
    if (job_count > 0xFFFFu) return set_error(BF_INVALID_ARGUMENT, "at most 65535 frames per batch");
    return trace_jobs(w, jobs, nullptr, job_count, o, report, traversal_counters, (hipStream_t)stream_handle);
}

int blackframe_hip_autotune_launch(bf_hip_workspace* w, const bf_frame_job* jobs, uint32_t job_count,
                                   const bf_transport_options* o, uint32_t trial_passes, bf_launch_tuning* result,
                                   void* stream_handle) {
    if (!w || !jobs || !o || !result || job_count == 0 || trial_passes == 0)
        return set_error(BF_INVALID_ARGUMENT, "NULL or empty autotune arguments");
    if (job_count > 0xFFFFu) return set_error(BF_INVALID_ARGUMENT, "at most 65535 frames per batch");
    hipStream_t stream = (hipStream_t)stream_handle;
    /* trial jobs: the first passes of each frame, rendered into scratch films */
    std::vector<bf_frame_job> trial(jobs, jobs + job_count);
    size_t film_floats = 0, film_doubles = 0, film_words = 0;
    for (bf_frame_job& J : trial) {
        const size_t pixels = (size_t)J.frame.camera.width * J.frame.camera.height;
        if (J.frame.samples_per_pixel > trial_passes) J.frame.samples_per_pixel = trial_passes;
        film_floats = std::max(film_floats, pixels * 3);
        film_doubles = std::max(film_doubles, pixels * 6);
        film_words = std::max(film_words, pixels);
        J.paths = nullptr;
        J.tape_radiance = nullptr;
        J.tape_visibility = nullptr;
        if (J.survival_mode == BF_SURVIVAL_RECORD) J.survival_mode = BF_SURVIVAL_NONE; /* trials leave it untouched */
    }
    void* scratch = nullptr;
    HIP_TRY(hipMalloc(&scratch, film_doubles * 8 + film_floats * 4 + film_words * 4));
    HIP_TRY(hipMemset(scratch, 0, film_doubles * 8 + film_floats * 4 + film_words * 4));
    for (bf_frame_job& J : trial) {
        J.film_compensated = (double*)scratch;
        J.film = (float*)((char*)scratch + film_doubles * 8);
        J.film_counts = (uint32_t*)((char*)scratch + film_doubles * 8 + film_floats * 4);
    }
    hipEvent_t start = nullptr, stop = nullptr;
    int status = BF_OK;
    bf_launch_tuning out{};
    auto run = [&](uint32_t launch, float& ms) -> int {
        bf_transport_options opt = *o;
        opt.launch = launch;
        opt.instrumentation = BF_INSTRUMENT_DISABLED;
        opt.record_queues = 0;
        opt.failure_policy = BF_FAILURES_COUNT;
        int st = trace_jobs(w, trial.data(), nullptr, job_count, &opt, nullptr, nullptr, stream); /* warm: capture */
        if (st != BF_OK) return st;
        HIP_TRY(hipEventRecord(start, stream));
        st = trace_jobs(w, trial.data(), nullptr, job_count, &opt, nullptr, nullptr, stream);
        if (st != BF_OK) return st;
        HIP_TRY(hipEventRecord(stop, stream));
        HIP_TRY(hipEventSynchronize(stop));
        HIP_TRY(hipEventElapsedTime(&ms, start, stop));
        return BF_OK;
    };
    const hipError_t e1 = hipEventCreate(&start), e2 = hipEventCreate(&stop);
    if (e1 != hipSuccess || e2 != hipSuccess) {
        status = hip_error(e1 != hipSuccess ? e1 : e2, "hipEventCreate");
    } else {
        const bool graphs = !o->record_queues && o->instrumentation != BF_INSTRUMENT_STAGE_EVENTS;
        status = run(0u, out.default_ms);
        out.best_ms = out.default_ms;
        out.launch = 0u;
        auto consider = [&](uint32_t launch) -> int {
            float ms = 0.0f;
            const int st = run(launch, ms);
            if (st != BF_OK) return st;
            if (out.candidates < 48) {
                out.tried[out.candidates] = launch;
                out.tried_ms[out.candidates] = ms;
            }
            ++out.candidates;
            if (ms < out.best_ms) {
                out.best_ms = ms;
                out.launch = launch;
            }
            return BF_OK;
        };
        /* stage 1: launch shapes; stage 2: the work orders on the best shape */
        const uint32_t blocks[3] = {6u, 7u, 8u}, groups[3] = {2u, 4u, 8u};
        for (uint32_t b = 0; b < 3 && status == BF_OK; ++b)
            for (uint32_t g = 0; g < 3 && status == BF_OK; ++g)
                for (uint32_t graph = 0; graph < (graphs ? 2u : 1u) && status == BF_OK; ++graph)
                    status = consider(groups[g] | BF_LAUNCH_TRACE_BLOCK(blocks[b]) | (graph ? BF_LAUNCH_GRAPH : 0u));
        const uint32_t shape = out.launch;
        const uint32_t orders[3] = {BF_LAUNCH_SHADE_ORDER, BF_LAUNCH_TRACE_ORDER, BF_LAUNCH_SHADE_ORDER | BF_LAUNCH_TRACE_ORDER};
        for (uint32_t k = 0; k < 3 && status == BF_OK; ++k) status = consider(shape | orders[k]);
        /* stage 3: paths in flight per launch (the whole workspace, a half, a quarter), while every
         * trial frame's sample pass still fits */
        uint64_t largest_pass = 0;
        for (const bf_frame_job& J : trial)
            largest_pass = std::max<uint64_t>(largest_pass, (uint64_t)J.frame.camera.width * J.frame.camera.height);
        const uint32_t best = out.launch;
        for (uint32_t k = 1; k <= 2 && status == BF_OK; ++k)
            if ((w->capacity >> k) >= largest_pass) status = consider(best | BF_LAUNCH_INFLIGHT(k));
    }
    if (start) (void)hipEventDestroy(start);
    if (stop) (void)hipEventDestroy(stop);
    const hipError_t fe = hipFree(scratch);
    if (status != BF_OK) return status;
    if (fe != hipSuccess) return hip_error(fe, "hipFree(autotune films)");
    *result = out;
    return BF_OK;
}

int blackframe_hip_trace_wavefront_adjoint(bf_hip_workspace* w, const bf_frame_job* jobs, const bf_adjoint_job* adjoints,
                                           uint32_t job_count, const bf_transport_options* o, bf_transport_report* report,
                                           void* stream_handle) {
    if (!adjoints) return set_error(BF_INVALID_ARGUMENT, "NULL adjoint jobs");
    if (job_count > 0xFFFFu) return set_error(BF_INVALID_ARGUMENT, "at most 65535 frames per batch");
    return trace_jobs(w, jobs, adjoints, job_count, o, report, nullptr, (hipStream_t)stream_handle);
}

int blackframe_hip_launch_scene_soa_hash(const uint8_t* scene_bytes, size_t scene_byte_count, uint64_t* output_hash) {
    if (!scene_bytes || !output_hash || scene_byte_count < BF_SCENE_SOA_HEADER_SIZE)
        return set_error(BF_INVALID_ARGUMENT, "scene hash needs a blob of at least %u bytes and an output word",
                         BF_SCENE_SOA_HEADER_SIZE);
    HIP_TRY(bfk_scene_hash(scene_bytes, scene_byte_count, output_hash, nullptr));
    return BF_OK;
}

int blackframe_hip_launch_scene_soa_validate(const uint8_t* scene_bytes, size_t scene_byte_count, uint32_t* output_status) {
    if (!scene_bytes || !output_status || scene_byte_count < BF_SCENE_SOA_HEADER_SIZE)
        return set_error(BF_INVALID_ARGUMENT, "scene validation needs a blob of at least %u bytes and an output word",
                         BF_SCENE_SOA_HEADER_SIZE);
    if ((reinterpret_cast<uintptr_t>(scene_bytes) & 15u) != 0)
        return set_error(BF_INVALID_ARGUMENT, "the device blob must be 16-byte aligned, as its header is");
    HIP_TRY(bfk_scene_soa_validate(scene_bytes, scene_byte_count, output_status, nullptr));
    return BF_OK;
}

int blackframe_hip_query_transport_abi_layout(bf_layout_manifest* output, int* device_count) {
    if (!output || !device_count) return set_error(BF_INVALID_ARGUMENT, "NULL manifest or device count");
    std::memset(output, 0, sizeof *output);
    *device_count = 0;
    HIP_TRY(hipGetDeviceCount(device_count));
    if (*device_count == 0) return set_error(BF_UNAVAILABLE, "no HIP device");
    int previous = 0;
    HIP_TRY(hipGetDevice(&previous));
    HIP_TRY(hipSetDevice(0));
    bf_layout_manifest* d = nullptr;
    hipError_t e = hipMalloc(&d, sizeof *d);
    if (e == hipSuccess) e = bfk_layout_manifest(d, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(output, d, sizeof *output, hipMemcpyDeviceToHost);
    const hipError_t fe = d ? hipFree(d) : hipSuccess;
    const hipError_t se = hipSetDevice(previous);
    if (e != hipSuccess) return hip_error(e, "transport layout probe");
    if (fe != hipSuccess) return hip_error(fe, "hipFree(layout manifest)");
    if (se != hipSuccess) return hip_error(se, "hipSetDevice(previous)");
    return BF_OK;
}

int blackframe_hip_launch_wavefront_queue_pushes(bf_queue_header* headers, uint32_t* path_slots, uint32_t queue_count,
                                                 uint32_t slot_stride, const bf_queue_push* requests,
                                                 uint32_t request_count, uint32_t* outcomes) {
    if (!headers || queue_count != BF_QUEUE_COUNT || (slot_stride != 0u && !path_slots) ||
        (request_count != 0u && (!requests || !outcomes)))
        return set_error(BF_INVALID_ARGUMENT, "invalid queue view or push arrays");
    const bf_queue_soa q{headers, path_slots, queue_count, slot_stride};
    HIP_TRY(bfk_queue_pushes(&q, requests, request_count, outcomes, nullptr));
    return BF_OK;
}

int blackframe_hip_query_wavefront_queue_compaction_scratch_bytes(uint32_t max_input_count, size_t* scratch_bytes) {
    if (!scratch_bytes) return set_error(BF_INVALID_ARGUMENT, "NULL scratch_bytes");
    *scratch_bytes = bfk_queue_compaction_scratch(max_input_count);
    return BF_OK;
}

int blackframe_hip_launch_wavefront_queue_compaction(bf_queue_soa queues, const bf_stage_outcome* outcomes,
                                                     uint32_t input_count, uint32_t route, void* scratch,
                                                     size_t scratch_bytes, bf_compaction_result* device_result,
                                                     void* stream_handle) {
    /* CudaWavefrontQueueCompaction.cu:253-273: every rejection happens before any launch */
    if (!queues.headers || queues.queue_count != BF_QUEUE_COUNT || (queues.slot_stride != 0u && !queues.path_slots))
        return set_error(BF_INVALID_ARGUMENT, "invalid queue view");
    if (route < 1u || route > 6u) return set_error(BF_INVALID_ARGUMENT, "route %u is not a compactable queue kind", route);
    if (input_count > queues.slot_stride)
        return set_error(BF_INVALID_ARGUMENT, "%u outcomes exceed the queue capacity %u", input_count, queues.slot_stride);
    if (!device_result || (input_count != 0u && !outcomes)) return set_error(BF_INVALID_ARGUMENT, "NULL outcomes or result");
    const size_t required = bfk_queue_compaction_scratch(input_count);
    if (scratch_bytes < required ||
        (required != 0u && (!scratch || reinterpret_cast<uintptr_t>(scratch) % BF_COMPACTION_SCRATCH_ALIGNMENT != 0u)))
        return set_error(BF_INVALID_ARGUMENT, "compaction scratch: %zu bytes at %p, %zu aligned bytes required",
                         scratch_bytes, scratch, required);
    HIP_TRY(bfk_queue_compaction(&queues, outcomes, input_count, route, scratch, device_result,
                                 (hipStream_t)stream_handle));
    return BF_OK;
}

namespace {
int blob_traverse(bool any, const uint8_t* scene_bytes, size_t scene_size, const uint8_t* bvh_bytes, size_t bvh_size,
                  const bf_transport_ray* rays, uint32_t ray_count, void* results, void* stream_handle) {
    if (!scene_bytes || scene_size < sizeof(bf_scene_soa_header) || !bvh_bytes || bvh_size < sizeof(bf_bvh_header) ||
        (ray_count != 0u && (!rays || !results)))
        return set_error(BF_INVALID_ARGUMENT,
                         "scene and BVH blobs of at least their header sizes, and rays and results when rays are given");
    if (ray_count == 0u) return BF_OK;
    HIP_TRY(bfk_blob_traverse(any, scene_bytes, scene_size, bvh_bytes, bvh_size, rays, ray_count, results,
                              (hipStream_t)stream_handle));
    return BF_OK;
}
} // namespace

int blackframe_hip_launch_scene_closest_hit_blob(const uint8_t* scene_bytes, size_t scene_size, const uint8_t* bvh_bytes,
                                                 size_t bvh_size, const bf_transport_ray* rays, uint32_t ray_count,
                                                 bf_closest_hit_result* results, void* stream) {
    return blob_traverse(false, scene_bytes, scene_size, bvh_bytes, bvh_size, rays, ray_count, results, stream);
}

int blackframe_hip_launch_scene_occlusion_blob(const uint8_t* scene_bytes, size_t scene_size, const uint8_t* bvh_bytes,
                                               size_t bvh_size, const bf_transport_ray* rays, uint32_t ray_count,
                                               bf_occlusion_result* results, void* stream) {
    return blob_traverse(true, scene_bytes, scene_size, bvh_bytes, bvh_size, rays, ray_count, results, stream);
}

int blackframe_hip_launch_wavefront_audit_stage(const bf_stage_outcome* outcomes, uint32_t work_count,
                                                uint32_t allowed_route_mask, uint32_t path_capacity,
                                                uint32_t stage_kind, bf_stage_audit* audit, void* stream_handle) {
    if (!audit || (work_count != 0u && !outcomes)) return set_error(BF_INVALID_ARGUMENT, "NULL audit or outcomes");
    if (stage_kind > BF_STAGE_KIND_CONTINUATION) return set_error(BF_INVALID_ARGUMENT, "unknown stage kind %u", stage_kind);
    HIP_TRY(bfk_audit_stage(outcomes, work_count, allowed_route_mask, path_capacity, stage_kind, audit,
                            (hipStream_t)stream_handle));
    return BF_OK;
}

int blackframe_hip_launch_wavefront_queue_compaction_static(bf_queue_soa queues, const bf_stage_outcome* outcomes,
                                                            uint32_t input_count, uint32_t route,
                                                            bf_compaction_result* device_result, void* stream_handle) {
    if (!queues.headers || queues.queue_count != BF_QUEUE_COUNT || (queues.slot_stride != 0u && !queues.path_slots))
        return set_error(BF_INVALID_ARGUMENT, "invalid queue view");
    if (route < 1u || route > 6u) return set_error(BF_INVALID_ARGUMENT, "route %u is not a compactable queue kind", route);
    if (input_count > queues.slot_stride)
        return set_error(BF_INVALID_ARGUMENT, "%u outcomes exceed the queue capacity %u", input_count, queues.slot_stride);
    if (input_count > BF_STATIC_COMPACTION_MAX_INPUT)
        return set_error(BF_INVALID_ARGUMENT, "%u outcomes exceed the library-owned scratch (at most %u)", input_count,
                         BF_STATIC_COMPACTION_MAX_INPUT);
    if (!device_result || (input_count != 0u && !outcomes)) return set_error(BF_INVALID_ARGUMENT, "NULL outcomes or result");
    void* scratch = nullptr;
    HIP_TRY(bfk_static_compaction_scratch(&scratch));
    HIP_TRY(bfk_queue_compaction(&queues, outcomes, input_count, route, scratch, device_result,
                                 (hipStream_t)stream_handle));
    return BF_OK;
}

} /* extern "C" */

namespace {

int validate_adam(const bf_adam_param* params, uint32_t param_count, const bf_adam_options* o) {
    if (!params || !o) return set_error(BF_INVALID_ARGUMENT, "NULL parameters or options");
    if (o->step == 0) return set_error(BF_INVALID_ARGUMENT, "Adam steps count from 1");
    if (o->reserved != 0) return set_error(BF_INVALID_ARGUMENT, "Adam options: reserved word must be zero");
    if (!(o->beta1 >= 0.0 && o->beta1 < 1.0) || !(o->beta2 >= 0.0 && o->beta2 < 1.0))
        return set_error(BF_INVALID_ARGUMENT, "Adam betas must lie in [0, 1)");
    if (!(o->epsilon > 0.0) || !std::isfinite(o->epsilon)) return set_error(BF_INVALID_ARGUMENT, "Adam epsilon must be positive");
    for (uint32_t k = 0; k < param_count; ++k) {
        const bf_adam_param& P = params[k];
        if (P.reserved != 0) return set_error(BF_INVALID_ARGUMENT, "Adam parameter %u: reserved word must be zero", k);
        if (P.count && (!P.value || !P.gradient || !P.first_moment || !P.second_moment))
            return set_error(BF_INVALID_ARGUMENT, "Adam parameter %u: NULL buffer", k);
        if (!std::isfinite(P.learning_rate) || !(P.lower <= P.upper))
            return set_error(BF_INVALID_ARGUMENT, "Adam parameter %u: bad learning rate or bounds", k);
    }
    return BF_OK;
}

/* The float coefficients of one parameter at step number `step`, from binary64 bias corrections. */
bool adam_coefficients(const bf_adam_options& o, const bf_adam_param& P, uint64_t step, bfk_adam_coefficients& c) {
    const double bias1 = 1.0 - std::pow(o.beta1, (double)step);
    const double bias2 = 1.0 - std::pow(o.beta2, (double)step);
    if (!(bias1 > 0.0) || !(bias2 > 0.0)) return false;
    c = bfk_adam_coefficients{(float)o.beta1,
                              (float)(1.0 - o.beta1),
                              (float)o.beta2,
                              (float)(1.0 - o.beta2),
                              (float)((double)P.learning_rate / bias1),
                              (float)std::sqrt(bias2),
                              (float)o.epsilon,
                              P.lower,
                              P.upper};
    return true;
}

} /* namespace */

extern "C" {

int blackframe_hip_film_l2_loss(const bf_film_loss_job* jobs, uint32_t job_count, float scale, double* losses,
                                void* stream_handle) {
    if (job_count == 0) return BF_OK;
    if (!jobs || !losses) return set_error(BF_INVALID_ARGUMENT, "NULL jobs or losses");
    if (!std::isfinite(scale)) return set_error(BF_INVALID_ARGUMENT, "non-finite loss scale");
    for (uint32_t j = 0; j < job_count; ++j) {
        const bf_film_loss_job& J = jobs[j];
        if (J.reserved != 0) return set_error(BF_INVALID_ARGUMENT, "film loss job %u: reserved word must be zero", j);
        if (J.pixel_count && (!J.film || !J.film_counts || !J.target || !J.film_gradient))
            return set_error(BF_INVALID_ARGUMENT, "film loss job %u: NULL buffer", j);
    }
    const hipStream_t s = (hipStream_t)stream_handle;
    HIP_TRY(bfk_zero_doubles(losses, job_count, s));
    const float two_scale = (float)(2.0 * (double)scale);
    for (uint32_t j = 0; j < job_count; ++j) {
        const bf_film_loss_job& J = jobs[j];
        HIP_TRY(bfk_film_l2(J.film, J.film_counts, J.target, J.film_gradient, J.pixel_count, two_scale, (double)scale,
                            losses + j, s));
    }
    return BF_OK;
}

int blackframe_hip_adam_step(const bf_adam_param* params, uint32_t param_count, const bf_adam_options* o,
                             void* stream_handle) {
    if (param_count == 0) return BF_OK;
    const int st = validate_adam(params, param_count, o);
    if (st != BF_OK) return st;
    const hipStream_t s = (hipStream_t)stream_handle;
    for (uint32_t k = 0; k < param_count; ++k) {
        const bf_adam_param& P = params[k];
        bfk_adam_coefficients c{};
        if (!adam_coefficients(*o, P, o->step, c)) return set_error(BF_INVALID_ARGUMENT, "Adam bias correction vanishes");
        HIP_TRY(bfk_adam(P.value, P.gradient, P.first_moment, P.second_moment, P.count, &c, nullptr, nullptr, s));
    }
    return BF_OK;
}

/* ---- one fit step as one graph ------------------------------------------------------------ */

struct bf_hip_fit_graph {
    bf_hip_workspace* w = nullptr;
    int device = 0;
    hipStream_t capture = nullptr;
    hipGraphExec_t exec = nullptr;
    bf_dev_job* staging = nullptr;           /* pinned job tables, one row per launch of either render */
    uint32_t* words = nullptr;               /* device: forward failure[4] samples[2], adjoint failure[4] samples[2], step */
    unsigned long long* totals = nullptr;    /* device: forward totals[4], adjoint totals[4] */
    float2* schedule = nullptr;              /* device: [param][max_steps] (step size, sqrt(bias2)) */
    uint32_t max_steps = 0, steps = 0, frame_count = 0;
    uint64_t paths = 0;
    uint32_t failure_policy = 0;
};

namespace {

int fit_graph_release(bf_hip_fit_graph* g) {
    if (!g) return BF_OK;
    hipError_t first = hipSuccess;
    const auto keep = [&](hipError_t e) {
        if (first == hipSuccess) first = e;
    };
    if (g->exec) keep(hipGraphExecDestroy(g->exec));
    if (g->capture) keep(hipStreamDestroy(g->capture));
    if (g->staging) keep(hipHostFree(g->staging));
    if (g->words) keep(hipFree(g->words));
    if (g->totals) keep(hipFree(g->totals));
    if (g->schedule) keep(hipFree(g->schedule));
    delete g;
    return first == hipSuccess ? BF_OK : hip_error(first, "fit graph release");
}

/* Everything one step issues, on the capturing stream. */
int record_fit_step(bf_hip_fit_graph* g, const bf_fit_step_desc& d, const bf_transport_options* o, uint32_t rows) {
    bf_hip_workspace* w = g->w;
    const hipStream_t s = g->capture;
    HIP_TRY(bfk_set_words((uint32_t*)g->totals, 16, 0u, s));
    for (uint32_t k = 0; k < d.frame_count; ++k) {
        const bf_frame_job& J = d.frames[k];
        const uint32_t pixels = J.frame.camera.width * J.frame.camera.height;
        HIP_TRY(bfk_set_words((uint32_t*)J.film, 3u * pixels, 0u, s));
        HIP_TRY(bfk_set_words((uint32_t*)J.film_compensated, 12u * pixels, 0u, s));
        HIP_TRY(bfk_set_words(J.film_counts, pixels, 0u, s));
    }
    device_loop forward;
    forward.staging = g->staging;
    forward.launches = rows;
    forward.totals = g->totals;
    int st = trace_jobs(w, d.frames, nullptr, d.frame_count, o, nullptr, nullptr, s, nullptr, nullptr, &forward);
    if (st != BF_OK) return st;
    HIP_TRY(bfk_copy_words(w->failure, g->words, 4, s));
    HIP_TRY(bfk_copy_words(w->sample_counts, g->words + 4, 2, s));
    for (uint32_t k = 0; k < d.frame_count; ++k) {
        st = blackframe_hip_film_l2_loss(&d.losses[k], 1, d.loss_scales[k], d.loss_values + k, s);
        if (st != BF_OK) return st;
    }
    for (uint32_t k = 0; k < d.param_count; ++k) {
        const bf_adam_param& P = d.params[k];
        for (uint64_t done = 0; done < P.count;) {
            const uint32_t n = (uint32_t)std::min<uint64_t>(P.count - done, 1u << 30);
            HIP_TRY(bfk_set_words((uint32_t*)(P.gradient + done), n, 0u, s));
            done += n;
        }
    }
    device_loop adjoint;
    adjoint.staging = g->staging + (size_t)rows * BF_MAX_BATCH_JOBS;
    adjoint.launches = rows;
    adjoint.totals = g->totals + 4;
    st = trace_jobs(w, d.frames, d.adjoints, d.frame_count, o, nullptr, nullptr, s, nullptr, nullptr, &adjoint);
    if (st != BF_OK) return st;
    HIP_TRY(bfk_copy_words(w->failure, g->words + 6, 4, s));
    HIP_TRY(bfk_copy_words(w->sample_counts, g->words + 10, 2, s));
    for (uint32_t k = 0; k < d.param_count; ++k) {
        const bf_adam_param& P = d.params[k];
        bfk_adam_coefficients c{};
        if (!adam_coefficients(d.adam, P, d.adam.step, c)) return set_error(BF_INVALID_ARGUMENT, "Adam bias correction vanishes");
        HIP_TRY(bfk_adam(P.value, P.gradient, P.first_moment, P.second_moment, P.count, &c,
                         g->schedule + (size_t)k * g->max_steps, g->words + 12, s));
    }
    HIP_TRY(bfk_tick(g->words + 12, s));
    return BF_OK;
}

void report_from_words(const uint32_t* words, const unsigned long long* totals, uint64_t paths, bf_transport_report& r) {
    r = bf_transport_report{};
    r.path_count = paths;
    r.closest_hit_queries = totals[0];
    r.shade_lanes = totals[1];
    r.miss_lanes = totals[0] - totals[1];
    r.shadow_queries = totals[2];
    r.bounces = (uint32_t)totals[3];
    r.failed_paths = words[0];
    r.first_failure_slot = BF_INVALID_INDEX;
    if (r.failed_paths != 0) {
        const uint64_t key = (uint64_t)words[2] | ((uint64_t)words[3] << 32);
        r.first_failure_job = (uint32_t)(key >> 48);
        r.first_failure_slot = (uint32_t)((key >> 8) & 0xFFFFFFFFFFull);
        r.first_failure_code = (uint32_t)(key & 0xFFu);
    }
    r.closure_samples = words[4];
    r.light_samples = words[5];
}

} /* namespace */

int blackframe_hip_fit_graph_create(bf_hip_workspace* w, const bf_fit_step_desc* d, const bf_transport_options* o,
                                    bf_hip_fit_graph** out) {
    if (!out) return set_error(BF_INVALID_ARGUMENT, "NULL output");
    *out = nullptr;
    if (!w || !d || !o) return set_error(BF_INVALID_ARGUMENT, "NULL workspace, step descriptor or options");
    if (d->reserved != 0) return set_error(BF_INVALID_ARGUMENT, "step descriptor: reserved word must be zero");
    if (d->frame_count == 0 || !d->frames || !d->adjoints || !d->losses || !d->loss_scales || !d->loss_values)
        return set_error(BF_INVALID_ARGUMENT, "a fit step needs frames, adjoint jobs, loss jobs, scales and loss values");
    if (d->max_steps == 0) return set_error(BF_INVALID_ARGUMENT, "max_steps must be positive");
    if (d->param_count) {
        const int st = validate_adam(d->params, d->param_count, &d->adam);
        if (st != BF_OK) return st;
    }
    for (uint32_t k = 0; k < d->frame_count; ++k) {
        const bf_film_loss_job& L = d->losses[k];
        const bf_frame_job& J = d->frames[k];
        if (!std::isfinite(d->loss_scales[k])) return set_error(BF_INVALID_ARGUMENT, "frame %u: non-finite loss scale", k);
        if (!J.film || !J.film_compensated || !J.film_counts)
            return set_error(BF_INVALID_ARGUMENT, "frame %u: film buffers are required", k);
        if (L.film != J.film || L.film_counts != J.film_counts || L.film_gradient != d->adjoints[k].film_gradient ||
            L.pixel_count != J.frame.camera.width * J.frame.camera.height || !L.target || L.reserved != 0)
            return set_error(BF_INVALID_ARGUMENT, "frame %u: the loss job must read the frame's film and write its film gradient", k);
    }
    const uint64_t in_flight = w->capacity >> ((o->launch & BF_LAUNCH_INFLIGHT_MASK) >> 22);
    const size_t rows = pack_launches(d->frames, d->frame_count, in_flight).size();
    if (rows == 0) return set_error(BF_RESOURCE_EXHAUSTED, "the frames do not fit the workspace's paths in flight");
    const uint64_t* directions = nullptr;
    int st = sobol_table_for_current_device(&directions); /* uploaded before the capture */
    if (st != BF_OK) return st;

    bf_hip_fit_graph* g = new bf_hip_fit_graph;
    g->w = w;
    g->device = w->device;
    g->max_steps = d->max_steps;
    g->frame_count = d->frame_count;
    g->failure_policy = o->failure_policy;
    for (uint32_t k = 0; k < d->frame_count; ++k)
        g->paths += (uint64_t)d->frames[k].frame.camera.width * d->frames[k].frame.camera.height *
                    d->frames[k].frame.samples_per_pixel;
    const auto fail = [&](int status) {
        (void)fit_graph_release(g);
        return status;
    };
    const auto hip_fail = [&](hipError_t e, const char* what) { return fail(hip_error(e, what)); };
    hipError_t e = hipHostMalloc((void**)&g->staging, 2 * rows * BF_MAX_BATCH_JOBS * sizeof(bf_dev_job), hipHostMallocDefault);
    if (e != hipSuccess) return hip_fail(e, "hipHostMalloc (job tables)");
    if ((e = hipMalloc((void**)&g->words, 16 * sizeof(uint32_t))) != hipSuccess) return hip_fail(e, "hipMalloc (step words)");
    if ((e = hipMalloc((void**)&g->totals, 8 * sizeof(unsigned long long))) != hipSuccess)
        return hip_fail(e, "hipMalloc (step totals)");
    std::vector<float2> schedule((size_t)std::max<uint32_t>(d->param_count, 1u) * d->max_steps);
    for (uint32_t k = 0; k < d->param_count; ++k)
        for (uint32_t i = 0; i < d->max_steps; ++i) {
            bfk_adam_coefficients c{};
            if (!adam_coefficients(d->adam, d->params[k], (uint64_t)d->adam.step + i, c))
                return fail(set_error(BF_INVALID_ARGUMENT, "Adam bias correction vanishes at step %u",
                                      d->adam.step + i));
            schedule[(size_t)k * d->max_steps + i] = make_float2(c.step_size, c.bias2_sqrt);
        }
    if ((e = hipMalloc((void**)&g->schedule, schedule.size() * sizeof(float2))) != hipSuccess)
        return hip_fail(e, "hipMalloc (Adam schedule)");
    if ((e = hipMemcpy(g->schedule, schedule.data(), schedule.size() * sizeof(float2), hipMemcpyHostToDevice)) != hipSuccess)
        return hip_fail(e, "hipMemcpy (Adam schedule)");
    const uint32_t zero_words[16] = {};
    if ((e = hipMemcpy(g->words, zero_words, sizeof zero_words, hipMemcpyHostToDevice)) != hipSuccess)
        return hip_fail(e, "hipMemcpy (step words)");
    if ((e = hipStreamCreateWithFlags(&g->capture, hipStreamNonBlocking)) != hipSuccess)
        return hip_fail(e, "hipStreamCreateWithFlags");
    if ((e = hipStreamBeginCapture(g->capture, hipStreamCaptureModeThreadLocal)) != hipSuccess)
        return hip_fail(e, "hipStreamBeginCapture");
    st = record_fit_step(g, *d, o, (uint32_t)rows);
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(g->capture, &graph);
    if (st != BF_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return fail(st);
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    e = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
        g->exec = nullptr;
        return hip_fail(e, "hipGraphInstantiate");
    }
    *out = g;
    return BF_OK;
}

int blackframe_hip_fit_graph_launch(bf_hip_fit_graph* g, void* stream_handle) {
    if (!g) return set_error(BF_INVALID_ARGUMENT, "NULL fit graph");
    if (g->steps >= g->max_steps)
        return set_error(BF_RESOURCE_EXHAUSTED, "the fit graph's Adam schedule covers %u steps", g->max_steps);
    int device = 0;
    HIP_TRY(hipGetDevice(&device));
    if (device != g->device) return set_error(BF_INCOMPATIBLE, "fit graph lives on device %d, current device is %d", g->device, device);
    HIP_TRY(hipGraphLaunch(g->exec, (hipStream_t)stream_handle));
    ++g->steps;
    return BF_OK;
}

int blackframe_hip_fit_graph_report(bf_hip_fit_graph* g, bf_transport_report* forward, bf_transport_report* adjoint,
                                    uint32_t* steps, void* stream_handle) {
    if (!g) return set_error(BF_INVALID_ARGUMENT, "NULL fit graph");
    const hipStream_t s = (hipStream_t)stream_handle;
    uint32_t words[16];
    unsigned long long totals[8];
    HIP_TRY(hipStreamSynchronize(s));
    HIP_TRY(hipMemcpy(words, g->words, sizeof words, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(totals, g->totals, sizeof totals, hipMemcpyDeviceToHost));
    bf_transport_report f{}, a{};
    report_from_words(words, totals, g->paths, f);
    report_from_words(words + 6, totals + 4, g->paths, a);
    if (forward) *forward = f;
    if (adjoint) *adjoint = a;
    if (steps) *steps = words[12];
    if (g->failure_policy == BF_FAILURES_ABORT && (f.failed_paths != 0 || a.failed_paths != 0))
        return set_error(BF_TRANSPORT_FAILURE, "%llu forward and %llu adjoint paths failed in the last step",
                         (unsigned long long)f.failed_paths, (unsigned long long)a.failed_paths);
    return BF_OK;
}

int blackframe_hip_fit_graph_destroy(bf_hip_fit_graph* g) { return fit_graph_release(g); }

} /* extern "C" */

/* The error channel of blackframe_hip_last_error for the library's other host sources. */
int bf_report_error(int status, const char* fmt, ...) {
    char buffer[512];
    va_list args;
    va_start(args, fmt);
    std::vsnprintf(buffer, sizeof buffer, fmt, args);
    va_end(args);
    g_last_error = buffer;
    return status;
}
