/*
 * The reference's serialized scene ("BLKFSOA4", ABI 4.0) at the C ABI: header validation and the
 * blob -> bf_scene_desc mapping behind blackframe_hip_scene_create_from_soa. Host code only; the
 * device scene is then built by blackframe_hip_scene_create.
 *
 * Column schema and header validation: bf_scene_soa.h.
 * Column contents: Backends/GPU/CUDA/Source/SceneSoA.cpp:333-1178.
 */
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <new>
#include <vector>

#include "../../include/blackframe_hip.h"
#include "bf_scene_soa.h"

using namespace bf_soa;

int bf_report_error(int status, const char* fmt, ...);
/* Records the content hash of the blob a scene was made from (bf_host.cpp). */
void bf_scene_set_source_hash(bf_hip_scene* s, uint64_t hash);

namespace {

/* FNV-1a-64 with the content-hash bytes read as zero (SceneSoA.cpp:226-236). */
uint64_t normalized_hash(const uint8_t* bytes, uint64_t n) {
    uint64_t h = BF_FNV1A64_OFFSET_BASIS;
    for (uint64_t i = 0; i < n; ++i) {
        const uint8_t v = (i >= BF_SCENE_SOA_HASH_OFFSET && i < BF_SCENE_SOA_HASH_OFFSET + 8u) ? 0u : bytes[i];
        h = (h ^ v) * BF_FNV1A64_PRIME;
    }
    return h;
}

uint32_t check_blob(const void* blob, size_t byte_count, uint32_t flags, bf_scene_soa_header& h) {
    if (byte_count < BF_SCENE_SOA_HEADER_SIZE) return BF_SCENE_SOA_TRUNCATED;
    std::memcpy(&h, blob, sizeof h);
    const uint32_t st = validate_header(h);
    if (st != BF_SCENE_SOA_VALID) return st;
    if (h.total_size_bytes > byte_count) return BF_SCENE_SOA_TRUNCATED;
    if ((flags & BF_SCENE_SOA_VERIFY_HASH) &&
        normalized_hash(static_cast<const uint8_t*>(blob), h.total_size_bytes) != h.content_hash)
        return BF_SCENE_SOA_HASH_MISMATCH;
    return BF_SCENE_SOA_VALID;
}

const char* status_name(uint32_t s) {
    static const char* names[] = {"valid", "bad magic", "incompatible version", "bad header size", "bad column count",
                                  "unsupported hash algorithm", "nonzero reserved", "invalid environment count",
                                  "invalid column descriptor", "size overflow", "bad total size", "truncated blob",
                                  "content hash mismatch"};
    return s < sizeof names / sizeof names[0] ? names[s] : "unknown";
}

/* A validated blob's columns, copied out element by element (the blob may be unaligned). */
struct Blob {
    const uint8_t* bytes;
    const bf_scene_soa_header& h;
    template <class T>
    std::vector<T> column(uint32_t c) const {
        std::vector<T> out(h.columns[c].element_count);
        if (!out.empty()) std::memcpy(out.data(), bytes + h.columns[c].offset_bytes, out.size() * sizeof(T));
        return out;
    }
};

/* Closure selection probabilities on the 2^-24 grid, identical to the Python scene builder's
 * project_probabilities: float64 shares of the float64 sum (sequential, or numpy's eight-way
 * pairwise order at eight terms), floors, then the missing units to the largest remainders, the
 * lowest index first on ties. */
void project_probabilities(const float* weights, uint32_t n, float* probabilities, float* cdf) {
    double sum = 0.0;
    if (n == 8) {
        double r[8];
        for (uint32_t i = 0; i < 8; ++i) r[i] = weights[i];
        sum = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
    } else {
        for (uint32_t i = 0; i < n; ++i) sum += (double)weights[i];
    }
    const double scale = 16777216.0;
    int64_t masses[BF_MAX_CLOSURES];
    double remainders[BF_MAX_CLOSURES];
    int64_t total = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const double scaled = (double)weights[i] / sum * scale;
        masses[i] = (int64_t)std::floor(scaled);
        remainders[i] = scaled - (double)masses[i];
        total += masses[i];
    }
    uint32_t order[BF_MAX_CLOSURES];
    for (uint32_t i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order, order + n, [&](uint32_t a, uint32_t b) { return remainders[a] > remainders[b]; });
    for (int64_t u = 0; u < (int64_t)scale - total && u < (int64_t)n; ++u) masses[order[u]] += 1;
    double running = 0.0;
    cdf[0] = 0.0f;
    for (uint32_t i = 0; i < n; ++i) {
        probabilities[i] = (float)((double)masses[i] / scale);
        running += (double)masses[i] / scale;
        cdf[i + 1] = (float)running;
    }
}

struct Spectra {
    std::vector<bf_spectrum_desc> descs;
    std::vector<float> values;
    float wavelengths[BF_SPECTRUM_LANES];
    uint32_t add(const float lanes[BF_SPECTRUM_LANES]) {
        const uint32_t id = (uint32_t)descs.size();
        descs.push_back(bf_spectrum_desc{(uint32_t)values.size(), BF_SPECTRUM_SAMPLED | BF_SPECTRUM_LANES});
        values.insert(values.end(), wavelengths, wavelengths + BF_SPECTRUM_LANES);
        values.insert(values.end(), lanes, lanes + BF_SPECTRUM_LANES);
        return id;
    }
};

template <class T>
bool strictly_increasing(const std::vector<T>& v) {
    for (size_t i = 1; i < v.size(); ++i)
        if (!(v[i - 1] < v[i])) return false;
    return true;
}

int index_of(const std::vector<uint32_t>& sorted_ids, uint32_t id, uint32_t& index) {
    const auto it = std::lower_bound(sorted_ids.begin(), sorted_ids.end(), id);
    if (it == sorted_ids.end() || *it != id) return 0;
    index = (uint32_t)(it - sorted_ids.begin());
    return 1;
}

int build_from_blob(const uint8_t* bytes, const bf_scene_soa_header& h, uint32_t flags, bf_hip_scene** out) {
    (void)flags;
    const Blob B{bytes, h};
    if (h.geometry_count == 0 || h.instance_count == 0 || h.material_count == 0)
        return bf_report_error(BF_INVALID_ARGUMENT, "scene blob: a renderable scene needs geometries, materials and instances");
    if (h.geometry_count > 0xFFFFFFFFull || h.material_count > 0xFFFFFFFFull || h.instance_count > 0xFFFFFFFFull ||
        h.closure_count > 0xFFFFFFFFull)
        return bf_report_error(BF_RESOURCE_EXHAUSTED, "scene blob: record counts exceed 32 bits");

    /* canonical identifier order (FrameScene.cpp:413-420) */
    const auto object_ids = B.column<uint32_t>(C_OBJECT_ID);
    const auto geometry_ids = B.column<uint32_t>(C_GEOMETRY_ID);
    const auto material_ids = B.column<uint32_t>(C_MATERIAL_ID);
    const auto instance_ids = B.column<uint32_t>(C_INSTANCE_ID);
    if (!strictly_increasing(object_ids) || !strictly_increasing(geometry_ids) || !strictly_increasing(material_ids) ||
        !strictly_increasing(instance_ids))
        return bf_report_error(BF_INVALID_ARGUMENT, "scene blob: identifiers are not in strictly increasing order");

    /* geometry ranges: consecutive, covering every vertex and triangle (SceneSoA.cpp:496-508) */
    const uint32_t G = (uint32_t)h.geometry_count;
    const auto vo = B.column<uint64_t>(C_GEOMETRY_VERTEX_OFFSET), vc = B.column<uint64_t>(C_GEOMETRY_VERTEX_COUNT);
    const auto to = B.column<uint64_t>(C_GEOMETRY_TRIANGLE_OFFSET), tc = B.column<uint64_t>(C_GEOMETRY_TRIANGLE_COUNT);
    std::vector<bf_geometry_desc> geometries(G);
    uint64_t vrun = 0, trun = 0;
    for (uint32_t g = 0; g < G; ++g) {
        if (vo[g] != vrun || to[g] != trun || vc[g] > h.vertex_count - vrun || tc[g] > h.triangle_count - trun)
            return bf_report_error(BF_INVALID_ARGUMENT, "scene blob: geometry %u has an inconsistent range", g);
        geometries[g] = bf_geometry_desc{geometry_ids[g], 0u, vo[g], vc[g], to[g], tc[g]};
        vrun += vc[g];
        trun += tc[g];
    }
    if (vrun != h.vertex_count || trun != h.triangle_count)
        return bf_report_error(BF_INVALID_ARGUMENT, "scene blob: geometry ranges do not cover the mesh columns");
    const uint64_t V = h.vertex_count, T = h.triangle_count;
    std::vector<float> positions(3 * V), normals(3 * V), uvs(2 * V);
    std::vector<uint32_t> indices(3 * T);
    for (uint32_t a = 0; a < 3; ++a) {
        const auto p = B.column<float>(C_POSITION_X + a), n = B.column<float>(C_NORMAL_X + a);
        const auto t = B.column<uint32_t>(C_TRIANGLE_VERTEX_0 + a);
        for (uint64_t v = 0; v < V; ++v) {
            positions[3 * v + a] = p[v];
            normals[3 * v + a] = n[v];
        }
        for (uint64_t i = 0; i < T; ++i) indices[3 * i + a] = t[i];
    }
    for (uint32_t a = 0; a < 2; ++a) {
        const auto u = B.column<float>(C_TEXCOORD_X + a);
        for (uint64_t v = 0; v < V; ++v) uvs[2 * v + a] = u[v];
    }

    /* one wavelength packet for every spectral material and the environment (spectra_match,
     * CudaWavefrontStages.cu:1170-1187, against the uniform visible sampler's pdf) */
    Spectra spectra;
    bool have_packet = false;
    const float uniform_pdf = 1.0f / 470.0f;
    auto admit_packet = [&](const float* nm, const float* pdf, const uint8_t* measure, const char* what) -> int {
        for (uint32_t k = 0; k < BF_SPECTRUM_LANES; ++k)
            if (!std::isfinite(nm[k]) || !(nm[k] > 0.0f) || pdf[k] != uniform_pdf || measure[k] != kMeasureWavelength)
                return bf_report_error(BF_INVALID_ARGUMENT,
                                       "scene blob: %s wavelengths are not a uniform visible packet (pdf 1/470, wavelength measure)", what);
        if (!have_packet) {
            std::memcpy(spectra.wavelengths, nm, sizeof spectra.wavelengths);
            have_packet = true;
        } else if (std::memcmp(spectra.wavelengths, nm, sizeof spectra.wavelengths) != 0) {
            return bf_report_error(BF_INVALID_ARGUMENT, "scene blob: %s wavelengths differ from the scene's packet", what);
        }
        return BF_OK;
    };

    /* materials and closures (SceneSoA.cpp:594-734) */
    const uint32_t M = (uint32_t)h.material_count, K = (uint32_t)h.closure_count;
    const auto present = B.column<uint8_t>(C_MATERIAL_SPECTRAL_PRESENT);
    std::vector<float> nm[4], pdf[4], emitted[4];
    std::vector<uint8_t> measure[4];
    for (uint32_t k = 0; k < 4; ++k) {
        nm[k] = B.column<float>(C_MATERIAL_WAVELENGTH_NM + k);
        pdf[k] = B.column<float>(C_MATERIAL_WAVELENGTH_PDF + k);
        measure[k] = B.column<uint8_t>(C_MATERIAL_WAVELENGTH_MEASURE + k);
        emitted[k] = B.column<float>(C_MATERIAL_EMITTED_RADIANCE + k);
    }
    const auto closure_offset = B.column<uint64_t>(C_MATERIAL_CLOSURE_OFFSET);
    const auto closure_count = B.column<uint64_t>(C_MATERIAL_CLOSURE_COUNT);
    const auto frame_mode = B.column<uint8_t>(C_MATERIAL_FRAME_MODE);
    const auto rotation = B.column<float>(C_MATERIAL_TANGENT_ROTATION);
    const auto kinds = B.column<uint32_t>(C_CLOSURE_KIND), lobes = B.column<uint32_t>(C_CLOSURE_LOBES);
    const auto probability = B.column<float>(C_CLOSURE_PROBABILITY);
    std::vector<float> weight[4], params[BF_CLOSURE_PARAMETERS];
    for (uint32_t k = 0; k < 4; ++k) weight[k] = B.column<float>(C_CLOSURE_WEIGHT + k);
    for (uint32_t p = 0; p < BF_CLOSURE_PARAMETERS; ++p) params[p] = B.column<float>(C_CLOSURE_PARAMETERS + p);

    std::vector<bf_material_desc> materials(M);
    std::vector<bf_closure_desc> closures(K);
    uint64_t crun = 0;
    for (uint32_t m = 0; m < M; ++m) {
        bf_material_desc& md = materials[m];
        md = bf_material_desc{};
        md.id = material_ids[m];
        md.emission_spectrum = BF_INVALID_INDEX;
        if (closure_offset[m] != crun || closure_count[m] > h.closure_count - crun)
            return bf_report_error(BF_INVALID_ARGUMENT, "scene blob: material %u has an inconsistent closure range", m);
        if (!present[m]) {
            if (closure_count[m] != 0)
                return bf_report_error(BF_INVALID_ARGUMENT, "scene blob: material %u has closures but no spectral data", m);
            md.closure_offset = (uint32_t)crun;
            continue;
        }
        if (closure_count[m] > BF_MAX_CLOSURES)
            return bf_report_error(BF_UNAVAILABLE, "scene blob: material %u has more than %u closures", m, BF_MAX_CLOSURES);
        const float lane_nm[4] = {nm[0][m], nm[1][m], nm[2][m], nm[3][m]};
        const float lane_pdf[4] = {pdf[0][m], pdf[1][m], pdf[2][m], pdf[3][m]};
        const uint8_t lane_measure[4] = {measure[0][m], measure[1][m], measure[2][m], measure[3][m]};
        if (const int st = admit_packet(lane_nm, lane_pdf, lane_measure, "material"); st != BF_OK) return st;
        md.closure_offset = (uint32_t)crun;
        md.closure_count = (uint32_t)closure_count[m];
        md.frame_mode = frame_mode[m];
        md.tangent_rotation = rotation[m];
        const float e[4] = {emitted[0][m], emitted[1][m], emitted[2][m], emitted[3][m]};
        if (e[0] != 0.0f || e[1] != 0.0f || e[2] != 0.0f || e[3] != 0.0f) md.emission_spectrum = spectra.add(e);
        float w[BF_MAX_CLOSURES];
        for (uint32_t j = 0; j < md.closure_count; ++j) {
            const uint64_t c = crun + j;
            bf_closure_desc& cd = closures[c];
            cd = bf_closure_desc{};
            cd.kind = kinds[c];
            cd.lobes = lobes[c];
            const float lane_weight[4] = {weight[0][c], weight[1][c], weight[2][c], weight[3][c]};
            cd.weight_spectrum = spectra.add(lane_weight);
            cd.eta_spectrum = cd.k_spectrum = BF_INVALID_INDEX;
            for (uint32_t p = 0; p < BF_CLOSURE_PARAMETERS; ++p) cd.parameters[p] = params[p][c];
            if (cd.kind == BF_CLOSURE_ROUGH_CONDUCTOR) {
                /* ClosureSet.hpp:203-209: parameters[0..3] relative eta, [4..7] relative k */
                cd.eta_spectrum = spectra.add(cd.parameters);
                cd.k_spectrum = spectra.add(cd.parameters + BF_SPECTRUM_LANES);
            }
            w[j] = probability[c];
        }
        project_probabilities(w, md.closure_count, md.probabilities, md.cdf);
        crun += md.closure_count;
    }
    if (crun != h.closure_count)
        return bf_report_error(BF_INVALID_ARGUMENT, "scene blob: material closure ranges do not cover the closure columns");

    /* instances: world transforms as serialized (SceneSoA.cpp:736-815) */
    const uint32_t I = (uint32_t)h.instance_count;
    const auto inst_object = B.column<uint32_t>(C_INSTANCE_OBJECT_ID);
    const auto inst_geometry = B.column<uint32_t>(C_INSTANCE_GEOMETRY_ID);
    const auto inst_material = B.column<uint32_t>(C_INSTANCE_MATERIAL_ID);
    const auto inst_mask = B.column<uint32_t>(C_INSTANCE_VISIBILITY_MASK);
    std::vector<float> l2w[16], w2l[16];
    for (uint32_t e = 0; e < 16; ++e) {
        l2w[e] = B.column<float>(C_INSTANCE_LOCAL_TO_WORLD + e);
        w2l[e] = B.column<float>(C_INSTANCE_WORLD_TO_LOCAL + e);
    }
    std::vector<bf_instance_desc> instances(I);
    std::vector<uint32_t> emissive_ids;
    for (uint32_t i = 0; i < I; ++i) {
        bf_instance_desc& id = instances[i];
        id = bf_instance_desc{};
        id.id = instance_ids[i];
        id.object_id = inst_object[i];
        id.visibility_mask = inst_mask[i];
        uint32_t unused = 0;
        if (!index_of(object_ids, inst_object[i], unused) || !index_of(geometry_ids, inst_geometry[i], id.geometry_index) ||
            !index_of(material_ids, inst_material[i], id.material_index))
            return bf_report_error(BF_INVALID_ARGUMENT, "scene blob: instance %u references an unknown object, geometry or material", i);
        if (!present[id.material_index])
            return bf_report_error(BF_UNAVAILABLE, "scene blob: instance %u uses a material without spectral data", i);
        if (l2w[12][i] != 0.0f || l2w[13][i] != 0.0f || l2w[14][i] != 0.0f || l2w[15][i] != 1.0f ||
            w2l[12][i] != 0.0f || w2l[13][i] != 0.0f || w2l[14][i] != 0.0f || w2l[15][i] != 1.0f)
            return bf_report_error(BF_INVALID_ARGUMENT, "scene blob: instance %u transform is not affine", i);
        for (uint32_t e = 0; e < 12; ++e) {
            id.local_to_world[e] = l2w[e][i];
            id.world_to_local[e] = w2l[e][i];
        }
        if (materials[id.material_index].emission_spectrum != BF_INVALID_INDEX) emissive_ids.push_back(id.id);
    }
    /* the serialized light registry is the emissive instances in instance order (FrameScene.cpp:201-262) */
    if (B.column<uint32_t>(C_MESH_AREA_LIGHT_INSTANCE_ID) != emissive_ids)
        return bf_report_error(BF_INVALID_ARGUMENT, "scene blob: mesh-area-light ids disagree with the emissive instances");

    /* environment (SceneSoA.cpp:880-906) */
    bf_scene_desc d{};
    if (h.environment_count == 1) {
        float env_nm[4], env_pdf[4], env_radiance[4];
        uint8_t env_measure[4];
        for (uint32_t k = 0; k < 4; ++k) {
            env_nm[k] = B.column<float>(C_ENVIRONMENT_WAVELENGTH_NM + k)[0];
            env_pdf[k] = B.column<float>(C_ENVIRONMENT_WAVELENGTH_PDF + k)[0];
            env_measure[k] = B.column<uint8_t>(C_ENVIRONMENT_WAVELENGTH_MEASURE + k)[0];
            env_radiance[k] = B.column<float>(C_ENVIRONMENT_RADIANCE + k)[0];
        }
        if (const int st = admit_packet(env_nm, env_pdf, env_measure, "environment"); st != BF_OK) return st;
        d.has_environment = 1;
        d.environment_spectrum = spectra.add(env_radiance);
    } else {
        d.environment_spectrum = BF_INVALID_INDEX;
    }

    /* punctual lights, in registry order (SceneSoA.cpp:818-878); their spectra are the path's lanes */
    const uint64_t P = h.punctual_light_count;
    if (P > 0xFFFFFFFFull) return bf_report_error(BF_RESOURCE_EXHAUSTED, "scene blob: punctual light count exceeds 32 bits");
    if (P != 0 && !have_packet)
        return bf_report_error(BF_INVALID_ARGUMENT, "scene blob: punctual lights need the scene's wavelength packet");
    std::vector<bf_punctual_desc> punctuals(P);
    {
        const auto kind = B.column<uint32_t>(C_PUNCTUAL_KIND);
        std::vector<float> pos[3], err[3], dir[3], lanes[4];
        for (uint32_t a = 0; a < 3; ++a) {
            pos[a] = B.column<float>(C_PUNCTUAL_POSITION_X + a);
            err[a] = B.column<float>(C_PUNCTUAL_POSITION_ERROR_X + a);
            dir[a] = B.column<float>(C_PUNCTUAL_DIRECTION_X + a);
        }
        for (uint32_t k = 0; k < 4; ++k) lanes[k] = B.column<float>(C_PUNCTUAL_SPECTRUM + k);
        const auto inner = B.column<float>(C_PUNCTUAL_INNER), outer = B.column<float>(C_PUNCTUAL_OUTER);
        for (uint64_t i = 0; i < P; ++i) {
            bf_punctual_desc& L = punctuals[i];
            L = bf_punctual_desc{};
            L.kind = kind[i];
            for (uint32_t a = 0; a < 3; ++a) {
                L.position[a] = pos[a][i];
                L.position_error[a] = err[a][i];
                L.direction[a] = dir[a][i];
            }
            L.inner_half_angle = inner[i];
            L.outer_half_angle = outer[i];
            const float v[4] = {lanes[0][i], lanes[1][i], lanes[2][i], lanes[3][i]};
            L.spectrum = spectra.add(v);
        }
    }

    if (spectra.values.empty()) spectra.values.push_back(0.0f); /* the builder wants a value array */
    if (spectra.descs.empty()) spectra.descs.push_back(bf_spectrum_desc{0u, 1u});
    d.geometry_count = G;
    d.instance_count = I;
    d.material_count = M;
    d.closure_count = K;
    d.spectrum_count = (uint32_t)spectra.descs.size();
    d.vertex_count = V;
    d.triangle_count = T;
    d.spectrum_value_count = spectra.values.size();
    d.positions = positions.data();
    d.normals = normals.data();
    d.uvs = uvs.data();
    d.indices = indices.data();
    d.geometries = geometries.data();
    d.instances = instances.data();
    d.materials = materials.data();
    d.closures = closures.empty() ? nullptr : closures.data();
    d.spectra = spectra.descs.data();
    d.spectrum_values = spectra.values.data();
    d.texture_count = 0;
    d.textures = nullptr;
    d.punctual_count = (uint32_t)P;
    d.punctuals = punctuals.empty() ? nullptr : punctuals.data();

    /* -- normal and bump maps over the image textures (SceneSoA.cpp image / mip / texel columns;
     * checked as load_image_texture, load_image_mip and the map loaders check them,
     * CudaWavefrontStages.cu:2876-2990, 3395-3560, once here instead of per lookup) -- */
    if (h.image_texture_count > 0xFFFFFFFFull || h.image_mip_count > 0xFFFFFFFFull)
        return bf_report_error(BF_RESOURCE_EXHAUSTED, "scene blob: image counts exceed 32 bits");
    const uint32_t IT = (uint32_t)h.image_texture_count;
    const auto image_ids = B.column<uint32_t>(C_IMAGE_TEXTURE_ID);
    const auto mip_offsets = B.column<uint64_t>(C_IMAGE_TEXTURE_MIP_OFFSET);
    const auto mip_counts = B.column<uint64_t>(C_IMAGE_TEXTURE_MIP_COUNT);
    const auto channel_counts = B.column<uint32_t>(C_IMAGE_TEXTURE_MIP_COUNT + 1);
    const auto color_spaces = B.column<uint32_t>(C_IMAGE_TEXTURE_MIP_COUNT + 2);
    const auto mip_widths = B.column<uint32_t>(C_IMAGE_MIP_WIDTH);
    const auto mip_heights = B.column<uint32_t>(C_IMAGE_MIP_WIDTH + 1);
    const auto texel_offsets = B.column<uint64_t>(C_IMAGE_MIP_TEXEL_OFFSET);
    const auto texel_counts = B.column<uint64_t>(C_IMAGE_MIP_TEXEL_COUNT);
    const auto texels = B.column<float>(C_IMAGE_TEXEL_VALUE);
    if (!strictly_increasing(image_ids)) return bf_report_error(BF_INVALID_ARGUMENT, "scene blob: image texture ids are not sorted");
    std::vector<bf_image_texture_desc> images(IT);
    uint64_t next_mip = 0, next_texel = 0;
    for (uint32_t t = 0; t < IT; ++t) {
        if (mip_counts[t] == 0 || mip_counts[t] > BF_MAX_IMAGE_MIPS || mip_offsets[t] != next_mip ||
            mip_counts[t] > h.image_mip_count - mip_offsets[t] || channel_counts[t] == 0)
            return bf_report_error(BF_INVALID_ARGUMENT, "scene blob: image texture %u has an invalid mip range", t);
        const uint64_t first = mip_offsets[t];
        for (uint64_t k = first; k < first + mip_counts[t]; ++k) {
            const uint64_t w = mip_widths[k], hh = mip_heights[k];
            const bool halves = k == first || (w == (mip_widths[k - 1] > 1 ? mip_widths[k - 1] / 2 : 1) &&
                                               hh == (mip_heights[k - 1] > 1 ? mip_heights[k - 1] / 2 : 1));
            if (w == 0 || hh == 0 || w > (1u << 24) || hh > (1u << 24) || !halves || texel_offsets[k] != next_texel ||
                texel_counts[k] != w * hh * channel_counts[t] || texel_counts[k] > h.image_texel_count - next_texel)
                return bf_report_error(BF_INVALID_ARGUMENT, "scene blob: image texture %u level %llu is malformed", t,
                                       (unsigned long long)(k - first));
            next_texel += texel_counts[k];
        }
        images[t] = bf_image_texture_desc{mip_widths[first], mip_heights[first], channel_counts[t], (uint32_t)mip_counts[t],
                                          texels.data() + texel_offsets[first], next_texel - texel_offsets[first]};
        next_mip += mip_counts[t];
    }
    if (next_mip != h.image_mip_count || next_texel != h.image_texel_count)
        return bf_report_error(BF_INVALID_ARGUMENT, "scene blob: image mips or texels are not all referenced");
    std::vector<bf_surface_maps_desc> maps(M);
    bool mapped = false;
    std::vector<std::vector<uint32_t>> ncol, bcol;
    for (uint32_t c = 0; c < 9; ++c) ncol.push_back(B.column<uint32_t>(C_NORMAL_MAP_PRESENT + 1 + c));
    for (uint32_t c = 0; c < 7; ++c) bcol.push_back(B.column<uint32_t>(C_BUMP_MAP_PRESENT + 1 + c));
    const auto normal_present = B.column<uint8_t>(C_NORMAL_MAP_PRESENT);
    const auto bump_present = B.column<uint8_t>(C_BUMP_MAP_PRESENT);
    const auto image_index = [&](uint32_t id, uint32_t& index) {
        return index_of(image_ids, id, index) && color_spaces[index] == 0u; /* data textures only */
    };
    for (uint32_t m = 0; m < M; ++m) {
        bf_surface_maps_desc& S = maps[m];
        S.normal_texture = S.bump_texture = BF_INVALID_INDEX;
        if (normal_present[m] > 1 || bump_present[m] > 1)
            return bf_report_error(BF_INVALID_ARGUMENT, "scene blob: material %u has an invalid map flag", m);
        if (!normal_present[m]) {
            for (uint32_t c = 0; c < 9; ++c)
                if (ncol[c][m]) return bf_report_error(BF_INVALID_ARGUMENT, "scene blob: material %u: absent normal map is not zero", m);
        } else {
            if (!image_index(ncol[0][m], S.normal_texture))
                return bf_report_error(BF_INVALID_ARGUMENT, "scene blob: material %u: normal map texture %u is not a data image", m,
                                       ncol[0][m]);
            for (uint32_t c = 0; c < 3; ++c) S.normal_channels[c] = ncol[1 + c][m];
            S.normal_u_wrap = ncol[4][m];
            S.normal_v_wrap = ncol[5][m];
            S.normal_max_anisotropy = ncol[6][m];
            S.normal_max_texel_visits = ncol[7][m];
            S.normal_y_convention = ncol[8][m];
            mapped = true;
        }
        if (!bump_present[m]) {
            for (uint32_t c = 0; c < 7; ++c)
                if (bcol[c][m]) return bf_report_error(BF_INVALID_ARGUMENT, "scene blob: material %u: absent bump map is not zero", m);
        } else {
            if (!image_index(bcol[0][m], S.bump_texture))
                return bf_report_error(BF_INVALID_ARGUMENT, "scene blob: material %u: bump map texture %u is not a data image", m,
                                       bcol[0][m]);
            S.bump_channel = bcol[1][m];
            std::memcpy(&S.bump_scale, &bcol[2][m], sizeof(float));
            S.bump_u_wrap = bcol[3][m];
            S.bump_v_wrap = bcol[4][m];
            S.bump_max_anisotropy = bcol[5][m];
            S.bump_max_texel_visits = bcol[6][m];
            mapped = true;
        }
    }
    d.image_texture_count = mapped ? IT : 0u;
    d.image_textures = mapped && IT ? images.data() : nullptr;
    d.surface_maps = mapped ? maps.data() : nullptr;
    const int st = blackframe_hip_scene_create(&d, out);
    if (st == BF_OK) bf_scene_set_source_hash(*out, h.content_hash);
    return st;
}

} // namespace

extern "C" int blackframe_hip_scene_soa_validate(const void* blob, size_t byte_count, uint32_t flags, uint32_t* status) {
    if (!blob || !status) return bf_report_error(BF_INVALID_ARGUMENT, "scene blob or status is NULL");
    if (flags & ~BF_SCENE_SOA_VERIFY_HASH) return bf_report_error(BF_INVALID_ARGUMENT, "unknown scene blob flags 0x%x", flags);
    bf_scene_soa_header h;
    *status = check_blob(blob, byte_count, flags, h);
    return BF_OK;
}

extern "C" int blackframe_hip_scene_create_from_soa(const void* blob, size_t byte_count, uint32_t flags, bf_hip_scene** out) {
    if (!blob || !out) return bf_report_error(BF_INVALID_ARGUMENT, "scene blob or output is NULL");
    *out = nullptr;
    if (flags & ~BF_SCENE_SOA_VERIFY_HASH) return bf_report_error(BF_INVALID_ARGUMENT, "unknown scene blob flags 0x%x", flags);
    bf_scene_soa_header h;
    const uint32_t st = check_blob(blob, byte_count, flags, h);
    if (st != BF_SCENE_SOA_VALID)
        return bf_report_error(BF_INVALID_ARGUMENT, "scene blob rejected: %s (status %u)", status_name(st), st);
    try {
        return build_from_blob(static_cast<const uint8_t*>(blob), h, flags, out);
    } catch (const std::bad_alloc&) {
        return bf_report_error(BF_RESOURCE_EXHAUSTED, "host allocation failed while reading the scene blob");
    }
}
