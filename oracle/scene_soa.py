"""ORACLE — test infrastructure only: a numpy restatement of the reference's scene-blob serializer.

Only tests/ import this module; the product never serializes through it.

It restates, for the "BLKFSOA4" scene blob (ABI 4.0):
  * the frozen column schema and header validation (XPU/Shared/.../SceneSoaAbi.hpp:11-175,
    179-351): per-column element count domain, element size, 16-byte column alignment;
  * the serializer (Backends/GPU/CUDA/Source/SceneSoA.cpp:333-1178): column order and the value
    each column takes from the canonical frame scene, then FNV-1a-64 over the blob with the
    content-hash field zeroed (SceneSoA.cpp:226-236);
  * the canonical frame scene it reads (Engine/Source/FrameScene.cpp:37-39,201-262,413-522):
    constant textures, image textures, objects, geometries, materials and instances sorted by
    identifier; punctual lights kept in registry order; mesh-area-light instance ids = instances
    (in sorted order) whose material has non-black emitted radiance;
  * instance transforms (Engine/Source/SceneHierarchyValidation.cpp:42-140): world = parent
    world * local (Renderer/.../MatrixTypes.hpp:41-52, float32, k-ordered accumulation from +0)
    and every inverse by float32 Gauss-Jordan with partial pivoting
    (Renderer/.../MatrixOperations.hpp:31-85);
  * wavelength sampling (Renderer/.../WavelengthSampling.hpp:62-84) and closure packing
    (Renderer/.../ClosureSet.hpp:150-214) for building test scenes.

Pinned by the reference's frozen rich-scene hash (Tests/Backends/GPU/CUDA/SceneSoATest.cpp:41,
1052-1075), which covers every non-image column, signed zeros in positions and matrices, the
parent chain, the Gauss-Jordan inverses and the 8-bit columns.
"""

import struct

import numpy as np

F32 = np.float32

MAGIC = 0x34414F53464B4C42  # "BLKFSOA4"
ABI_MAJOR, ABI_MINOR = 4, 0
HASH_FNV1A64 = 1
HEADER_SIZE = 4848
HASH_OFFSET = 32
COLUMN_COUNT = 194
ALIGN = 16
FNV_OFFSET, FNV_PRIME = 0xCBF29CE484222325, 0x100000001B3

# Column index of each group's first column (SceneSoaAbi.hpp:24-132).
OBJECT_ID = 0
GEOMETRY_ID, GEOMETRY_VERTEX_OFFSET, GEOMETRY_VERTEX_COUNT = 1, 2, 3
GEOMETRY_TRIANGLE_OFFSET, GEOMETRY_TRIANGLE_COUNT = 4, 5
POSITION_X, NORMAL_X, TEXCOORD_X, TRIANGLE_VERTEX_0 = 6, 9, 12, 14
MATERIAL_ID, MATERIAL_SPECTRAL_PRESENT = 17, 18
MATERIAL_WAVELENGTH_NM, MATERIAL_WAVELENGTH_PDF, MATERIAL_WAVELENGTH_MEASURE = 19, 23, 27
MATERIAL_CLOSURE_OFFSET, MATERIAL_CLOSURE_COUNT, MATERIAL_FRAME_MODE = 31, 32, 33
MATERIAL_TANGENT_ROTATION, MATERIAL_EMITTED_RADIANCE = 34, 35
CLOSURE_KIND, CLOSURE_LOBES, CLOSURE_WEIGHT, CLOSURE_PARAMETERS, CLOSURE_PROBABILITY = 39, 40, 41, 45, 55
INSTANCE_ID, INSTANCE_PARENT_PRESENT, INSTANCE_PARENT_ID, INSTANCE_OBJECT_ID = 56, 57, 58, 59
INSTANCE_GEOMETRY_ID, INSTANCE_MATERIAL_ID, INSTANCE_VISIBILITY_MASK = 60, 61, 62
INSTANCE_LOCAL_TO_PARENT, INSTANCE_PARENT_TO_LOCAL = 63, 79
INSTANCE_LOCAL_TO_WORLD, INSTANCE_WORLD_TO_LOCAL = 95, 111
PUNCTUAL_KIND, PUNCTUAL_POSITION_X, PUNCTUAL_POSITION_ERROR_X, PUNCTUAL_DIRECTION_X = 127, 128, 131, 134
PUNCTUAL_INNER, PUNCTUAL_OUTER, PUNCTUAL_SPECTRUM = 137, 138, 139
MESH_AREA_LIGHT_INSTANCE_ID = 143
ENVIRONMENT_WAVELENGTH_NM, ENVIRONMENT_WAVELENGTH_PDF = 144, 148
ENVIRONMENT_WAVELENGTH_MEASURE, ENVIRONMENT_RADIANCE = 152, 156
TEXTURE_ID, TEXTURE_KIND, TEXTURE_VALUE = 160, 161, 162
NORMAL_MAP_PRESENT, BUMP_MAP_PRESENT, BUMP_MAP_SCALE = 166, 176, 179
IMAGE_TEXTURE_ID, IMAGE_TEXTURE_MIP_OFFSET, IMAGE_TEXTURE_MIP_COUNT = 184, 185, 186
IMAGE_MIP_WIDTH, IMAGE_MIP_TEXEL_OFFSET, IMAGE_MIP_TEXEL_COUNT, IMAGE_TEXEL_VALUE = 189, 191, 192, 193

# Header count fields in header order (offset 40 + 8 * position).
COUNT_FIELDS = ("object", "geometry", "vertex", "triangle", "material", "closure", "instance",
                "punctual_light", "mesh_area_light", "environment", "texture", "image_texture",
                "image_mip", "image_texel")

# Header validation statuses (SceneSoaAbi.hpp:177-189).
VALID, BAD_MAGIC, INCOMPATIBLE_VERSION, BAD_HEADER_SIZE, BAD_COLUMN_COUNT = 0, 1, 2, 3, 4
UNSUPPORTED_HASH_ALGORITHM, NONZERO_RESERVED, INVALID_ENVIRONMENT_COUNT = 5, 6, 7
INVALID_COLUMN_DESCRIPTOR, SIZE_OVERFLOW, BAD_TOTAL_SIZE = 8, 9, 10

PUNCTUAL_POINT, PUNCTUAL_DIRECTIONAL, PUNCTUAL_SPOT = 0, 1, 2
TEXTURE_FLOAT, TEXTURE_COLOR, TEXTURE_SPECTRUM = 0, 1, 2
MEASURE_WAVELENGTH = 5  # TransportConventions.hpp:11-18


def column_domain(c):
    """Which header count sizes column c (SceneSoaAbi.hpp:193-240)."""
    if c == OBJECT_ID:
        return "object"
    if GEOMETRY_ID <= c <= GEOMETRY_TRIANGLE_COUNT:
        return "geometry"
    if POSITION_X <= c < TRIANGLE_VERTEX_0:
        return "vertex"
    if TRIANGLE_VERTEX_0 <= c < MATERIAL_ID:
        return "triangle"
    if MATERIAL_ID <= c < CLOSURE_KIND:
        return "material"
    if CLOSURE_KIND <= c < INSTANCE_ID:
        return "closure"
    if INSTANCE_ID <= c < PUNCTUAL_KIND:
        return "instance"
    if PUNCTUAL_KIND <= c < MESH_AREA_LIGHT_INSTANCE_ID:
        return "punctual_light"
    if c == MESH_AREA_LIGHT_INSTANCE_ID:
        return "mesh_area_light"
    if ENVIRONMENT_WAVELENGTH_NM <= c < TEXTURE_ID:
        return "environment"
    if TEXTURE_ID <= c < NORMAL_MAP_PRESENT:
        return "texture"
    if NORMAL_MAP_PRESENT <= c < IMAGE_TEXTURE_ID:
        return "material"
    if IMAGE_TEXTURE_ID <= c < IMAGE_MIP_WIDTH:
        return "image_texture"
    if IMAGE_MIP_WIDTH <= c < IMAGE_TEXEL_VALUE:
        return "image_mip"
    if c == IMAGE_TEXEL_VALUE:
        return "image_texel"
    return None


def column_dtype(c):
    """Element type of column c (SceneSoaAbi.hpp:242-276)."""
    if GEOMETRY_VERTEX_OFFSET <= c <= GEOMETRY_TRIANGLE_COUNT or c in (MATERIAL_CLOSURE_OFFSET, MATERIAL_CLOSURE_COUNT):
        return np.uint64
    if (c in (MATERIAL_SPECTRAL_PRESENT, MATERIAL_FRAME_MODE, INSTANCE_PARENT_PRESENT, NORMAL_MAP_PRESENT,
              BUMP_MAP_PRESENT) or MATERIAL_WAVELENGTH_MEASURE <= c < MATERIAL_CLOSURE_OFFSET
            or ENVIRONMENT_WAVELENGTH_MEASURE <= c < ENVIRONMENT_RADIANCE):
        return np.uint8
    if (POSITION_X <= c < TRIANGLE_VERTEX_0 or MATERIAL_WAVELENGTH_NM <= c < MATERIAL_WAVELENGTH_MEASURE
            or c == MATERIAL_TANGENT_ROTATION or MATERIAL_EMITTED_RADIANCE <= c < CLOSURE_KIND
            or CLOSURE_WEIGHT <= c < INSTANCE_ID or INSTANCE_LOCAL_TO_PARENT <= c < PUNCTUAL_KIND
            or PUNCTUAL_POSITION_X <= c < MESH_AREA_LIGHT_INSTANCE_ID
            or ENVIRONMENT_WAVELENGTH_NM <= c < ENVIRONMENT_WAVELENGTH_MEASURE
            or ENVIRONMENT_RADIANCE <= c < TEXTURE_ID or TEXTURE_VALUE <= c < NORMAL_MAP_PRESENT
            or c in (BUMP_MAP_SCALE, IMAGE_TEXEL_VALUE)):
        return np.float32
    if c in (IMAGE_TEXTURE_MIP_OFFSET, IMAGE_TEXTURE_MIP_COUNT, IMAGE_MIP_TEXEL_OFFSET, IMAGE_MIP_TEXEL_COUNT):
        return np.uint64
    return np.uint32


def align_up(v):
    return v if v % ALIGN == 0 else v + (ALIGN - v % ALIGN)


def fnv1a64_normalized(blob):
    """FNV-1a-64 with the content-hash bytes read as zero (SceneSoA.cpp:226-236)."""
    data = bytearray(blob)
    data[HASH_OFFSET:HASH_OFFSET + 8] = b"\0" * 8
    h = FNV_OFFSET
    for b in data:
        h = ((h ^ b) * FNV_PRIME) & 0xFFFFFFFFFFFFFFFF
    return h


# ---- header ---------------------------------------------------------------------------------

HEADER_HEAD = struct.Struct("<QHHIIIQQ")        # magic .. content_hash: 40 bytes
HEADER_COUNTS = struct.Struct("<14Q")           # 112 bytes
DESCRIPTOR = struct.Struct("<QQII")             # 24 bytes
RESERVED = struct.Struct("<5Q")


def pack_header(h):
    out = bytearray(HEADER_HEAD.pack(h["magic"], h["abi_major"], h["abi_minor"], h["header_size"],
                                     h["column_count"], h["hash_algorithm"], h["total_size_bytes"],
                                     h["content_hash"]))
    out += HEADER_COUNTS.pack(*[h["counts"][k] for k in COUNT_FIELDS])
    for d in h["columns"]:
        out += DESCRIPTOR.pack(*d)
    out += RESERVED.pack(*h["reserved"])
    assert len(out) == HEADER_SIZE
    return bytes(out)


def unpack_header(blob):
    magic, major, minor, hsize, ccount, halg, total, chash = HEADER_HEAD.unpack_from(blob, 0)
    counts = dict(zip(COUNT_FIELDS, HEADER_COUNTS.unpack_from(blob, 40)))
    columns = [DESCRIPTOR.unpack_from(blob, 152 + 24 * c) for c in range(COLUMN_COUNT)]
    return {"magic": magic, "abi_major": major, "abi_minor": minor, "header_size": hsize,
            "column_count": ccount, "hash_algorithm": halg, "total_size_bytes": total,
            "content_hash": chash, "counts": counts, "columns": columns,
            "reserved": list(RESERVED.unpack_from(blob, 4808))}


def validate_header(h):
    """validate_scene_soa_header (SceneSoaAbi.hpp:283-341), in the same check order."""
    if h["magic"] != MAGIC:
        return BAD_MAGIC
    if h["abi_major"] != ABI_MAJOR or h["abi_minor"] != ABI_MINOR:
        return INCOMPATIBLE_VERSION
    if h["header_size"] != HEADER_SIZE:
        return BAD_HEADER_SIZE
    if h["column_count"] != COLUMN_COUNT:
        return BAD_COLUMN_COUNT
    if h["hash_algorithm"] != HASH_FNV1A64:
        return UNSUPPORTED_HASH_ALGORITHM
    if any(h["reserved"]):
        return NONZERO_RESERVED
    if h["counts"]["environment"] > 1:
        return INVALID_ENVIRONMENT_COUNT
    if h["total_size_bytes"] < HEADER_SIZE:
        return BAD_TOTAL_SIZE
    cursor = align_up(HEADER_SIZE)
    top = (1 << 64) - 1
    for c in range(COLUMN_COUNT):
        off, count, size, reserved = h["columns"][c]
        expected = h["counts"][column_domain(c)]
        esize = np.dtype(column_dtype(c)).itemsize
        if count != expected or size != esize or reserved != 0:
            return INVALID_COLUMN_DESCRIPTOR
        if expected == 0:
            if off != 0:
                return INVALID_COLUMN_DESCRIPTOR
            continue
        if cursor > top - (ALIGN - 1):
            return SIZE_OVERFLOW
        cursor = align_up(cursor)
        if off != cursor:
            return INVALID_COLUMN_DESCRIPTOR
        if expected > top // esize:
            return SIZE_OVERFLOW
        nbytes = expected * esize
        if cursor > top - nbytes:
            return SIZE_OVERFLOW
        cursor += nbytes
    if h["total_size_bytes"] != cursor:
        return BAD_TOTAL_SIZE
    return VALID


def read_column(blob, h, c):
    off, count, size, _ = h["columns"][c]
    dt = column_dtype(c)
    if count == 0:
        return np.zeros(0, dt)
    return np.frombuffer(blob, dtype=dt, count=count, offset=off).copy()


# ---- float32 matrix arithmetic ------------------------------------------------------------------

def mat_mul(a, b):
    """Matrix4 operator* (MatrixTypes.hpp:41-52): result(r, c) += a(r, k) * b(k, c), k = 0..3."""
    out = [F32(0.0)] * 16
    for r in range(4):
        for c in range(4):
            acc = F32(0.0)
            for k in range(4):
                acc = F32(acc + F32(a[r * 4 + k] * b[k * 4 + c]))
            out[r * 4 + c] = acc
    return out


def mat_inverse(m):
    """Gauss-Jordan with partial pivoting on a 4x8 augmented float32 matrix
    (MatrixOperations.hpp:31-85); returns None for a singular or non-finite matrix."""
    m = [F32(v) for v in m]
    if not all(np.isfinite(v) for v in m):
        return None
    aug = [F32(0.0)] * 32
    for r in range(4):
        for c in range(4):
            aug[r * 8 + c] = m[r * 4 + c]
            aug[r * 8 + c + 4] = F32(1.0) if r == c else F32(0.0)
    with np.errstate(all="ignore"):
        for p in range(4):
            prow, pmag = p, abs(aug[p * 8 + p])
            for cand in range(p + 1, 4):
                mag = abs(aug[cand * 8 + p])
                if mag > pmag:
                    pmag, prow = mag, cand
            if pmag == 0:
                return None
            if prow != p:
                for c in range(8):
                    aug[p * 8 + c], aug[prow * 8 + c] = aug[prow * 8 + c], aug[p * 8 + c]
            pivot = aug[p * 8 + p]
            for c in range(8):
                aug[p * 8 + c] = F32(aug[p * 8 + c] / pivot)
            for r in range(4):
                if r == p:
                    continue
                factor = aug[r * 8 + p]
                for c in range(8):
                    aug[r * 8 + c] = F32(aug[r * 8 + c] - F32(factor * aug[p * 8 + c]))
    out = [aug[r * 8 + c + 4] for r in range(4) for c in range(4)]
    if not all(np.isfinite(v) for v in out):
        return None
    return out


def affine_ok(m):
    return m[12] == 0 and m[13] == 0 and m[14] == 0 and m[15] == 1


# ---- scene building helpers ------------------------------------------------------------------

def uniform_visible_wavelengths(u):
    """sample_uniform_visible_wavelengths (WavelengthSampling.hpp:62-84): [(nm, pdf, measure)] * 4."""
    u = F32(u)
    pdf = F32(F32(1.0) / F32(470.0))
    out = []
    for lane in range(4):
        offset = F32(F32(lane) / F32(4.0))
        wrap = F32(F32(1.0) - offset)
        shifted = F32(u - wrap) if (lane != 0 and u >= wrap) else F32(u + offset)
        nm = F32(470.0 * float(shifted) + 360.0)  # fma: the double product and sum are exact
        out.append((nm, pdf, MEASURE_WAVELENGTH))
    return out


def lambertian(reflectance):
    """ClosureSet::append_lambertian_reflection (ClosureSet.hpp:145-156)."""
    return {"kind": 1, "lobes": 0x1 | 0x8, "weight": [F32(v) for v in reflectance], "parameters": [F32(0.0)] * 10}


def rough_conductor(coefficient, eta, k, alpha_x, alpha_y):
    """ClosureSet::append_rough_conductor_reflection (ClosureSet.hpp:195-216)."""
    params = [F32(v) for v in eta] + [F32(v) for v in k] + [F32(alpha_x), F32(alpha_y)]
    return {"kind": 3, "lobes": 0x2 | 0x8, "weight": [F32(v) for v in coefficient], "parameters": params}


# ---- canonical frame scene ---------------------------------------------------------------------

def frame_scene(desc):
    """FrameScene::create's canonical form (FrameScene.cpp:413-522) of a description dict with keys
    constant_textures, image_textures, objects, geometries, materials, instances, punctual_lights,
    environment. Raises ValueError where the reference rejects the scene."""
    by_id = lambda r: r["id"]
    s = {k: sorted(desc.get(k, []), key=by_id)
         for k in ("constant_textures", "image_textures", "objects", "geometries", "materials", "instances")}
    for k, v in s.items():
        ids = [r["id"] for r in v]
        if len(set(ids)) != len(ids):
            raise ValueError("duplicate %s identifier" % k)
    s["punctual_lights"] = list(desc.get("punctual_lights", []))
    s["environment"] = desc.get("environment")
    index = {k: {r["id"]: i for i, r in enumerate(s[k])} for k in ("objects", "geometries", "materials", "instances")}
    for inst in s["instances"]:
        for key, dom in (("object", "objects"), ("geometry", "geometries"), ("material", "materials")):
            if inst[key] not in index[dom]:
                raise ValueError("instance references an unknown %s" % key)
        if inst.get("parent") is not None and inst["parent"] not in index["instances"]:
            raise ValueError("unknown parent instance")
    # SceneHierarchyValidation.cpp:42-140: locals, then worlds along each parent chain.
    locals_, worlds = [], [None] * len(s["instances"])
    for inst in s["instances"]:
        m = [F32(v) for v in inst["local_to_parent"]]
        inv = mat_inverse(m) if affine_ok(m) else None
        if inv is None:
            raise ValueError("local transform must be finite, affine and invertible")
        locals_.append((m, inv))
    state = [0] * len(s["instances"])
    for start in range(len(s["instances"])):
        if state[start] == 2:
            continue
        chain, cur = [], start
        while state[cur] != 2:
            if state[cur] == 1:
                raise ValueError("instance hierarchy contains a cycle")
            state[cur] = 1
            chain.append(cur)
            parent = s["instances"][cur].get("parent")
            if parent is None:
                break
            cur = index["instances"][parent]
        while chain:
            i = chain.pop()
            parent = s["instances"][i].get("parent")
            if parent is None:
                worlds[i] = locals_[i]
            else:
                m = mat_mul(worlds[index["instances"][parent]][0], locals_[i][0])
                inv = mat_inverse(m) if affine_ok(m) else None
                if inv is None:
                    raise ValueError("hierarchy produced an invalid world transform")
                worlds[i] = (m, inv)
            state[i] = 2
    s["local_transforms"], s["world_transforms"] = locals_, worlds
    # FrameScene.cpp:201-262: emissive instances in sorted instance order.
    s["mesh_area_light_instance_ids"] = [
        inst["id"] for inst in s["instances"]
        if (mat := s["materials"][index["materials"][inst["material"]]].get("spectral")) is not None
        and any(F32(v) != 0 for v in mat.get("emission", [0, 0, 0, 0]))]
    return s


# ---- serializer ------------------------------------------------------------------------------

def serialize(scene):
    """serialize_scene (SceneSoA.cpp:333-1178) over a canonical scene from frame_scene()."""
    mats, geos, insts = scene["materials"], scene["geometries"], scene["instances"]
    lights, env, texs, imgs = scene["punctual_lights"], scene["environment"], scene["constant_textures"], scene["image_textures"]
    spectral = [m.get("spectral") for m in mats]
    closures = [c for sp in spectral if sp is not None for c in sp["closures"]]
    mips = [lvl for t in imgs for lvl in t["levels"]]
    counts = {
        "object": len(scene["objects"]), "geometry": len(geos),
        "vertex": sum(len(g["positions"]) for g in geos), "triangle": sum(len(g["triangles"]) for g in geos),
        "material": len(mats), "closure": len(closures), "instance": len(insts),
        "punctual_light": len(lights), "mesh_area_light": len(scene["mesh_area_light_instance_ids"]),
        "environment": 1 if env is not None else 0, "texture": len(texs), "image_texture": len(imgs),
        "image_mip": len(mips), "image_texel": sum(len(l["pixels"]) for l in mips),
    }
    cols = [None] * COLUMN_COUNT

    def put(c, values):
        cols[c] = np.asarray(values, dtype=column_dtype(c)).reshape(-1)

    put(OBJECT_ID, [o["id"] for o in scene["objects"]])
    put(GEOMETRY_ID, [g["id"] for g in geos])
    voff = np.cumsum([0] + [len(g["positions"]) for g in geos])[:-1]
    toff = np.cumsum([0] + [len(g["triangles"]) for g in geos])[:-1]
    put(GEOMETRY_VERTEX_OFFSET, voff)
    put(GEOMETRY_VERTEX_COUNT, [len(g["positions"]) for g in geos])
    put(GEOMETRY_TRIANGLE_OFFSET, toff)
    put(GEOMETRY_TRIANGLE_COUNT, [len(g["triangles"]) for g in geos])
    cat = lambda key, width: (np.concatenate([np.asarray(g[key], np.float32).reshape(-1, width) for g in geos])
                              if geos else np.zeros((0, width), np.float32))
    pos, nrm, uv = cat("positions", 3), cat("normals", 3), cat("uvs", 2)
    tri = (np.concatenate([np.asarray(g["triangles"], np.uint32).reshape(-1, 3) for g in geos])
           if geos else np.zeros((0, 3), np.uint32))
    for a in range(3):
        put(POSITION_X + a, pos[:, a])
    for a in range(3):
        put(NORMAL_X + a, nrm[:, a])
    for a in range(2):
        put(TEXCOORD_X + a, uv[:, a])
    for a in range(3):
        put(TRIANGLE_VERTEX_0 + a, tri[:, a])
    put(MATERIAL_ID, [m["id"] for m in mats])
    put(MATERIAL_SPECTRAL_PRESENT, [1 if sp is not None else 0 for sp in spectral])
    for lane in range(4):
        put(MATERIAL_WAVELENGTH_NM + lane, [sp["wavelengths"][lane][0] if sp else 0.0 for sp in spectral])
    for lane in range(4):
        put(MATERIAL_WAVELENGTH_PDF + lane, [sp["wavelengths"][lane][1] if sp else 0.0 for sp in spectral])
    for lane in range(4):
        put(MATERIAL_WAVELENGTH_MEASURE + lane, [sp["wavelengths"][lane][2] if sp else 0 for sp in spectral])
    offsets, running = [], 0
    for sp in spectral:
        offsets.append(running)
        running += len(sp["closures"]) if sp else 0
    put(MATERIAL_CLOSURE_OFFSET, offsets)
    put(MATERIAL_CLOSURE_COUNT, [len(sp["closures"]) if sp else 0 for sp in spectral])
    put(MATERIAL_FRAME_MODE, [sp["frame_mode"] if sp else 0 for sp in spectral])
    put(MATERIAL_TANGENT_ROTATION, [sp["tangent_rotation"] if sp else 0.0 for sp in spectral])
    for lane in range(4):
        put(MATERIAL_EMITTED_RADIANCE + lane,
            [sp.get("emission", [0.0] * 4)[lane] if sp else 0.0 for sp in spectral])
    put(CLOSURE_KIND, [c["kind"] for c in closures])
    put(CLOSURE_LOBES, [c["lobes"] for c in closures])
    for lane in range(4):
        put(CLOSURE_WEIGHT + lane, [c["weight"][lane] for c in closures])
    for p in range(10):
        put(CLOSURE_PARAMETERS + p, [c["parameters"][p] for c in closures])
    put(CLOSURE_PROBABILITY, [p for sp in spectral if sp is not None for p in sp["probabilities"]])
    put(INSTANCE_ID, [i["id"] for i in insts])
    put(INSTANCE_PARENT_PRESENT, [0 if i.get("parent") is None else 1 for i in insts])
    put(INSTANCE_PARENT_ID, [0 if i.get("parent") is None else i["parent"] for i in insts])
    put(INSTANCE_OBJECT_ID, [i["object"] for i in insts])
    put(INSTANCE_GEOMETRY_ID, [i["geometry"] for i in insts])
    put(INSTANCE_MATERIAL_ID, [i["material"] for i in insts])
    put(INSTANCE_VISIBILITY_MASK, [i["visibility_mask"] for i in insts])
    for e in range(16):
        put(INSTANCE_LOCAL_TO_PARENT + e, [F32(i["local_to_parent"][e]) for i in insts])
    for e in range(16):
        put(INSTANCE_PARENT_TO_LOCAL + e, [t[1][e] for t in scene["local_transforms"]])
    for e in range(16):
        put(INSTANCE_LOCAL_TO_WORLD + e, [t[0][e] for t in scene["world_transforms"]])
    for e in range(16):
        put(INSTANCE_WORLD_TO_LOCAL + e, [t[1][e] for t in scene["world_transforms"]])
    kinds = {"point": PUNCTUAL_POINT, "directional": PUNCTUAL_DIRECTIONAL, "spot": PUNCTUAL_SPOT}
    put(PUNCTUAL_KIND, [kinds[l["kind"]] for l in lights])
    zero3 = [0.0, 0.0, 0.0]
    for a in range(3):
        put(PUNCTUAL_POSITION_X + a, [l.get("position", zero3)[a] if l["kind"] != "directional" else 0.0 for l in lights])
    for a in range(3):
        put(PUNCTUAL_POSITION_ERROR_X + a,
            [l.get("position_error", zero3)[a] if l["kind"] != "directional" else 0.0 for l in lights])
    for a in range(3):
        put(PUNCTUAL_DIRECTION_X + a, [l["direction"][a] if l["kind"] != "point" else 0.0 for l in lights])
    put(PUNCTUAL_INNER, [l["inner"] if l["kind"] == "spot" else 0.0 for l in lights])
    put(PUNCTUAL_OUTER, [l["outer"] if l["kind"] == "spot" else 0.0 for l in lights])
    for lane in range(4):
        put(PUNCTUAL_SPECTRUM + lane, [l["spectrum"][lane] for l in lights])
    put(MESH_AREA_LIGHT_INSTANCE_ID, scene["mesh_area_light_instance_ids"])
    envs = [env] if env is not None else []
    for lane in range(4):
        put(ENVIRONMENT_WAVELENGTH_NM + lane, [e["wavelengths"][lane][0] for e in envs])
    for lane in range(4):
        put(ENVIRONMENT_WAVELENGTH_PDF + lane, [e["wavelengths"][lane][1] for e in envs])
    for lane in range(4):
        put(ENVIRONMENT_WAVELENGTH_MEASURE + lane, [e["wavelengths"][lane][2] for e in envs])
    for lane in range(4):
        put(ENVIRONMENT_RADIANCE + lane, [e["radiance"][lane] for e in envs])
    put(TEXTURE_ID, [t["id"] for t in texs])
    put(TEXTURE_KIND, [t["kind"] for t in texs])

    def tex_component(t, i):  # SceneSoA.cpp:306-330
        v = list(t["value"]) if t["kind"] != TEXTURE_FLOAT else [t["value"]]
        return v[i] if i < len(v) else 0.0
    for i in range(4):
        put(TEXTURE_VALUE + i, [tex_component(t, i) for t in texs])
    nmap = [sp.get("normal_map") if sp else None for sp in spectral]
    bmap = [sp.get("bump_map") if sp else None for sp in spectral]
    put(NORMAL_MAP_PRESENT, [0 if b is None else 1 for b in nmap])
    for j, key in enumerate(("texture", "red_channel", "green_channel", "blue_channel", "u_wrap", "v_wrap",
                             "maximum_anisotropy", "maximum_texel_visits", "y_convention")):
        put(NORMAL_MAP_PRESENT + 1 + j, [0 if b is None else b[key] for b in nmap])
    put(BUMP_MAP_PRESENT, [0 if b is None else 1 for b in bmap])
    for j, key in enumerate(("texture", "channel", "scale", "u_wrap", "v_wrap", "maximum_anisotropy",
                             "maximum_texel_visits")):
        put(BUMP_MAP_PRESENT + 1 + j, [(0.0 if key == "scale" else 0) if b is None else b[key] for b in bmap])
    put(IMAGE_TEXTURE_ID, [t["id"] for t in imgs])
    mip_offsets, running = [], 0
    for t in imgs:
        mip_offsets.append(running)
        running += len(t["levels"])
    put(IMAGE_TEXTURE_MIP_OFFSET, mip_offsets)
    put(IMAGE_TEXTURE_MIP_COUNT, [len(t["levels"]) for t in imgs])
    put(IMAGE_TEXTURE_MIP_OFFSET + 2, [t["levels"][0]["channels"] for t in imgs])
    put(IMAGE_TEXTURE_MIP_OFFSET + 3, [t["storage_color_space"] for t in imgs])
    put(IMAGE_MIP_WIDTH, [l["width"] for l in mips])
    put(IMAGE_MIP_WIDTH + 1, [l["height"] for l in mips])
    texel_offsets, running = [], 0
    for l in mips:
        texel_offsets.append(running)
        running += len(l["pixels"])
    put(IMAGE_MIP_TEXEL_OFFSET, texel_offsets)
    put(IMAGE_MIP_TEXEL_COUNT, [len(l["pixels"]) for l in mips])
    put(IMAGE_TEXEL_VALUE, [v for l in mips for v in l["pixels"]])

    descriptors, body, cursor = [], bytearray(b"\0" * HEADER_SIZE), align_up(HEADER_SIZE)
    for c in range(COLUMN_COUNT):
        col = cols[c]
        expected = counts[column_domain(c)]
        if col is None or len(col) != expected:
            raise AssertionError("column %d has %s elements, schema wants %d" % (c, None if col is None else len(col), expected))
        esize = np.dtype(column_dtype(c)).itemsize
        if expected == 0:
            descriptors.append((0, 0, esize, 0))
            continue
        cursor = align_up(cursor)
        body += b"\0" * (cursor - len(body))
        descriptors.append((cursor, expected, esize, 0))
        body += col.tobytes()
        cursor += expected * esize
    header = {"magic": MAGIC, "abi_major": ABI_MAJOR, "abi_minor": ABI_MINOR, "header_size": HEADER_SIZE,
              "column_count": COLUMN_COUNT, "hash_algorithm": HASH_FNV1A64, "total_size_bytes": len(body),
              "content_hash": 0, "counts": counts, "columns": descriptors, "reserved": [0] * 5}
    assert validate_header(header) == VALID
    body[:HEADER_SIZE] = pack_header(header)
    header["content_hash"] = fnv1a64_normalized(body)
    body[:HEADER_SIZE] = pack_header(header)
    return bytes(body)
