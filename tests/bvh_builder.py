"""An independent restatement, in Python, of the reference's deterministic scene BVH builder, used
only to pin the product's and the oracle's builders on small scenes.

It follows Backends/GPU/CUDA/Source/SceneBvh.cpp and nothing in this repository's C/C++:
  - primitive bounds: the first vertex, then min/max with the other two, every bound through
    canonical_zero (:323-353);
  - centroid: (double(min) + double(max)) * 0.5 per axis (:133-139);
  - build_node (:355-427): pre-order node numbering (a node takes its index before its
    children), leaf capacity 1 (a leaf emits its reference when it is built), split axis = the
    largest double centroid extent (the first axis on ties), items ordered by (centroid on that
    axis, stable key), split at begin + (end - begin) / 2, internal bounds the canonical-zero
    union of the two children (:99-106);
  - BLAS per geometry in geometry order, stable key = the primitive index (:477-521);
  - TLAS over instances in instance order, stable key = the instance id, instance bounds the union
    of the conservative point bounds of every transformed vertex of every triangle (:555-601,
    :603-660): per row the binary64 sum of the exact products, widened by
    fma(gamma7, magnitude, 7 * denorm_min) and rounded outward to binary32.
Binary64 fma is evaluated exactly with fractions (math.fma needs Python 3.13); Python floats are
binary64, numpy.float32 is binary32.
"""
from fractions import Fraction

import numpy as np

F32 = np.float32
INVALID = 0xFFFFFFFF
EPS32 = float(np.finfo(np.float32).eps)
GAMMA7 = (7.0 * EPS32) / (1.0 - 7.0 * EPS32)
UNDERFLOW = 7.0 * float(np.float32(2.0 ** -149))
F32_MAX = float(np.finfo(np.float32).max)


def canonical_zero(v):
    v = F32(v)
    return F32(0.0) if v == 0 else v


def cmin(a, b):
    """std::min(a, b): b if b < a else a, then canonical_zero."""
    return canonical_zero(b if b < a else a)


def cmax(a, b):
    """std::max(a, b): b if a < b else a, then canonical_zero."""
    return canonical_zero(b if a < b else a)


def fma64(a, b, c):
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def point_bounds(matrix, point):
    """Conservative binary32 bounds of matrix (3x4 row-major, binary32) applied to point."""
    lo, hi = [], []
    for row in range(3):
        exact = float(F32(matrix[row * 4 + 3]))
        magnitude = abs(exact)
        for col in range(3):
            product = float(F32(matrix[row * 4 + col])) * float(F32(point[col]))
            exact += product
            magnitude += abs(product)
        error = fma64(GAMMA7, magnitude, UNDERFLOW)
        lower, upper = exact - error, exact + error
        if not (np.isfinite(lower) and np.isfinite(upper)) or lower < -F32_MAX or upper > F32_MAX:
            raise ValueError("point bounds are not representable")
        rl, ru = F32(lower), F32(upper)
        if float(rl) > lower:
            rl = np.nextafter(rl, F32(-np.inf))
        if float(ru) < upper:
            ru = np.nextafter(ru, F32(np.inf))
        lo.append(canonical_zero(rl))
        hi.append(canonical_zero(ru))
    return lo + hi


def triangle_bounds(v0, v1, v2):
    box = [canonical_zero(v0[a]) for a in range(3)] * 2
    for v in (v1, v2):
        for a in range(3):
            box[a] = cmin(box[a], F32(v[a]))
            box[3 + a] = cmax(box[3 + a], F32(v[a]))
    return box


def unite(a, b):
    return [cmin(a[k], b[k]) for k in range(3)] + [cmax(a[3 + k], b[3 + k]) for k in range(3)]


def centroid(box):
    return [(float(box[a]) + float(box[3 + a])) * 0.5 for a in range(3)]


def node(box, kind, axis=INVALID, first=INVALID, second=INVALID, ref_offset=0, ref_count=0):
    return {"box": [F32(v) for v in box], "kind": kind, "axis": axis, "first": first, "second": second,
            "ref_offset": ref_offset, "ref_count": ref_count}


def build(items, nodes, emit, begin, end):
    """items: list of dicts {box, centroid, key, payload}; returns the node index."""
    index = len(nodes)
    nodes.append(None)
    if end - begin == 1:
        offset = emit(items[begin])
        nodes[index] = node(items[begin]["box"], 1, ref_offset=offset, ref_count=1)
        return index
    lo = list(items[begin]["centroid"])
    hi = list(items[begin]["centroid"])
    for it in items[begin + 1:end]:
        for a in range(3):
            lo[a] = min(lo[a], it["centroid"][a])
            hi[a] = max(hi[a], it["centroid"][a])
    axis, largest = 0, hi[0] - lo[0]
    for a in (1, 2):
        if hi[a] - lo[a] > largest:
            axis, largest = a, hi[a] - lo[a]
    items[begin:end] = sorted(items[begin:end], key=lambda it: (it["centroid"][axis], it["key"]))
    middle = begin + (end - begin) // 2
    first = build(items, nodes, emit, begin, middle)
    second = build(items, nodes, emit, middle, end)
    nodes[index] = node(unite(nodes[first]["box"], nodes[second]["box"]), 0, axis, first, second)
    return index


def scene_bvh(geometries, instances):
    """geometries: list of (id, positions (n, 3) float32, triangles (m, 3) int) in geometry order;
    instances: list of (id, geometry index, local_to_world 12 floats, visibility mask) in instance
    order. Returns the canonical arrays as bytes, keyed like the product's export."""
    blases, blas_nodes, prim_refs = [], [], []
    for gid, positions, triangles in geometries:
        items = []
        for p, tri in enumerate(triangles):
            box = triangle_bounds(*(positions[int(i)] for i in tri))
            items.append({"box": box, "centroid": centroid(box), "key": p})
        node_offset, ref_offset = len(blas_nodes), len(prim_refs)

        def emit(it):
            prim_refs.append(it["key"])
            return len(prim_refs) - 1

        root = build(items, blas_nodes, emit, 0, len(items))
        blases.append((gid, root, node_offset, len(blas_nodes) - node_offset, ref_offset, len(triangles)))
    tlas_nodes, inst_refs = [], []
    items = []
    for i, (iid, g, matrix, mask) in enumerate(instances):
        gid, positions, triangles = geometries[g]
        box = None
        for tri in triangles:
            for v in tri:
                pb = point_bounds(matrix, positions[int(v)])
                box = pb if box is None else unite(box, pb)
        items.append({"box": box, "centroid": centroid(box), "key": iid,
                      "payload": (i, iid, g, gid, g, mask)})

    def emit_instance(it):
        inst_refs.append(it["payload"])
        return len(inst_refs) - 1

    root = build(items, tlas_nodes, emit_instance, 0, len(items)) if items else INVALID
    node_dtype = np.dtype([("lo", "<f4", 3), ("kind", "<u4"), ("hi", "<f4", 3), ("axis", "<u4"), ("first", "<u4"),
                           ("second", "<u4"), ("ref_offset", "<u4"), ("ref_count", "<u4")])

    def pack_nodes(ns):
        out = np.zeros(len(ns), node_dtype)
        for k, n in enumerate(ns):
            out[k] = (n["box"][:3], n["kind"], n["box"][3:], n["axis"], n["first"], n["second"], n["ref_offset"],
                      n["ref_count"])
        return out.tobytes()

    blas_bytes = np.array([list(b) + [0, 0] for b in blases], dtype="<u4").tobytes()
    inst_bytes = np.array([list(r) + [0, 0] for r in inst_refs], dtype="<u4").tobytes()
    return {"counts": [len(blases), len(blas_nodes), len(prim_refs), len(tlas_nodes), len(inst_refs)],
            "blas": blas_bytes, "blas_nodes": pack_nodes(blas_nodes),
            "primitive_refs": np.array(prim_refs, dtype="<u4").tobytes(), "tlas_nodes": pack_nodes(tlas_nodes),
            "instance_refs": inst_bytes, "tlas_root": root}


# ---- the serialized blob (SceneBvhAbi.hpp:11-229, SceneBvh.cpp:882-1071) ---------------------
MAGIC = 0x31485642464B4C42
HEADER_SIZE = 256
ELEMENT_SIZES = (32, 48, 4, 48, 32)  # BLAS, BLAS node, primitive reference, TLAS node, instance reference
ARRAY_KEYS = ("blas", "blas_nodes", "primitive_refs", "tlas_nodes", "instance_refs")
(VALID, BAD_MAGIC, INCOMPATIBLE_VERSION, BAD_HEADER_SIZE, BAD_ARRAY_COUNT, UNSUPPORTED_HASH_ALGORITHM, NONZERO_RESERVED,
 BAD_LEAF_CAPACITY, INVALID_TOPOLOGY_COUNTS, INVALID_TLAS_ROOT, INVALID_ARRAY_DESCRIPTOR, SIZE_OVERFLOW,
 BAD_TOTAL_SIZE) = range(13)
TOP = (1 << 64) - 1


def align16(v):
    return (v + 15) // 16 * 16


def header(counts, source_hash=0):
    """A header dict laid out as prepare_header_layout does (SceneBvh.cpp:882-925)."""
    h = {"magic": MAGIC, "abi_major": 1, "abi_minor": 0, "header_size": HEADER_SIZE, "array_count": 5,
         "hash_algorithm": 1, "total_size_bytes": 0, "content_hash": 0, "source_scene_hash": source_hash,
         "counts": list(counts), "tlas_root_node": INVALID if counts[3] == 0 else 0, "leaf_capacity": 1,
         "arrays": [], "reserved": [0] * 5}
    cursor = align16(HEADER_SIZE)
    for a in range(5):
        if counts[a] == 0:
            h["arrays"].append([0, 0, ELEMENT_SIZES[a], 0])
            continue
        cursor = align16(cursor)
        h["arrays"].append([cursor, counts[a], ELEMENT_SIZES[a], 0])
        cursor += counts[a] * ELEMENT_SIZES[a]
    h["total_size_bytes"] = cursor
    return h


def pack_header(h):
    import struct
    out = struct.pack("<QHHIIIQQQ", h["magic"], h["abi_major"], h["abi_minor"], h["header_size"], h["array_count"],
                      h["hash_algorithm"], h["total_size_bytes"], h["content_hash"], h["source_scene_hash"])
    out += struct.pack("<5QII", *h["counts"], h["tlas_root_node"], h["leaf_capacity"])
    for d in h["arrays"]:
        out += struct.pack("<QQII", *d)
    out += struct.pack("<5Q", *h["reserved"])
    assert len(out) == HEADER_SIZE
    return out


def fnv1a64_normalized(blob):
    h = 0xCBF29CE484222325
    for i, byte in enumerate(blob):
        h = ((h ^ (0 if 32 <= i < 40 else byte)) * 0x100000001B3) & TOP
    return h


def serialize(bvh, source_hash=0):
    """The blob of the canonical arrays (as scene_bvh returns them)."""
    h = header(bvh["counts"], source_hash)
    blob = bytearray(h["total_size_bytes"])
    for a, key in enumerate(ARRAY_KEYS):
        if h["counts"][a]:
            off = h["arrays"][a][0]
            blob[off:off + len(bvh[key])] = bvh[key]
    blob[:HEADER_SIZE] = pack_header(h)
    h["content_hash"] = fnv1a64_normalized(blob)
    blob[:HEADER_SIZE] = pack_header(h)
    return bytes(blob)


def validate_header(h):
    """validate_scene_bvh_header (SceneBvhAbi.hpp:146-229) in its check order."""
    if h["magic"] != MAGIC:
        return BAD_MAGIC
    if h["abi_major"] != 1 or h["abi_minor"] != 0:
        return INCOMPATIBLE_VERSION
    if h["header_size"] != HEADER_SIZE:
        return BAD_HEADER_SIZE
    if h["array_count"] != 5:
        return BAD_ARRAY_COUNT
    if h["hash_algorithm"] != 1:
        return UNSUPPORTED_HASH_ALGORITHM
    if any(h["reserved"]):
        return NONZERO_RESERVED
    if h["leaf_capacity"] != 1:
        return BAD_LEAF_CAPACITY
    blas, blas_nodes, prims, tlas_nodes, irefs = h["counts"]
    if (max(h["counts"]) > INVALID or (blas == 0 and (blas_nodes or prims)) or
            (blas and (prims < blas or blas_nodes != prims * 2 - blas)) or (irefs == 0 and tlas_nodes) or
            (irefs and tlas_nodes != irefs * 2 - 1)):
        return INVALID_TOPOLOGY_COUNTS
    if (tlas_nodes == 0 and h["tlas_root_node"] != INVALID) or (tlas_nodes and h["tlas_root_node"] != 0):
        return INVALID_TLAS_ROOT
    if h["total_size_bytes"] < HEADER_SIZE:
        return BAD_TOTAL_SIZE
    cursor = align16(HEADER_SIZE)
    for a in range(5):
        off, count, size, reserved = h["arrays"][a]
        if count != h["counts"][a] or size != ELEMENT_SIZES[a] or reserved:
            return INVALID_ARRAY_DESCRIPTOR
        if count == 0:
            if off:
                return INVALID_ARRAY_DESCRIPTOR
            continue
        if cursor > TOP - 15:
            return SIZE_OVERFLOW
        cursor = align16(cursor)
        if off != cursor:
            return INVALID_ARRAY_DESCRIPTOR
        if count > TOP // size:
            return SIZE_OVERFLOW
        if cursor > TOP - count * size:
            return SIZE_OVERFLOW
        cursor += count * size
    return VALID if h["total_size_bytes"] == cursor else BAD_TOTAL_SIZE
