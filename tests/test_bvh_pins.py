"""Pins of the deterministic scene BVH, for both builders under test (the library's host builder
and the oracle's), on the CPU.

Two independent checks:
  1. bytes: tests/bvh_builder.py restates Backends/GPU/CUDA/Source/SceneBvh.cpp in Python
     (written from the reference, not from either C builder); on the reference test's scene and
     on 60 seeded scenes of 1-20 triangles (random affine instances, duplicated triangles, equal
     centroids, signed zeros) and three crafted ones (a binary64-only centroid order, three equal
     centroids, cancellation) the canonical arrays of all three agree byte for byte;
  2. the invariants of Tests/Backends/GPU/CUDA/SceneBvhTest.cpp:502-700, checked against the
     scene's own geometry with exact rational arithmetic:
       - BLAS bounds exactly cover the mesh and every leaf box is its triangle's box (:502-535);
       - TLAS leaf and root boxes contain every exactly transformed vertex (:537-576);
       - the TLAS bound survives fused and unfused cancellation (x' = 4097 x - 16785408 = 1) (:578-597);
       - three triangles with one centroid give 5 nodes and references {0, 1, 2} (:599-610);
       - the topology is acyclic, every node reachable, every leaf reference used once, BLAS
         ranges contiguous, instance references carry the scene's indices and masks (:612-680);
       - the bytes do not depend on insertion order (the FrameScene canonical form sorts every
         table by id, SceneBvhTest.cpp:682-700 with make_description(permuted)).
A planted mistake in the restated builder (a binary32 centroid, the wrong tie-break, post-order
numbering or the smallest extent) changes the bytes, which test_planted_builder_mistakes_are_caught
shows.
"""
from fractions import Fraction

import numpy as np
import pytest

import bvh_builder as ref
from soa_scenes import soa

F32 = np.float32
NODE = np.dtype([("lo", "<f4", 3), ("kind", "<u4"), ("hi", "<f4", 3), ("axis", "<u4"), ("first", "<u4"),
                 ("second", "<u4"), ("ref_offset", "<u4"), ("ref_count", "<u4")])
BLAS = np.dtype([("geometry_id", "<u4"), ("root", "<u4"), ("node_offset", "<u4"), ("node_count", "<u4"),
                 ("ref_offset", "<u4"), ("ref_count", "<u4"), ("reserved", "<u4", 2)])
IREF = np.dtype([("instance_index", "<u4"), ("instance_id", "<u4"), ("geometry_index", "<u4"), ("geometry_id", "<u4"),
                 ("blas_index", "<u4"), ("mask", "<u4"), ("reserved", "<u4", 2)])


def build_scene(bf, geometries, instances):
    """geometries: [(id, positions (n,3), triangles (m,3))] in id order; instances: [(id, geometry
    index, local_to_world 12 floats, mask)] in id order. A bf_scene_desc through the Scene builder,
    with the world matrices given exactly (binary32)."""
    abi = bf.abi
    s = bf.scenes.Scene()
    material = s.add_material(1, [bf.scenes.lambert(0.5)])
    for gid, positions, triangles in geometries:
        n = len(positions)
        s.add_geometry(gid, positions, np.tile([0.0, 0.0, 1.0], (n, 1)), np.zeros((n, 2)), triangles)
    for k, (iid, g, m12, mask) in enumerate(instances):
        rec = abi.InstanceDesc()
        rec.id, rec.object_id, rec.geometry_index, rec.material_index, rec.visibility_mask = iid, k + 1, g, material, mask
        m = np.vstack([np.asarray(m12, np.float64).reshape(3, 4), [0.0, 0.0, 0.0, 1.0]])
        inv = np.linalg.inv(m)
        for e in range(12):
            rec.local_to_world[e] = float(F32(m12[e]))
            rec.world_to_local[e] = float(F32(inv.reshape(-1)[e]))
        s.instances.append(rec)
    return s.finalize()


def reference_scene(permuted):
    """make_description(permuted) of SceneBvhTest.cpp:170-250 in the FrameScene canonical form
    (sorted tables, world = parent world * local in binary32)."""
    def mesh(gid, points, triangles):
        p = np.asarray(points, np.float32)
        return {"id": gid, "positions": p, "normals": np.tile(F32([0, 0, 1]), (len(p), 1)),
                "uvs": np.zeros((len(p), 2), np.float32), "triangles": np.asarray(triangles, np.uint32)}

    mesh_a = mesh(7, [[-4, -2, -1], [-3, -2, -1], [-4, -1, -1], [-1, 1, 0], [0, 1, 0], [-1, 2, 0], [2, -3, 1], [3, -3, 1],
                      [2, -2, 1], [5, 4, 2], [6, 4, 2], [5, 5, 2], [8, 0, 3], [9, 0, 3], [8, 1, 3]],
                  [[3 * t, 3 * t + 1, 3 * t + 2] for t in range(5)])
    mesh_b = mesh(3, [[-2, -1, -2], [1, -1, -2], [-2, 2, -2], [2, 2, 1], [4, 2, 1], [2, 3, 1]], [[0, 1, 2], [3, 4, 5]])

    def affine(rows):
        return [F32(v) for v in rows] + [F32(0), F32(0), F32(0), F32(1)]

    d = {"objects": [{"id": 300}, {"id": 100}, {"id": 200}], "geometries": [mesh_a, mesh_b],
         "materials": [{"id": 11, "spectral": None}],
         "instances": [
             {"id": 20, "parent": 10, "object": 200, "geometry": 7, "material": 11,
              "local_to_parent": affine([1, 0, 0, -4, 0, 3, 0, 2, 0, 0, 1, -1]), "visibility_mask": 0x0000FFFF},
             {"id": 30, "parent": None, "object": 300, "geometry": 7, "material": 11,
              "local_to_parent": affine([0, -1, 0, 10, 1, 0, 0, 1, 0, 0, 1, 2]), "visibility_mask": 0xFFFFFFFF},
             {"id": 10, "parent": 30, "object": 100, "geometry": 3, "material": 11,
              "local_to_parent": affine([2, 0, 0, 1, 0, 1, 0, -2, 0, 0, -1, 3]), "visibility_mask": 0xFF00FF00}]}
    if permuted:
        for k in ("objects", "geometries", "materials", "instances"):
            d[k] = list(reversed(d[k]))
    canon = soa.frame_scene(d)
    gindex = {g["id"]: k for k, g in enumerate(canon["geometries"])}
    geometries = [(g["id"], g["positions"], g["triangles"]) for g in canon["geometries"]]
    instances = [(inst["id"], gindex[inst["geometry"]], [F32(v) for v in world[0][:12]], inst["visibility_mask"])
                 for inst, world in zip(canon["instances"], canon["world_transforms"])]
    return geometries, instances


def random_scene(seed):
    rng = np.random.default_rng(seed)
    geometries = []
    for g in range(int(rng.integers(1, 4))):
        n_tri = int(rng.integers(1, 8))
        positions = rng.uniform(-3, 3, (3 * n_tri, 3)).astype(np.float32)
        if seed % 3 == 0:
            positions = np.round(positions * 2) / 2  # shared coordinates, equal centroids, signed zeros
            positions[rng.random(positions.shape) < 0.1] = -0.0
        triangles = np.arange(3 * n_tri).reshape(n_tri, 3)
        if n_tri > 2 and seed % 4 == 1:
            triangles[1] = triangles[0]  # a duplicated triangle: equal centroids, tie broken by index
        if n_tri > 3 and seed % 5 == 2:
            triangles[2:] = triangles[2:][::-1]
        geometries.append((10 * (g + 1), positions.astype(np.float32), triangles))
    instances = []
    for i in range(int(rng.integers(1, 6))):
        q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
        linear = q @ np.diag(rng.uniform(0.3, 2.5, 3))
        if seed % 3 == 0 and i == 0:
            linear = np.eye(3)
        m12 = np.hstack([linear, rng.uniform(-5, 5, (3, 1))]).reshape(-1).astype(np.float32)
        instances.append((100 + 7 * i, int(rng.integers(0, len(geometries))), list(m12), int(rng.integers(0, 2 ** 32))))
    return geometries, instances


def product_bvh(bf, hip, geometries, instances):
    return hip.XpuScene(build_scene(bf, geometries, instances), upload=False).bvh()


def oracle_bvh(bf, oracle, geometries, instances):
    return oracle.scene(build_scene(bf, geometries, instances)).bvh()


KEYS = ("counts", "blas", "blas_nodes", "primitive_refs", "tlas_nodes", "instance_refs")


@pytest.mark.parametrize("seed", list(range(60)) + ["precision", "triple", "cancellation"])
def test_builders_equal_the_restated_reference_builder(bf, hip, oracle, seed):
    geometries, instances = random_scene(seed) if isinstance(seed, int) else invariant_scene(seed)
    want = ref.scene_bvh(geometries, instances)
    got = product_bvh(bf, hip, geometries, instances)
    mine = oracle_bvh(bf, oracle, geometries, instances)
    for key in KEYS:
        assert got[key] == want[key], (seed, key)
        assert mine[key] == want[key], (seed, key)
    assert got["tlas_root"] == want["tlas_root"] == 0


@pytest.mark.parametrize("permuted", [False, True])
def test_reference_test_scene_bytes(bf, hip, oracle, permuted):
    geometries, instances = reference_scene(permuted)
    want = ref.scene_bvh(geometries, instances)
    got = product_bvh(bf, hip, geometries, instances)
    assert oracle_bvh(bf, oracle, geometries, instances)["tlas_nodes"] == want["tlas_nodes"]
    for key in KEYS:
        assert got[key] == want[key], key
    other = product_bvh(bf, hip, *reference_scene(not permuted))
    for key in KEYS:  # CanonicalBytesIgnoreInsertionAndAllocationIdentity
        assert other[key] == got[key], key


def test_planted_builder_mistakes_are_caught(bf, hip):
    """The byte comparison is sensitive to each plausible builder mistake: every variant of the
    restated builder below changes the bytes of at least one of the seeded scenes."""
    original_centroid, original_build = ref.centroid, ref.build

    def f32_centroid(box):
        return [float(F32(F32(box[a]) + F32(box[3 + a])) * F32(0.5)) for a in range(3)]

    def variant(mistake):
        def build(items, nodes, emit, begin, end):
            index = len(nodes)
            nodes.append(None)
            if end - begin == 1:
                nodes[index] = ref.node(items[begin]["box"], 1, ref_offset=emit(items[begin]), ref_count=1)
                return index
            lo = [min(it["centroid"][a] for it in items[begin:end]) for a in range(3)]
            hi = [max(it["centroid"][a] for it in items[begin:end]) for a in range(3)]
            ext = [hi[a] - lo[a] for a in range(3)]
            axis = int(np.argmin(ext)) if mistake == "smallest" else int(np.argmax(ext))
            key = (lambda it: (it["centroid"][axis], -it["key"])) if mistake == "tie" else \
                (lambda it: (it["centroid"][axis], it["key"]))
            items[begin:end] = sorted(items[begin:end], key=key)
            middle = begin + (end - begin) // 2
            if mistake == "postorder":
                nodes.pop()
                first = build(items, nodes, emit, begin, middle)
                second = build(items, nodes, emit, middle, end)
                nodes.append(None)
                index = len(nodes) - 1
            else:
                first = build(items, nodes, emit, begin, middle)
                second = build(items, nodes, emit, middle, end)
            nodes[index] = ref.node(ref.unite(nodes[first]["box"], nodes[second]["box"]), 0, axis, first, second)
            return index
        return build

    scenes = [random_scene(seed) for seed in range(24)] + [invariant_scene("precision")]
    products = [product_bvh(bf, hip, *sc) for sc in scenes]
    try:
        for mistake in ("f32-centroid", "tie", "postorder", "smallest"):
            ref.centroid = f32_centroid if mistake == "f32-centroid" else original_centroid
            ref.build = variant(mistake) if mistake != "f32-centroid" else original_build
            differs = any(any(ref.scene_bvh(*sc)[k] != p[k] for k in KEYS) for sc, p in zip(scenes, products))
            assert differs, mistake
    finally:
        ref.centroid, ref.build = original_centroid, original_build


def parse(bvh):
    return (np.frombuffer(bvh["blas"], BLAS), np.frombuffer(bvh["blas_nodes"], NODE),
            np.frombuffer(bvh["primitive_refs"], "<u4"), np.frombuffer(bvh["tlas_nodes"], NODE),
            np.frombuffer(bvh["instance_refs"], IREF))


def triangle_box(positions, tri):
    p = positions[np.asarray(tri)]
    return np.concatenate([p.min(axis=0), p.max(axis=0)])


def node_box(n):
    return np.concatenate([n["lo"], n["hi"]])


def exact_transform(m12, point, row):
    acc = Fraction(float(F32(m12[row * 4 + 3])))
    for col in range(3):
        acc += Fraction(float(F32(m12[row * 4 + col]))) * Fraction(float(F32(point[col])))
    return acc


def validate_tree(nodes, offset, count, root, ref_offset, ref_count):
    """Acyclic, every node of [offset, offset + count) reachable from root, every reference of the
    range used by exactly one leaf, internal boxes the exact union of their children."""
    state = {}
    used = np.zeros(ref_count, int)
    stack = [root]
    while stack:
        i = stack.pop()
        assert offset <= i < offset + count and i not in state, "cycle or out-of-range child"
        state[i] = True
        n = nodes[i]
        if n["kind"] == 1:
            assert n["ref_count"] == 1 and n["first"] == ref.INVALID and n["second"] == ref.INVALID
            assert n["axis"] == ref.INVALID
            used[n["ref_offset"] - ref_offset] += 1
        else:
            assert n["kind"] == 0 and n["axis"] < 3 and n["ref_count"] == 0
            a, b = nodes[n["first"]], nodes[n["second"]]
            assert np.array_equal(n["lo"], np.minimum(a["lo"], b["lo"])) and np.array_equal(n["hi"], np.maximum(a["hi"], b["hi"]))
            stack += [n["second"], n["first"]]
    assert len(state) == count
    assert np.all(used == 1)


INVARIANT_SCENES = ["reference", "reference-permuted", "triple", "cancellation", "precision"] + [f"random-{s}" for s in range(0, 60, 7)]


def invariant_scene(name):
    if name.startswith("reference"):
        return reference_scene(name.endswith("permuted"))
    if name == "triple":  # make_single_triangle_mesh(3), identity
        pos = np.array([[4097, 0, 0], [4097, 1, 0], [4097, 0, 1]], np.float32)
        return [(2, pos, np.array([[0, 1, 2]] * 3))], [(4, 0, [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0], 0xFFFFFFFF)]
    if name == "precision":
        # centroids 1 + 2^-24 and 1 differ in binary64 only (binary32 rounds the first to 1): the
        # builder's binary64 centroid orders triangle 1 first, a binary32 one would tie on the key
        e = float(np.float32(2.0 ** -23))
        pos = np.array([[1, 0, 0], [1 + e, 1, 0], [1, 0, 1], [1, 0, 0], [1, 1, 0], [1, 0, 1]], np.float32)
        return [(2, pos, np.array([[0, 1, 2], [3, 4, 5]]))], [(4, 0, [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0], 0xFFFFFFFF)]
    if name == "cancellation":  # TlasBoundsCoverFusedAndUnfusedCancellation
        pos = np.array([[4097, 0, 0], [4097, 1, 0], [4097, 0, 1]], np.float32)
        return [(2, pos, np.array([[0, 1, 2]]))], [(4, 0, [4097, 0, 0, -16785408, 0, 1, 0, 0, 0, 0, 1, 0], 0xFFFFFFFF)]
    return random_scene(int(name.split("-")[1]))


@pytest.mark.parametrize("name", INVARIANT_SCENES)
@pytest.mark.parametrize("builder", ["product", "oracle"])
def test_reference_bvh_invariants(bf, hip, oracle, builder, name):
    geometries, instances = invariant_scene(name)
    bvh = product_bvh(bf, hip, geometries, instances) if builder == "product" else oracle_bvh(bf, oracle, geometries, instances)
    blases, blas_nodes, prims, tlas_nodes, irefs = parse(bvh)
    assert len(blases) == len(geometries)
    next_node = next_ref = 0
    for g, ((gid, positions, triangles), b) in enumerate(zip(geometries, blases)):
        # BinaryTopologyIsAcyclicAndReferencesEveryLeafExactlyOnce: contiguous ranges
        assert (b["geometry_id"], b["node_offset"], b["ref_offset"], b["ref_count"]) == (gid, next_node, next_ref, len(triangles))
        assert b["node_count"] == 2 * len(triangles) - 1 and b["root"] == b["node_offset"] and not b["reserved"].any()
        validate_tree(blas_nodes, b["node_offset"], b["node_count"], b["root"], b["ref_offset"], b["ref_count"])
        # BlasBoundsExactlyCoverSourceTriangles
        used = np.unique(np.asarray(triangles).reshape(-1))
        mesh_box = np.concatenate([positions[used].min(axis=0), positions[used].max(axis=0)])
        assert np.array_equal(node_box(blas_nodes[b["root"]]), mesh_box)
        leaves = 0
        for k in range(b["node_offset"], b["node_offset"] + b["node_count"]):
            n = blas_nodes[k]
            if n["kind"] == 1:
                leaves += 1
                prim = prims[n["ref_offset"]]
                assert prim < len(triangles)
                assert np.array_equal(node_box(n), triangle_box(positions, triangles[prim]))
        assert leaves == len(triangles)
        next_node += b["node_count"]
        next_ref += b["ref_count"]
    assert next_node == len(blas_nodes) and next_ref == len(prims)
    # TLAS: acyclic, root 0, instance references carry the scene's indices and masks once each
    assert len(tlas_nodes) == 2 * len(instances) - 1 and bvh["tlas_root"] == 0
    validate_tree(tlas_nodes, 0, len(tlas_nodes), 0, 0, len(irefs))
    assert sorted(irefs["instance_id"].tolist()) == sorted(i[0] for i in instances)
    for r in irefs:
        iid, g, m12, mask = instances[r["instance_index"]]
        assert (r["instance_id"], r["geometry_index"], r["blas_index"], r["mask"]) == (iid, g, g, mask)
        assert r["geometry_id"] == geometries[g][0] and not r["reserved"].any()
    # TlasBoundsConservativelyCoverHierarchicalInstances: exact rational containment
    root_box = node_box(tlas_nodes[0])
    for n in tlas_nodes:
        if n["kind"] != 1:
            continue
        r = irefs[n["ref_offset"]]
        _, g, m12, _ = instances[r["instance_index"]]
        positions, triangles = geometries[g][1], geometries[g][2]
        for v in np.unique(np.asarray(triangles).reshape(-1)):
            for row in range(3):
                exact = exact_transform(m12, positions[v], row)
                assert Fraction(float(n["lo"][row])) <= exact <= Fraction(float(n["hi"][row]))
                assert Fraction(float(root_box[row])) <= exact <= Fraction(float(root_box[3 + row]))
    if name == "triple":  # EqualCentroidsUsePrimitiveIdentifiersAsTieBreaks
        assert blases[0]["node_count"] == 5 and prims.tolist() == [0, 1, 2]
    if name == "precision":
        assert prims.tolist() == [1, 0] and blas_nodes[0]["axis"] == 0
    if name == "cancellation":
        assert len(tlas_nodes) == 1 and float(tlas_nodes[0]["lo"][0]) <= 0.0 and float(tlas_nodes[0]["hi"][0]) >= 1.0
        for v in range(3):
            assert exact_transform(instances[0][2], geometries[0][1][v], 0) == 1


@pytest.mark.parametrize("seed", list(range(0, 60, 5)) + ["reference", "reference-permuted", "precision"])
def test_serialized_blob_equals_the_restated_serializer(bf, hip, seed):
    """blackframe_hip_scene_bvh_serialize against the restated layout of SceneBvh.cpp:882-1071
    (header fields, aligned array offsets, array order, FNV-1a-64 with the hash field zeroed) over
    the independently built arrays: the whole blob, byte for byte."""
    geometries, instances = random_scene(seed) if isinstance(seed, int) else invariant_scene(seed)
    blob = hip.XpuScene(build_scene(bf, geometries, instances), upload=False).bvh_blob()
    assert blob == ref.serialize(ref.scene_bvh(geometries, instances), source_hash=0)
    assert hip.scene_bvh_validate(blob) == ref.VALID


def test_blob_bytes_ignore_insertion_order(bf, hip):
    """CanonicalBytesIgnoreInsertionAndAllocationIdentity (SceneBvhTest.cpp:682-700)."""
    a = hip.XpuScene(build_scene(bf, *reference_scene(False)), upload=False).bvh_blob()
    b = hip.XpuScene(build_scene(bf, *reference_scene(True)), upload=False).bvh_blob()
    assert a == b and ref.fnv1a64_normalized(a) == int.from_bytes(a[32:40], "little")


def bvh_header_cases():
    """The reference's EncodesEmptySceneAndRejectsBudgetOrAbiMismatch checks (SceneBvhTest.cpp:702-731)
    plus one case per remaining status reachable with 32-bit counts (size_overflow needs byte
    counts beyond 2^64, which counts below 2^32 cannot reach)."""
    empty = ref.header([0, 0, 0, 0, 0])
    cases = [(empty, ref.VALID)]
    h = dict(empty, counts=[0, 0, 0, 0, 1])
    cases.append((h, ref.INVALID_TOPOLOGY_COUNTS))
    cases.append((dict(empty, tlas_root_node=0), ref.INVALID_TLAS_ROOT))
    full = ref.header([2, 12, 7, 5, 3])
    cases += [(full, ref.VALID), (dict(full, magic=0), ref.BAD_MAGIC), (dict(full, abi_minor=1), ref.INCOMPATIBLE_VERSION),
              (dict(full, header_size=255), ref.BAD_HEADER_SIZE), (dict(full, array_count=4), ref.BAD_ARRAY_COUNT),
              (dict(full, hash_algorithm=2), ref.UNSUPPORTED_HASH_ALGORITHM),
              (dict(full, reserved=[0, 0, 1, 0, 0]), ref.NONZERO_RESERVED), (dict(full, leaf_capacity=2), ref.BAD_LEAF_CAPACITY),
              (dict(full, counts=[2, 13, 7, 5, 3]), ref.INVALID_TOPOLOGY_COUNTS),
              (dict(full, tlas_root_node=3), ref.INVALID_TLAS_ROOT),
              (dict(full, total_size_bytes=100), ref.BAD_TOTAL_SIZE),
              (dict(full, total_size_bytes=full["total_size_bytes"] + 16), ref.BAD_TOTAL_SIZE)]
    shifted = dict(full, arrays=[list(d) for d in full["arrays"]])
    shifted["arrays"][2][0] += 16
    cases.append((shifted, ref.INVALID_ARRAY_DESCRIPTOR))
    resized = dict(full, arrays=[list(d) for d in full["arrays"]])
    resized["arrays"][0][2] = 16
    cases.append((resized, ref.INVALID_ARRAY_DESCRIPTOR))
    return cases


@pytest.mark.parametrize("case", range(17))
def test_bvh_header_validation_order(hip, case):
    h, expected = bvh_header_cases()[case]
    assert ref.validate_header(h) == expected
    blob = ref.pack_header(h)
    status = hip.scene_bvh_validate(blob + bytes(max(0, h["total_size_bytes"] - len(blob))), verify_hash=False)
    assert status == expected, (case, status, expected)


def test_bvh_blob_integrity_statuses(bf, hip, abi):
    blob = hip.XpuScene(build_scene(bf, *reference_scene(False)), upload=False).bvh_blob()
    assert hip.scene_bvh_validate(blob) == abi.BVH_VALID
    assert hip.scene_bvh_validate(blob[:-1]) == abi.BVH_TRUNCATED
    assert hip.scene_bvh_validate(blob[:255]) == abi.BVH_TRUNCATED
    flipped = bytearray(blob)
    flipped[300] ^= 1
    assert hip.scene_bvh_validate(bytes(flipped)) == abi.BVH_HASH_MISMATCH
    assert hip.scene_bvh_validate(bytes(flipped), verify_hash=False) == abi.BVH_VALID


@pytest.mark.parametrize("shape", ["one large geometry", "many geometries"])
def test_parallel_build_writes_the_sequential_bytes(bf, hip, oracle, shape):
    """The library builds BLASes concurrently (one geometry per thread) and splits large ranges onto
    two threads down to 16 384 items, into node and reference ranges fixed in advance; the oracle
    builds sequentially in pre-order. A 60 000-triangle mesh (split over several levels, with
    shared coordinates so that centroid ties cross the splits) and 40 instanced geometries of up to
    3 000 triangles give identical bytes."""
    rng = np.random.default_rng(11)
    if shape == "one large geometry":
        n = 60000
        positions = (np.round(rng.uniform(-20, 20, (3 * n, 3)) * 4) / 4).astype(np.float32)
        geometries = [(10, positions, np.arange(3 * n).reshape(n, 3))]
        instances = [(100, 0, [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0], 0xFFFFFFFF),
                     (101, 0, [0, 0, 1, 5, 0, 1, 0, -2, -1, 0, 0, 1], 0xFFFFFFFF)]
    else:
        geometries = []
        for g in range(40):
            n = int(rng.integers(1, 3000))
            geometries.append((10 * (g + 1), rng.uniform(-3, 3, (3 * n, 3)).astype(np.float32), np.arange(3 * n).reshape(n, 3)))
        instances = []
        for i in range(300):
            m = np.hstack([np.eye(3) * rng.uniform(0.5, 2.0), rng.uniform(-50, 50, (3, 1))]).reshape(-1)
            instances.append((1000 + i, int(rng.integers(0, 40)), [float(x) for x in m.astype(np.float32)], 0xFFFFFFFF))
    got = product_bvh(bf, hip, geometries, instances)
    want = oracle_bvh(bf, oracle, geometries, instances)
    for key in KEYS + ("tlas_root",):
        assert got[key] == want[key], key
