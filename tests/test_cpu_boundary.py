"""CPU checks of the drop-in boundary: the C-ABI library loads and exports every entry point
declared in include/blackframe_hip.h, record sizes match the reference's, the host BVH builder
reproduces the oracle's canonical BVH byte for byte, and the 64-byte wide-node relayout is a
bijection onto the canonical internal nodes with the children's exact boxes."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
WIDE_LEAF = 0x80000000


def declared_entry_points():
    names = []
    for header in sorted((ROOT / "include").glob("*.h")):
        names += re.findall(r"^\w[\w\s\*]*?\b(blackframe_hip_\w+|blackframe_query_extension)\s*\(", header.read_text(),
                            flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_entry_point(hip):
    lib = hip.library()
    names = declared_entry_points()
    assert len(names) >= 14 and "blackframe_query_extension" in names
    raw = C.CDLL(str(hip.LIB_PATH))
    missing = [n for n in names if not hasattr(raw, n)]
    assert not missing, missing
    assert lib.blackframe_hip_last_error() is not None


def test_record_sizes(abi):
    sizes = {abi.SampleStreamIndex: 24, abi.TransportRay: 48, abi.ClosestHitResult: 80, abi.OcclusionResult: 16,
             abi.BvhNode: 48, abi.BvhBlas: 32, abi.BvhInstanceRef: 32, abi.ClosureRecord: 64, abi.StageAudit: 64}
    for record, size in sizes.items():
        assert C.sizeof(record) == size, record.__name__


def scenes(bf):
    yield "cornell", bf.scenes.cornell()[0]
    yield "cornell-mixed", bf.scenes.cornell(materials="mixed")[0]
    yield "grid4", bf.scenes.icosphere_grid(grid=4, subdivisions=2)[0]


@pytest.mark.parametrize("index", range(3))
def test_product_bvh_is_the_oracle_bvh(bf, oracle, hip, index):
    name, scene = list(scenes(bf))[index]
    expected = oracle.scene(scene).bvh()
    got = hip.XpuScene(scene, upload=False).bvh()
    for key in expected:
        assert got[key] == expected[key], (name, key)
    # closed-form node counts of a leaf-size-1 binary BVH (SceneBvhAbi.hpp:181-193)
    blases, blas_nodes, prims, tlas_nodes, instance_refs = got["counts"]
    assert blas_nodes == 2 * prims - blases and tlas_nodes == 2 * instance_refs - 1


def nodes_array(raw, count):
    return np.frombuffer(raw, dtype=np.uint32).reshape(count, 12)


@pytest.mark.parametrize("index", range(3))
def test_wide_nodes_are_a_relayout_of_the_canonical_bvh(bf, hip, index):
    _, scene = list(scenes(bf))[index]
    xs = hip.XpuScene(scene, upload=False)
    bvh = xs.bvh()
    blas_wide, tlas_wide = xs.wide_nodes()
    for wide, raw, count in ((blas_wide, bvh["blas_nodes"], bvh["counts"][1]),
                             (tlas_wide, bvh["tlas_nodes"], bvh["counts"][3])):
        canonical = nodes_array(raw, count)
        floats = canonical.view(np.float32)
        internal = np.flatnonzero(canonical[:, 3] == 0)
        assert sorted(wide[:, 14].tolist()) == internal.tolist()  # bijection onto internal nodes
        ref_of = {int(c): w for w, c in enumerate(wide[:, 14])}
        for w, row in enumerate(wide):
            node = canonical[row[14]]
            assert row[15] == node[7]  # split axis
            for k, child in enumerate((node[8], node[9])):
                box = floats[child][[0, 1, 2, 4, 5, 6]]
                got = row.view(np.float32)[[0 + k, 2 + k, 4 + k, 6 + k, 8 + k, 10 + k]]
                assert np.array_equal(got.view(np.uint32), box.view(np.uint32))
                if canonical[child][3] == 0:
                    assert row[12 + k] == ref_of[int(child)]
                else:
                    assert row[12 + k] == (WIDE_LEAF | int(canonical[child][10]))


def near_tie(a, b):
    """Both hit and their parameters lie within 4 ulps: a coplanar pair (a box resting on the floor)
    that the traversal resolves by visit order, since its slab test prunes against the exact plane
    distance while the triangle test returns a rounded parameter (CudaSceneClosestHit.cu)."""
    if a.status != 1 or b.status != 1:
        return False
    ta, tb = np.float32(a.hit.parameter), np.float32(b.hit.parameter)
    return abs(int(ta.view(np.int32)) - int(tb.view(np.int32))) <= 4


@pytest.mark.parametrize("seed,grazing", [(1, 0.0), (2, 0.2)])
def test_oracle_traversal_equals_brute_force(bf, abi, oracle, seed, grazing):
    """The canonical near-first traversal finds the candidate a linear scan of every triangle of
    every instance selects by (t, instance id, primitive id), up to near-coplanar ties."""
    from rays import random_rays

    cases = [(bf.scenes.cornell(materials="mixed")[0], (0.0, 0.0, 1.0), 3.0),
             (bf.scenes.icosphere_grid(grid=3, subdivisions=1)[0], (0.0, 0.0, 0.0), 6.0)]
    for scene, center, radius in cases:
        o = oracle.scene(scene)
        rays = random_rays(abi, 3000, seed, center=center, radius=radius, grazing=grazing)
        got, want = o.closest_hit(rays), o.closest_hit_brute_force(rays)
        ties = 0
        for a, b in zip(got, want):
            if bytes(a) != bytes(b):
                assert near_tie(a, b)
                ties += 1
        assert ties <= 10
        assert sum(r.status == 1 for r in got) > 300


def test_material_probabilities_on_the_sampler_grid(bf):
    probs = bf.scenes.project_probabilities([0.7, 0.3])
    assert probs.sum() == 1.0 and np.all(probs * (1 << 24) == np.round(probs * (1 << 24)))
    probs = bf.scenes.project_probabilities([1.0, 1.0, 1.0])
    assert list(probs * (1 << 24)) == [5592406.0, 5592405.0, 5592405.0]


def test_extension_negotiation(abi, hip):
    """ExtensionAbi.h / XpuBackendAbi.h negotiation rules: struct sizes, ABI major, interface ids."""
    lib = hip.library()
    ext = abi.ExtensionApi()
    ext.struct_size = C.sizeof(abi.ExtensionApi)
    assert lib.blackframe_query_extension(C.byref(abi.ExtensionQuery(8, 1, 0, 0)), C.byref(ext)) == 1
    assert lib.blackframe_query_extension(C.byref(abi.ExtensionQuery(16, 2, 0, 0)), C.byref(ext)) == 2
    assert lib.blackframe_query_extension(C.byref(abi.ExtensionQuery(16, 1, 0, 0)), C.byref(ext)) == 0
    assert (ext.abi_major, ext.struct_size) == (1, 40)
    backend = abi.XpuBackendApi()
    bad = abi.InterfaceId(1, 2)
    good = abi.InterfaceId(*abi.XPU_BACKEND_INTERFACE_ID)
    assert ext.get_interface(None, bad, 1, 0, C.addressof(backend), C.sizeof(backend)) == 3
    assert ext.get_interface(None, good, 2, 0, C.addressof(backend), C.sizeof(backend)) == 2
    assert ext.get_interface(None, good, 1, 0, C.addressof(backend), 8) == 1
    assert ext.get_interface(None, good, 1, 0, C.addressof(backend), C.sizeof(backend)) == 0
    assert backend.backend_name.text().startswith("Blackframe HIP")
    desc = abi.DeviceDescriptor()
    desc.struct_size = C.sizeof(desc)
    assert backend.get_device_descriptor(None, 1 << 20, C.byref(desc)) == 5
    desc.struct_size = 4
    assert backend.get_device_descriptor(None, 0, C.byref(desc)) == 1
