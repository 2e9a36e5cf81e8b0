"""The "BLKFSOA4" scene blob: the oracle serializer pinned by the reference's frozen hash, and the
header validation order (Tests/Backends/GPU/CUDA/SceneSoATest.cpp)."""

import struct

import numpy as np
import pytest

from soa_scenes import soa, cornell_description, scene_from_canonical

MAX = 0xFFFFFFFF
FROZEN_RICH_SCENE_HASH = 0x9763897EF2337C69  # SceneSoATest.cpp:41


def affine(scale, translation):
    """affine_matrix of SceneSoATest.cpp:113-124: identity with a diagonal scale, a translation
    column and a negative zero at (0, 1)."""
    m = [1.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 1.0]
    m[0], m[5], m[10] = scale
    m[1] = -0.0
    m[3], m[7], m[11] = translation
    return [np.float32(v) for v in m]


def quad_mesh(geometry_id):
    """make_mesh (SceneSoATest.cpp:126-149)."""
    return {"id": geometry_id,
            "positions": [[-0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [1.0, 1.0, 0.0], [0.0, 1.0, 0.0]],
            "normals": [[0.0, 0.0, 1.0]] * 4,
            "uvs": [[-0.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.0, 1.0]],
            "triangles": [[0, 1, 2], [0, 2, 3]]}


def material(wavelengths, reflectance, emission, conductor):
    """make_material (SceneSoATest.cpp:159-200)."""
    closures = [soa.lambertian(reflectance)]
    probabilities, frame_mode, rotation = [1.0], 0, 0.25
    if conductor:
        closures.append(soa.rough_conductor([0.2, 0.3, 0.4, 0.5], [0.25, 0.5, 1.0, 1.5], [3.0, 2.5, 2.0, 1.5], 0.2, 0.7))
        probabilities, frame_mode, rotation = [0.25, 0.75], 1, -0.375
    return {"wavelengths": wavelengths, "closures": closures, "probabilities": probabilities,
            "frame_mode": frame_mode, "tangent_rotation": rotation, "emission": emission}


def rich_description(permute=False):
    """make_rich_description (SceneSoATest.cpp:202-327)."""
    wl = soa.uniform_visible_wavelengths(0.375)
    d = {
        "constant_textures": [
            {"id": MAX, "kind": soa.TEXTURE_SPECTRUM, "value": [-2.0, -0.5, 0.75, 8.0]},
            {"id": 7, "kind": soa.TEXTURE_COLOR, "value": [-1.0, 0.25, 3.5]},
            {"id": 0, "kind": soa.TEXTURE_FLOAT, "value": -0.75},
        ],
        "objects": [{"id": MAX}, {"id": 0}],
        "geometries": [quad_mesh(MAX), quad_mesh(0)],
        "materials": [
            {"id": MAX, "spectral": material(wl, [0.15, 0.25, 0.35, 0.45], [1.0, 2.0, 3.0, 4.0], True)},
            {"id": 0, "spectral": material(wl, [0.65, 0.55, 0.45, 0.35], [0.0] * 4, False)},
        ],
        "instances": [
            {"id": 17, "parent": MAX, "object": 0, "geometry": 0, "material": 0,
             "local_to_parent": affine((2.0, 1.5, 1.0), (0.0, 1.0, 0.0)), "visibility_mask": 0x00FF00FF},
            {"id": 0, "parent": 17, "object": MAX, "geometry": MAX, "material": MAX,
             "local_to_parent": affine((1.0, 1.0, 0.5), (-2.0, 0.5, 3.0)), "visibility_mask": 0xA5A55A5A},
            {"id": MAX, "parent": None, "object": MAX, "geometry": MAX, "material": MAX,
             "local_to_parent": affine((1.0, 2.0, 1.0), (4.0, -1.0, 2.0)), "visibility_mask": MAX},
        ],
        "punctual_lights": [
            {"kind": "spot", "position": [-2.0, 1.0, 4.0], "position_error": [0.03, 0.02, 0.01],
             "direction": [0.0, 0.0, -1.0], "inner": 0.25, "outer": 0.5, "spectrum": [3.0, 6.0, 9.0, 12.0]},
            {"kind": "point", "position": [1.0, 2.0, 3.0], "position_error": [0.01, 0.02, 0.03],
             "spectrum": [1.0, 2.0, 3.0, 4.0]},
            {"kind": "directional", "direction": [0.0, 0.0, -1.0], "spectrum": [2.0, 4.0, 6.0, 8.0]},
        ],
        "environment": {"wavelengths": wl, "radiance": [0.1, 0.2, 0.3, 0.4]},
    }
    if permute:
        for key in ("constant_textures", "objects", "geometries", "materials", "instances"):
            d[key] = d[key][::-1]
    return d


def rich_blob(permute=False):
    return soa.serialize(soa.frame_scene(rich_description(permute)))


def test_rich_scene_matches_the_frozen_hash():
    blob = rich_blob()
    header = soa.unpack_header(blob)
    assert soa.validate_header(header) == soa.VALID
    assert header["total_size_bytes"] == len(blob)
    assert header["content_hash"] == soa.fnv1a64_normalized(blob)
    assert header["content_hash"] == FROZEN_RICH_SCENE_HASH


def test_canonical_bytes_ignore_insertion_order():
    """SceneSoATest.cpp:1077-1090: reversed registries serialize to the same bytes."""
    assert rich_blob(permute=True) == rich_blob()


def test_punctual_registry_order_is_semantic():
    """SceneSoATest.cpp:1092-1106: punctual lights keep their registry order."""
    d = rich_description()
    d["punctual_lights"] = d["punctual_lights"][::-1]
    other = soa.serialize(soa.frame_scene(d))
    assert soa.unpack_header(other)["content_hash"] != FROZEN_RICH_SCENE_HASH


def test_empty_scene_is_a_bare_header():
    """SceneSoATest.cpp:1108-1124."""
    blob = soa.serialize(soa.frame_scene({}))
    header = soa.unpack_header(blob)
    assert len(blob) == soa.HEADER_SIZE and soa.validate_header(header) == soa.VALID
    assert all(off == 0 and count == 0 and size != 0 for off, count, size, _ in header["columns"])


def test_absent_spectral_data_encodes_zeros():
    """SceneSoATest.cpp:1142-1204: a material without spectral data, no environment invented."""
    quad = rich_description()["geometries"][0]
    d = {"objects": [{"id": 0}], "geometries": [quad], "materials": [{"id": 0, "spectral": None}],
         "instances": [{"id": MAX, "parent": None, "object": 0, "geometry": MAX, "material": 0,
                        "local_to_parent": [np.float32(v) for v in np.eye(4).reshape(-1)], "visibility_mask": MAX}]}
    blob = soa.serialize(soa.frame_scene(d))
    h = soa.unpack_header(blob)
    assert soa.validate_header(h) == soa.VALID
    for key in ("environment", "closure", "punctual_light", "mesh_area_light", "texture"):
        assert h["counts"][key] == 0
    col = lambda c: soa.read_column(blob, h, c)
    assert list(col(soa.MATERIAL_SPECTRAL_PRESENT)) == [0]
    assert list(col(soa.MATERIAL_CLOSURE_OFFSET)) == [0] and list(col(soa.MATERIAL_CLOSURE_COUNT)) == [0]
    assert list(col(soa.MATERIAL_FRAME_MODE)) == [0] and list(col(soa.MATERIAL_TANGENT_ROTATION)) == [0.0]
    for lane in range(4):
        assert list(col(soa.MATERIAL_WAVELENGTH_NM + lane)) == [0.0]
        assert list(col(soa.MATERIAL_WAVELENGTH_PDF + lane)) == [0.0]
        assert list(col(soa.MATERIAL_WAVELENGTH_MEASURE + lane)) == [0]
        assert list(col(soa.MATERIAL_EMITTED_RADIANCE + lane)) == [0.0]


def test_column_schema_matches_the_frozen_table():
    """SceneSoATest.cpp:482-558: indices and element sizes the reference freezes."""
    sizes = {31: 8, 32: 8, 33: 1, 34: 4, 39: 4, 40: 4, 41: 4, 45: 4, 55: 4, 160: 4, 161: 4, 162: 4, 166: 1,
             179: 4, 185: 8, 192: 8, 193: 4}
    for column, size in sizes.items():
        assert np.dtype(soa.column_dtype(column)).itemsize == size, column
    assert soa.COLUMN_COUNT == 194 and soa.HEADER_SIZE == 4848 and soa.HASH_OFFSET == 32


def bare_header(**overrides):
    """RejectsIncompatibleAndMalformedHeaders' starting header (SceneSoATest.cpp:561-572)."""
    h = {"magic": soa.MAGIC, "abi_major": 4, "abi_minor": 0, "header_size": soa.HEADER_SIZE,
         "column_count": soa.COLUMN_COUNT, "hash_algorithm": 1, "total_size_bytes": soa.HEADER_SIZE,
         "content_hash": 0, "counts": {k: 0 for k in soa.COUNT_FIELDS},
         "columns": [(0, 0, np.dtype(soa.column_dtype(c)).itemsize, 0) for c in range(soa.COLUMN_COUNT)],
         "reserved": [0] * 5}
    h.update(overrides)
    return h


def malformed_headers():
    """The reference's rejection cases (SceneSoATest.cpp:574-617) plus one per remaining status."""
    cases = [(bare_header(), soa.VALID),
             (bare_header(abi_major=1), soa.INCOMPATIBLE_VERSION),
             (bare_header(reserved=[0, 0, 0, 0, 1]), soa.NONZERO_RESERVED),
             (bare_header(magic=0), soa.BAD_MAGIC),
             (bare_header(header_size=4847), soa.BAD_HEADER_SIZE),
             (bare_header(column_count=193), soa.BAD_COLUMN_COUNT),
             (bare_header(hash_algorithm=2), soa.UNSUPPORTED_HASH_ALGORITHM),
             (bare_header(total_size_bytes=16), soa.BAD_TOTAL_SIZE)]
    h = bare_header()
    h["counts"]["environment"] = 2
    cases.append((h, soa.INVALID_ENVIRONMENT_COUNT))
    h = bare_header()
    h["columns"][0] = (0, 0, 4, 1)
    cases.append((h, soa.INVALID_COLUMN_DESCRIPTOR))
    h = bare_header()
    h["counts"]["object"], h["counts"]["geometry"] = 4611686018427387006, 1
    h["columns"][soa.OBJECT_ID] = (soa.HEADER_SIZE, 4611686018427387006, 4, 0)
    h["columns"][soa.GEOMETRY_ID] = (0, 1, 4, 0)
    cases.append((h, soa.SIZE_OVERFLOW))
    h = bare_header()
    h["counts"]["closure"] = (2**64 - 1) // 4 + 1
    h["columns"][soa.CLOSURE_KIND] = (soa.HEADER_SIZE, h["counts"]["closure"], 4, 0)
    cases.append((h, soa.SIZE_OVERFLOW))
    return cases


@pytest.mark.parametrize("case", range(12))
def test_header_validation_order(case, hip):
    """Oracle restatement and the C-ABI validator agree on every status (host-only call)."""
    header, expected = malformed_headers()[case]
    assert soa.validate_header(header) == expected
    assert hip.scene_soa_validate(soa.pack_header(header), verify_hash=False) == expected


def test_blob_integrity_statuses(hip, abi):
    blob = rich_blob()
    assert hip.scene_soa_validate(blob) == abi.SCENE_SOA_VALID
    assert hip.scene_soa_validate(blob[:-1]) == abi.SCENE_SOA_TRUNCATED
    assert hip.scene_soa_validate(blob[: soa.HEADER_SIZE - 1]) == abi.SCENE_SOA_TRUNCATED
    corrupt = bytearray(blob)
    corrupt[-1] ^= 0x01
    assert hip.scene_soa_validate(bytes(corrupt)) == abi.SCENE_SOA_HASH_MISMATCH
    assert hip.scene_soa_validate(bytes(corrupt), verify_hash=False) == abi.SCENE_SOA_VALID
    # the stored hash field itself is excluded from the hash
    rehashed = bytearray(blob)
    rehashed[soa.HASH_OFFSET:soa.HASH_OFFSET + 8] = b"\0" * 8
    assert soa.fnv1a64_normalized(bytes(rehashed)) == FROZEN_RICH_SCENE_HASH


def patched(blob, column, element, value):
    h = soa.unpack_header(blob)
    off, count, size, _ = h["columns"][column]
    out = bytearray(blob)
    np.frombuffer(out, dtype=soa.column_dtype(column), count=count, offset=off)[element] = value
    return bytes(out)


def ingest_status(hip, blob, verify_hash=False):
    try:
        hip.XpuScene.from_soa(blob, verify_hash=verify_hash)
    except hip.XpuError as e:
        return e.status, str(e)
    return 0, ""


def test_blob_ingest_rejections_before_any_device_work(bf, hip, abi):
    """Each case is refused by host checks that run before the device scene is built."""
    rich = rich_blob()
    # the spot light's axis is no longer a unit vector
    status, message = ingest_status(hip, patched(rich, soa.PUNCTUAL_DIRECTION_X + 2, 0, -0.5))
    assert status == abi.INVALID_ARGUMENT and "unit vector" in message
    # inner cone wider than the outer one
    assert ingest_status(hip, patched(rich, soa.PUNCTUAL_INNER, 0, 0.75))[0] == abi.INVALID_ARGUMENT
    # negative position error
    assert ingest_status(hip, patched(rich, soa.PUNCTUAL_POSITION_ERROR_X, 1, -0.01))[0] == abi.INVALID_ARGUMENT
    d, _ = cornell_description(bf)
    blob = soa.serialize(soa.frame_scene(d))
    corrupt = bytearray(blob)
    corrupt[-1] ^= 0x01
    assert ingest_status(hip, bytes(corrupt), verify_hash=True)[0] == abi.INVALID_ARGUMENT
    h = soa.unpack_header(blob)
    assert ingest_status(hip, patched(blob, soa.MATERIAL_WAVELENGTH_PDF + 2, 0, 0.5))[0] == abi.INVALID_ARGUMENT
    assert ingest_status(hip, patched(blob, soa.MATERIAL_WAVELENGTH_NM + 1, 3, 500.0))[0] == abi.INVALID_ARGUMENT
    assert ingest_status(hip, patched(blob, soa.MESH_AREA_LIGHT_INSTANCE_ID, 0, 31))[0] == abi.INVALID_ARGUMENT
    assert ingest_status(hip, patched(blob, soa.INSTANCE_MATERIAL_ID, 0, 99))[0] == abi.INVALID_ARGUMENT
    assert ingest_status(hip, patched(blob, soa.MATERIAL_ID, 0, 0xFFFFFFFF))[0] == abi.INVALID_ARGUMENT
    # a normal map naming an image the blob does not hold, and an absent bump map with a field set
    status, message = ingest_status(hip, patched(blob, soa.NORMAL_MAP_PRESENT, 0, 1))
    assert status == abi.INVALID_ARGUMENT and "not a data image" in message
    status, message = ingest_status(hip, patched(blob, soa.BUMP_MAP_PRESENT + 2, 0, 3))
    assert status == abi.INVALID_ARGUMENT and "absent bump map" in message
    assert ingest_status(hip, patched(blob, soa.INSTANCE_LOCAL_TO_WORLD + 12, 0, 1.0))[0] == abi.INVALID_ARGUMENT
    assert h["counts"]["mesh_area_light"] == 1


def test_cornell_blob_round_trip_of_the_description(bf):
    """The test scenes serialize validly and the Scene-builder twin holds the same records."""
    for lane_constant in (False, True):
        d, wl = cornell_description(bf, lane_constant=lane_constant)
        canon = soa.frame_scene(d)
        blob = soa.serialize(canon)
        h = soa.unpack_header(blob)
        assert soa.validate_header(h) == soa.VALID
        assert h["counts"]["instance"] == len(d["instances"])
        scene = scene_from_canonical(bf, canon)
        assert scene.desc.instance_count == h["counts"]["instance"]
        assert scene.desc.material_count == h["counts"]["material"]
        assert scene.desc.closure_count == h["counts"]["closure"]
        packet = [w[0] for w in wl]
        assert packet == [np.float32(v) for v in (536.25, 653.75, 771.25, 418.75)]
