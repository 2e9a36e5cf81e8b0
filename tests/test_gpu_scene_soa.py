"""Scenes that arrive as the reference's serialized blob ("BLKFSOA4"): the device hash of the
frozen rich scene, and renders of a blob scene against the oracle's CPU wavefront on the same scene
built as a bf_scene_desc (bar: the transport parity of test_gpu_transport.check_parity)."""
import numpy as np
import pytest

from soa_scenes import soa, cornell_description, scene_from_canonical
from test_gpu_transport import check_parity, path_arrays
from test_scene_soa import FROZEN_RICH_SCENE_HASH, rich_blob

pytestmark = pytest.mark.gpu

W, H, SPP = 32, 32, 4


def test_device_hash_of_the_rich_scene_is_the_frozen_value(hip):
    """SceneSoATest.cpp:1052-1075: device hash == host hash == 0x9763897EF2337C69."""
    assert hip.scene_soa_hash(rich_blob()) == FROZEN_RICH_SCENE_HASH


def camera_frame(bf, spp=SPP, seed=0):
    cam = bf.scenes.look_at_camera((0.0, 0.0, 4.0), (0.0, -0.05, 0.92), W, H, 0.70)
    return bf.scenes.frame_desc(cam, spp, seed=seed)


def packet_options(bf, wl, **kw):
    options = bf.scenes.transport_options(wavelength_mode=0, packet=[float(w[0]) for w in wl], **kw)
    options.record_queues = 1
    return options


@pytest.mark.parametrize("opts", [dict(), dict(mis=1, max_depth=12, sampler=0)])
def test_blob_scene_matches_the_oracle(bf, oracle, hip, opts):
    """Per-lane spectra, a Lambert + GGX-conductor mixture in the tangent frame, a parent chain
    and an environment, rendered from the blob and by the oracle from the Scene-builder twin."""
    d, wl = cornell_description(bf)
    canon = soa.frame_scene(d)
    blob = soa.serialize(canon)
    twin = scene_from_canonical(bf, canon)
    frame, options = camera_frame(bf), packet_options(bf, wl, **opts)
    print(f"blob scene {len(blob)} bytes, {W}x{H}x{SPP}, options {opts}", flush=True)
    expected = oracle.scene(twin).render(frame, options, threads=8)
    ws = hip.XpuWavefrontTransportWorkspace(W * H * SPP)
    blob_scene = hip.XpuScene.from_soa(blob)
    got = hip.trace_xpu_wavefront_transport(ws, blob_scene, frame, options, want_paths=True, want_counters=True)
    check_parity(expected, got)
    # the blob and the desc build the same BVH
    assert blob_scene.bvh() == hip.XpuScene(twin).bvh()


def test_flat_blob_spectra_render_like_constant_spectra(bf, hip):
    """Lane-constant blob spectra (point-sampled) give the same paths and film as one-value
    spectra of the same scene built as a bf_scene_desc."""
    d, wl = cornell_description(bf, lane_constant=True)
    canon = soa.frame_scene(d)
    frame, options = camera_frame(bf), packet_options(bf, wl)
    ws = hip.XpuWavefrontTransportWorkspace(W * H * SPP)
    print("flat blob scene vs constant spectra", flush=True)
    a = hip.trace_xpu_wavefront_transport(ws, hip.XpuScene.from_soa(soa.serialize(canon)), frame, options, want_paths=True)
    b = hip.trace_xpu_wavefront_transport(ws, hip.XpuScene(scene_from_canonical(bf, canon, constant_spectra=True)), frame,
                                          options, want_paths=True)
    assert np.array_equal(path_arrays(a["paths"]), path_arrays(b["paths"]))
    fa, fb = a["film"].host(), b["film"].host()
    for key in ("film", "film_compensated", "film_counts"):
        assert np.array_equal(fa[key], fb[key]), key


def test_blob_scene_renders_only_at_its_packet(bf, hip, abi):
    d, wl = cornell_description(bf)
    scene = hip.XpuScene.from_soa(soa.serialize(soa.frame_scene(d)))
    ws = hip.XpuWavefrontTransportWorkspace(W * H)
    frame = camera_frame(bf, spp=1)
    print("packet checks", flush=True)
    for options in (bf.scenes.transport_options(), packet_options(bf, soa.uniform_visible_wavelengths(0.5))):
        with pytest.raises(hip.XpuError) as e:
            hip.trace_xpu_wavefront_transport(ws, scene, frame, options)
        assert e.value.status == abi.INVALID_ARGUMENT
    hip.trace_xpu_wavefront_transport(ws, scene, frame, packet_options(bf, wl))


def test_device_header_validation_matches_the_oracle(hip, abi):
    """The device-side header check returns the status the oracle's restatement of
    validate_scene_soa_header gives, for the reference's rejection cases and the rich blob."""
    from test_scene_soa import malformed_headers
    for header, expected in malformed_headers():
        assert soa.validate_header(header) == expected
        assert hip.scene_soa_validate_on_device(soa.pack_header(header)) == expected
    blob = rich_blob()
    assert hip.scene_soa_validate_on_device(blob) == abi.SCENE_SOA_VALID
    assert hip.scene_soa_validate_on_device(blob[:-16]) == abi.SCENE_SOA_TRUNCATED
    # the stage check's further conditions (CudaWavefrontStages.cu:818-862): a padded blob, a record
    # count above 2^32 - 1 and more than 2^24 mesh-area lights, each on a header whose columns are
    # consistent with its counts (the check reads the header only)
    assert hip.scene_soa_validate_on_device(blob + bytes(16)) == abi.SCENE_SOA_SIZE_MISMATCH
    for field, count, expected in (("object", 1 << 32, abi.SCENE_SOA_COUNT_OVERFLOW),
                                   ("mesh_area_light", (1 << 24) + 1, abi.SCENE_SOA_TOO_MANY_MESH_LIGHTS),
                                   ("mesh_area_light", 1 << 24, abi.SCENE_SOA_VALID)):
        header = consistent_header({field: count})
        assert soa.validate_header(header) == soa.VALID
        assert hip.scene_soa_validate_on_device(soa.pack_header(header), header["total_size_bytes"]) == expected


def consistent_header(counts):
    """A header whose column descriptors and total size follow `counts` (others zero)."""
    import numpy as np

    from test_scene_soa import bare_header
    h = bare_header()
    h["counts"].update(counts)
    cursor = soa.align_up(soa.HEADER_SIZE)
    for c in range(soa.COLUMN_COUNT):
        n = h["counts"][soa.column_domain(c)]
        size = np.dtype(soa.column_dtype(c)).itemsize
        if n == 0:
            h["columns"][c] = (0, 0, size, 0)
            continue
        cursor = soa.align_up(cursor)
        h["columns"][c] = (cursor, n, size, 0)
        cursor += n * size
    h["total_size_bytes"] = cursor
    return h


def test_bvh_blob_of_a_blob_scene_names_its_source(hip, abi):
    """The serialized BVH of a scene made from a scene blob carries that blob's content hash as its
    source_scene_hash (SceneBvh.cpp:1081, CanonicalBytesIgnoreInsertionAndAllocationIdentity) and
    validates with its own content hash."""
    import struct

    blob = rich_blob()
    bvh = hip.XpuScene.from_soa(blob).bvh_blob()
    assert hip.scene_bvh_validate(bvh) == abi.BVH_VALID
    assert struct.unpack_from("<Q", bvh, 40)[0] == struct.unpack_from("<Q", blob, 32)[0]
