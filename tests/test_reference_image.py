"""The reference's own rendered image as a fixture: tests/golden holds scenes/S02_CornellDiffuse_64x64
.json and the 4096-spp image the reference's scalar backend rendered from it
(scenes/references/S02_CornellDiffuse_64x64_scalar_ref_4096spp.exr, copied with its recorded
SHA-256 by tests/golden/copy_reference_fixtures.py). Its integrator is BSDF-only (no light
sampling); the oracle's and the GPU's transport use light sampling with MIS, which converges to the
same image, so the comparison is statistical:

  - the EXR reader decodes the OpenEXR file (ZIP blocks, FLOAT B/G/R, full data window, default
    chromaticities, the run metadata) and the writer round-trips it bit for bit with the header the
    reference's ExrWriterTest.cpp checks (crop data window inside the display window);
  - the oracle at 256 spp (seed 0): per-channel image means within 2e-3 relative of the reference,
    8x8-pixel block means within 1.5e-2 RMS relative, pixels within 6e-2 RMS relative (the 256-spp
    noise is about 2.8e-2); a row flip fails them by more than ten times, and dropping one
    diffuse bounce fails the mean bar ten times over and the block bar five times over;
  - on the GPU: the same scene at 1024 spp, bit-identical to the oracle at 16 spp first, then
    within the same bars with the block bar at 1e-2.
"""
import numpy as np
import pytest

from s02_scene import IMAGE, s02

MEAN_BAR, BLOCK_BAR, PIXEL_BAR = 2e-3, 1.5e-2, 6e-2


def reference(bf):
    from blackframe_renderer_amd import exr

    return exr.rgb_of(exr.read_exr(str(IMAGE))).astype(np.float64)


def errors(image, ref):
    block = lambda a: a.reshape(8, 8, 8, 8, 3).mean(axis=(1, 3))
    mean = np.abs(image.mean(axis=(0, 1)) / ref.mean(axis=(0, 1)) - 1.0).max()
    blocks = np.sqrt((((block(image) - block(ref)) / block(ref)) ** 2).mean())
    pixels = np.sqrt((((image - ref) / ref) ** 2).mean())
    return mean, blocks, pixels


def film_image(film_compensated, counts):
    return film_compensated.astype(np.float64) / np.maximum(counts.astype(np.float64), 1.0)[..., None]


def test_exr_reader_and_writer(bf, tmp_path):
    from blackframe_renderer_amd import exr

    image = exr.read_exr(str(IMAGE))
    assert image["compression"] == exr.COMPRESSION["zip"]
    assert image["data_window"] == (0, 0, 63, 63) and image["display_window"] == (0, 0, 63, 63)
    assert sorted(image["channels"]) == ["B", "G", "R"]
    assert np.allclose(image["chromaticities"], exr.DEFAULT_CHROMATICITIES, atol=1e-7)
    meta = image["metadata"]
    assert meta["scene"] == "scenes/S02_CornellDiffuse_64x64.json" and meta["color_space"] == "scene-linear-srgb"
    assert meta["seed"] == str(0x243F6A8885A308D3) and "spp=4096" in meta["options"]
    rgb = exr.rgb_of(image)
    assert rgb.shape == (64, 64, 3) and np.isfinite(rgb).all() and rgb.min() > 0
    # round trip, as ExrWriterTest.cpp:45-140 checks it: a 3x2 crop at (1, 1) of a 4x3 frame
    crop = np.array([[[0.1, -0.25, 1.0], [1 / 3, 42.5, -8.0], [100000.25, 0.00001, 3.1415927]],
                     [[-1024.5, 65536.5, 0.0], [2.0, 4.0, 8.0], [0.0078125, -0.03125, 16.0]]], np.float32)
    metadata = {"scene": "scenes/S00_ImagesSynthetic.json", "seed": 2 ** 64 - 1, "commit": "0123456789abcdef",
                "options": "spp=8;depth=4;crop=1,1,4,3", "backend": "scalar_ref",
                "capabilities": "film_float32,exr_openexr", "asset_hashes": "scene=4df6c03f"}
    for compression in ("zip", "none"):
        path = tmp_path / f"crop_{compression}.exr"
        exr.write_scene_linear_exr(str(path), crop, metadata, display=(4, 3), origin=(1, 1), compression=compression)
        back = exr.read_exr(str(path))
        assert back["data_window"] == (1, 1, 3, 2) and back["display_window"] == (0, 0, 3, 2)
        assert back["header"]["lineOrder"][1] == b"\0"
        assert np.array_equal(exr.rgb_of(back).view(np.uint32), crop.view(np.uint32))
        assert back["metadata"]["seed"] == "18446744073709551615" and back["metadata"]["run_metadata_version"] == "1"
    path = tmp_path / "reference.exr"
    exr.write_scene_linear_exr(str(path), rgb, {k: meta[k] for k in exr.METADATA_KEYS})
    assert np.array_equal(exr.rgb_of(exr.read_exr(str(path))).view(np.uint32), rgb.view(np.uint32))


def test_oracle_converges_to_the_reference_image(bf, oracle):
    ref = reference(bf)
    scene, cam, options, _ = s02(bf)
    o = oracle.scene(scene)
    r = o.render(bf.scenes.frame_desc(cam, 256, seed=0), options, threads=8)
    image = film_image(r["film_compensated"], r["film_counts"])
    mean, blocks, pixels = errors(image, ref)
    assert mean <= MEAN_BAR and blocks <= BLOCK_BAR and pixels <= PIXEL_BAR, (mean, blocks, pixels)
    # the bars see a dropped bounce and a row flip
    flipped = errors(image[::-1], ref)
    assert flipped[0] > 10 * MEAN_BAR or flipped[1] > 10 * BLOCK_BAR
    options.diffuse_depth_limit = 2
    r = o.render(bf.scenes.frame_desc(cam, 64, seed=0), options, threads=8)
    short = errors(film_image(r["film_compensated"], r["film_counts"]), ref)
    assert short[0] > 10 * MEAN_BAR and short[1] > 5 * BLOCK_BAR, short


@pytest.mark.gpu
def test_gpu_converges_to_the_reference_image(bf, oracle, hip):
    ref = reference(bf)
    scene, cam, options, _ = s02(bf)
    xs = hip.XpuScene(scene)
    ws = hip.XpuWavefrontTransportWorkspace(64 * 64 * 1024)
    frame = bf.scenes.frame_desc(cam, 16, seed=2)
    got = hip.trace_xpu_wavefront_transport(ws, xs, frame, options)
    want = oracle.scene(scene).render(frame, options, threads=8)
    host = got["film"].host()
    assert np.array_equal(host["film"].view(np.uint32), want["film"].view(np.uint32))
    got = hip.trace_xpu_wavefront_transport(ws, xs, bf.scenes.frame_desc(cam, 1024, seed=3), options)
    host = got["film"].host()
    image = host["film_compensated"] / np.maximum(host["film_counts"].astype(np.float64), 1.0)[..., None]
    mean, blocks, pixels = errors(image, ref)
    assert mean <= MEAN_BAR and blocks <= 1e-2 and pixels <= PIXEL_BAR, (mean, blocks, pixels)
