"""OpenEXR scanline files of scene-linear float RGB, as the reference writes its images
(Renderer/Source/ExrWriter.cpp, write_scene_linear_exr; read back by
Tests/Renderer/CornellDiffuseReferenceImage.hpp): single-part scanline files with float B, G, R
channels, ZIP compression (zlib over blocks of 16 scanlines after the byte split and delta
predictor of the OpenEXR file layout), increasing-y line order, the default Rec. 709 / D65
chromaticities, a display window of the full frame and a data window of the film's crop, and the
run metadata as "blackframe.*" string attributes.

    write_scene_linear_exr(path, rgb, metadata, display=(width, height), origin=(x0, y0))
    image = read_exr(path)   # {"header": {...}, "channels": {"R": array, ...}, "data_window": ...}
"""
import struct
import zlib

import numpy as np

MAGIC = 20000630
COMPRESSION = {"none": 0, "rle": 1, "zips": 2, "zip": 3}
LINES_PER_BLOCK = {0: 1, 2: 1, 3: 16}
FLOAT = 2
# Rec. 709 primaries and D65 white, the OpenEXR default chromaticities
DEFAULT_CHROMATICITIES = (0.6400, 0.3300, 0.3000, 0.6000, 0.1500, 0.0600, 0.3127, 0.3290)
METADATA_KEYS = ("scene", "seed", "commit", "options", "backend", "capabilities", "asset_hashes")


def _attribute(name, kind, payload):
    return name.encode() + b"\0" + kind.encode() + b"\0" + struct.pack("<i", len(payload)) + payload


def _predict(raw):
    """The byte split then delta predictor OpenEXR applies before zlib."""
    b = np.frombuffer(raw, dtype=np.uint8)
    split = np.concatenate([b[0::2], b[1::2]])
    out = split.copy()
    out[1:] = (split[1:].astype(np.int16) - split[:-1].astype(np.int16) + 128).astype(np.uint8)
    return out.tobytes()


def _unpredict(data):
    d = np.frombuffer(data, dtype=np.uint8).astype(np.int64)
    # running sum of (d - 128) from the first byte, modulo 256
    t = d.copy()
    t[1:] -= 128
    t = (np.cumsum(t) & 0xFF).astype(np.uint8)
    half = (len(t) + 1) // 2
    out = np.empty(len(t), dtype=np.uint8)
    out[0::2] = t[:half]
    out[1::2] = t[half:]
    return out.tobytes()


def write_scene_linear_exr(path, rgb, metadata, display=None, origin=(0, 0), compression="zip"):
    """rgb: (height, width, 3) float32 linear sRGB of the data window, whose top-left pixel is
    `origin` within a display window of `display` = (width, height) (the whole image by default).
    metadata: dict with the keys of METADATA_KEYS (values str or int)."""
    rgb = np.ascontiguousarray(rgb, dtype=np.float32)
    height, width = rgb.shape[:2]
    dw, dh = display if display is not None else (origin[0] + width, origin[1] + height)
    x0, y0 = origin
    code = COMPRESSION[compression]
    channels = b"".join(name + b"\0" + struct.pack("<iB3xii", FLOAT, 0, 1, 1) for name in (b"B", b"G", b"R")) + b"\0"
    attributes = {
        "blackframe.run_metadata_version": ("string", b"1"),
        "blackframe.color_space": ("string", b"scene-linear-srgb"),
        "channels": ("chlist", channels),
        "chromaticities": ("chromaticities", struct.pack("<8f", *DEFAULT_CHROMATICITIES)),
        "compression": ("compression", bytes([code])),
        "dataWindow": ("box2i", struct.pack("<4i", x0, y0, x0 + width - 1, y0 + height - 1)),
        "displayWindow": ("box2i", struct.pack("<4i", 0, 0, dw - 1, dh - 1)),
        "lineOrder": ("lineOrder", b"\0"),
        "pixelAspectRatio": ("float", struct.pack("<f", 1.0)),
        "screenWindowCenter": ("v2f", struct.pack("<2f", 0.0, 0.0)),
        "screenWindowWidth": ("float", struct.pack("<f", 1.0)),
    }
    for key in METADATA_KEYS:
        attributes["blackframe." + key] = ("string", str(metadata[key]).encode())
    header = struct.pack("<ii", MAGIC, 2)
    header += b"".join(_attribute(name, *attributes[name]) for name in sorted(attributes)) + b"\0"
    lines = LINES_PER_BLOCK[code]
    blocks = []
    for first in range(0, height, lines):
        rows = rgb[first:first + lines]
        # per scanline: all B, then all G, then all R (channels in name order)
        raw = b"".join(np.ascontiguousarray(row[:, c]).tobytes() for row in rows for c in (2, 1, 0))
        data = raw
        if code in (2, 3):
            packed = zlib.compress(_predict(raw), 6)
            if len(packed) < len(raw):
                data = packed
        blocks.append(struct.pack("<ii", y0 + first, len(data)) + data)
    offsets, cursor = [], len(header) + 8 * len(blocks)
    for block in blocks:
        offsets.append(cursor)
        cursor += len(block)
    with open(path, "wb") as f:
        f.write(header + struct.pack("<%dQ" % len(offsets), *offsets) + b"".join(blocks))


def read_exr(path):
    """A single-part scanline file with FLOAT channels and NONE, ZIPS or ZIP compression."""
    with open(path, "rb") as f:
        b = f.read()
    magic, version = struct.unpack_from("<ii", b, 0)
    if magic != MAGIC or (version & 0xFF) != 2 or version & 0x1E00:
        raise ValueError("not a single-part scanline OpenEXR file")
    header, i = {}, 8
    while b[i] != 0:
        j = b.index(b"\0", i)
        k = b.index(b"\0", j + 1)
        name, kind = b[i:j].decode(), b[j + 1:k].decode()
        size = struct.unpack_from("<i", b, k + 1)[0]
        value = b[k + 5:k + 5 + size]
        header[name] = (kind, value)
        i = k + 5 + size
    i += 1
    channels, c = [], 0
    raw = header["channels"][1]
    while raw[c] != 0:
        e = raw.index(b"\0", c)
        pixel_type, _, xs, ys = struct.unpack_from("<iB3xii", raw, e + 1)
        if pixel_type != FLOAT or xs != 1 or ys != 1:
            raise ValueError("only full-resolution FLOAT channels are read")
        channels.append(raw[c:e].decode())
        c = e + 17
    code = header["compression"][1][0]
    if code not in LINES_PER_BLOCK:
        raise ValueError(f"compression {code} is not read")
    x0, y0, x1, y1 = struct.unpack("<4i", header["dataWindow"][1])
    width, height = x1 - x0 + 1, y1 - y0 + 1
    lines = LINES_PER_BLOCK[code]
    blocks = (height + lines - 1) // lines
    offsets = struct.unpack_from("<%dQ" % blocks, b, i)
    planes = {name: np.zeros((height, width), np.float32) for name in channels}
    for off in offsets:
        y, size = struct.unpack_from("<ii", b, off)
        data = b[off + 8:off + 8 + size]
        rows = min(lines, y1 + 1 - y)
        expected = rows * width * 4 * len(channels)
        if size < expected:
            data = _unpredict(zlib.decompress(data))
        values = np.frombuffer(data, dtype="<f4").reshape(rows, len(channels), width)
        for k, name in enumerate(channels):
            planes[name][y - y0:y - y0 + rows] = values[:, k]
    strings = {name[len("blackframe."):]: value.decode() for name, (kind, value) in header.items()
               if kind == "string" and name.startswith("blackframe.")}
    return {"header": header, "channels": planes, "data_window": (x0, y0, x1, y1),
            "display_window": struct.unpack("<4i", header["displayWindow"][1]), "compression": code,
            "chromaticities": struct.unpack("<8f", header["chromaticities"][1]) if "chromaticities" in header else None,
            "metadata": strings}


def rgb_of(image):
    """(height, width, 3) float32 R, G, B of a read image."""
    c = image["channels"]
    return np.stack([c["R"], c["G"], c["B"]], axis=-1)
