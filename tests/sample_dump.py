"""The reference's sample-stream dump requests and its canonical text form
(Tests/CUDA/SampleStreamTest.cpp:197-414), whose FNV-1a-64 the reference freezes as
0x59DC894B1E63155D (:35). Shared by the oracle (CPU) and device (GPU) dump tests."""
import struct

FROZEN_CANONICAL_DUMP_FNV1A = 0x59DC894B1E63155D
MAX_U32, MAX_U64, HIGH_BIT = 0xFFFFFFFF, 0xFFFFFFFFFFFFFFFF, 1 << 63
RASTER_X, RASTER_Y = 0xA24BAED4963EE407, 0x9FB21C651E98DF25


def requests():
    """(pixel_x, pixel_y, sample_index, seed, dimension, bounce_index) in the reference's order."""
    base = (0x01234567, 0x89ABCDEF, 0x0123456789ABCDEF, 0xFEDCBA9876543210)
    out = [
        (0, 0, 0, 0, 0, 0),
        (17, 29, 0, 0x0123456789ABCDEF, 0, 1),
        (17, 29, 4095, 0x0123456789ABCDEF, 1, MAX_U32),
        (MAX_U32, MAX_U32, MAX_U64, MAX_U64, MAX_U64, MAX_U32),
        (0, 0, 0, 0, RASTER_X, MAX_U32 + 1),
        base + (HIGH_BIT | 0x25, 0),
    ]
    for field, flip in ((0, 0x80000000), (1, 0x80000000), (2, HIGH_BIT), (3, HIGH_BIT)):
        changed = list(base)
        changed[field] ^= flip
        out.append(tuple(changed) + (HIGH_BIT | 0x25, 0))
    out.append(base + (0x25, 0))
    return out


def manifest_values():
    """(schema, primary dimensions, first bounce dimension, dimensions per bounce, maximum bounce,
    bounce offsets) of the current dimension map."""
    return 1, (RASTER_X, RASTER_Y, 0, 1, 2, 3), 4, 10, MAX_U32, tuple(range(10))


def canonical_text(manifest, reqs, results):
    """manifest as manifest_values(); results: (status, indexed bits, sample float bits, 10 dims)."""
    schema, primary, first, per, maximum, offsets = manifest
    hx = lambda v, digits: "0x" + format(v, "0%dx" % digits)
    lines = ["schema=%d" % schema, "primary=" + ",".join(hx(v, 16) for v in primary),
             "bounce=" + ",".join([hx(first, 16), str(per), str(maximum)] + [str(v) for v in offsets])]
    for (px, py, s, seed, dim, bounce), (status, bits, sample_bits, dims) in zip(reqs, results):
        lines.append("request=" + ",".join([hx(px, 8), hx(py, 8), hx(s, 16), hx(seed, 16), hx(dim, 16), str(bounce)]) +
                     ";result=" + ",".join([str(status), hx(bits, 16), hx(sample_bits, 8)] + [hx(d, 16) for d in dims]))
    return "\n".join(lines) + "\n"


def fnv1a64(text):
    h = 0xCBF29CE484222325
    for byte in text.encode():
        h = ((h ^ byte) * 0x100000001B3) & MAX_U64
    return h


def float_bits(value):
    return struct.unpack("<I", struct.pack("<f", value))[0]
