"""Scene blob hash (blackframe_hip_launch_scene_soa_hash): FNV-1a-64 with the blob's
content-hash field (bytes 32..39 of the 4848-byte header) read as zero, as the reference's
CudaSceneSoaHash.cu:13-36 and SceneSoATest.cpp's normalized_hash define it.

The oracle's FNV-1a is pinned by the published FNV-1a-64 test vectors; the device must equal the
oracle on blobs of several sizes and alignments, and the hash must not depend on the hash field's
own contents. (The reference's frozen rich-scene hash needs its scene blob serializer, which this
build does not restate; see the scope notes of the design document.)"""
import ctypes as C

import numpy as np
import pytest

HEADER = 4848


def test_fnv1a64_published_vectors(oracle_module):
    assert oracle_module.fnv1a64(b"") == 0xCBF29CE484222325
    assert oracle_module.fnv1a64(b"a") == 0xAF63DC4C8601EC8C
    assert oracle_module.fnv1a64(b"foobar") == 0x85944171F73967E8


def test_hash_field_is_read_as_zero(oracle_module):
    rng = np.random.default_rng(3)
    blob = bytearray(rng.integers(0, 256, HEADER + 40, dtype=np.uint8).tobytes())
    a = oracle_module.fnv1a64(blob, 32, 40)
    blob[32:40] = b"\xff" * 8
    assert oracle_module.fnv1a64(blob, 32, 40) == a
    blob[32:40] = b"\x00" * 8
    assert oracle_module.fnv1a64(blob) == a
    blob[31] ^= 1
    assert oracle_module.fnv1a64(blob, 32, 40) != a


def test_launcher_rejects_short_blobs_and_null_pointers(hip):
    lib = hip.library()
    fake = C.c_void_p(0x10000)
    assert lib.blackframe_hip_launch_scene_soa_hash(fake, HEADER - 1, fake) == 1
    assert lib.blackframe_hip_launch_scene_soa_hash(None, HEADER, fake) == 1
    assert lib.blackframe_hip_launch_scene_soa_hash(fake, HEADER, None) == 1


@pytest.mark.gpu
@pytest.mark.parametrize("size,offset", [(HEADER, 0), (HEADER + 1, 0), (HEADER + 15, 3), (100003, 0), (100003, 1)])
def test_device_hash_equals_the_oracle(hip, oracle_module, size, offset):
    rng = np.random.default_rng(size + offset)
    blob = rng.integers(0, 256, size, dtype=np.uint8).tobytes()
    assert hip.scene_soa_hash(blob, offset) == oracle_module.fnv1a64(blob, 32, 40)
