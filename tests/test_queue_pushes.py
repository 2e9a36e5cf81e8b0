"""Queue pushes (blackframe_hip_launch_wavefront_queue_pushes), the reference's
try_push_wavefront_queue_warp contract (WavefrontQueueDevice.cuh:118-197, launched by
CudaWavefrontQueues.cu:22-68; cases after Tests/Backends/GPU/CUDA/WavefrontQueuesTest.cpp).

Which slot lands where, and which requests win under overflow, are unspecified (atomics), so the
checks are the contract's invariants: with room, every request is pushed and each queue holds
exactly its requests' slots (as a multiset) after its earlier contents; without room, exactly the
remaining capacity is pushed, the rest report capacity_exhausted and the header's overflow and
rejected counters grow by their number; unknown kinds and broken headers report
invalid_contract and change nothing."""
import ctypes as C

import numpy as np
import pytest

PUSHED, FULL, INVALID = 0, 1, 2
SENTINEL = 0xF17EC0DE


def check(oracle_module, h0, s0, req, h, s, out):
    stride = s0.size // 7
    req = np.asarray(req, dtype=np.uint32).reshape(-1, 2)
    for k in range(7):
        mine = req[:, 0] == k
        ok = (h0[k, 0] == 1) and h0[k, 1] == 32 and h0[k, 2] == k and h0[k, 3] == stride and h0[k, 7] == 0
        if not ok or h0[k, 4] > h0[k, 3]:
            assert np.all(out[mine] == INVALID), k
            assert np.array_equal(h[k], h0[k]) and np.array_equal(s[k * stride:(k + 1) * stride], s0[k * stride:(k + 1) * stride])
            continue
        room = int(h0[k, 3] - h0[k, 4])
        wanted = int(mine.sum())
        pushed = min(room, wanted)
        assert int((out[mine] == PUSHED).sum()) == pushed, k
        assert int((out[mine] == FULL).sum()) == wanted - pushed, k
        assert h[k, 4] == h0[k, 4] + pushed
        assert h[k, 5] == h0[k, 5] + (wanted - pushed) and h[k, 6] == h0[k, 6] + (wanted - pushed)
        begin = k * stride + int(h0[k, 4])
        assert np.array_equal(s[k * stride:begin], s0[k * stride:begin])  # earlier contents kept
        landed = np.sort(s[begin:begin + pushed])
        assert np.array_equal(landed, np.sort(req[mine & (out == PUSHED), 1]))
        assert np.all(s[begin + pushed:(k + 1) * stride] == s0[begin + pushed:(k + 1) * stride])
    assert np.all(out[req[:, 0] >= 7] == INVALID)


def test_launcher_rejects_bad_arguments_before_any_launch(hip):
    lib = hip.library()
    fake = C.c_void_p(0x10000)
    call = lib.blackframe_hip_launch_wavefront_queue_pushes
    assert call(None, fake, 7, 4, fake, 1, fake) == 1
    assert call(fake, fake, 6, 4, fake, 1, fake) == 1
    assert call(fake, None, 7, 4, fake, 1, fake) == 1
    assert call(fake, fake, 7, 4, None, 1, fake) == 1
    assert call(fake, fake, 7, 4, fake, 1, None) == 1
    assert call(fake, fake, 7, 4, None, 0, None) == 0  # nothing to push


@pytest.mark.gpu
@pytest.mark.parametrize("count,capacity", [(1, 4), (64, 64), (65, 64), (1000, 300), (5000, 8192)])
def test_pushes_keep_the_queue_contract(hip, oracle_module, count, capacity):
    rng = np.random.default_rng(count + capacity)
    prefix = [3, 0, 1, 2, 0, 5, 0]
    h0 = oracle_module.queue_headers(capacity, prefix)
    s0 = np.full(7 * capacity, SENTINEL, dtype=np.uint32)
    for k, p in enumerate(prefix):
        s0[k * capacity:k * capacity + p] = np.arange(p) + 100 * k
    kinds = rng.integers(0, 8, count)  # kind 7 is not a queue
    req = np.stack([kinds, rng.integers(0, 1 << 20, count)], axis=1).astype(np.uint32)
    h, s, out = hip.queue_pushes(h0, s0, req)
    check(oracle_module, h0, s0, req, h, s, out)


@pytest.mark.gpu
def test_broken_headers_reject_their_pushes(hip, oracle_module):
    capacity = 16
    h0 = oracle_module.queue_headers(capacity)
    h0[2, 3] = capacity + 1  # capacity differs from the slot stride
    h0[4, 4] = capacity + 2  # size beyond capacity
    h0[5, 7] = 1             # reserved word set
    s0 = np.full(7 * capacity, SENTINEL, dtype=np.uint32)
    req = np.array([[k % 7, k] for k in range(70)], dtype=np.uint32)
    h, s, out = hip.queue_pushes(h0, s0, req)
    check(oracle_module, h0, s0, req, h, s, out)
    assert np.all(out[req[:, 0] == 2] == INVALID) and np.all(out[req[:, 0] == 5] == INVALID)
