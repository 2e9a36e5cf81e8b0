"""Queue compaction (blackframe_hip_launch_wavefront_queue_compaction), the reference's
WavefrontQueueCompaction contract.

The cases restate Tests/Backends/GPU/CUDA/WavefrontQueueCompactionTest.cpp:
  - stable prefix scan over sizes 0, 1, 31, 32, 33, 255, 256, 257, 65 536 and the selection
    patterns none / all / alternating / sparse (:304-340);
  - every compactable route 1..6 appends to its own queue (:342-365);
  - an append after an existing prefix keeps the prefix and the outcome order (:367-409);
  - a scratch sized for a larger maximum serves a smaller input (:411-433);
  - capacity exhaustion rejects the whole append and counts it in the header (:435-463);
  - only successful outcomes of the requested route are selected (:465-505);
  - a selected path slot outside the queue rejects the append unchanged (:507-539);
  - short scratch and routes 0 / 7 are rejected before any launch (:541-593).
Their expected slots are the closed form of the contract: the path slots of the selected
outcomes in outcome order (a stable partition). CPU tests check the oracle against those
expectations and the library's synchronous rejections; GPU tests check the device against the
oracle bit for bit (slots, headers and the result record).
"""
import ctypes as C

import numpy as np
import pytest

SENTINEL = 0xF17EC0DE
RAY, HIT, MISS, SHADE, SHADOW, CONTINUATION, TERMINATED = 1, 2, 3, 4, 5, 6, 7


def selected(pattern, index, count):
    return {"none": False, "all": True, "alternating": index % 2 == 0,
            "sparse": index % 17 == 0 or (count != 0 and index + 1 == count)}[pattern]


def make_outcomes(count, pattern, route):
    rejected = RAY if route == MISS else MISS
    return np.array([[0, route if selected(pattern, i, count) else rejected, count - i - 1, 0] for i in range(count)],
                    dtype=np.uint32).reshape(-1, 4)


def make_state(oracle_module, capacity, route, prefix=()):
    slots = np.full(7 * capacity, SENTINEL, dtype=np.uint32)
    slots[route * capacity:route * capacity + len(prefix)] = prefix
    sizes = [0] * 7
    sizes[route] = len(prefix)
    return oracle_module.queue_headers(capacity, sizes), slots


def expected_slots(outcomes, route):
    return [int(o[2]) for o in outcomes if o[0] == 0 and o[1] == route]


def cases(oracle_module):
    """(name, headers, slots, outcomes, route, expected (status, initial, selected, published, rejected), tail)"""
    out = []
    for size in (0, 1, 31, 32, 33, 255, 256, 257, 65536):
        for pattern in ("none", "all", "alternating", "sparse"):
            cap = max(size, 1)
            h, s = make_state(oracle_module, cap, RAY)
            o = make_outcomes(size, pattern, RAY)
            n_sel = len(expected_slots(o, RAY))
            out.append((f"scan-{size}-{pattern}", h, s, o, RAY, (0, 0, n_sel, n_sel, 0)))
    for route in range(1, 7):
        h, s = make_state(oracle_module, 33, route)
        o = make_outcomes(33, "alternating", route)
        out.append((f"route-{route}", h, s, o, route, (0, 0, 17, 17, 0)))
    o = make_outcomes(257, "sparse", CONTINUATION)
    sel = len(expected_slots(o, CONTINUATION))
    h, s = make_state(oracle_module, max(257, 3 + sel + 5), CONTINUATION, (4, 2, 9))
    out.append(("append-after-prefix", h, s, o, CONTINUATION, (0, 3, sel, sel, 0)))
    h, s = make_state(oracle_module, 8, SHADE, (0, 1, 2, 3, 4, 5))
    out.append(("capacity-exhausted", h, s, make_outcomes(3, "all", SHADE), SHADE, (2, 6, 3, 0, 3)))
    h, s = make_state(oracle_module, 4, RAY)
    o = np.array([[0, RAY, 0, 0], [1, RAY, 1, 0], [0, MISS, 2, 0], [0, RAY, 3, 0]], dtype=np.uint32)
    out.append(("successful-only", h, s, o, RAY, (0, 0, 2, 2, 0)))
    h, s = make_state(oracle_module, 4, RAY, (1,))
    o = np.array([[0, RAY, 2, 0], [0, RAY, 4, 0]], dtype=np.uint32)
    out.append(("invalid-path-slot", h, s, o, RAY, (1, 1, 2, 0, 0)))
    return out


def check_expectations(name, h0, s0, o, route, exp, h, s, r):
    status, initial, sel, published, rejected = exp
    assert list(r[:8]) == [status, route, route, len(o), initial, sel, published, rejected], (name, r)
    assert not r[8:].any()
    cap = s0.size // 7
    if status == 0:
        want = s0.copy()
        begin = route * cap + initial
        want[begin:begin + sel] = expected_slots(o, route)
        assert np.array_equal(s, want), name
        assert h[route, 4] == initial + sel and h[route, 5] == 0 and h[route, 6] == 0
    else:
        assert np.array_equal(s, s0), name
        assert h[route, 4] == h0[route, 4]
        assert (h[route, 5], h[route, 6]) == ((rejected, rejected) if status == 2 else (0, 0))
    others = [k for k in range(7) if k != route]
    assert np.array_equal(h[others], h0[others]), name


def test_oracle_meets_the_reference_expectations(oracle_module):
    for name, h0, s0, o, route, exp in cases(oracle_module):
        h, s, r = oracle_module.queue_compaction(h0, s0, o, route)
        check_expectations(name, h0, s0, o, route, exp, h, s, r)


def test_planted_unstable_order_is_caught(oracle_module):
    """A compaction that appended the selected slots in reverse order would fail the stable-order
    expectation above (the check is sensitive to order, not only to the set)."""
    h0, s0 = make_state(oracle_module, 33, RAY)
    o = make_outcomes(33, "alternating", RAY)
    h, s, r = oracle_module.queue_compaction(h0, s0, o, RAY)
    s_wrong = s.copy()
    s_wrong[33:33 + 17] = s_wrong[33:33 + 17][::-1]
    with pytest.raises(AssertionError):
        check_expectations("planted", h0, s0, o, RAY, (0, 0, 17, 17, 0), h, s_wrong, r)


def test_launcher_rejects_bad_arguments_before_any_launch(hip, abi):
    lib = hip.library()
    fake = C.c_void_p(0x10000)  # never dereferenced: every case below is rejected on the host
    view = abi.QueueSoa(0x10000, 0x20000, 7, 257)
    need = hip.queue_compaction_scratch_bytes(257)
    assert need > 0 and need % 256 == 0
    assert hip.queue_compaction_scratch_bytes(0) == 0
    call = lib.blackframe_hip_launch_wavefront_queue_compaction
    assert call(view, fake, 257, RAY, C.c_void_p(0x40000), need - 1, fake, None) == 1  # short scratch
    for bad_route in (0, TERMINATED):
        assert call(view, fake, 257, bad_route, C.c_void_p(0x40000), need, fake, None) == 1
    assert call(view, fake, 258, RAY, C.c_void_p(0x40000), need + 4096, fake, None) == 1  # input > capacity
    assert call(view, fake, 257, RAY, C.c_void_p(0x40000 + 4), need, fake, None) == 1  # misaligned scratch
    assert call(view, fake, 257, RAY, C.c_void_p(0x40000), need, None, None) == 1  # no result
    assert call(abi.QueueSoa(0x10000, 0x20000, 6, 257), fake, 1, RAY, C.c_void_p(0x40000), need, fake, None) == 1


@pytest.mark.gpu
def test_device_compaction_equals_the_oracle(hip, oracle_module):
    for name, h0, s0, o, route, exp in cases(oracle_module):
        h, s, r = hip.queue_compaction(h0, s0, o, route)
        hw, sw, rw = oracle_module.queue_compaction(h0, s0, o, route)
        assert np.array_equal(r, rw), (name, r, rw)
        assert np.array_equal(h, hw), name
        assert np.array_equal(s, sw), name
        check_expectations(name, h0, s0, o, route, exp, h, s, r)


@pytest.mark.gpu
def test_capacity_sized_scratch_serves_a_smaller_input(hip, oracle_module):
    h0, s0 = make_state(oracle_module, 33, HIT)
    o = make_outcomes(33, "alternating", HIT)
    big = hip.queue_compaction_scratch_bytes(257)
    h, s, r = hip.queue_compaction(h0, s0, o, HIT, scratch_bytes=big)
    check_expectations("bigger-scratch", h0, s0, o, HIT, (0, 0, 17, 17, 0), h, s, r)


@pytest.mark.gpu
def test_compaction_replays_bit_exactly(hip, oracle_module):
    o = make_outcomes(257, "sparse", CONTINUATION)
    h0, s0 = make_state(oracle_module, 300, CONTINUATION, (4, 2, 9))
    first = hip.queue_compaction(h0, s0, o, CONTINUATION)
    second = hip.queue_compaction(h0, s0, o, CONTINUATION)
    for a, b in zip(first, second):
        assert np.array_equal(a, b)
