"""The allocation-free compaction entry point (blackframe_hip_launch_wavefront_queue_compaction_static)
against the caller-scratch one (blackframe_hip_launch_wavefront_queue_compaction), both called
through the C ABI.

The library-owned scratch keeps whatever the previous call left in it, so the GPU test runs two
calls in a row, the second on a smaller input, and each must equal the caller-scratch launcher bit
for bit (queue slots, headers and the result record). Inputs above the fixed maximum are rejected
on the host before any launch.
"""
import ctypes as C

import numpy as np
import pytest

from test_queue_compaction import CONTINUATION, HIT, RAY, check_expectations, expected_slots, make_outcomes, make_state


def test_static_launcher_rejects_inputs_beyond_its_scratch(hip, abi):
    lib = hip.library()
    fake = C.c_void_p(0x10000)  # never dereferenced: every case below is rejected on the host
    too_many = abi.STATIC_COMPACTION_MAX_INPUT + 1
    view = abi.QueueSoa(0x10000, 0x20000, 7, too_many)
    call = lib.blackframe_hip_launch_wavefront_queue_compaction_static
    assert call(view, fake, too_many, RAY, fake, None) == 1
    assert b"library-owned scratch" in lib.blackframe_hip_last_error()
    assert call(view, fake, 1, 0, fake, None) == 1  # route 0 is not compactable
    assert call(view, fake, 1, RAY, None, None) == 1  # no result


@pytest.mark.gpu
def test_static_scratch_equals_caller_scratch_on_consecutive_calls(hip, oracle_module):
    """A large sparse append, then a small alternating one: the second call runs on scratch the
    first left behind and must still match the caller-scratch launcher."""
    cases = []
    big = make_outcomes(300_000, "sparse", HIT)
    h, s = make_state(oracle_module, 300_000, HIT)
    cases.append(("large-sparse", h, s, big, HIT))
    small = make_outcomes(33, "alternating", CONTINUATION)
    h, s = make_state(oracle_module, 64, CONTINUATION, (4, 2, 9))
    cases.append(("small-after-large", h, s, small, CONTINUATION))
    for name, h0, s0, o, route in cases:
        print(f"case {name}: {len(o)} outcomes, route {route}", flush=True)
        got = hip.queue_compaction(h0, s0, o, route, library_scratch=True)
        want = hip.queue_compaction(h0, s0, o, route)
        for a, b in zip(got, want):
            assert np.array_equal(a, b), name
        sel = len(expected_slots(o, route))
        initial = int(h0[route, 4])
        check_expectations(name, h0, s0, o, route, (0, initial, sel, sel, 0), *got)
