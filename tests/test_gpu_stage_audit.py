"""blackframe_hip_launch_wavefront_audit_stage on the GPU against a restatement of the reference's
audit (CudaWavefrontStages.cu:4945-5033): an outcome fails on a non-success status, an unknown
route, a route outside the allowed mask or a path slot at or above the capacity; the audit keeps the
smallest failing index and that outcome, and for the shade stage the closure- and light-sample
detail bits. Seeded outcome streams of 0 to 300 000 entries with failures at the ends, in the
middle and nowhere; an invalid stage kind is rejected before any launch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def expected_audit(abi, outcomes, mask, capacity, kind):
    status, route, slot, detail = (outcomes[:, k].astype(np.uint64) for k in range(4))
    failed = (status != 0) | (route > abi.ROUTE_TERMINATED) | (slot >= capacity)
    known = route <= abi.ROUTE_TERMINATED
    failed |= known & ((np.uint64(mask) >> np.where(known, route, 0)) & 1 == 0)
    idx = np.nonzero(failed)[0]
    first = int(idx[0]) if idx.size else 0xFFFFFFFF
    shade = kind == abi.STAGE_KIND_SHADE
    closure = int(((detail & abi.SHADE_DETAIL_CLOSURE_SAMPLED) != 0).sum()) if shade else 0
    light = int(((detail & abi.SHADE_DETAIL_LIGHT_SAMPLED) != 0).sum()) if shade else 0
    return first, closure, light


@pytest.mark.parametrize("n,where,kind", [(0, None, 6), (1, 0, 6), (1000, None, 6), (300000, 299999, 6),
                                          (300000, 150001, 9), (65536, 64, 3), (257, None, 4)])
def test_audit_matches_the_restatement(abi, hip, n, where, kind):
    rng = np.random.default_rng(n + (where or 0))
    capacity = 1 << 20
    mask = (1 << abi.ROUTE_SHADE) | (1 << abi.ROUTE_TERMINATED) | (1 << abi.ROUTE_SHADOW)
    outcomes = np.zeros((n, 4), dtype=np.uint32)
    if n:
        outcomes[:, 1] = rng.choice([abi.ROUTE_SHADE, abi.ROUTE_TERMINATED, abi.ROUTE_SHADOW], n)
        outcomes[:, 2] = rng.integers(0, capacity, n)
        outcomes[:, 3] = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
    if where is not None:
        failure = where % 4
        if failure == 0:
            outcomes[where, 0] = 7          # numerical failure
        elif failure == 1:
            outcomes[where, 1] = 12         # unknown route
        elif failure == 2:
            outcomes[where, 1] = abi.ROUTE_HIT  # not allowed
        else:
            outcomes[where, 2] = capacity   # slot out of range
        outcomes[where + 1:where + 40, 0] = 3  # later failures do not move the first
    got = hip.audit_stage(outcomes, mask, capacity, kind)
    first, closure, light = expected_audit(abi, outcomes, mask, capacity, kind)
    assert (got.abi_major, got.abi_minor, got.struct_size) == (1, 0, 64)
    assert got.stage_kind == kind and got.expected_work_count == n and got.inspected_work_count == n
    assert got.first_failure_work_index == first
    assert (got.closure_samples, got.light_samples) == (closure, light)
    want_first = list(outcomes[first]) if first < n else [0, 0, 0, 0]
    assert list(got.first_failure) == want_first and list(got.reserved) == [0, 0, 0, 0]


def test_invalid_stage_kind_is_rejected(abi, hip):
    with pytest.raises(hip.XpuError):
        hip.audit_stage(np.zeros((4, 4), dtype=np.uint32), 0xFF, 16, 10)
