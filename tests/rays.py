"""Seeded ray sets for traversal parity (host only)."""
import numpy as np


def random_rays(abi, n, seed, center=(0.0, 0.0, 1.0), radius=3.0, t_max=float("inf"), grazing=0.0):
    rng = np.random.default_rng(seed)
    rays = (abi.TransportRay * n)()
    origins = np.asarray(center) + rng.uniform(-radius, radius, size=(n, 3))
    targets = np.asarray(center) + rng.uniform(-1.0, 1.0, size=(n, 3))
    directions = targets - origins
    if grazing:
        # snap a fraction of directions onto axis-aligned planes to exercise zero components and edges
        mask = rng.uniform(size=(n, 3)) < grazing
        directions[mask] = 0.0
    for i in range(n):
        d = directions[i]
        if not np.any(d):
            d = np.array([0.0, 0.0, -1.0])
        d = d / np.linalg.norm(d)
        r = rays[i]
        r.origin_x, r.origin_y, r.origin_z = origins[i]
        r.direction_x, r.direction_y, r.direction_z = d
        r.t_min = 0.0
        r.t_max = t_max
        r.time = 0.5
        r.visibility_mask = 0xFFFFFFFF
    return rays


def result_bytes(results):
    return bytes(memoryview(results).cast("B"))
