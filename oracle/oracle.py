"""ORACLE — test infrastructure only: ctypes binding of oracle/liboracle.so (the C restatement).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import this module.
"""

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"


def build(force=False):
    if force or not LIB_PATH.exists():
        subprocess.run(["make", "-s", "-C", str(HERE), "liboracle.so"] + (["-B"] if force else []), check=True)


def load(abi):
    if not LIB_PATH.exists():
        build()
    lib = C.CDLL(str(LIB_PATH))
    lib.bfo_mix_bits.restype = C.c_uint64
    lib.bfo_mix_bits.argtypes = [C.c_uint64]
    lib.bfo_indexed_bits.restype = C.c_uint64
    lib.bfo_indexed_bits.argtypes = [C.POINTER(abi.SampleStreamIndex), C.c_uint64]
    lib.bfo_sobol_bits.restype = C.c_uint64
    lib.bfo_sobol_bits.argtypes = [C.c_uint64, C.c_uint64]
    lib.bfo_owen_bits.restype = C.c_uint64
    lib.bfo_owen_bits.argtypes = [C.POINTER(abi.SampleStreamIndex), C.c_uint64]
    lib.bfo_sample_1d.restype = C.c_float
    lib.bfo_sample_1d.argtypes = [C.c_uint32, C.POINTER(abi.SampleStreamIndex), C.c_uint64]
    lib.bfo_sobol_dimension_count.restype = C.c_uint64
    lib.bfo_cie_y_integral.restype = C.c_double
    lib.bfo_cie_fingerprint.restype = C.c_uint64
    lib.bfo_cie_xyz.argtypes = [C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_float, C.POINTER(C.c_float)]
    lib.bfo_xyz_to_rgb.argtypes = [C.POINTER(C.c_float), C.POINTER(C.c_float)]
    lib.bfo_scene_build.argtypes = [C.POINTER(abi.SceneDesc), C.POINTER(C.c_void_p)]
    fp = C.POINTER(C.c_float)
    lib.bfo_map_filter.argtypes = [C.POINTER(abi.ImageTextureDesc), C.c_float, C.c_float, fp, C.c_uint32, C.c_uint32,
                                   C.c_uint32, C.c_uint32, C.c_uint32, fp]
    lib.bfo_map_cone_differentials.argtypes = [fp, fp, fp, fp, fp, fp]
    lib.bfo_map_apply.argtypes = [C.c_void_p, C.c_uint32, fp, fp, fp, fp, fp, fp, fp]
    for name in ("bfo_expf", "bfo_expm1f", "bfo_logf"):
        getattr(lib, name).restype = C.c_float
        getattr(lib, name).argtypes = [C.c_float]
    lib.bfo_scene_free.argtypes = [C.c_void_p]
    lib.bfo_scene_bvh.argtypes = [C.c_void_p, C.POINTER(C.c_uint64), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_void_p]
    lib.bfo_scene_tlas_root.restype = C.c_uint32
    lib.bfo_scene_tlas_root.argtypes = [C.c_void_p]
    lib.bfo_render_scalar.argtypes = [C.c_void_p, C.POINTER(abi.FrameDesc), C.POINTER(abi.TransportOptions), C.c_uint64,
                                      C.c_uint64, C.c_void_p, C.POINTER(abi.TransportReport), C.c_void_p]
    lib.bfo_light_power_cdf.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p]
    lib.bfo_scene_set_light_powers.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32]
    lib.bfo_light_cdf_sample.argtypes = [C.c_void_p, C.c_uint32, C.c_float, C.POINTER(C.c_uint32), C.POINTER(C.c_float)]
    lib.bfo_light_cdf_probability.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(C.c_float)]
    lib.bfo_scene_light_count.restype = C.c_uint32
    lib.bfo_scene_light_count.argtypes = [C.c_void_p]
    lib.bfo_scene_light.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                    C.POINTER(C.c_float), C.c_void_p, C.c_void_p, C.c_void_p]
    lib.bfo_triangle_area.argtypes = [C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float)]
    lib.bfo_closest_hit.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]
    lib.bfo_occlusion.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]
    lib.bfo_closest_hit_brute_force.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]
    lib.bfo_camera_ray.argtypes = [C.POINTER(abi.CameraDesc), C.c_uint32, C.POINTER(abi.SampleStreamIndex),
                                   C.POINTER(abi.TransportRay)]
    lib.bfo_camera_cone.argtypes = [C.POINTER(abi.CameraDesc), C.c_uint32, C.POINTER(abi.SampleStreamIndex),
                                    C.POINTER(abi.TransportRay), C.POINTER(abi.RayCone)]
    lib.bfo_render.argtypes = [C.c_void_p, C.POINTER(abi.FrameDesc), C.POINTER(abi.TransportOptions), C.c_uint64,
                               C.c_uint64, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                               C.POINTER(abi.TransportReport), C.c_void_p]
    f3 = C.POINTER(C.c_float)
    rec = C.POINTER(abi.ClosureRecord)
    lib.bfo_closure_eval.argtypes = [rec, C.c_uint32, f3, f3, f3]
    lib.bfo_closure_pdf.argtypes = [rec, C.c_uint32, f3, f3, f3]
    lib.bfo_closure_sample.argtypes = [rec, C.c_uint32, f3, C.c_float, C.c_float, C.c_float, f3, C.POINTER(C.c_uint32)]
    lib.bfo_dielectric_fresnel.argtypes = [C.c_float, C.c_float, C.c_float, f3]
    lib.bfo_conductor_fresnel.argtypes = [C.c_float, C.c_float, C.c_float, f3]
    for name in ("bfo_ggx_d", "bfo_smith_g1", "bfo_smith_lambda"):
        getattr(lib, name).argtypes = [C.c_float, C.c_float, f3, f3]
    lib.bfo_smith_g2.argtypes = [C.c_float, C.c_float, f3, f3, f3]
    lib.bfo_vndf_pdf.argtypes = [C.c_float, C.c_float, f3, f3, f3]
    lib.bfo_sample_vndf.argtypes = [C.c_float, C.c_float, f3, C.c_float, C.c_float, f3, f3]
    lib.bfo_ggx_widths_representable.argtypes = [C.c_float, C.c_float]
    lib.bfo_mis_weight.argtypes = [C.c_uint32, C.c_float, C.c_float, f3]
    lib.bfo_russian_roulette.argtypes = [f3, C.c_float, C.c_float, C.c_float, C.c_float, C.POINTER(C.c_int)]
    lib.bfo_sample_wavelengths.argtypes = [C.c_float, f3]
    lib.bfo_hypotf.restype = C.c_float
    lib.bfo_hypotf.argtypes = [C.c_float, C.c_float]
    lib.bfo_atan2f.restype = C.c_float
    lib.bfo_atan2f.argtypes = [C.c_float, C.c_float]
    lib.bfo_sincosf.argtypes = [C.c_float, f3, f3]
    lib.bfo_render_taped.argtypes = [C.c_void_p, C.POINTER(abi.FrameDesc), C.POINTER(abi.TransportOptions), C.c_uint64,
                                     C.c_uint64, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.bfo_bounce_dimensions.argtypes = [C.c_uint64, C.POINTER(C.c_uint64)]
    lib.bfo_render_survival.argtypes = [C.c_void_p, C.POINTER(abi.FrameDesc), C.POINTER(abi.TransportOptions), C.c_uint64,
                                        C.c_uint64, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_void_p, C.c_void_p, C.c_void_p]
    lib.bfo_render_paths.argtypes = [C.c_void_p, C.POINTER(abi.TransportOptions), C.c_uint64, C.c_uint32] + [C.c_void_p] * 8 + \
        [C.POINTER(abi.TransportReport)]
    lib.bfo_render_adjoint.argtypes = [C.c_void_p, C.POINTER(abi.FrameDesc), C.POINTER(abi.TransportOptions), C.c_uint64,
                                       C.c_uint64, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    d2 = C.POINTER(C.c_double)
    lib.bfo_closure_alpha_log_gradient.argtypes = [rec, f3, f3, d2]
    lib.bfo_closure_weight_derivative.argtypes = [rec, C.c_uint32, f3, f3, d2]
    lib.bfo_cie_cmf64.argtypes = [C.c_float, d2]
    lib.bfo_texture_sample.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p, C.c_float, C.c_float, f3,
                                       C.POINTER(C.c_uint32), f3]
    lib.bfo_texel_reflectance.restype = C.c_float
    lib.bfo_texel_reflectance.argtypes = [f3, C.c_uint32, C.c_float]
    return lib


# ---- scene blob hash: Blackframe/XPU/CUDA/Source/CudaSceneSoaHash.cu:13-36 -----------------------
def fnv1a64(data, zero_begin=None, zero_end=None):
    """FNV-1a-64 (offset basis 14695981039346656037, prime 1099511628211) of the bytes, those in
    [zero_begin, zero_end) read as zero."""
    h = 14695981039346656037
    for i, b in enumerate(bytes(data)):
        if zero_begin is not None and zero_begin <= i < zero_end:
            b = 0
        h ^= b
        h = (h * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


# ---- queue compaction: Blackframe/XPU/CUDA/Source/CudaWavefrontQueueCompaction.cu:55-330 ---------
QUEUE_COUNT = 7


def queue_headers(capacity, sizes=None):
    """Canonical QueueHeader words [7, 8] (CudaWavefrontQueues::create): abi 1.0, struct size 32,
    kind k, capacity, size, zero overflow / rejected / reserved."""
    h = np.zeros((QUEUE_COUNT, 8), dtype=np.uint32)
    for k in range(QUEUE_COUNT):
        h[k] = [1 | (0 << 16), 32, k, capacity, 0 if sizes is None else sizes[k], 0, 0, 0]
    return h


def queue_compaction(headers, slots, outcomes, route):
    """The compaction contract on host arrays (words as in the device records). Returns new
    (headers, slots, result[12]); arguments the launcher rejects synchronously are not modelled."""
    headers = np.array(headers, dtype=np.uint32).reshape(QUEUE_COUNT, 8)
    slots = np.array(slots, dtype=np.uint32)
    outcomes = np.asarray(outcomes, dtype=np.uint32).reshape(-1, 4)
    stride = slots.size // QUEUE_COUNT
    n = len(outcomes)

    def header_ok(h, kind):
        return (h[0] & 0xFFFF) == 1 and (h[0] >> 16) == 0 and h[1] == 32 and h[2] == kind and h[3] == stride and h[7] == 0

    result = np.zeros(12, dtype=np.uint32)
    h = headers[route]
    result[1] = result[2] = route
    result[3] = n
    result[4] = h[4]
    status = 0
    if not header_ok(h, route) or h[4] > h[3] or h[5] != 0 or h[6] != 0:
        status = 1
    if n == 0:
        result[0] = status
        return headers, slots, result
    selected = (outcomes[:, 0] == 0) & (outcomes[:, 1] == route)
    if np.any(selected & (outcomes[:, 2] >= stride)):
        status = 1
    count = int(selected.sum())
    result[5] = count
    if status == 0:
        if count > int(h[3]) - int(h[4]):
            headers[route, 5] = min(0xFFFFFFFF, int(headers[route, 5]) + count)
            headers[route, 6] = min(0xFFFFFFFF, int(headers[route, 6]) + count)
            status = 2
            result[7] = count
        else:
            begin = route * stride + int(h[4])
            slots[begin:begin + count] = outcomes[selected, 2]
            headers[route, 4] += count
            result[6] = count
    result[0] = status
    return headers, slots, result


# ---- texture fitting: no reference counterpart (the reference renders forward only) -----------
# These restate the binary32 operation order of blackframe_hip_film_l2_loss and
# blackframe_hip_adam_step. They are pinned by torch's own implementations (torch.nn.functional
# .mse_loss and torch.optim.Adam) within a few ulps, and by closed forms (tests/test_fit_oracle.py).


def film_l2_loss(film, counts, target, scale):
    """film, target: float32 [pixels * 3]; counts: uint32 [pixels]. Returns (loss as float64,
    film gradient as float32) with d = film / max(count, 1) - target:
    loss = scale * sum(d^2) (binary64), gradient = ((2 * scale) * d) / max(count, 1) (binary32)."""
    scale = float(np.float32(scale))  # the ABI takes the scale as a binary32
    c = np.repeat(np.maximum(np.asarray(counts, dtype=np.uint32), 1).astype(np.float32), 3)
    d = np.asarray(film, dtype=np.float32) / c - np.asarray(target, dtype=np.float32)
    two_scale = np.float32(2.0 * scale)
    gradient = (two_scale * d) / c
    dd = d.astype(np.float64)
    return scale * float(np.sum(dd * dd)), gradient


def adam_step(value, gradient, first, second, learning_rate, step, beta1=0.9, beta2=0.999, epsilon=1e-8,
              lower=-np.inf, upper=np.inf):
    """One projected Adam step on float32 arrays; returns new (value, first, second)."""
    f32 = np.float32
    beta1, beta2 = float(beta1), float(beta2)
    b1, b2 = f32(beta1), f32(beta2)
    bias1 = 1.0 - beta1 ** step
    bias2 = 1.0 - beta2 ** step
    omb1, omb2 = f32(1.0 - beta1), f32(1.0 - beta2)
    step_size, bias2_sqrt = f32(float(f32(learning_rate)) / bias1), f32(np.sqrt(bias2))
    g = np.asarray(gradient, dtype=np.float32)
    m = b1 * np.asarray(first, dtype=np.float32) + omb1 * g
    v = b2 * np.asarray(second, dtype=np.float32) + omb2 * (g * g)
    den = np.sqrt(v) / bias2_sqrt + f32(epsilon)
    p = np.asarray(value, dtype=np.float32) - step_size * (m / den)
    p = np.minimum(np.maximum(p, f32(lower)), f32(upper))
    return p.astype(np.float32), m, v


class Oracle:
    def __init__(self, abi):
        self.abi = abi
        self.lib = load(abi)

    def map_filter(self, levels, u, v, differentials, channel=0, u_wrap=0, v_wrap=0, anisotropy=8, visits=1 << 20):
        """EWA filter of one channel of an image given as its mip levels ((h, w, c) arrays):
        (status, value)."""
        texels = np.ascontiguousarray(np.concatenate([np.asarray(l, np.float32).reshape(-1) for l in levels]))
        h, w = np.shape(levels[0])[:2]
        desc = self.abi.ImageTextureDesc(w, h, np.shape(levels[0])[2], len(levels), texels.ctypes.data, texels.size)
        out = C.c_float()
        st = self.lib.bfo_map_filter(C.byref(desc), float(u), float(v), (C.c_float * 4)(*differentials), channel, u_wrap,
                                     v_wrap, anisotropy, visits, C.byref(out))
        return st, out.value

    def cone_differentials(self, cone, direction, ng, dpdu, dpdv):
        f3 = lambda v: (C.c_float * 3)(*[float(x) for x in v])
        out = (C.c_float * 4)()
        ok = self.lib.bfo_map_cone_differentials((C.c_float * 2)(*cone), f3(direction), f3(ng), f3(dpdu), f3(dpdv), out)
        return list(out) if ok else None

    def power_sampler(self, powers):
        return PowerLightSampler(self, powers)

    def scene(self, scene):
        handle = C.c_void_p()
        status = self.lib.bfo_scene_build(C.byref(scene.desc), C.byref(handle))
        if status != 0:
            raise RuntimeError(f"oracle scene build failed with status {status}")
        return OracleScene(self, handle)


class PowerLightSampler:
    """The reference's power-weighted LightSampler restated (liboracle): boundaries from four-lane
    spectral powers, sample(u) -> (index, probability), probability(i)."""

    def __init__(self, oracle, powers):
        self.lib = oracle.lib
        self.powers = np.ascontiguousarray(powers, dtype=np.float32).reshape(-1, 4)
        self.count = len(self.powers)
        self.cdf = np.zeros(self.count + 1, dtype=np.float32)
        self.status = self.lib.bfo_light_power_cdf(self.powers.ctypes.data, self.count, self.cdf.ctypes.data)

    def sample(self, u):
        index, probability = C.c_uint32(), C.c_float()
        if not self.lib.bfo_light_cdf_sample(self.cdf.ctypes.data, self.count, float(np.float32(u)), C.byref(index),
                                             C.byref(probability)):
            return None
        return index.value, probability.value

    def probability(self, i):
        probability = C.c_float()
        if not self.lib.bfo_light_cdf_probability(self.cdf.ctypes.data, self.count, i, C.byref(probability)):
            return None
        return probability.value


class OracleScene:
    def __init__(self, oracle, handle):
        self.oracle = oracle
        self.lib = oracle.lib
        self.abi = oracle.abi
        self.handle = handle

    def __del__(self):
        if getattr(self, "handle", None):
            self.lib.bfo_scene_free(self.handle)
            self.handle = None

    def apply_maps(self, material, cone, direction, uv, ng, dpdu, dpdv, ns):
        """A material's normal / bump maps applied to shading normal ns: (status, new ns)."""
        f3 = lambda v: (C.c_float * 3)(*[float(x) for x in v])
        n = f3(ns)
        st = self.lib.bfo_map_apply(self.handle, material, (C.c_float * 2)(*cone), f3(direction),
                                    (C.c_float * 2)(*uv), f3(ng), f3(dpdu), f3(dpdv), n)
        return st, np.array(list(n), dtype=np.float32)

    def render_scalar(self, frame, options, path_begin=0, path_count=None):
        """The scalar per-path integrator (one path at a time, brute-force scene queries checked against
        the BVH): returns (PathResult array, report, per-path count of near-coplanar ties)."""
        cam = frame.camera
        n = path_count if path_count is not None else cam.width * cam.height * frame.samples_per_pixel
        paths = (self.abi.PathResult * n)()
        report = self.abi.TransportReport()
        ties = np.zeros(n, dtype=np.uint32)
        status = self.lib.bfo_render_scalar(self.handle, C.byref(frame), C.byref(options), path_begin, n,
                                            C.addressof(paths), C.byref(report), ties.ctypes.data)
        if status != 0:
            raise RuntimeError(f"oracle scalar render failed with status {status}")
        return paths, report, ties

    def light_count(self):
        """Mesh-area lights of the scene (punctual lights are registered before them)."""
        return self.lib.bfo_scene_light_count(self.handle)

    def set_light_powers(self, powers):
        """Power-weighted light selection from [lights][4] spectral powers (None: uniform)."""
        if powers is None:
            status = self.lib.bfo_scene_set_light_powers(self.handle, None, 0)
        else:
            p = np.ascontiguousarray(powers, dtype=np.float32).reshape(-1, 4)
            status = self.lib.bfo_scene_set_light_powers(self.handle, p.ctypes.data, len(p))
        if status != 0:
            raise RuntimeError(f"oracle light powers rejected with status {status}")
        return self

    def bvh(self):
        counts = (C.c_uint64 * 5)()
        self.lib.bfo_scene_bvh(self.handle, counts, None, None, None, None, None)
        abi = self.abi
        blases = (abi.BvhBlas * max(1, counts[0]))()
        blas_nodes = (abi.BvhNode * max(1, counts[1]))()
        prims = np.zeros(max(1, counts[2]), dtype=np.uint32)
        tlas = (abi.BvhNode * max(1, counts[3]))()
        irefs = (abi.BvhInstanceRef * max(1, counts[4]))()
        self.lib.bfo_scene_bvh(self.handle, counts, blases, blas_nodes, prims.ctypes.data, tlas, irefs)
        as_bytes = lambda arr, n, size: bytes(memoryview(arr).cast("B")[: n * size])
        return {
            "counts": list(counts),
            "blas": as_bytes(blases, counts[0], 32),
            "blas_nodes": as_bytes(blas_nodes, counts[1], 48),
            "primitive_refs": prims[: counts[2]].tobytes(),
            "tlas_nodes": as_bytes(tlas, counts[3], 48),
            "instance_refs": as_bytes(irefs, counts[4], 32),
            "tlas_root": self.lib.bfo_scene_tlas_root(self.handle),
        }

    def closest_hit(self, rays, counters=False):
        n = len(rays)
        out = (self.abi.ClosestHitResult * n)()
        ctr = np.zeros(3 * n, dtype=np.uint32) if counters else None
        self.lib.bfo_closest_hit(self.handle, C.addressof(rays), n, C.addressof(out),
                                 ctr.ctypes.data if counters else None)
        return (out, ctr.reshape(n, 3)) if counters else out

    def occlusion(self, rays):
        n = len(rays)
        out = (self.abi.OcclusionResult * n)()
        self.lib.bfo_occlusion(self.handle, C.addressof(rays), n, C.addressof(out), None)
        return out

    def closest_hit_brute_force(self, rays):
        n = len(rays)
        out = (self.abi.ClosestHitResult * n)()
        self.lib.bfo_closest_hit_brute_force(self.handle, C.addressof(rays), n, C.addressof(out))
        return out

    def render(self, frame, options, path_begin=0, path_count=None, threads=1, want_paths=True, want_film=True):
        cam = frame.camera
        total = cam.width * cam.height * frame.samples_per_pixel
        if path_count is None:
            path_count = total - path_begin
        paths = (self.abi.PathResult * path_count)() if want_paths else None
        pixels = cam.width * cam.height
        film = np.zeros(pixels * 3, dtype=np.float32) if want_film else None
        film_c = np.zeros(pixels * 6, dtype=np.float64) if want_film else None
        counts = np.zeros(pixels, dtype=np.uint32) if want_film else None
        report = self.abi.TransportReport()
        tc = np.zeros(6, dtype=np.uint64)
        status = self.lib.bfo_render(self.handle, C.byref(frame), C.byref(options), path_begin, path_count, threads,
                                     C.addressof(paths) if want_paths else None,
                                     film.ctypes.data if want_film else None,
                                     film_c.ctypes.data if want_film else None,
                                     counts.ctypes.data if want_film else None, C.byref(report), tc.ctypes.data)
        if status != 0:
            raise RuntimeError(f"oracle render failed with status {status}")
        result = {"report": report, "traversal_counters": tc}
        if want_paths:
            result["paths"] = paths
        if want_film:
            result["film"] = film.reshape(cam.height, cam.width, 3)
            comp = film_c.reshape(cam.height, cam.width, 3, 2)
            result["film_compensated"] = comp[..., 0] + comp[..., 1]
            result["film_counts"] = counts.reshape(cam.height, cam.width)
        return result

    def render_taped(self, frame, options, threads=1):
        """Forward render of the whole frame plus its adjoint tape (radiance, visibility bits)."""
        cam = frame.camera
        pixels = cam.width * cam.height
        total = pixels * frame.samples_per_pixel
        film = np.zeros(pixels * 3, dtype=np.float32)
        film_c = np.zeros(pixels * 6, dtype=np.float64)
        counts = np.zeros(pixels, dtype=np.uint32)
        radiance = np.zeros(total * 4, dtype=np.float32)
        visibility = np.zeros(total, dtype=np.uint64)
        status = self.lib.bfo_render_taped(self.handle, C.byref(frame), C.byref(options), 0, total, threads,
                                           film.ctypes.data, film_c.ctypes.data, counts.ctypes.data,
                                           radiance.ctypes.data, visibility.ctypes.data)
        if status != 0:
            raise RuntimeError(f"oracle taped render failed with status {status}")
        comp = film_c.reshape(cam.height, cam.width, 3, 2)
        return {"film": film.reshape(cam.height, cam.width, 3), "film_compensated": comp[..., 0] + comp[..., 1],
                "film_counts": counts.reshape(cam.height, cam.width), "tape_radiance": radiance,
                "tape_visibility": visibility}

    def render_survival(self, frame, options, survival, replay, taped=False, threads=1):
        """Forward render of the whole frame whose Russian-roulette survival probabilities are
        recorded into (replay=False) or decided by (replay=True) `survival`, a float32 array of
        paths * (diffuse + glossy + specular limits + 1) entries; taped=True adds the adjoint tape."""
        cam = frame.camera
        pixels = cam.width * cam.height
        total = pixels * frame.samples_per_pixel
        film = np.zeros(pixels * 3, dtype=np.float32)
        film_c = np.zeros(pixels * 6, dtype=np.float64)
        counts = np.zeros(pixels, dtype=np.uint32)
        abi = self.abi
        paths = (abi.PathResult * total)()
        radiance = np.zeros(total * 4, dtype=np.float32) if taped else None
        visibility = np.zeros(total, dtype=np.uint64) if taped else None
        assert survival.dtype == np.float32 and survival.flags.c_contiguous
        status = self.lib.bfo_render_survival(self.handle, C.byref(frame), C.byref(options), 0, total, threads,
                                              abi.SURVIVAL_REPLAY if replay else abi.SURVIVAL_RECORD,
                                              survival.ctypes.data, paths, film.ctypes.data, film_c.ctypes.data,
                                              counts.ctypes.data, radiance.ctypes.data if taped else None,
                                              visibility.ctypes.data if taped else None)
        if status != 0:
            raise RuntimeError(f"oracle survival render failed with status {status}")
        comp = film_c.reshape(cam.height, cam.width, 3, 2)
        out = {"film": film.reshape(cam.height, cam.width, 3), "film_compensated": comp[..., 0] + comp[..., 1],
               "film_counts": counts.reshape(cam.height, cam.width), "paths": paths}
        if taped:
            out.update(tape_radiance=radiance, tape_visibility=visibility)
        return out

    def render_paths(self, inputs, options, threads=1):
        """The path-input transport: inputs = (samples, rays, cones, states) ctypes arrays of equal
        length. Returns (states, rays, cones, controls, report)."""
        abi = self.abi
        samples, rays, cones, states = inputs
        n = len(rays)
        outs = ((abi.PathStateLane * n)(), (abi.TransportRay * n)(), (abi.RayCone * n)(), (abi.LaneControl * n)())
        report = abi.TransportReport()
        status = self.lib.bfo_render_paths(self.handle, C.byref(options), n, threads, C.addressof(samples), C.addressof(rays),
                                           C.addressof(cones), C.addressof(states), *[C.addressof(o) for o in outs],
                                           C.byref(report))
        if status != 0:
            raise RuntimeError(f"oracle path-input render failed with status {status}")
        return (*outs, report)

    def adjoint(self, frame, options, tape, film_gradient, texture_shapes, threads=1):
        """Binary64 texel gradients of sum(film_gradient * film) (film = RGB sums), one array per
        texture shape (None: not wanted)."""
        cam = frame.camera
        total = cam.width * cam.height * frame.samples_per_pixel
        grads = [np.zeros(shape, dtype=np.float64) if shape is not None else None for shape in texture_shapes]
        table = (C.c_void_p * 8)(*[g.ctypes.data if g is not None else None for g in grads] + [None] * (8 - len(grads)))
        fg = np.ascontiguousarray(film_gradient, dtype=np.float32).ravel()
        status = self.lib.bfo_render_adjoint(self.handle, C.byref(frame), C.byref(options), 0, total, threads,
                                             tape["tape_radiance"].ctypes.data, tape["tape_visibility"].ctypes.data,
                                             fg.ctypes.data, table)
        if status != 0:
            raise RuntimeError(f"oracle adjoint failed with status {status}")
        return grads
