"""Python host mirror of the HIP backend (ctypes over libblackframe_hip.so; torch for device memory).

The names follow the reference's C++ API for the CUDA wavefront path
(Blackframe/Backends/GPU/CUDA/Include/Blackframe/Backends/GPU/CUDA/WavefrontTransport.hpp:147-208):
XpuSceneSoA/XpuSceneBvh collapse into XpuScene (scene upload + BVH build), XpuWavefrontTransportWorkspace
owns the fixed-capacity device scratch, and trace_xpu_wavefront_transport runs one frame. Errors raise
XpuError carrying the library status and message; there is no fallback path.
"""

import ctypes as C
from pathlib import Path

import numpy as np

from . import abi

LIB_PATH = Path(__file__).resolve().parent / "libblackframe_hip.so"

STATUS_NAMES = {0: "ok", 1: "invalid_argument", 2: "incompatible", 3: "resource_exhausted", 4: "unavailable",
                5: "device_error", 6: "internal_error", 7: "transport_failure"}


class XpuError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


_lib = None


def library():
    """Loads the in-tree HIP library; raises if it has not been built.

    torch is imported first so that the library binds to the HIP runtime torch already loaded
    (both carry the SONAME libamdhip64.so.7): one runtime per process, shared device memory and
    streams."""
    global _lib
    if _lib is not None:
        return _lib
    import torch  # noqa: F401
    if not LIB_PATH.exists():
        raise XpuError(4, f"{LIB_PATH} is missing; run __graft_entry__.build()")
    lib = C.CDLL(str(LIB_PATH))
    vp, u32, u64 = C.c_void_p, C.c_uint32, C.c_uint64
    lib.blackframe_hip_last_error.restype = C.c_char_p
    lib.blackframe_hip_device_count.argtypes = [C.POINTER(u32)]
    lib.blackframe_hip_scene_create.argtypes = [C.POINTER(abi.SceneDesc), C.POINTER(vp)]
    lib.blackframe_hip_scene_prepare.argtypes = [C.POINTER(abi.SceneDesc), C.POINTER(vp)]
    lib.blackframe_hip_scene_destroy.argtypes = [vp]
    lib.blackframe_hip_scene_bvh_export.argtypes = [vp, C.POINTER(u64), vp, vp, vp, vp, vp, C.POINTER(u32)]
    lib.blackframe_hip_scene_wide_export.argtypes = [vp, C.POINTER(u64), vp, vp]
    lib.blackframe_hip_scene_bvh_serialize.argtypes = [vp, vp, C.c_size_t, C.POINTER(C.c_size_t)]
    lib.blackframe_hip_scene_bvh_validate.argtypes = [vp, C.c_size_t, u32, C.POINTER(u32)]
    lib.blackframe_hip_launch_scene_closest_hit.argtypes = [vp, vp, u32, vp, vp, vp]
    lib.blackframe_hip_launch_scene_occlusion.argtypes = [vp, vp, u32, vp, vp]
    lib.blackframe_hip_launch_sample_dump.argtypes = [u32, vp, vp, u32, vp, vp]
    lib.blackframe_hip_launch_sample_stream_dump.argtypes = [u32, vp, u32, vp, vp]
    lib.blackframe_hip_workspace_create.argtypes = [u64, C.POINTER(vp)]
    lib.blackframe_hip_workspace_destroy.argtypes = [vp]
    lib.blackframe_hip_trace_wavefront_batch.argtypes = [vp, C.POINTER(abi.FrameJob), u32, C.POINTER(abi.TransportOptions),
                                                         C.POINTER(abi.TransportReport), vp, vp]
    lib.blackframe_hip_trace_wavefront_adjoint.argtypes = [vp, C.POINTER(abi.FrameJob), C.POINTER(abi.AdjointJob), u32,
                                                           C.POINTER(abi.TransportOptions), C.POINTER(abi.TransportReport),
                                                           vp]
    lib.blackframe_hip_autotune_launch.argtypes = [vp, C.POINTER(abi.FrameJob), u32, C.POINTER(abi.TransportOptions), u32,
                                                   C.POINTER(abi.LaunchTuning), vp]
    lib.blackframe_hip_film_l2_loss.argtypes = [C.POINTER(abi.FilmLossJob), u32, C.c_float, vp, vp]
    lib.blackframe_hip_query_wavefront_queue_compaction_scratch_bytes.argtypes = [u32, C.POINTER(C.c_size_t)]
    lib.blackframe_hip_query_transport_abi_layout.argtypes = [C.POINTER(abi.LayoutManifest), C.POINTER(C.c_int)]
    lib.blackframe_hip_launch_scene_soa_hash.argtypes = [vp, C.c_size_t, vp]
    lib.blackframe_hip_scene_soa_validate.argtypes = [vp, C.c_size_t, u32, C.POINTER(C.c_uint32)]
    lib.blackframe_hip_launch_scene_soa_validate.argtypes = [vp, C.c_size_t, vp]
    lib.blackframe_hip_scene_create_from_soa.argtypes = [vp, C.c_size_t, u32, C.POINTER(vp)]
    lib.blackframe_hip_launch_wavefront_queue_pushes.argtypes = [vp, vp, u32, u32, vp, u32, vp]
    lib.blackframe_hip_launch_wavefront_queue_compaction.argtypes = [abi.QueueSoa, vp, u32, u32, vp, C.c_size_t, vp, vp]
    lib.blackframe_hip_launch_wavefront_queue_compaction_static.argtypes = [abi.QueueSoa, vp, u32, u32, vp, vp]
    lib.blackframe_hip_adam_step.argtypes = [C.POINTER(abi.AdamParam), u32, C.POINTER(abi.AdamOptions), vp]
    lib.blackframe_hip_fit_graph_create.argtypes = [vp, C.POINTER(abi.FitStepDesc), C.POINTER(abi.TransportOptions),
                                                    C.POINTER(vp)]
    lib.blackframe_hip_fit_graph_launch.argtypes = [vp, vp]
    lib.blackframe_hip_fit_graph_report.argtypes = [vp, C.POINTER(abi.TransportReport), C.POINTER(abi.TransportReport),
                                                    C.POINTER(u32), vp]
    lib.blackframe_hip_fit_graph_destroy.argtypes = [vp]
    lib.blackframe_query_extension.restype = C.c_uint32
    lib.blackframe_query_extension.argtypes = [C.POINTER(abi.ExtensionQuery), C.POINTER(abi.ExtensionApi)]
    lib.blackframe_hip_trace_wavefront_transport.argtypes = [vp, vp, C.POINTER(abi.FrameDesc),
                                                             C.POINTER(abi.TransportOptions), vp, vp, vp, vp,
                                                             C.POINTER(abi.TransportReport), vp, vp]
    lib.blackframe_hip_launch_wavefront_audit_stage.argtypes = [vp, u32, u32, u32, u32, vp, vp]
    lib.blackframe_hip_launch_scene_closest_hit_blob.argtypes = [vp, C.c_size_t, vp, C.c_size_t, vp, u32, vp, vp]
    lib.blackframe_hip_launch_scene_occlusion_blob.argtypes = [vp, C.c_size_t, vp, C.c_size_t, vp, u32, vp, vp]
    lib.blackframe_hip_light_power_cdf.argtypes = [vp, u32, vp]
    lib.blackframe_hip_scene_set_light_powers.argtypes = [vp, vp, u32]
    lib.blackframe_hip_trace_wavefront_paths.argtypes = [vp, vp, C.POINTER(abi.PathInputs),
                                                         C.POINTER(abi.TransportOptions), C.POINTER(abi.PathOutputs),
                                                         C.POINTER(abi.TransportReport), vp]
    _lib = lib
    return lib


def _check(status):
    if status != 0:
        raise XpuError(status, library().blackframe_hip_last_error().decode())


def _stream_handle(stream=None):
    import torch

    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


class XpuScene:
    """Device scene + deterministic BVH (CudaSceneSoA::upload + CudaSceneBvh::build)."""

    def __init__(self, scene, upload=True):
        self._lib = library()
        self.handle = C.c_void_p()
        self.scene = scene  # keeps host arrays alive for the duration of the build
        # texel counts per texture slot (gradient buffers must match them)
        self.texture_sizes = [w * h * c for (w, h, c, _, _) in getattr(scene, "textures", [])]
        create = self._lib.blackframe_hip_scene_create if upload else self._lib.blackframe_hip_scene_prepare
        _check(create(C.byref(scene.desc), C.byref(self.handle)))

    @classmethod
    def from_soa(cls, blob, verify_hash=True):
        """A device scene from the reference's serialized scene blob ("BLKFSOA4" bytes):
        blackframe_hip_scene_create_from_soa. Render it in frame-packet mode at the blob's
        wavelength packet (soa_wavelengths)."""
        self = cls.__new__(cls)
        self._lib = library()
        self.handle = C.c_void_p()
        self.scene = None
        self.texture_sizes = []
        data = bytes(blob)
        _check(self._lib.blackframe_hip_scene_create_from_soa(data, len(data), abi.SCENE_SOA_VERIFY_HASH if verify_hash else 0,
                                                              C.byref(self.handle)))
        return self

    def set_light_powers(self, powers):
        """Power-weighted light selection (blackframe_hip_scene_set_light_powers) from [lights][4]
        spectral powers in registry order (punctual lights, then mesh-area lights); None returns
        the scene to uniform selection."""
        if powers is None:
            _check(self._lib.blackframe_hip_scene_set_light_powers(self.handle, None, 0))
        else:
            p = np.ascontiguousarray(powers, dtype=np.float32).reshape(-1, 4)
            _check(self._lib.blackframe_hip_scene_set_light_powers(self.handle, C.c_void_p(p.ctypes.data), len(p)))
        return self

    def close(self):
        if self.handle:
            _check(self._lib.blackframe_hip_scene_destroy(self.handle))
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def bvh(self):
        counts = (C.c_uint64 * 5)()
        root = C.c_uint32()
        _check(self._lib.blackframe_hip_scene_bvh_export(self.handle, counts, None, None, None, None, None, C.byref(root)))
        blases = (abi.BvhBlas * max(1, counts[0]))()
        nodes = (abi.BvhNode * max(1, counts[1]))()
        prims = np.zeros(max(1, counts[2]), dtype=np.uint32)
        tlas = (abi.BvhNode * max(1, counts[3]))()
        irefs = (abi.BvhInstanceRef * max(1, counts[4]))()
        _check(self._lib.blackframe_hip_scene_bvh_export(self.handle, counts, blases, nodes, prims.ctypes.data, tlas, irefs,
                                                         C.byref(root)))
        as_bytes = lambda arr, n, size: bytes(memoryview(arr).cast("B")[: n * size])
        return {
            "counts": list(counts),
            "blas": as_bytes(blases, counts[0], 32),
            "blas_nodes": as_bytes(nodes, counts[1], 48),
            "primitive_refs": prims[: counts[2]].tobytes(),
            "tlas_nodes": as_bytes(tlas, counts[3], 48),
            "instance_refs": as_bytes(irefs, counts[4], 32),
            "tlas_root": root.value,
        }

    def bvh_blob(self):
        """The canonical BVH as the reference's serialized "BLKFBVH1" blob (blackframe_hip_scene_bvh_serialize)."""
        size = C.c_size_t()
        _check(self._lib.blackframe_hip_scene_bvh_serialize(self.handle, None, 0, C.byref(size)))
        out = (C.c_uint8 * size.value)()
        _check(self._lib.blackframe_hip_scene_bvh_serialize(self.handle, out, size.value, C.byref(size)))
        return bytes(out)

    def wide_nodes(self):
        counts = (C.c_uint64 * 2)()
        _check(self._lib.blackframe_hip_scene_wide_export(self.handle, counts, None, None))
        blas = np.zeros((max(1, counts[0]), 16), dtype=np.uint32)
        tlas = np.zeros((max(1, counts[1]), 16), dtype=np.uint32)
        _check(self._lib.blackframe_hip_scene_wide_export(self.handle, counts, blas.ctypes.data, tlas.ctypes.data))
        return blas[: counts[0]], tlas[: counts[1]]

    def closest_hit(self, rays, counters=False, stream=None):
        """rays: ctypes array of abi.TransportRay (host). Returns host results (and counters)."""
        import torch

        n = len(rays)
        d_rays = torch.frombuffer(bytearray(memoryview(rays).cast("B")), dtype=torch.uint8).cuda()
        d_out = torch.zeros(n * C.sizeof(abi.ClosestHitResult), dtype=torch.uint8, device="cuda")
        d_ctr = torch.zeros(3 * n, dtype=torch.int32, device="cuda") if counters else None
        _check(self._lib.blackframe_hip_launch_scene_closest_hit(
            self.handle, C.c_void_p(d_rays.data_ptr()), n, C.c_void_p(d_out.data_ptr()),
            C.c_void_p(d_ctr.data_ptr()) if counters else None, _stream_handle(stream)))
        torch.cuda.synchronize()
        out = (abi.ClosestHitResult * n).from_buffer_copy(d_out.cpu().numpy().tobytes())
        if counters:
            return out, d_ctr.cpu().numpy().view(np.uint32).reshape(n, 3)
        return out

    def occlusion(self, rays, stream=None):
        import torch

        n = len(rays)
        d_rays = torch.frombuffer(bytearray(memoryview(rays).cast("B")), dtype=torch.uint8).cuda()
        d_out = torch.zeros(n * C.sizeof(abi.OcclusionResult), dtype=torch.uint8, device="cuda")
        _check(self._lib.blackframe_hip_launch_scene_occlusion(self.handle, C.c_void_p(d_rays.data_ptr()), n,
                                                               C.c_void_p(d_out.data_ptr()), _stream_handle(stream)))
        torch.cuda.synchronize()
        return (abi.OcclusionResult * n).from_buffer_copy(d_out.cpu().numpy().tobytes())


class XpuWavefrontTransportWorkspace:
    """Fixed-capacity device scratch (CudaWavefrontTransportWorkspace::create / close)."""

    def __init__(self, path_capacity):
        self._lib = library()
        self.handle = C.c_void_p()
        self.capacity = int(path_capacity)
        _check(self._lib.blackframe_hip_workspace_create(self.capacity, C.byref(self.handle)))

    def close(self):
        if self.handle:
            _check(self._lib.blackframe_hip_workspace_destroy(self.handle))
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Film:
    """Device film buffers: float sums, compensated double sums and sample counts."""

    def __init__(self, width, height):
        import torch

        self.width, self.height = width, height
        self.float_sum = torch.zeros(width * height * 3, dtype=torch.float32, device="cuda")
        self.compensated = torch.zeros(width * height * 6, dtype=torch.float64, device="cuda")
        self.counts = torch.zeros(width * height, dtype=torch.int32, device="cuda")

    def resolved(self):
        counts = self.counts.cpu().numpy().reshape(self.height, self.width, 1).astype(np.float64)
        return self.float_sum.cpu().numpy().reshape(self.height, self.width, 3) / np.maximum(counts, 1)

    def host(self):
        comp = self.compensated.cpu().numpy().reshape(self.height, self.width, 3, 2)
        return {"film": self.float_sum.cpu().numpy().reshape(self.height, self.width, 3),
                "film_compensated": comp[..., 0] + comp[..., 1],
                "film_counts": self.counts.cpu().numpy().view(np.uint32).reshape(self.height, self.width)}


class Tape:
    """Per-path record of a forward render for the adjoint: final spectral radiance and the bit
    mask of light samples that reached their light, in frame path order (device memory)."""

    def __init__(self, path_count):
        import torch

        self.path_count = path_count
        self.radiance = torch.zeros(path_count * 4, dtype=torch.float32, device="cuda")
        self.visibility = torch.zeros(path_count, dtype=torch.int64, device="cuda")


class Survival:
    """Russian-roulette survival probabilities by path vertex (bf_frame_job.tape_survival), device
    memory of path_count * (diffuse + glossy + specular depth limits + 1) floats. mode is
    abi.SURVIVAL_RECORD (a render writes them) or abi.SURVIVAL_REPLAY (a render decides with them)."""

    def __init__(self, path_count, options, mode=abi.SURVIVAL_RECORD):
        import torch

        self.path_count = path_count
        self.stride = options.diffuse_depth_limit + options.glossy_depth_limit + options.specular_depth_limit + 1
        self.values = torch.zeros(path_count * self.stride, dtype=torch.float32, device="cuda")
        self.mode = mode

    def replaying(self):
        """The same probabilities, decided with instead of recorded."""
        other = Survival.__new__(Survival)
        other.path_count, other.stride, other.values, other.mode = self.path_count, self.stride, self.values, abi.SURVIVAL_REPLAY
        return other


def _bind_survival(job, survival, n_paths, k):
    if survival is None:
        return
    if survival.path_count != n_paths:
        raise XpuError(1, f"survival tape holds {survival.path_count} paths, frame {k} has {n_paths}")
    job.tape_survival = survival.values.data_ptr()
    job.survival_mode = survival.mode


def trace_xpu_wavefront_transport(workspace, scene, frame, options, film=None, want_paths=False, want_counters=False,
                                  stream=None, synchronize=True, tape=None, survival=None):
    """Renders one frame; mirrors engine::trace_cuda_wavefront_transport (WavefrontTransport.hpp:197-208).

    want_paths: False, True (host ctypes array of PathResult) or "device" (a torch uint8 tensor of
    n_paths * 64 bytes left in device memory, for frames whose results are only sampled)."""
    import torch

    if tape is not None or survival is not None:
        # the single-frame entry point has the reference's signature; taped frames use the batch one
        r = trace_xpu_wavefront_batch(workspace, [(scene, frame, film, tape, survival)], options, want_paths=want_paths,
                                      want_counters=want_counters, stream=stream, synchronize=synchronize)
        out = {"film": r["films"][0], "report": r["report"]}
        if want_counters:
            out["traversal_counters"] = r["traversal_counters"]
        if want_paths:
            out["paths"] = r["paths"][0]
        return out
    cam = frame.camera
    film = film if film is not None else Film(cam.width, cam.height)
    n_paths = cam.width * cam.height * frame.samples_per_pixel
    d_paths = torch.zeros(n_paths * C.sizeof(abi.PathResult), dtype=torch.uint8, device="cuda") if want_paths else None
    report = abi.TransportReport()
    counters = np.zeros(6, dtype=np.uint64)
    status = library().blackframe_hip_trace_wavefront_transport(
        workspace.handle, scene.handle, C.byref(frame), C.byref(options), C.c_void_p(film.float_sum.data_ptr()),
        C.c_void_p(film.compensated.data_ptr()), C.c_void_p(film.counts.data_ptr()),
        C.c_void_p(d_paths.data_ptr()) if want_paths else None, C.byref(report),
        C.c_void_p(counters.ctypes.data) if want_counters else None, _stream_handle(stream))
    _check(status)
    if synchronize:
        torch.cuda.synchronize()
    result = {"film": film, "report": report}
    if want_counters:
        result["traversal_counters"] = counters
    if want_paths == "device":
        result["paths"] = d_paths
    elif want_paths:
        result["paths"] = (abi.PathResult * n_paths).from_buffer_copy(d_paths.cpu().numpy().tobytes())
    return result


def light_power_cdf(powers):
    """blackframe_hip_light_power_cdf: the power-weighted selection CDF (count + 1 boundaries) of
    [count][4] spectral powers; raises XpuError with the reference's status on rejection."""
    p = np.ascontiguousarray(powers, dtype=np.float32).reshape(-1, 4)
    out = np.zeros(len(p) + 1, dtype=np.float32)
    _check(library().blackframe_hip_light_power_cdf(C.c_void_p(p.ctypes.data) if len(p) else C.c_void_p(out.ctypes.data),
                                                     len(p), C.c_void_p(out.ctypes.data)))
    return out


def _device_records(records):
    """A ctypes record array copied into a CUDA uint8 tensor, or a CUDA tensor passed through."""
    import torch

    if isinstance(records, torch.Tensor):
        if not records.is_cuda or not records.is_contiguous():
            raise ValueError("device record columns must be contiguous CUDA tensors")
        return records
    raw = np.frombuffer(bytes(memoryview(records).cast("B")), dtype=np.uint8)
    return torch.from_numpy(raw.copy()).cuda()


def trace_xpu_wavefront_paths(workspace, scene, inputs, options, stream=None, synchronize=True, want_host=True):
    """The path-input transport (blackframe_hip_trace_wavefront_paths); mirrors
    engine::trace_cuda_wavefront_transport(workspace, scene, bvh, inputs, light_sampler, options)
    (WavefrontTransport.hpp:197-208).

    inputs: (sample streams, rays, ray cones, path states), each a ctypes array of SampleStreamIndex /
    TransportRay / RayCone / PathStateLane (copied to the device) or a CUDA uint8 tensor of those
    records. Returns {"states", "rays", "cones", "controls"} (ctypes arrays when want_host, else CUDA
    uint8 tensors), "device" (the four output tensors) and "report"."""
    import torch

    samples, rays, cones, states = inputs
    sizes = (C.sizeof(abi.SampleStreamIndex), C.sizeof(abi.TransportRay), C.sizeof(abi.RayCone),
             C.sizeof(abi.PathStateLane))
    columns = [_device_records(x) for x in (samples, rays, cones, states)]
    n = columns[1].numel() // sizes[1]
    for col, size in zip(columns, sizes):
        if col.numel() != n * size:
            raise ValueError("path input columns differ in length")
    record = (abi.PathStateLane, abi.TransportRay, abi.RayCone, abi.LaneControl)
    outs = [torch.empty(n * C.sizeof(t), dtype=torch.uint8, device="cuda") for t in record]
    pin = abi.PathInputs(columns[0].data_ptr(), columns[1].data_ptr(), columns[2].data_ptr(), columns[3].data_ptr(), n)
    pout = abi.PathOutputs(outs[0].data_ptr(), outs[1].data_ptr(), outs[2].data_ptr(), outs[3].data_ptr(), n)
    report = abi.TransportReport()
    _check(library().blackframe_hip_trace_wavefront_paths(workspace.handle, scene.handle, C.byref(pin), C.byref(options),
                                                          C.byref(pout), C.byref(report), _stream_handle(stream)))
    if synchronize or want_host:
        torch.cuda.synchronize()
    result = {"device": outs, "report": report}
    for name, t, out in zip(("states", "rays", "cones", "controls"), record, outs):
        result[name] = (t * n).from_buffer_copy(out.cpu().numpy().tobytes()) if want_host else out
    return result


def trace_xpu_wavefront_batch(workspace, items, options, want_paths=False, want_counters=False, stream=None,
                              synchronize=True):
    """Renders many frames through one set of queues (blackframe_hip_trace_wavefront_batch).

    items: sequence of (XpuScene, FrameDesc[, Film[, Tape[, Survival]]]). want_paths as in
    trace_xpu_wavefront_transport ("device": one torch uint8 tensor of n_paths * 64 bytes per frame).
    Returns {"films": [...], "report": ...,
    "paths": [...] (per frame, as in trace_xpu_wavefront_transport), "traversal_counters": ...}."""
    import torch

    jobs = (abi.FrameJob * len(items))()
    films, path_buffers = [], []
    for k, item in enumerate(items):
        scene, frame = item[0], item[1]
        cam = frame.camera
        film = item[2] if len(item) > 2 and item[2] is not None else Film(cam.width, cam.height)
        films.append(film)
        n_paths = cam.width * cam.height * frame.samples_per_pixel
        d_paths = torch.zeros(n_paths * C.sizeof(abi.PathResult), dtype=torch.uint8, device="cuda") if want_paths else None
        path_buffers.append((d_paths, n_paths))
        jobs[k].scene = scene.handle.value
        jobs[k].frame = frame
        jobs[k].film = film.float_sum.data_ptr()
        jobs[k].film_compensated = film.compensated.data_ptr()
        jobs[k].film_counts = film.counts.data_ptr()
        jobs[k].paths = d_paths.data_ptr() if want_paths else None
        tape = item[3] if len(item) > 3 else None
        if tape is not None:
            if tape.path_count != n_paths:
                raise XpuError(1, f"tape holds {tape.path_count} paths, frame {k} has {n_paths}")
            jobs[k].tape_radiance = tape.radiance.data_ptr()
            jobs[k].tape_visibility = tape.visibility.data_ptr()
        _bind_survival(jobs[k], item[4] if len(item) > 4 else None, n_paths, k)
    report = abi.TransportReport()
    counters = np.zeros(6, dtype=np.uint64)
    _check(library().blackframe_hip_trace_wavefront_batch(
        workspace.handle, jobs, len(items), C.byref(options), C.byref(report),
        C.c_void_p(counters.ctypes.data) if want_counters else None, _stream_handle(stream)))
    if synchronize:
        torch.cuda.synchronize()
    result = {"films": films, "report": report}
    if want_counters:
        result["traversal_counters"] = counters
    if want_paths == "device":
        result["paths"] = [d for d, _ in path_buffers]
    elif want_paths:
        result["paths"] = [(abi.PathResult * n).from_buffer_copy(d.cpu().numpy().tobytes()) for d, n in path_buffers]
    return result


def trace_xpu_wavefront_adjoint(workspace, items, options, stream=None, synchronize=True):
    """Path-replay adjoint (blackframe_hip_trace_wavefront_adjoint).

    items: sequence of (XpuScene, FrameDesc, Tape, film_gradient, texture_gradients[, Survival]) with
    film_gradient a float32 device tensor of dLoss/d(film RGB sums) [pixels * 3] and
    texture_gradients a sequence (per scene texture slot) of float32 device tensors or None;
    gradients are accumulated into them. Returns the report."""
    import torch

    jobs = (abi.FrameJob * len(items))()
    adjoints = (abi.AdjointJob * len(items))()
    for k, item in enumerate(items):
        scene, frame, tape, film_gradient, texture_gradients = item[:5]
        cam = frame.camera
        if film_gradient.dtype != torch.float32 or film_gradient.numel() != cam.width * cam.height * 3:
            raise XpuError(1, f"frame {k}: film gradient must be float32 [{cam.width * cam.height * 3}]")
        if tape.path_count != cam.width * cam.height * frame.samples_per_pixel:
            raise XpuError(1, f"frame {k}: tape holds {tape.path_count} paths")
        jobs[k].scene = scene.handle.value
        jobs[k].frame = frame
        jobs[k].tape_radiance = tape.radiance.data_ptr()
        jobs[k].tape_visibility = tape.visibility.data_ptr()
        _bind_survival(jobs[k], item[5] if len(item) > 5 else None, tape.path_count, k)
        adjoints[k].film_gradient = film_gradient.data_ptr()
        gradients = list(texture_gradients or [])
        if len(gradients) > len(scene.texture_sizes):
            raise XpuError(1, f"frame {k}: {len(gradients)} gradient buffers for {len(scene.texture_sizes)} textures")
        for slot, g in enumerate(gradients):
            if g is not None:
                if g.dtype != torch.float32 or not g.is_contiguous() or g.numel() != scene.texture_sizes[slot]:
                    raise XpuError(1, f"frame {k}: texture gradient {slot} must be contiguous float32 "
                                      f"[{scene.texture_sizes[slot]}]")
                adjoints[k].texture_gradients[slot] = g.data_ptr()
    report = abi.TransportReport()
    _check(library().blackframe_hip_trace_wavefront_adjoint(workspace.handle, jobs, adjoints, len(items), C.byref(options),
                                                            C.byref(report), _stream_handle(stream)))
    if synchronize:
        torch.cuda.synchronize()
    return report


def scene_soa_hash(blob, offset=0):
    """blackframe_hip_launch_scene_soa_hash over host bytes copied to the device (starting `offset`
    bytes into the device copy, to exercise unaligned blobs). Returns the hash."""
    import torch

    data = np.frombuffer(bytes(blob), dtype=np.uint8)
    d = torch.zeros(len(data) + offset, dtype=torch.uint8, device="cuda")
    d[offset:] = torch.from_numpy(data.copy()).cuda()
    out = torch.zeros(1, dtype=torch.int64, device="cuda")
    _check(library().blackframe_hip_launch_scene_soa_hash(C.c_void_p(d.data_ptr() + offset), len(data),
                                                           C.c_void_p(out.data_ptr())))
    torch.cuda.synchronize()
    return int(out.cpu().numpy().view(np.uint64)[0])


def scene_soa_validate(blob, verify_hash=True):
    """blackframe_hip_scene_soa_validate: the bf_scene_soa_status of host blob bytes (host only)."""
    data = bytes(blob)
    status = C.c_uint32()
    _check(library().blackframe_hip_scene_soa_validate(data, len(data), abi.SCENE_SOA_VERIFY_HASH if verify_hash else 0,
                                                       C.byref(status)))
    return status.value


def scene_bvh_validate(blob, verify_hash=True):
    """blackframe_hip_scene_bvh_validate on host bytes: a bf_bvh_status value."""
    data = bytes(blob)
    status = C.c_uint32()
    _check(library().blackframe_hip_scene_bvh_validate(data, len(data), abi.BVH_VERIFY_HASH if verify_hash else 0,
                                                       C.byref(status)))
    return status.value


def scene_soa_validate_on_device(blob, byte_count=None):
    """blackframe_hip_launch_scene_soa_validate over host bytes copied to the device. byte_count
    (default: the bytes given) is the device byte count the call is told; the check reads the
    header only, so a header can be checked against the size its columns claim."""
    import torch

    data = bytes(blob)
    d = torch.frombuffer(bytearray(data), dtype=torch.uint8).cuda()
    status = torch.full((1,), 0xFFFF, dtype=torch.int32, device="cuda")
    _check(library().blackframe_hip_launch_scene_soa_validate(C.c_void_p(d.data_ptr()),
                                                              len(data) if byte_count is None else byte_count,
                                                              C.c_void_p(status.data_ptr())))
    torch.cuda.synchronize()
    return int(status.item())


def transport_abi_layout():
    """blackframe_hip_query_transport_abi_layout: (LayoutManifest, device count)."""
    manifest = abi.LayoutManifest()
    count = C.c_int()
    _check(library().blackframe_hip_query_transport_abi_layout(C.byref(manifest), C.byref(count)))
    return manifest, count.value


def queue_pushes(headers, slots, requests):
    """Runs blackframe_hip_launch_wavefront_queue_pushes on host arrays copied to the device:
    headers uint32 [7, 8], slots uint32 [7 * slot_stride], requests uint32 [n, 2] (queue kind,
    slot). Returns (headers, slots, outcomes) read back."""
    import torch

    headers = np.ascontiguousarray(headers, dtype=np.uint32).reshape(abi.QUEUE_COUNT, 8)
    slots = np.ascontiguousarray(slots, dtype=np.uint32)
    requests = np.ascontiguousarray(requests, dtype=np.uint32).reshape(-1, 2)
    stride = slots.size // abi.QUEUE_COUNT
    d_headers = torch.from_numpy(headers.view(np.int32).copy()).cuda()
    d_slots = torch.from_numpy(slots.view(np.int32).copy()).cuda()
    d_requests = torch.from_numpy(requests.view(np.int32).copy()).cuda() if requests.size else None
    d_outcomes = torch.full((max(len(requests), 1),), -1, dtype=torch.int32, device="cuda")
    _check(library().blackframe_hip_launch_wavefront_queue_pushes(
        C.c_void_p(d_headers.data_ptr()), C.c_void_p(d_slots.data_ptr()), abi.QUEUE_COUNT, stride,
        C.c_void_p(d_requests.data_ptr()) if d_requests is not None else None, len(requests),
        C.c_void_p(d_outcomes.data_ptr())))
    torch.cuda.synchronize()
    return (d_headers.cpu().numpy().view(np.uint32).reshape(abi.QUEUE_COUNT, 8), d_slots.cpu().numpy().view(np.uint32),
            d_outcomes.cpu().numpy().view(np.uint32)[:len(requests)])


def queue_compaction_scratch_bytes(max_input_count):
    """blackframe_hip_query_wavefront_queue_compaction_scratch_bytes."""
    out = C.c_size_t()
    _check(library().blackframe_hip_query_wavefront_queue_compaction_scratch_bytes(max_input_count, C.byref(out)))
    return out.value


def blob_closest_hit(scene_blob, bvh_blob, rays, occlusion=False, scene_size=None, bvh_size=None, stream=None):
    """blackframe_hip_launch_scene_closest_hit_blob (or _occlusion_blob): the two blobs (bytes, copied
    to the device) and host rays; scene_size / bvh_size override the byte counts passed (to test
    mismatches). Returns host results."""
    import torch

    n = len(rays)
    d_scene = torch.frombuffer(bytearray(scene_blob), dtype=torch.uint8).cuda()
    d_bvh = torch.frombuffer(bytearray(bvh_blob), dtype=torch.uint8).cuda()
    d_rays = torch.frombuffer(bytearray(memoryview(rays).cast("B")), dtype=torch.uint8).cuda()
    record = abi.OcclusionResult if occlusion else abi.ClosestHitResult
    d_out = torch.zeros(max(1, n) * C.sizeof(record), dtype=torch.uint8, device="cuda")
    fn = (library().blackframe_hip_launch_scene_occlusion_blob if occlusion
          else library().blackframe_hip_launch_scene_closest_hit_blob)
    _check(fn(C.c_void_p(d_scene.data_ptr()), len(scene_blob) if scene_size is None else scene_size,
              C.c_void_p(d_bvh.data_ptr()), len(bvh_blob) if bvh_size is None else bvh_size,
              C.c_void_p(d_rays.data_ptr()), n, C.c_void_p(d_out.data_ptr()), _stream_handle(stream)))
    torch.cuda.synchronize()
    return (record * n).from_buffer_copy(d_out.cpu().numpy().tobytes()[: n * C.sizeof(record)])


def audit_stage(outcomes, allowed_route_mask, path_capacity, stage_kind):
    """blackframe_hip_launch_wavefront_audit_stage on host outcomes (uint32 [n, 4]: status, route,
    path_slot, detail) copied to the device; returns the StageAudit read back."""
    import torch

    outcomes = np.ascontiguousarray(outcomes, dtype=np.uint32).reshape(-1, 4)
    d_out = torch.from_numpy(outcomes.view(np.int32).copy()).cuda() if outcomes.size else None
    d_audit = torch.full((C.sizeof(abi.StageAudit) // 4,), -1, dtype=torch.int32, device="cuda")
    _check(library().blackframe_hip_launch_wavefront_audit_stage(
        C.c_void_p(d_out.data_ptr()) if d_out is not None else None, len(outcomes), allowed_route_mask, path_capacity,
        stage_kind, C.c_void_p(d_audit.data_ptr()), _stream_handle()))
    torch.cuda.synchronize()
    return abi.StageAudit.from_buffer_copy(d_audit.cpu().numpy().tobytes())


def queue_compaction(headers, slots, outcomes, route, scratch_bytes=None, library_scratch=False):
    """Runs blackframe_hip_launch_wavefront_queue_compaction on host arrays copied to the device:
    headers uint32 [7, 8] (QueueHeader words), slots uint32 [7 * slot_stride], outcomes uint32
    [n, 4] (status, route, path_slot, detail). Returns (headers, slots, result) read back, result a
    uint32 [12] array (WavefrontQueueCompactionResult). library_scratch=True calls
    blackframe_hip_launch_wavefront_queue_compaction_static instead (no scratch argument)."""
    import torch

    headers = np.ascontiguousarray(headers, dtype=np.uint32).reshape(abi.QUEUE_COUNT, 8)
    slots = np.ascontiguousarray(slots, dtype=np.uint32)
    outcomes = np.ascontiguousarray(outcomes, dtype=np.uint32).reshape(-1, 4)
    stride = slots.size // abi.QUEUE_COUNT
    d_headers = torch.from_numpy(headers.view(np.int32).copy()).cuda()
    d_slots = torch.from_numpy(slots.view(np.int32).copy()).cuda() if slots.size else None
    d_outcomes = torch.from_numpy(outcomes.view(np.int32).copy()).cuda() if outcomes.size else None
    d_result = torch.full((12,), -1, dtype=torch.int32, device="cuda")
    need = queue_compaction_scratch_bytes(len(outcomes))
    size = need if scratch_bytes is None else scratch_bytes
    view = abi.QueueSoa(d_headers.data_ptr(), d_slots.data_ptr() if d_slots is not None else None, abi.QUEUE_COUNT, stride)
    d_out = C.c_void_p(d_outcomes.data_ptr()) if d_outcomes is not None else None
    if library_scratch:
        _check(library().blackframe_hip_launch_wavefront_queue_compaction_static(
            view, d_out, len(outcomes), route, C.c_void_p(d_result.data_ptr()), _stream_handle()))
    else:
        d_scratch = torch.zeros(max(size, 1), dtype=torch.uint8, device="cuda")
        _check(library().blackframe_hip_launch_wavefront_queue_compaction(
            view, d_out, len(outcomes), route, C.c_void_p(d_scratch.data_ptr()), size, C.c_void_p(d_result.data_ptr()),
            _stream_handle()))
    torch.cuda.synchronize()
    out_slots = d_slots.cpu().numpy().view(np.uint32) if d_slots is not None else slots.copy()
    return (d_headers.cpu().numpy().view(np.uint32).reshape(abi.QUEUE_COUNT, 8), out_slots,
            d_result.cpu().numpy().view(np.uint32))


def film_l2_loss(items, scale, losses=None, stream=None):
    """L2 loss of films against target images and dLoss/d(film sums) (blackframe_hip_film_l2_loss).

    items: sequence of (Film, target, film_gradient) with target and film_gradient float32 CUDA
    tensors of width * height * 3 elements. Returns the float64 CUDA tensor of per-film losses
    (`losses` if given), without synchronising."""
    import torch

    jobs = (abi.FilmLossJob * len(items))()
    for k, (film, target, gradient) in enumerate(items):
        n = film.width * film.height
        for t in (target, gradient):
            if t.dtype != torch.float32 or not t.is_cuda or not t.is_contiguous() or t.numel() != n * 3:
                raise XpuError(1, f"film loss item {k}: target and gradient must be contiguous float32 CUDA tensors of {n * 3} elements")
        jobs[k].film = film.float_sum.data_ptr()
        jobs[k].film_counts = film.counts.data_ptr()
        jobs[k].target = target.data_ptr()
        jobs[k].film_gradient = gradient.data_ptr()
        jobs[k].pixel_count = n
    if losses is None:
        losses = torch.empty(len(items), dtype=torch.float64, device="cuda")
    _check(library().blackframe_hip_film_l2_loss(jobs, len(items), float(scale), C.c_void_p(losses.data_ptr()),
                                                 _stream_handle(stream)))
    return losses


def adam_step(params, step, beta1=0.9, beta2=0.999, epsilon=1e-8, stream=None):
    """One projected Adam step (blackframe_hip_adam_step). params: sequence of (value, gradient,
    first_moment, second_moment, learning_rate, lower, upper) with float32 contiguous CUDA tensors
    of one shape per entry; step counts from 1. Does not synchronise."""
    import torch

    records = (abi.AdamParam * len(params))()
    for k, (value, gradient, m, v, lr, lower, upper) in enumerate(params):
        for t in (value, gradient, m, v):
            if t.dtype != torch.float32 or not t.is_cuda or not t.is_contiguous() or t.numel() != value.numel():
                raise XpuError(1, f"Adam parameter {k}: tensors must be contiguous float32 CUDA tensors of one size")
        records[k].value, records[k].gradient = value.data_ptr(), gradient.data_ptr()
        records[k].first_moment, records[k].second_moment = m.data_ptr(), v.data_ptr()
        records[k].count = value.numel()
        records[k].learning_rate, records[k].lower, records[k].upper = lr, lower, upper
    options = abi.AdamOptions(beta1, beta2, epsilon, step, 0)
    _check(library().blackframe_hip_adam_step(records, len(params), C.byref(options), _stream_handle(stream)))


class FitGraph:
    """One texture-fit step recorded once as one graph (blackframe_hip_fit_graph_*): clear films,
    taped forward, per-frame L2 loss, clear gradients, adjoint, projected Adam; launch() replays it
    with no host work.

    frames: sequence of (XpuScene, FrameDesc, Film, Tape, target, film_gradient, texture_gradients,
    loss_scale[, Survival]); params as adam_step's (value, gradient, first, second, lr, lower,
    upper); loss_values: float64 CUDA tensor of len(frames); step: the Adam step number of the
    first replay. Every tensor must outlive the graph."""

    def __init__(self, workspace, frames, params, loss_values, options, step=1, betas=(0.9, 0.999), epsilon=1e-8,
                 max_steps=4096):
        import torch

        n = len(frames)
        jobs = (abi.FrameJob * n)()
        adjoints = (abi.AdjointJob * n)()
        losses = (abi.FilmLossJob * n)()
        scales = (C.c_float * n)()
        for k, item in enumerate(frames):
            scene, frame, film, tape, target, film_gradient, texture_gradients, scale = item[:8]
            cam = frame.camera
            pixels = cam.width * cam.height
            if tape.path_count != pixels * frame.samples_per_pixel:
                raise XpuError(1, f"frame {k}: tape holds {tape.path_count} paths")
            for t in (target, film_gradient):
                if t.dtype != torch.float32 or not t.is_cuda or not t.is_contiguous() or t.numel() != pixels * 3:
                    raise XpuError(1, f"frame {k}: target and film gradient must be contiguous float32 CUDA tensors "
                                      f"of {pixels * 3} elements")
            jobs[k].scene = scene.handle.value
            jobs[k].frame = frame
            jobs[k].film = film.float_sum.data_ptr()
            jobs[k].film_compensated = film.compensated.data_ptr()
            jobs[k].film_counts = film.counts.data_ptr()
            jobs[k].tape_radiance = tape.radiance.data_ptr()
            jobs[k].tape_visibility = tape.visibility.data_ptr()
            _bind_survival(jobs[k], item[8] if len(item) > 8 else None, tape.path_count, k)
            adjoints[k].film_gradient = film_gradient.data_ptr()
            gradients = list(texture_gradients or [])
            if len(gradients) > len(scene.texture_sizes):
                raise XpuError(1, f"frame {k}: {len(gradients)} gradient buffers for {len(scene.texture_sizes)} textures")
            for slot, g in enumerate(gradients):
                if g is not None:
                    if g.dtype != torch.float32 or not g.is_contiguous() or g.numel() != scene.texture_sizes[slot]:
                        raise XpuError(1, f"frame {k}: texture gradient {slot} must be contiguous float32 "
                                          f"[{scene.texture_sizes[slot]}]")
                    adjoints[k].texture_gradients[slot] = g.data_ptr()
            losses[k].film = film.float_sum.data_ptr()
            losses[k].film_counts = film.counts.data_ptr()
            losses[k].target = target.data_ptr()
            losses[k].film_gradient = film_gradient.data_ptr()
            losses[k].pixel_count = pixels
            scales[k] = scale
        records = (abi.AdamParam * max(1, len(params)))()
        for k, (value, gradient, m, v, lr, lower, upper) in enumerate(params):
            for t in (value, gradient, m, v):
                if t.dtype != torch.float32 or not t.is_cuda or not t.is_contiguous() or t.numel() != value.numel():
                    raise XpuError(1, f"Adam parameter {k}: tensors must be contiguous float32 CUDA tensors of one size")
            records[k].value, records[k].gradient = value.data_ptr(), gradient.data_ptr()
            records[k].first_moment, records[k].second_moment = m.data_ptr(), v.data_ptr()
            records[k].count = value.numel()
            records[k].learning_rate, records[k].lower, records[k].upper = lr, lower, upper
        if loss_values.dtype != torch.float64 or not loss_values.is_cuda or loss_values.numel() != n:
            raise XpuError(1, f"loss values must be a float64 CUDA tensor of {n} elements")
        desc = abi.FitStepDesc()
        desc.frames, desc.adjoints, desc.losses, desc.loss_scales = jobs, adjoints, losses, scales
        desc.loss_values = loss_values.data_ptr()
        desc.params = records
        desc.frame_count, desc.param_count = n, len(params)
        desc.adam = abi.AdamOptions(betas[0], betas[1], epsilon, step, 0)
        desc.max_steps = max_steps
        self._lib = library()
        self.handle = C.c_void_p()
        torch.cuda.synchronize()
        _check(self._lib.blackframe_hip_fit_graph_create(workspace.handle, C.byref(desc), C.byref(options),
                                                         C.byref(self.handle)))

    def launch(self, stream=None):
        _check(self._lib.blackframe_hip_fit_graph_launch(self.handle, _stream_handle(stream)))

    def report(self, stream=None):
        """Waits for the stream; returns (forward report, adjoint report, replays so far)."""
        forward, adjoint, steps = abi.TransportReport(), abi.TransportReport(), C.c_uint32()
        _check(self._lib.blackframe_hip_fit_graph_report(self.handle, C.byref(forward), C.byref(adjoint), C.byref(steps),
                                                         _stream_handle(stream)))
        return forward, adjoint, steps.value

    def close(self):
        if self.handle:
            _check(self._lib.blackframe_hip_fit_graph_destroy(self.handle))
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune_launch(workspace, items, options, trial_passes=1, stream=None):
    """Times launch shapes on the first sample passes of (XpuScene, FrameDesc) items
    (blackframe_hip_autotune_launch); returns the LaunchTuning record. Set options.launch to
    its .launch to use the choice."""
    jobs = (abi.FrameJob * len(items))()
    for k, (scene, frame) in enumerate(items):
        jobs[k].scene = scene.handle.value
        jobs[k].frame = frame
    tuning = abi.LaunchTuning()
    _check(library().blackframe_hip_autotune_launch(workspace.handle, jobs, len(items), C.byref(options), trial_passes,
                                                    C.byref(tuning), _stream_handle(stream)))
    return tuning


def sample_stream_dump(sampler, indices, dimensions, stream=None):
    """indices: ctypes array of SampleStreamIndex; dimensions: uint64 numpy array. Returns float32 numpy."""
    import torch

    n = len(indices)
    d_idx = torch.frombuffer(bytearray(memoryview(indices).cast("B")), dtype=torch.uint8).cuda()
    d_dim = torch.from_numpy(np.ascontiguousarray(dimensions, dtype=np.uint64).view(np.int64)).cuda()
    d_out = torch.zeros(n, dtype=torch.float32, device="cuda")
    _check(library().blackframe_hip_launch_sample_dump(sampler, C.c_void_p(d_idx.data_ptr()), C.c_void_p(d_dim.data_ptr()),
                                                       n, C.c_void_p(d_out.data_ptr()), _stream_handle(stream)))
    torch.cuda.synchronize()
    return d_out.cpu().numpy()


def sample_stream_dump_requests(requests, schema=abi.SAMPLE_DIMENSION_SCHEMA):
    """blackframe_hip_launch_sample_stream_dump on a ctypes array of SampleDumpRequest copied to the
    device. Returns (SampleDimensionManifest, ctypes array of SampleDumpResult)."""
    import torch

    n = len(requests)
    d_req = torch.frombuffer(bytearray(memoryview(requests).cast("B")), dtype=torch.uint8).cuda()
    d_manifest = torch.zeros(C.sizeof(abi.SampleDimensionManifest), dtype=torch.uint8, device="cuda")
    d_results = torch.zeros(n * C.sizeof(abi.SampleDumpResult), dtype=torch.uint8, device="cuda")
    _check(library().blackframe_hip_launch_sample_stream_dump(schema, C.c_void_p(d_req.data_ptr()), n,
                                                              C.c_void_p(d_manifest.data_ptr()),
                                                              C.c_void_p(d_results.data_ptr())))
    torch.cuda.synchronize()
    manifest = abi.SampleDimensionManifest.from_buffer_copy(d_manifest.cpu().numpy().tobytes())
    results = (abi.SampleDumpResult * n).from_buffer_copy(d_results.cpu().numpy().tobytes())
    return manifest, results


def xpu_devices():
    """Enumerates devices through the plugin ABI (blackframe_query_extension -> XPU backend v1),
    as the reference's extension loader does. Returns (backend name, [(kind, name, vendor, id)])."""
    lib = library()
    query = abi.ExtensionQuery(C.sizeof(abi.ExtensionQuery), 1, 0, 0)
    ext = abi.ExtensionApi()
    ext.struct_size = C.sizeof(abi.ExtensionApi)
    status = lib.blackframe_query_extension(C.byref(query), C.byref(ext))
    if status:
        raise XpuError(status, abi.EXT_STATUS.get(status, "extension query failed"))
    backend = abi.XpuBackendApi()
    status = ext.get_interface(ext.extension_context, abi.InterfaceId(*abi.XPU_BACKEND_INTERFACE_ID), 1, 0,
                               C.addressof(backend), C.sizeof(backend))
    if status:
        raise XpuError(status, abi.EXT_STATUS.get(status, "interface query failed"))
    count = C.c_uint32()
    backend.get_device_count(backend.backend_context, C.byref(count))
    devices = []
    for d in range(count.value):
        desc = abi.DeviceDescriptor()
        desc.struct_size = C.sizeof(desc)
        if backend.get_device_descriptor(backend.backend_context, d, C.byref(desc)) == 0:
            devices.append((desc.kind, desc.name.text(), desc.vendor.text(),
                            (desc.identifier_high, desc.identifier_low)))
    return backend.backend_name.text(), devices
