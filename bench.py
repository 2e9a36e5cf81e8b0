#!/usr/bin/env python3
"""Throughput benchmark of the MI355X wavefront path tracer (plain forward).

One step = one full frame through trace_xpu_wavefront_transport: device camera rays, every bounce
(closest hit, shade with NEE/MIS, shadow rays, stable queue compaction) and device film
accumulation, with the scene and BVH already resident in HBM. The default workload is the plain
forward at scale: a 32x32 grid of displaced subdivision-3 icospheres (1 310 720 triangles) with
per-object Lambert / GGX conductor / GGX dielectric materials, 64 emissive triangles plus a
constant environment, 1920x1080 at 64 spp, 4 wavelengths per path, max depth 16, power MIS.

Printed: one JSON line. `value` is Mrays/s = (closest-hit + shadow queries) / frame time.
`roofline` prices the dominant kernel (closest-hit traversal) by its algorithmic bytes: per ray
48 B ray + 80 B result + 64 B per BVH node visited + 48 B per triangle tested + 48 B per instance
entered, counted by the traversal counters of an untimed warmup frame, divided by the kernel's
summed HIP-event duration over the timed frames (events recorded on the launch stream by the
library's stage instrumentation). `cpu_baseline` times the CPU wavefront restatement (the oracle,
threaded with contiguous partitions) on a bounded band of the same frame on this host.
"""
import argparse
import importlib
import importlib.util
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
METRIC = "Mrays/s plain vs batched; adjoint/forward time at grad rel err ≤1e-2"
HBM_PEAK_GBS = 8000.0
TRAFFIC_SUMMARY = ROOT / "profiles" / "traffic_trace_closest_grid_c1.json"
# Launch word of the timed frames: 64-lane traversal workgroups, the closure-kind shade order and
# groups of 4 bounces, the autotuner's choice on this workload (its candidates lie within 2 % of
# each other on a one-pass trial, so a fixed word keeps the traced kernel the same from run to run).
FRAME_LAUNCH = 0x00160004


def load_package():
    name = "blackframe_renderer_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(
        name, ROOT / "blackframe-renderer_amd" / "__init__.py",
        submodule_search_locations=[str(ROOT / "blackframe-renderer_amd")])
    module = importlib.util.module_from_spec(spec)
    sys.modules[name] = module
    spec.loader.exec_module(module)
    return module


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_workload(bf, args):
    if args.workload == "cornell":
        scene, cam = bf.scenes.cornell(args.width or 64, args.height or 64)
        spp = args.spp or 16
        options = bf.scenes.transport_options(max_depth=args.depth or 8, mis=0)
        desc = f"Cornell box {cam.width}x{cam.height}x{spp} spp, depth {options.diffuse_depth_limit}, balance MIS"
    else:
        scene, cam = bf.scenes.icosphere_grid(args.width or 1920, args.height or 1080, materials=args.materials)
        spp = args.spp or 64
        options = bf.scenes.transport_options(max_depth=args.depth or 16, mis=1)
        desc = (f"32x32 displaced icosphere grid ({scene.triangle_count} tris, {args.materials} materials), "
                f"64 emissive tris + environment, {cam.width}x{cam.height}x{spp} spp, "
                f"depth {options.diffuse_depth_limit}, power MIS")
    frame = bf.scenes.frame_desc(cam, spp, seed=args.seed)
    # a numerically failed path is counted and reported instead of aborting the frame
    options.failure_policy = bf.abi.FAILURES_COUNT
    return scene, cam, frame, options, desc


def algorithmic_bytes(report, tc, path_count):
    """Per-kernel and whole-frame algorithmic bytes (ray/result records, node/triangle/instance
    fetches, per-lane state records, queue slots)."""
    closest = 128 * report.closest_hit_queries + 64 * int(tc[0]) + 48 * int(tc[1]) + 48 * int(tc[2])
    shadow = 64 * report.shadow_queries + 64 * int(tc[3]) + 48 * int(tc[4]) + 48 * int(tc[5])
    slots = report.closest_hit_queries + report.shade_lanes + report.shadow_queries + (report.closest_hit_queries - path_count)
    frame = (closest + shadow + 680 * report.shade_lanes + 192 * report.miss_lanes + 208 * report.shadow_queries
             + 416 * path_count + 8 * slots)
    return closest, shadow, frame


def host_cpu():
    """The host CPU the baseline ran on: model name, logical CPUs of the machine, and the CPUs this
    process may use (a one-GPU job is given a share of 16 CPUs; the worker count is that share, the
    reference scheduler's hardware_concurrency() of the job)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "machine_logical_cpus": os.cpu_count(), "job_cpus": len(os.sched_getaffinity(0))}


def cpu_baseline(bf, scene, frame, options, target_seconds, threads):
    """Times the oracle's threaded CPU wavefront on a contiguous range of path slots starting at the
    middle pixel row of the first sample pass (no film), growing the range until the run lasts about
    target_seconds."""
    sys.path.insert(0, str(ROOT / "oracle"))
    oracle_module = importlib.import_module("oracle")
    t0 = time.perf_counter()
    oscene = oracle_module.Oracle(bf.abi).scene(scene)
    build_s = time.perf_counter() - t0
    cam = frame.camera
    pixels = cam.width * cam.height
    total = pixels * frame.samples_per_pixel
    # a band of pixel rows centred in the image, continued over the following sample passes
    begin = (cam.height // 2) * cam.width
    count = min(4096, total - begin)
    while True:
        t0 = time.perf_counter()
        result = oscene.render(frame, options, path_begin=begin, path_count=count, threads=threads, want_paths=False,
                               want_film=False)
        elapsed = time.perf_counter() - t0
        if elapsed >= target_seconds * 0.5 or count == total - begin:
            break
        count = min(total - begin, max(2 * count, int(count * target_seconds / max(elapsed, 1e-3))))
    rep = result["report"]
    rays = rep.closest_hit_queries + rep.shadow_queries
    return {
        "value": rays / elapsed / 1e6,
        "unit": "Mrays/s",
        "cores": threads,
        "kind": "port",
        "host": host_cpu(),
        "sample": (f"path slots {begin}..{begin + count - 1} of the frame (from pixel row {cam.height // 2} of pass 0 "
                   f"on; {count} paths, {rays} rays) in {elapsed:.1f} s; paths/s {count / elapsed:.0f}; "
                   f"oracle BVH build {build_s:.1f} s"),
    }


def cornell_section(bf, hip, args, log):
    """The reference's own CPU-runnable case: Cornell box 64x64, 16 spp (65 536 paths), depth 8,
    NEE with balance MIS, Russian roulette from depth 3, Sobol-Owen seed 0. GPU frames are timed
    back to back; the oracle renders the same whole frame on the host as the threaded CPU
    wavefront and as the single-threaded scalar path. The float film of one GPU frame is compared
    with the compensated binary64 film of the same frame (mean and largest relative error over
    pixels with radiance)."""
    import numpy as np
    import torch

    scene, cam = bf.scenes.cornell(64, 64)
    frame = bf.scenes.frame_desc(cam, 16, seed=0)
    options = bf.scenes.transport_options(max_depth=8, mis=bf.abi.MIS_BALANCE)
    options.failure_policy = bf.abi.FAILURES_COUNT
    xs = hip.XpuScene(scene)
    paths = cam.width * cam.height * frame.samples_per_pixel
    ws = hip.XpuWavefrontTransportWorkspace(paths)
    film = hip.Film(cam.width, cam.height)
    r = hip.trace_xpu_wavefront_transport(ws, xs, frame, options, film=film)
    rays = r["report"].closest_hit_queries + r["report"].shadow_queries
    if not args.no_autotune:
        # a frame this small is bound by launches: the tuner times the whole frame
        options.launch = hip.autotune_launch(ws, [(xs, frame)], options, trial_passes=frame.samples_per_pixel).launch
    host = film.host()
    f32, f64 = host["film"].astype(np.float64), host["film_compensated"]
    lit = np.abs(f64) > 0
    rel = np.abs(f32[lit] - f64[lit]) / np.abs(f64[lit])
    steps = 50
    for _ in range(3):
        hip.trace_xpu_wavefront_transport(ws, xs, frame, options, film=film, synchronize=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        hip.trace_xpu_wavefront_transport(ws, xs, frame, options, film=film, synchronize=False)
    torch.cuda.synchronize()
    gpu_s = (time.perf_counter() - t0) / steps
    result = {"workload": "Cornell box 64x64x16 spp (65536 paths), depth 8, balance MIS, roulette from depth 3, "
                          "Sobol-Owen seed 0",
              "launch": f"0x{options.launch:08x}", "rays_per_frame": rays, "gpu_ms_per_frame": round(gpu_s * 1e3, 3),
              "gpu_mrays_per_s": round(rays / gpu_s / 1e6, 2),
              "film_float_vs_compensated_mean_rel_err": float(f"{rel.mean():.3e}"),
              "film_float_vs_compensated_max_rel_err": float(f"{rel.max():.3e}")}
    log(f"cornell: GPU {gpu_s * 1e3:.2f} ms per frame ({rays / gpu_s / 1e6:.1f} Mrays/s); float film vs "
        f"compensated film: mean rel err {rel.mean():.2e}, max {rel.max():.2e}")
    if not args.no_cpu_baseline:
        sys.path.insert(0, str(ROOT / "oracle"))
        oracle_module = importlib.import_module("oracle")
        oscene = oracle_module.Oracle(bf.abi).scene(scene)
        threads = min(16, len(os.sched_getaffinity(0)))
        for key, n in (("cpu_wavefront", threads), ("cpu_scalar", 1)):
            t0 = time.perf_counter()
            if key == "cpu_wavefront":
                rep = oscene.render(frame, options, threads=n, want_paths=False, want_film=False)["report"]
                sample = "the whole frame (the oracle's queued wavefront)"
            else:
                rep = oscene.render_scalar(frame, options)[1]
                sample = ("the whole frame (the oracle's scalar per-path loop: brute-force scene queries, each "
                          "checked against the BVH)")
            cpu_s = time.perf_counter() - t0
            cpu_rays = rep.closest_hit_queries + rep.shadow_queries
            result[key] = {"value": round(cpu_rays / cpu_s / 1e6, 3), "unit": "Mrays/s", "cores": n, "kind": "port",
                           "sample": sample, "seconds": round(cpu_s, 3)}
            log(f"cornell: oracle {key} on {n} host thread(s): {cpu_s:.2f} s, {cpu_rays / cpu_s / 1e6:.2f} Mrays/s")
    return result


def batched_section(bf, hip, args, log):
    """Plain vs batched on the uneven-batch workload: 32 seeded scenes (10k-500k triangles,
    128^2-1024^2, 4-256 spp, the same material and light generator), rendered once frame by frame
    through trace_xpu_wavefront_transport and once through trace_xpu_wavefront_batch (shared
    queues, bounce groups replayed as hipGraphs). One step of each mode renders every frame."""
    import torch

    t0 = time.perf_counter()
    items = bf.scenes.batch_scenes()
    log(f"batch: {len(items)} scenes generated in {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    xscenes = [hip.XpuScene(scene) for scene, _, _ in items]
    torch.cuda.synchronize()
    log(f"batch: BVH builds + uploads {time.perf_counter() - t0:.1f} s")
    frames = [bf.scenes.frame_desc(cam, spp, seed=args.seed) for _, cam, spp in items]
    films = [hip.Film(cam.width, cam.height) for _, cam, _ in items]
    paths = sum(f.camera.width * f.camera.height * f.samples_per_pixel for f in frames)
    options = bf.scenes.transport_options(max_depth=16, mis=bf.abi.MIS_POWER)
    options.failure_policy = bf.abi.FAILURES_COUNT
    capacity = args.batch_capacity
    ws = hip.XpuWavefrontTransportWorkspace(capacity)
    # one launch word for both modes (tuned on the batch's first sample passes), so that the
    # comparison isolates batching
    launch = bf.abi.LAUNCH_GRAPH | 4
    if not args.no_autotune:
        t0 = time.perf_counter()
        tuning = hip.autotune_launch(ws, list(zip(xscenes, frames)), options, trial_passes=1)
        launch = tuning.launch
        log(f"batch: autotune launch 0x{launch:08x} ({tuning.best_ms:.1f} ms vs default {tuning.default_ms:.1f} ms, "
            f"{time.perf_counter() - t0:.1f} s)")
    options.launch = launch

    def plain():
        rays = 0
        for x, f, film in zip(xscenes, frames, films):
            r = hip.trace_xpu_wavefront_transport(ws, x, f, options, film=film, synchronize=False)
            rays += r["report"].closest_hit_queries + r["report"].shadow_queries
        return rays

    def batched():
        r = hip.trace_xpu_wavefront_batch(ws, [(x, f, film) for x, f, film in zip(xscenes, frames, films)], options,
                                          synchronize=False)
        return r["report"].closest_hit_queries + r["report"].shadow_queries

    result = {"workload": (f"{len(items)} seeded scenes, {sum(s.triangle_count for s, _, _ in items)} tris in total "
                           f"(10k-500k each), 128^2-1024^2, 4-256 spp, depth 16, power MIS"),
              "paths": paths, "paths_in_flight": capacity, "launch": f"0x{launch:08x}"}
    for name, fn in (("plain", plain), ("batched", batched)):
        rays = fn()  # warmup
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.batch_steps):
            assert fn() == rays, "non-deterministic ray count"
        torch.cuda.synchronize()
        elapsed = (time.perf_counter() - t0) / args.batch_steps
        result[f"{name}_s"] = round(elapsed, 4)
        result[f"{name}_mrays_per_s"] = round(rays / elapsed / 1e6, 2)
        result["rays"] = rays
        log(f"batch: {name} {elapsed:.2f} s per batch, {rays / elapsed / 1e6:.1f} Mrays/s")
    result["speedup"] = round(result["batched_mrays_per_s"] / result["plain_mrays_per_s"], 3)
    ws.close()
    return result


def adjoint_section(bf, hip, args, log):
    """Adjoint vs forward on the differentiable Cornell box: 256^2 x 64 spp (4 194 304 paths),
    depth 8, balance MIS, roulette from depth 3, a 512^2 RGB albedo texture on the back wall and a
    GGX conductor box (alpha 0.3, roughness texture). Loss = L2 to a target rendered with the
    texture scaled by 0.8. Times: plain forward (no tape), taped forward, adjoint replay. The
    gradient of this same configuration is checked against central finite differences (eps 1e-3,
    binary64 film) along a random direction over all texels and on 64 seeded texels; the finite
    difference renders replay the roulette survival probabilities the taped render recorded, so
    the roulette is held fixed while the texels move. The ratios are reported only when both
    relative errors are <= 1e-2."""
    import numpy as np
    import torch

    from blackframe_renderer_amd.autodiff import DifferentiableFrame

    t0 = time.perf_counter()
    host_albedo = bf.scenes.textured_cornell(8, 8, texture_size=512)[2]["albedo"]
    albedo = torch.from_numpy(host_albedo).cuda().contiguous().requires_grad_(True)
    roughness = torch.full((1, 1, 1), 0.3, dtype=torch.float32, device="cuda", requires_grad=True)
    scene, cam, _ = bf.scenes.textured_cornell(256, 256, albedo=(albedo.data_ptr(), 512, 512, 3),
                                               roughness=(roughness.data_ptr(), 1, 1, 1))
    frame = bf.scenes.frame_desc(cam, 64, seed=args.seed)
    options = bf.scenes.transport_options(max_depth=8, mis=bf.abi.MIS_BALANCE)
    options.failure_policy = bf.abi.FAILURES_COUNT
    xs = hip.XpuScene(scene)
    df = DifferentiableFrame(xs, frame, options, [albedo, roughness])
    log(f"adjoint: scene + 512^2 texture ready in {time.perf_counter() - t0:.1f} s")
    with torch.no_grad():
        base = albedo.detach().clone()
        albedo.mul_(0.8)
        target = df.render().clone()
        albedo.copy_(base)

    def timed(fn, steps):
        fn()
        torch.cuda.synchronize()
        start = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - start) / steps

    plain_report = {}

    def plain():
        r = hip.trace_xpu_wavefront_transport(df.workspace, xs, frame, options, film=df.film, synchronize=False)
        plain_report["r"] = r["report"]

    image = df.render()
    grad_image = (2.0 * (image - target)).contiguous()
    steps = args.adjoint_steps
    t_plain = timed(plain, steps)
    t_taped = timed(df.render, steps)
    df.render()
    t_adjoint = timed(lambda: df.gradients(grad_image), steps)
    rays = plain_report["r"].closest_hit_queries + plain_report["r"].shadow_queries
    adj_report = df.last_adjoint_report
    log(f"adjoint: forward {t_plain * 1e3:.1f} ms (taped {t_taped * 1e3:.1f} ms), adjoint {t_adjoint * 1e3:.1f} ms")

    # the gradient check in the timed configuration (roulette on): the FD renders replay the roulette
    # survival probabilities the taped render recorded, which is the function the adjoint differentiates
    check = DifferentiableFrame(xs, frame, options, [albedo, roughness], workspace=df.workspace,
                                record_survival=True)
    albedo.grad = None
    roughness.grad = None
    ((check.image(albedo, roughness) - target) ** 2).sum().backward()
    grad = albedo.grad.detach().double()
    target64 = target.double()

    def value():
        return float(((check.render_frozen() - target64) ** 2).sum())

    eps = 1e-3
    v = torch.randn(albedo.shape, generator=torch.Generator(device="cuda").manual_seed(7), device="cuda")
    with torch.no_grad():
        albedo.copy_(base + eps * v)
        up = value()
        albedo.copy_(base - eps * v)
        dn = value()
        albedo.copy_(base)
    fd = (up - dn) / (2 * eps)
    ad = float((grad * v.double()).sum())
    err = abs(ad - fd) / abs(fd)
    texels = np.random.default_rng(64).choice(albedo.numel(), 64, replace=False)
    fds, ads = [], []
    for index in texels.tolist():
        with torch.no_grad():
            albedo.view(-1)[index] = base.view(-1)[index] + eps
            up = value()
            albedo.view(-1)[index] = base.view(-1)[index] - eps
            dn = value()
            albedo.copy_(base)
        fds.append((up - dn) / (2 * eps))
        ads.append(float(grad.view(-1)[index]))
    fds, ads = np.array(fds), np.array(ads)
    texel_err = float(np.linalg.norm(ads - fds) / np.linalg.norm(fds))
    log(f"adjoint: roulette frozen, directional derivative fd {fd:.6g} adjoint {ad:.6g} (rel err {err:.2e}); "
        f"64 texels: |fd - adjoint| / |fd| = {texel_err:.2e}")
    ok = err <= 1e-2 and texel_err <= 1e-2
    result = {
        "workload": ("textured Cornell box 256x256x64 spp (4194304 paths), 512^2 RGB albedo texture, GGX conductor "
                     "box alpha 0.3, depth 8, balance MIS, roulette from depth 3, L2 loss"),
        "forward_ms": round(t_plain * 1e3, 3),
        "forward_taped_ms": round(t_taped * 1e3, 3),
        "adjoint_ms": round(t_adjoint * 1e3, 3),
        "forward_mrays_per_s": round(rays / t_plain / 1e6, 2),
        "adjoint_closest_hit_queries": adj_report.closest_hit_queries,
        "grad_rel_err": float(f"{err:.3e}"),
        "grad_texel_err": float(f"{texel_err:.3e}"),
        "adjoint_over_forward": round(t_adjoint / t_plain, 4) if ok else None,
        "taped_forward_plus_adjoint_over_forward": round((t_taped + t_adjoint) / t_plain, 4) if ok else None,
    }
    return result


def fit_section(bf, hip, args, log):
    """Texture fitting over 16 seeded scenes: per scene an icosphere grid of 5e4-2e5 triangles
    (uniform), a square resolution uniform in {256..512}, 32 spp, depth 8, balance MIS, roulette
    off, and one shared material (Lambert + anisotropic gold GGX mixture) bound to its own 1024^2
    RGB albedo texture and 1024^2 (alpha_x, alpha_y) texture. Targets are rendered with seeded
    smooth target textures; the fit starts from albedo 0.5 and widths 0.25. One step = one batched
    taped forward of all 16 frames + film loss + one batched adjoint + one Adam step over the 32
    textures. Before the loop, the adjoint's directional derivative of the summed loss along a
    random direction over all albedo texels is checked against a central finite difference
    (eps 1e-2); the adjoint/forward ratio is reported only when that relative error is <= 1e-2."""
    import numpy as np
    import torch

    from blackframe_renderer_amd.autodiff import TextureFit

    from blackframe_renderer_amd import workloads

    options = bf.scenes.transport_options(max_depth=8, mis=bf.abi.MIS_BALANCE)
    options.russian_roulette_enabled = 0
    options.failure_policy = bf.abi.FAILURES_COUNT
    built, build_s = workloads.fit_frames(hip, seed=args.seed)
    frames = [[b["scene"], b["frame"], None, [b["albedo"], b["widths"]]] for b in built]
    textures = [t for b in built for t in (b["albedo"], b["widths"])]
    triangles = sum(b["triangles"] for b in built)
    paths = sum(f[1].camera.width * f[1].camera.height * f[1].samples_per_pixel for f in frames)
    ws = hip.XpuWavefrontTransportWorkspace(min(paths, args.batch_capacity))
    log(f"fit: 16 scenes ({triangles} tris, {paths} paths per render) + 32 textures ready in {build_s:.1f} s")
    # targets with the target textures, then the starting point
    films = [hip.Film(f[1].camera.width, f[1].camera.height) for f in frames]
    hip.trace_xpu_wavefront_batch(ws, [(f[0], f[1], film) for f, film in zip(frames, films)], options)
    for f, film in zip(frames, films):
        w, h = film.width, film.height
        counts = film.counts.to(torch.float32).clamp_min(1.0).view(h, w, 1)
        f[2] = (film.float_sum.view(h, w, 3) / counts).contiguous()
    del films
    with torch.no_grad():
        for t in textures:
            t.fill_(0.5 if t.shape[-1] == 3 else 0.25)
    if not args.no_autotune:
        tuning = hip.autotune_launch(ws, [(f[0], f[1]) for f in frames], options, trial_passes=1)
        options.launch = tuning.launch
    else:
        options.launch = bf.abi.LAUNCH_GRAPH | 4
    params = [(t, 0.02, 0.0, 1.0) if t.shape[-1] == 3 else (t, 0.01, 0.01, 1.0) for t in textures]
    fit = TextureFit([tuple(f) for f in frames], options, params, workspace=ws)

    # gradient check at the starting point
    fit.forward()
    fit.loss()
    fit.adjoint()
    albedos = [p for p in fit.parameters if p["value"].shape[-1] == 3]
    gen = torch.Generator(device="cuda").manual_seed(13)
    directions = [torch.randn(p["value"].shape, generator=gen, device="cuda") for p in albedos]
    ad = sum(float((p["gradient"].double() * d.double()).sum()) for p, d in zip(albedos, directions))
    base = [p["value"].clone() for p in albedos]
    eps = 1e-2

    def loss_at(sign):
        with torch.no_grad():
            for p, b, d in zip(albedos, base, directions):
                p["value"].copy_(b + sign * eps * d)
        fit.forward()
        return float(fit.loss().sum())

    fd = (loss_at(1.0) - loss_at(-1.0)) / (2 * eps)
    with torch.no_grad():
        for p, b in zip(albedos, base):
            p["value"].copy_(b)
    err = abs(ad - fd) / abs(fd)
    log(f"fit: directional derivative fd {fd:.6g} adjoint {ad:.6g} (rel err {err:.2e})")

    steps = args.fit_steps
    events = [[torch.cuda.Event(enable_timing=True) for _ in range(5)] for _ in range(steps)]
    losses = []
    torch.cuda.synchronize()
    start = time.perf_counter()
    for k in range(steps):
        e = events[k]
        e[0].record()
        fit.forward()
        e[1].record()
        step_losses = fit.loss()
        e[2].record()
        fit.adjoint()
        e[3].record()
        fit.update()
        e[4].record()
        if k == 0 or k == steps - 1:
            losses.append(step_losses.clone())
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - start
    phase = np.array([[events[k][i].elapsed_time(events[k][i + 1]) for i in range(4)] for k in range(steps)])
    forward_ms, loss_ms, adjoint_ms, adam_ms = phase.mean(axis=0)
    rays = fit.last_forward_report.closest_hit_queries + fit.last_forward_report.shadow_queries
    first, last = float(losses[0].sum()), float(losses[-1].sum())
    log(f"fit: {steps} steps in {elapsed:.1f} s ({elapsed / steps * 1e3:.1f} ms/step: forward {forward_ms:.1f}, "
        f"loss {loss_ms:.2f}, adjoint {adjoint_ms:.1f}, Adam {adam_ms:.2f} ms); loss {first:.5g} -> {last:.5g}")

    # the same fit continued with each step replayed as one recorded graph
    graph = None
    if args.fit_graph_steps > 0:
        fit.use_graph = True
        t0 = time.perf_counter()
        fit.step()
        torch.cuda.synchronize()
        record_s = time.perf_counter() - t0
        g_steps = args.fit_graph_steps
        torch.cuda.synchronize()
        start = time.perf_counter()
        for _ in range(g_steps):
            g_losses = fit.step()
        torch.cuda.synchronize()
        g_elapsed = time.perf_counter() - start
        forward_report, adjoint_report = fit.graph_reports()
        log(f"fit: {g_steps} graph steps in {g_elapsed:.1f} s ({g_elapsed / g_steps * 1e3:.1f} ms/step; recording and "
            f"first replay {record_s:.2f} s); loss {float(g_losses.sum()):.5g}")
        graph = {"steps": g_steps, "ms_per_step": round(g_elapsed / g_steps * 1e3, 3),
                 "record_and_first_replay_s": round(record_s, 3), "loss_last_step": float(g_losses.sum()),
                 "failed_paths": int(forward_report.failed_paths + adjoint_report.failed_paths)}
    return {
        "workload": (f"16 seeded scenes ({triangles} tris in total, 5e4-2e5 each), 256^2-512^2 x 32 spp ({paths} "
                     "paths), per scene 1024^2 RGB albedo + 1024^2 (alpha_x, alpha_y) textures on a Lambert + "
                     "anisotropic GGX mixture, depth 8, balance MIS, roulette off, L2 loss, Adam"),
        "steps": steps,
        "ms_per_step": round(elapsed / steps * 1e3, 3),
        "forward_ms": round(float(forward_ms), 3),
        "loss_ms": round(float(loss_ms), 3),
        "adjoint_ms": round(float(adjoint_ms), 3),
        "adam_ms": round(float(adam_ms), 3),
        "forward_mrays_per_s": round(rays / (forward_ms * 1e-3) / 1e6, 2),
        "launch": f"0x{options.launch:08x}",
        "grad_rel_err": float(f"{err:.3e}"),
        "adjoint_over_forward": round(float(adjoint_ms / forward_ms), 4) if err <= 1e-2 else None,
        "loss_first_step": first,
        "loss_last_step": last,
        "graph": graph,
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=("grid", "cornell"), default="grid")
    ap.add_argument("--materials", choices=("mixed", "lambert"), default="mixed")
    ap.add_argument("--width", type=int, default=0)
    ap.add_argument("--height", type=int, default=0)
    ap.add_argument("--spp", type=int, default=0)
    ap.add_argument("--depth", type=int, default=0)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--capacity", type=int, default=0, help="paths in flight per launch (0: the whole frame)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the plain-vs-batched measurement")
    ap.add_argument("--no-cornell", action="store_true", help="skip the 64x64 Cornell measurement")
    ap.add_argument("--batch-steps", type=int, default=1)
    ap.add_argument("--batch-capacity", type=int, default=1 << 27, help="paths in flight per batched launch")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint/forward measurement")
    ap.add_argument("--no-autotune", action="store_true",
                    help="batched and fit sections: keep a fixed launch word instead of tuning one")
    ap.add_argument("--launch", type=lambda v: int(v, 0), default=FRAME_LAUNCH,
                    help="bf_transport_options.launch word of the timed frames (default 0x%08x)" % FRAME_LAUNCH)
    ap.add_argument("--autotune-frame", action="store_true",
                    help="tune the timed frames' launch word with blackframe_hip_autotune_launch instead of --launch")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="run only the adjoint/forward measurement (prints its JSON object; for profiling)")
    ap.add_argument("--adjoint-steps", type=int, default=5)
    ap.add_argument("--no-fit", action="store_true", help="skip the 16-scene texture fit")
    ap.add_argument("--fit-steps", type=int, default=100, help="Adam steps of the texture fit")
    ap.add_argument("--fit-graph-steps", type=int, default=20,
                    help="steps of the fit replayed as one recorded graph each, after the direct steps (0: none)")
    ap.add_argument("--fit-only", action="store_true",
                    help="run only the texture fit (prints its JSON object; for profiling)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB mebibytes (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (the frame would shard by "
              "sample pass with one film reduction, which this single-GPU build does not implement)", file=sys.stderr)
        return 2
    import numpy as np
    import torch

    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        return 1
    torch.cuda.set_device(0)
    if args.calibrate_copy:
        src = torch.ones(args.calibrate_copy << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibration_copy_bytes": src.numel() * 4}), flush=True)
        return 0
    bf = load_package()
    hip = importlib.import_module("blackframe_renderer_amd.hip")
    if args.adjoint_only:
        print(json.dumps(adjoint_section(bf, hip, args, log)), flush=True)
        return 0
    if args.fit_only:
        print(json.dumps(fit_section(bf, hip, args, log)), flush=True)
        return 0

    t0 = time.perf_counter()
    scene, cam, frame, options, desc = build_workload(bf, args)
    log(f"scene generated in {time.perf_counter() - t0:.1f} s: {desc}")
    t0 = time.perf_counter()
    xscene = hip.XpuScene(scene)
    torch.cuda.synchronize()
    upload_s = time.perf_counter() - t0
    log(f"BVH build + upload {upload_s:.1f} s")
    pixels = cam.width * cam.height
    path_count = pixels * frame.samples_per_pixel
    capacity = args.capacity or min(path_count, (2 ** 31 - 1) // pixels * pixels)
    ws = hip.XpuWavefrontTransportWorkspace(capacity)
    film = hip.Film(cam.width, cam.height)

    # warmup: the first frame also gathers the deterministic traversal counters
    report = counters = None
    for w in range(max(1, args.warmup)):
        t0 = time.perf_counter()
        r = hip.trace_xpu_wavefront_transport(ws, xscene, frame, options, film=film, want_counters=(w == 0))
        if w == 0:
            report, counters = r["report"], r["traversal_counters"]
        log(f"warmup {w}: {time.perf_counter() - t0:.2f} s, {r['report'].bounces} bounces")
    closest_bytes, shadow_bytes, frame_bytes = algorithmic_bytes(report, counters, path_count)
    rays = report.closest_hit_queries + report.shadow_queries

    options.instrumentation = bf.abi.INSTRUMENT_STAGE_EVENTS
    options.launch = args.launch
    tuning = None
    if args.autotune_frame:
        # launch shape for this workload, timed on its first sample pass (stage events stay on in
        # the timed frames, which keeps graph replay out of the candidates)
        t0 = time.perf_counter()
        tuning = hip.autotune_launch(ws, [(xscene, frame)], options, trial_passes=1)
        options.launch = tuning.launch
        log(f"autotune: launch 0x{tuning.launch:08x} ({tuning.best_ms:.2f} ms vs default {tuning.default_ms:.2f} ms "
            f"per trial pass, {tuning.candidates} candidates, {time.perf_counter() - t0:.1f} s)")
    stage_ns = np.zeros(8, dtype=np.float64)
    stage_dispatches = np.zeros(8, dtype=np.float64)
    stream = torch.cuda.current_stream()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for step in range(args.steps):
        r = hip.trace_xpu_wavefront_transport(ws, xscene, frame, options, film=film, synchronize=False, stream=stream)
        rep = r["report"]
        for k in range(8):
            stage_ns[k] += rep.stage_metrics[k].gpu_elapsed_nanoseconds
            stage_dispatches[k] += rep.stage_metrics[k].kernel_dispatches
        assert rep.closest_hit_queries + rep.shadow_queries == rays, "non-deterministic ray count"
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed / args.steps * 1e3
    mrays = rays * args.steps / elapsed / 1e6
    log(f"timed {args.steps} frames: {ms_per_step:.1f} ms/frame, {mrays:.1f} Mrays/s")

    closest_launches = stage_dispatches[1] / args.steps
    closest_avg_s = stage_ns[1] / max(stage_dispatches[1], 1) * 1e-9
    achieved = closest_bytes / closest_launches / closest_avg_s / 1e9
    # HBM bytes per launch of the same kernel from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    # runs of this workload, scaled by the counters' reading for a plain copy of known size
    traffic = None
    if (TRAFFIC_SUMMARY.exists() and args.workload == "grid" and args.materials == "mixed" and not args.capacity
            and not (args.width or args.height or args.spp or args.depth) and options.launch == FRAME_LAUNCH):
        traffic = json.loads(TRAFFIC_SUMMARY.read_text()).get("hbm_bytes_per_launch")
    stages = {name: {"ms_per_frame": round(stage_ns[k] / args.steps * 1e-6, 3),
                     "dispatches_per_frame": int(stage_dispatches[k] / args.steps)}
              for k, name in enumerate(bf.abi.STAGES[:7])}
    line = {
        "metric": METRIC,
        "value": round(mrays, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": f"synthetic (seeded procedural scene, frame seed {args.seed})",
        "config": {"workload": desc, "mode": "plain", "paths_per_frame": path_count, "rays_per_frame": rays,
                   "launch": f"0x{options.launch:08x}",
                   "autotune": None if tuning is None else {
                       "trial": "first sample pass", "best_ms": round(tuning.best_ms, 3),
                       "default_ms": round(tuning.default_ms, 3),
                       "tried": {f"0x{tuning.tried[k]:08x}": round(tuning.tried_ms[k], 3)
                                 for k in range(min(tuning.candidates, 24))}},
                   "paths_in_flight": capacity, "bvh_build_and_upload_s": round(upload_s, 2),
                   "failed_paths_per_frame": report.failed_paths,
                   "first_failure": [report.first_failure_slot, report.first_failure_code] if report.failed_paths else None},
        "paths_per_s": round(path_count * args.steps / elapsed, 1),
        "roofline": {
            "bound": "hbm",
            "kernel": "k_trace_closest",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": traffic,
            "traffic_source": TRAFFIC_SUMMARY.name if traffic else None,
            "algorithmic_bytes_per_launch": round(closest_bytes / closest_launches),
            "avg_launch_ms": round(closest_avg_s * 1e3, 4),
            "frame_algorithmic_gbs": round(frame_bytes * args.steps / elapsed / 1e9, 2),
            # the other two large kernels, priced the same way over their stage time per frame:
            # any-hit = ray + result records, node / triangle / instance fetches and the 208-byte
            # pending-shadow record per lane; shade = 680 bytes of lane state per shade lane
            "others": {name: {"algorithmic_bytes_per_frame": int(b),
                              "ms_per_frame": round(stage_ns[k] / args.steps * 1e-6, 3),
                              "achieved": round(b / (stage_ns[k] / args.steps * 1e-9) / 1e9, 2),
                              "frac": round(b / (stage_ns[k] / args.steps * 1e-9) / 1e9 / HBM_PEAK_GBS, 5)}
                       for name, k, b in (("k_trace_shadow", 3, shadow_bytes + 208 * report.shadow_queries),
                                          ("k_shade", 2, 680 * report.shade_lanes))
                       if stage_ns[k] > 0},
        },
        "stages": stages,
        "cpu_baseline": None,
    }
    if not args.no_cornell and args.workload == "grid":
        line["cornell"] = cornell_section(bf, hip, args, log)
    if not args.no_batched and args.workload == "grid":
        del ws
        torch.cuda.empty_cache()
        line["batched"] = batched_section(bf, hip, args, log)
    if not args.no_adjoint and args.workload == "grid":
        line["adjoint"] = adjoint_section(bf, hip, args, log)
        torch.cuda.empty_cache()
    if not args.no_fit and args.workload == "grid":
        line["fit"] = fit_section(bf, hip, args, log)
        torch.cuda.empty_cache()
    if not args.no_cpu_baseline:
        threads = min(16, len(os.sched_getaffinity(0)))
        log(f"CPU baseline with {threads} threads")
        line["cpu_baseline"] = cpu_baseline(bf, scene, frame, options, args.cpu_seconds, threads)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
