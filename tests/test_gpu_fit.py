"""Texture fitting on the GPU: the film L2 loss and projected Adam entry points, and the
TextureFit driver (batched taped forward + loss + batched adjoint + Adam).

Bars:
  - Adam: bit-identical to the oracle's binary32 restatement over three consecutive steps, for
    16-byte-aligned arrays (float4 body + tail) and unaligned ones (scalar path);
  - film loss: film gradients bit-identical to the oracle, the binary64 loss within 1e-12
    relative (its sum runs in another order);
  - TextureFit: the directional derivative of the summed loss over two frames (random direction
    over both albedo textures) within 1e-2 relative of central finite differences of the GPU
    forward render (Russian roulette off, so every sampling decision is independent of the
    texels), and the loss falls over five Adam steps.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("count,aligned", [(1, True), (7, True), (4099, True), (1 << 20, True), (4099, False),
                                           (1 << 20, False)])
def test_adam_bit_exact(hip, oracle_module, count, aligned):
    import torch

    rng = np.random.default_rng(count + aligned)
    p = rng.uniform(0.0, 1.0, count).astype(np.float32)
    m, v = np.zeros_like(p), np.zeros_like(p)

    def device(a):
        if aligned:
            return torch.from_numpy(a).cuda()
        big = torch.zeros(count + 1, dtype=torch.float32, device="cuda")
        big[1:] = torch.from_numpy(a).cuda()
        return big[1:]

    dp, dm, dv = device(p), device(m), device(v)
    for step in range(1, 4):
        g = (rng.standard_normal(count) * 10.0 ** rng.uniform(-4, 2)).astype(np.float32)
        g[::17] = 0.0
        dg = device(g)
        hip.adam_step([(dp, dg, dm, dv, 0.03, 0.0, 1.0)], step, 0.9, 0.999, 1e-8)
        p, m, v = oracle_module.adam_step(p, g, m, v, 0.03, step, 0.9, 0.999, 1e-8, 0.0, 1.0)
        torch.cuda.synchronize()
        assert np.array_equal(dp.cpu().numpy().view(np.uint32), p.view(np.uint32)), step
        assert np.array_equal(dm.cpu().numpy().view(np.uint32), m.view(np.uint32)), step
        assert np.array_equal(dv.cpu().numpy().view(np.uint32), v.view(np.uint32)), step


def test_film_loss_bit_exact(hip, oracle_module):
    import torch

    rng = np.random.default_rng(5)
    items, expected = [], []
    for w, h in ((37, 23), (64, 64), (1, 1)):
        film = hip.Film(w, h)
        counts = rng.integers(0, 12, w * h).astype(np.int32)
        sums = (rng.uniform(0, 3, w * h * 3) * np.repeat(counts, 3)).astype(np.float32)
        film.counts.copy_(torch.from_numpy(counts))
        film.float_sum.copy_(torch.from_numpy(sums))
        target = rng.uniform(0, 1, w * h * 3).astype(np.float32)
        gradient = torch.full((w * h * 3,), np.nan, dtype=torch.float32, device="cuda")
        items.append((film, torch.from_numpy(target).cuda(), gradient))
        expected.append(oracle_module.film_l2_loss(sums, counts.view(np.uint32), target, 1.0 / (3 * w * h)))
    losses = torch.full((3,), np.nan, dtype=torch.float64, device="cuda")
    for k, item in enumerate(items):
        hip.film_l2_loss([item], 1.0 / (3 * item[0].width * item[0].height), losses=losses[k:k + 1])
    torch.cuda.synchronize()
    for k, ((_, _, gradient), (loss, want_gradient)) in enumerate(zip(items, expected)):
        assert np.array_equal(gradient.cpu().numpy().view(np.uint32), want_gradient.view(np.uint32)), k
        assert abs(float(losses[k]) - loss) <= 1e-12 * abs(loss), (k, float(losses[k]), loss)


def fit_frames(bf, hip, size=32, resolution=48, spp=8):
    """Two textured grid frames with their own textures at the fit's starting values, and their
    targets rendered with seeded target textures."""
    import torch

    options = bf.scenes.transport_options(max_depth=6)
    options.russian_roulette_enabled = 0
    frames, textures = [], []
    for k, grid in enumerate((2, 3)):
        albedo_target, widths_target = bf.scenes.fit_textures(100 + k, size=size)
        albedo = torch.from_numpy(albedo_target).cuda().contiguous()
        widths = torch.from_numpy(widths_target).cuda().contiguous()
        scene, cam = bf.scenes.icosphere_grid(resolution, resolution, grid=grid, subdivisions=2, seed=7 + k,
                                              materials="textured", albedo=(albedo.data_ptr(), size, size, 3),
                                              roughness=(widths.data_ptr(), size, size, 2))
        xs = hip.XpuScene(scene)
        frame = bf.scenes.frame_desc(cam, spp, seed=k)
        film = hip.Film(resolution, resolution)
        hip.trace_xpu_wavefront_transport(hip.XpuWavefrontTransportWorkspace(resolution * resolution * spp), xs, frame,
                                          options, film=film)
        counts = film.counts.to(torch.float32).clamp_min(1.0).view(resolution, resolution, 1)
        target = (film.float_sum.view(resolution, resolution, 3) / counts).contiguous()
        with torch.no_grad():
            albedo.fill_(0.5)
            widths.fill_(0.25)
        frames.append((xs, frame, target, [albedo, widths]))
        textures += [albedo, widths]
    return options, frames, textures


def test_texture_fit_gradient_matches_finite_differences(bf, hip):
    import torch

    from blackframe_renderer_amd.autodiff import TextureFit

    options, frames, textures = fit_frames(bf, hip)
    params = [(t, 0.05, 0.0, 1.0) if t.shape[-1] == 3 else (t, 0.02, 0.01, 1.0) for t in textures]
    fit = TextureFit(frames, options, params)
    fit.forward()
    fit.loss()
    fit.adjoint()
    torch.cuda.synchronize()
    assert fit.last_adjoint_report.failed_paths == 0 and fit.last_forward_report.failed_paths == 0
    albedos = [p for p in fit.parameters if p["value"].shape[-1] == 3]
    rng = torch.Generator(device="cuda").manual_seed(11)
    directions = [torch.randn(p["value"].shape, generator=rng, device="cuda") for p in albedos]
    ad = sum(float((p["gradient"].double() * d.double()).sum()) for p, d in zip(albedos, directions))
    base = [p["value"].clone() for p in albedos]
    eps = 1e-2

    def loss_at(sign):
        with torch.no_grad():
            for p, b, d in zip(albedos, base, directions):
                p["value"].copy_(b + sign * eps * d)
        fit.forward()
        total = float(fit.loss().sum())
        return total

    fd = (loss_at(1.0) - loss_at(-1.0)) / (2 * eps)
    assert abs(fd - ad) <= 1e-2 * abs(fd), (fd, ad)


def test_texture_fit_lowers_the_loss(bf, hip):
    import torch

    from blackframe_renderer_amd.autodiff import TextureFit

    options, frames, textures = fit_frames(bf, hip)
    params = [(t, 0.05, 0.0, 1.0) if t.shape[-1] == 3 else (t, 0.02, 0.01, 1.0) for t in textures]
    fit = TextureFit(frames, options, params)
    losses = [float(fit.step().sum()) for _ in range(5)]
    assert all(np.isfinite(losses)) and losses[-1] < 0.8 * losses[0], losses
    for t in textures:
        lo, hi = (0.0, 1.0) if t.shape[-1] == 3 else (0.01, 1.0)
        assert float(t.min()) >= lo and float(t.max()) <= hi
    torch.cuda.synchronize()
