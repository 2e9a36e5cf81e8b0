"""Textured transport and the path-replay adjoint on the GPU.

Bars:
  - textured forward renders and their tapes: bit-identical to the oracle (paths, film, tape);
  - adjoint texel gradients: within 1e-3 of the largest gradient (plus 1e-3 relative) of the
    oracle's binary64 accumulation of the same algorithm (the GPU accumulates float atomics in
    arbitrary order);
  - at the adjoint's full size (256^2 x 64 spp, 512^2 RGB texture): relative error <= 1e-2 against
    central finite differences of the GPU forward render, for single texels and for a random
    direction over all texels (Russian roulette off: every sampling decision is then independent
    of the texels, so the render's difference quotient is the quantity the adjoint computes);
  - the torch.autograd harness returns the adjoint's gradients and an optimizer step is seen by
    the next render (textures are read in place).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def path_arrays(paths):
    return np.frombuffer(bytes(memoryview(paths).cast("B")), dtype=np.uint32).reshape(len(paths), 16)


def render_both(bf, oracle, hip, mixed, width=32, height=32, spp=4, specular=None, **opts):
    scene, cam, textures = bf.scenes.textured_cornell(width, height, texture_size=16, mixed=mixed, specular=specular)
    frame = bf.scenes.frame_desc(cam, spp, seed=opts.pop("seed", 0))
    options = bf.scenes.transport_options(**opts)
    s = oracle.scene(scene)
    expected_tape = s.render_taped(frame, options, threads=8)
    expected = s.render(frame, options, threads=8)
    xs = hip.XpuScene(scene)
    ws = hip.XpuWavefrontTransportWorkspace(width * height * spp)
    tape = hip.Tape(width * height * spp)
    got = hip.trace_xpu_wavefront_transport(ws, xs, frame, options, want_paths=True, tape=tape)
    return scene, frame, options, textures, s, expected, expected_tape, xs, ws, tape, got


@pytest.mark.parametrize("mixed", [False, True])
def test_textured_forward_and_tape_bit_exact(bf, oracle, hip, mixed):
    _, _, _, _, _, expected, expected_tape, _, _, tape, got = render_both(bf, oracle, hip, mixed, max_depth=10,
                                                                         mis=bf.abi.MIS_POWER)
    assert got["report"].failed_paths == 0 and expected["report"].failed_paths == 0
    assert np.array_equal(path_arrays(got["paths"]), path_arrays(expected["paths"]))
    film = got["film"].host()
    assert np.array_equal(film["film"], expected["film"])
    assert np.array_equal(tape.radiance.cpu().numpy(), expected_tape["tape_radiance"])
    assert np.array_equal(tape.visibility.cpu().numpy().view(np.uint64), expected_tape["tape_visibility"])


@pytest.mark.parametrize("mixed,rr,specular", [(False, False, None), (True, False, None), (True, True, None),
                                                (False, False, "mirror"), (False, True, "mixture")])
def test_adjoint_matches_oracle(bf, oracle, hip, mixed, rr, specular):
    """specular: a textured mirror back wall, or a textured Lambert + textured mirror mixture, whose
    delta continuations move only the mirror's texture."""
    import torch

    opts = dict(max_depth=8)
    scene, frame, options, textures, s, _, expected_tape, xs, ws, tape, _ = render_both(bf, oracle, hip, mixed,
                                                                                       specular=specular, **opts)
    if not rr:
        options.russian_roulette_enabled = 0
        expected_tape = s.render_taped(frame, options, threads=8)
        hip.trace_xpu_wavefront_transport(ws, xs, frame, options, tape=tape)
    rng = np.random.default_rng(4)
    film_gradient = rng.standard_normal(32 * 32 * 3).astype(np.float32)
    shapes = [t.shape for t in textures.values()]
    expected = s.adjoint(frame, options, expected_tape, film_gradient, shapes, threads=8)
    grads = [torch.zeros(sh, dtype=torch.float32, device="cuda") for sh in shapes]
    report = hip.trace_xpu_wavefront_adjoint(ws, [(xs, frame, tape, torch.from_numpy(film_gradient).cuda(), grads)],
                                             options)
    assert report.failed_paths == 0 and report.shadow_queries == 0
    for k, (g, e) in enumerate(zip(grads, expected)):
        g = g.cpu().numpy().astype(np.float64)
        scale = np.abs(e).max()
        assert scale > 0.0, k
        assert np.all(np.abs(g - e) <= 1e-3 * scale + 1e-3 * np.abs(e)), (k, np.abs(g - e).max(), scale)


def test_adjoint_full_size_matches_finite_differences(bf, hip):
    """The differentiable Cornell box at full size: 256^2 x 64 spp (4 194 304 paths), 512^2 RGB
    albedo texture on the back wall, GGX box alpha 0.3, depth 8, balance MIS and Russian roulette
    from depth 3, L2 loss against a 0.9x target. The taped forward records every roulette survival
    probability; the finite differences replay them (render_frozen), so the roulette is held fixed
    while the texels move, which is what the adjoint differentiates. Central differences
    (eps 1e-3, binary64 film) on 64 seeded random texels and along a random direction over all
    texels agree with the adjoint to 1e-2 relative."""
    import torch

    albedo = torch.from_numpy(bf.scenes.textured_cornell(8, 8, texture_size=512)[2]["albedo"]).cuda().contiguous()
    albedo.requires_grad_(True)
    roughness = torch.full((1, 1, 1), 0.3, dtype=torch.float32, device="cuda", requires_grad=True)
    scene, cam, _ = bf.scenes.textured_cornell(256, 256, albedo=(albedo.data_ptr(), 512, 512, 3),
                                               roughness=(roughness.data_ptr(), 1, 1, 1))
    frame = bf.scenes.frame_desc(cam, 64, seed=0)
    options = bf.scenes.transport_options(max_depth=8, rr_first_depth=3)
    xs = hip.XpuScene(scene)
    from blackframe_renderer_amd.autodiff import DifferentiableFrame

    df = DifferentiableFrame(xs, frame, options, [albedo, roughness], record_survival=True)
    target = df.render().clone() * 0.9
    torch.cuda.synchronize()
    image = df.image(albedo, roughness)
    ((image - target) ** 2).sum().backward()
    grad = albedo.grad.clone()
    assert df.last_forward_report.failed_paths == 0 and df.last_adjoint_report.failed_paths == 0
    base = albedo.detach().clone()
    target64 = target.double()

    def loss():
        return float(((df.render_frozen() - target64) ** 2).sum().item())

    eps = 1e-3
    # a random direction over all texels
    v = torch.randn(albedo.shape, generator=torch.Generator(device="cuda").manual_seed(3), device="cuda")
    with torch.no_grad():
        albedo.copy_(base + eps * v)
        up = loss()
        albedo.copy_(base - eps * v)
        dn = loss()
        albedo.copy_(base)
    fd = (up - dn) / (2 * eps)
    ad = float((grad.double() * v.double()).sum())
    assert abs(fd - ad) <= 1e-2 * abs(fd), (fd, ad)
    # 64 seeded random texels
    rng = np.random.default_rng(64)
    indices = rng.choice(albedo.numel(), 64, replace=False)
    fds, ads = [], []
    for index in indices.tolist():
        with torch.no_grad():
            albedo.view(-1)[index] = base.view(-1)[index] + eps
            up = loss()
            albedo.view(-1)[index] = base.view(-1)[index] - eps
            dn = loss()
            albedo.copy_(base)
        fds.append((up - dn) / (2 * eps))
        ads.append(float(grad.view(-1)[index]))
    fds, ads = np.array(fds), np.array(ads)
    scale = np.abs(fds).max()
    assert scale > 0.0 and (np.abs(fds) > 1e-3 * scale).sum() >= 32  # most sampled texels are seen
    assert np.linalg.norm(ads - fds) <= 1e-2 * np.linalg.norm(fds), (fds, ads)
    assert np.all(np.abs(ads - fds) <= 1e-2 * np.abs(fds) + 1e-4 * scale), (fds, ads)


def test_autograd_step_is_seen_by_the_next_render(bf, hip):
    import torch

    from blackframe_renderer_amd.autodiff import DifferentiableFrame

    albedo = torch.full((16, 16, 3), 0.5, dtype=torch.float32, device="cuda", requires_grad=True)
    roughness = torch.full((1, 1, 1), 0.3, dtype=torch.float32, device="cuda", requires_grad=True)
    scene, cam, _ = bf.scenes.textured_cornell(32, 32, albedo=(albedo.data_ptr(), 16, 16, 3),
                                               roughness=(roughness.data_ptr(), 1, 1, 1))
    frame = bf.scenes.frame_desc(cam, 8, seed=0)
    options = bf.scenes.transport_options(max_depth=6)
    df = DifferentiableFrame(hip.XpuScene(scene), frame, options, [albedo, roughness])
    target = torch.zeros(32, 32, 3, device="cuda")
    opt = torch.optim.Adam([albedo, roughness], lr=0.05)
    losses = []
    for _ in range(3):
        opt.zero_grad()
        loss = ((df.image(albedo, roughness) - target) ** 2).sum()
        loss.backward()
        assert albedo.grad is not None and float(albedo.grad.abs().sum()) > 0.0
        assert roughness.grad is not None and np.isfinite(float(roughness.grad.sum()))
        losses.append(float(loss))
        opt.step()
        with torch.no_grad():
            albedo.clamp_(0.0, 1.0)
            roughness.clamp_(0.05, 1.0)
    # darkening the back wall toward a black target lowers the loss
    assert losses[2] < losses[0]


def test_batched_adjoint_equals_separate_frames(bf, hip):
    """Three textured frames in one adjoint call (shared queues, per-frame film gradients, one
    gradient buffer shared by two frames) give the gradients of three separate calls."""
    import torch

    rng = np.random.default_rng(9)
    albedo_a = torch.from_numpy((0.1 + 0.8 * rng.random((8, 8, 3))).astype(np.float32)).cuda()
    albedo_b = torch.from_numpy((0.1 + 0.8 * rng.random((16, 16, 3))).astype(np.float32)).cuda()
    rough = torch.full((1, 1, 1), 0.3, dtype=torch.float32, device="cuda")
    specs = [(albedo_a, 32, 24, 4, 0), (albedo_b, 16, 16, 8, 3), (albedo_a, 24, 24, 2, 5)]
    options = bf.scenes.transport_options(max_depth=8)
    frames = []
    for albedo, w, h, spp, seed in specs:
        scene, cam, _ = bf.scenes.textured_cornell(w, h, albedo=(albedo.data_ptr(),) + tuple(albedo.shape[1::-1]) + (3,),
                                                   roughness=(rough.data_ptr(), 1, 1, 1))
        xs = hip.XpuScene(scene)
        frame = bf.scenes.frame_desc(cam, spp, seed=seed)
        tape = hip.Tape(w * h * spp)
        film_gradient = torch.from_numpy(rng.standard_normal(w * h * 3).astype(np.float32)).cuda()
        frames.append((xs, frame, tape, film_gradient, albedo))
    ws = hip.XpuWavefrontTransportWorkspace(1 << 15)
    hip.trace_xpu_wavefront_batch(ws, [(xs, f, None, tape) for xs, f, tape, _, _ in frames], options)

    def grads_for(k):
        return [torch.zeros_like(frames[k][4]), torch.zeros_like(rough)]

    separate = []
    for k, (xs, f, tape, fg, _) in enumerate(frames):
        g = grads_for(k)
        hip.trace_xpu_wavefront_adjoint(ws, [(xs, f, tape, fg, g)], options)
        separate.append(g)
    # batched: frames 0 and 2 share the albedo texture and accumulate into one buffer
    shared_albedo, albedo_b_grad, shared_rough = torch.zeros_like(albedo_a), torch.zeros_like(albedo_b), torch.zeros_like(rough)
    items = [(frames[0][0], frames[0][1], frames[0][2], frames[0][3], [shared_albedo, shared_rough]),
             (frames[1][0], frames[1][1], frames[1][2], frames[1][3], [albedo_b_grad, shared_rough]),
             (frames[2][0], frames[2][1], frames[2][2], frames[2][3], [shared_albedo, shared_rough])]
    report = hip.trace_xpu_wavefront_adjoint(ws, items, options)
    assert report.failed_paths == 0

    def close(got, want):
        got, want = got.cpu().double().numpy(), want.cpu().double().numpy()
        scale = np.abs(want).max()
        assert scale > 0.0
        assert np.all(np.abs(got - want) <= 1e-4 * scale + 1e-4 * np.abs(want)), np.abs(got - want).max() / scale

    close(shared_albedo, separate[0][0] + separate[2][0])
    close(albedo_b_grad, separate[1][0])
    close(shared_rough, separate[0][1] + separate[1][1] + separate[2][1])


def test_survival_record_and_replay_match_the_oracle(bf, oracle, hip):
    """Roulette from depth 2 on the textured Cornell box: the recorded survival probabilities equal
    the oracle's bit for bit, recording leaves the render unchanged, and replaying them with the
    same texels reproduces the render bit for bit."""
    scene, cam, _ = bf.scenes.textured_cornell(32, 32, texture_size=16)
    frame = bf.scenes.frame_desc(cam, 4, seed=5)
    options = bf.scenes.transport_options(max_depth=8, rr_first_depth=2)
    n = 32 * 32 * 4
    s = oracle.scene(scene)
    stride = options.diffuse_depth_limit + options.glossy_depth_limit + options.specular_depth_limit + 1
    expected_survival = np.zeros(n * stride, np.float32)
    expected = s.render_survival(frame, options, expected_survival, replay=False, threads=8)
    xs = hip.XpuScene(scene)
    ws = hip.XpuWavefrontTransportWorkspace(n)
    survival = hip.Survival(n, options)
    got = hip.trace_xpu_wavefront_transport(ws, xs, frame, options, want_paths=True, survival=survival)
    assert got["report"].failed_paths == 0
    assert np.array_equal(survival.values.cpu().numpy().view(np.uint32), expected_survival.view(np.uint32))
    assert np.array_equal(path_arrays(got["paths"]), path_arrays(expected["paths"]))
    plain = hip.trace_xpu_wavefront_transport(ws, xs, frame, options, want_paths=True)
    replayed = hip.trace_xpu_wavefront_transport(ws, xs, frame, options, want_paths=True, survival=survival.replaying())
    assert np.array_equal(path_arrays(replayed["paths"]), path_arrays(plain["paths"]))
    assert np.array_equal(replayed["film"].host()["film"], plain["film"].host()["film"])
