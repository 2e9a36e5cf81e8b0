"""torch.autograd harness over the HIP path-replay adjoint.

A DifferentiableFrame binds float32 device textures (torch tensors, read in place by the scene)
to one frame. Its image is the film mean (film RGB sums / sample counts), and image() is a
torch.autograd.Function of the bound textures: backward replays the frame's paths from the tape
the forward pass recorded and returns texel gradients, with sampling, MIS weights and Russian
roulette held fixed (the estimator's derivative with respect to the closure values).

    frame = DifferentiableFrame(scene, frame_desc, options, textures=[albedo, roughness])
    loss = ((frame.image(albedo, roughness) - target) ** 2).sum()
    loss.backward()          # albedo.grad, roughness.grad
    optimizer.step()         # in place: the scene reads the updated texels on the next render

There is no reference counterpart: the reference renders forward only.
"""

import torch

from . import hip


class DifferentiableFrame:
    def __init__(self, scene, frame, options, textures, workspace=None, record_survival=False):
        """scene: an XpuScene whose texture slots k were created from textures[k] (device=(ptr, ...));
        textures: the float32 CUDA tensors bound to those slots, in slot order. record_survival:
        render() also records every Russian-roulette survival probability, so that render_frozen()
        can replay the roulette with moved texels (finite-difference checks of the adjoint)."""
        self.scene, self.frame, self.options = scene, frame, options
        self.textures = list(textures)
        cam = frame.camera
        self.width, self.height = cam.width, cam.height
        self.path_count = cam.width * cam.height * frame.samples_per_pixel
        self.workspace = workspace if workspace is not None else hip.XpuWavefrontTransportWorkspace(self.path_count)
        self.film = hip.Film(cam.width, cam.height)
        self.tape = hip.Tape(self.path_count)
        self.survival = hip.Survival(self.path_count, options) if record_survival else None
        self.last_forward_report = None
        self.last_adjoint_report = None

    def render(self):
        """Forward render with a tape; returns the (height, width, 3) film mean."""
        for t in (self.film.float_sum, self.film.compensated, self.film.counts):
            t.zero_()
        r = hip.trace_xpu_wavefront_transport(self.workspace, self.scene, self.frame, self.options, film=self.film,
                                              tape=self.tape, survival=self.survival, synchronize=False)
        self.last_forward_report = r["report"]
        counts = self.film.counts.to(torch.float32).clamp_min(1.0).view(self.height, self.width, 1)
        return self.film.float_sum.view(self.height, self.width, 3) / counts

    def render_frozen(self):
        """Forward render (no tape) whose roulette decides with the survival probabilities the last
        render() recorded, so that only the texels' values differ from that render; returns the
        binary64 (height, width, 3) mean of the compensated film. The adjoint differentiates exactly
        this function of the texels."""
        if self.survival is None:
            raise ValueError("render_frozen() needs a frame created with record_survival=True")
        film = hip.Film(self.width, self.height)
        hip.trace_xpu_wavefront_transport(self.workspace, self.scene, self.frame, self.options, film=film,
                                          survival=self.survival.replaying(), synchronize=False)
        compensated = film.compensated.view(self.height, self.width, 3, 2)
        counts = film.counts.to(torch.float64).clamp_min(1.0).view(self.height, self.width, 1)
        return (compensated[..., 0] + compensated[..., 1]) / counts

    def gradients(self, image_gradient):
        """Texel gradients of sum(image_gradient * image) for the last render."""
        counts = self.film.counts.to(torch.float32).clamp_min(1.0).view(self.height, self.width, 1)
        film_gradient = (image_gradient.to(torch.float32) / counts).contiguous().view(-1)
        grads = [torch.zeros_like(t) for t in self.textures]
        self.last_adjoint_report = hip.trace_xpu_wavefront_adjoint(
            self.workspace, [(self.scene, self.frame, self.tape, film_gradient, grads)], self.options, synchronize=False)
        return grads

    def image(self, *textures):
        """The film mean as a differentiable function of the bound textures (pass them in slot
        order; they must be the tensors the scene reads)."""
        if len(textures) != len(self.textures):
            raise ValueError(f"image() takes all {len(self.textures)} bound textures, got {len(textures)}")
        for given, bound in zip(textures, self.textures):
            if given.data_ptr() != bound.data_ptr():
                raise ValueError("image() takes the textures bound to the scene, in slot order")
        return _Render.apply(self, *textures)


class _Render(torch.autograd.Function):
    @staticmethod
    def forward(ctx, frame, *textures):
        ctx.frame = frame
        return frame.render().clone()

    @staticmethod
    def backward(ctx, grad_image):
        return (None, *ctx.frame.gradients(grad_image.contiguous()))


class TextureFit:
    """Fits scene textures to target images with projected Adam on the L2 film loss.

    One step(), all on the device:
      1. one batched taped forward render of every frame (blackframe_hip_trace_wavefront_batch);
      2. per frame, the loss mean((image - target)^2) and its film gradient
         (blackframe_hip_film_l2_loss, image = film sums / sample counts);
      3. one batched path-replay adjoint of every frame (blackframe_hip_trace_wavefront_adjoint);
      4. one projected Adam step over every fitted texture (blackframe_hip_adam_step).
    Direct calls wait on the host where the batched calls read live-path counts. With graph=True
    the first step() records the four phases once as one graph (hip.FitGraph) and every step()
    replays it: no host work and no host read inside a step (forward, loss and adjoint of every
    replay equal the direct calls', the adjoint's float atomics aside).

        fit = TextureFit(frames, options, [(albedo, 0.02, 0.0, 1.0), (widths, 0.01, 0.01, 1.0)])
        for _ in range(100):
            losses = fit.step()

    frames: sequence of (XpuScene, FrameDesc, target, textures) with target a (height, width, 3)
    float32 CUDA tensor of per-pixel means and textures the float32 CUDA tensors bound to the
    scene's texture slots, in slot order. parameters: (tensor, learning_rate, lower, upper) for
    each texture to fit; bound textures not listed stay fixed. A tensor bound to several frames
    gets the sum of their gradients.
    """

    def __init__(self, frames, options, parameters, workspace=None, betas=(0.9, 0.999), epsilon=1e-8, graph=False,
                 max_graph_steps=4096):
        self.options = options
        self.use_graph, self.max_graph_steps = graph, max_graph_steps
        self.graph = None
        self.betas, self.epsilon = betas, epsilon
        self.parameters = []
        by_pointer = {}
        for tensor, lr, lower, upper in parameters:
            if tensor.dtype != torch.float32 or not tensor.is_cuda or not tensor.is_contiguous():
                raise ValueError("fitted textures must be contiguous float32 CUDA tensors")
            entry = dict(value=tensor, gradient=torch.zeros_like(tensor), first=torch.zeros_like(tensor),
                         second=torch.zeros_like(tensor), lr=float(lr), lower=float(lower), upper=float(upper))
            self.parameters.append(entry)
            by_pointer[tensor.data_ptr()] = entry
        self.frames = []
        paths = 0
        for scene, frame, target, textures in frames:
            cam = frame.camera
            n = cam.width * cam.height * frame.samples_per_pixel
            paths = max(paths, n)
            if tuple(target.shape) != (cam.height, cam.width, 3) or target.dtype != torch.float32:
                raise ValueError("targets are (height, width, 3) float32 CUDA tensors")
            grads = [by_pointer[t.data_ptr()]["gradient"] if t.data_ptr() in by_pointer else None for t in textures]
            self.frames.append(dict(scene=scene, frame=frame, target=target.contiguous().view(-1),
                                    film=hip.Film(cam.width, cam.height), tape=hip.Tape(n),
                                    film_gradient=torch.zeros(cam.width * cam.height * 3, dtype=torch.float32,
                                                              device="cuda"),
                                    gradients=grads, pixels=cam.width * cam.height))
        self.workspace = workspace if workspace is not None else hip.XpuWavefrontTransportWorkspace(paths)
        self.losses = torch.zeros(len(self.frames), dtype=torch.float64, device="cuda")
        self.step_count = 0
        self.last_forward_report = None
        self.last_adjoint_report = None

    def forward(self):
        """Phase 1: the taped forward render of every frame (film buffers cleared first)."""
        for f in self.frames:
            for t in (f["film"].float_sum, f["film"].compensated, f["film"].counts):
                t.zero_()
        r = hip.trace_xpu_wavefront_batch(self.workspace, [(f["scene"], f["frame"], f["film"], f["tape"])
                                                           for f in self.frames], self.options, synchronize=False)
        self.last_forward_report = r["report"]

    def loss(self):
        """Phase 2: per-frame mean squared error and film gradients; returns the loss tensor (one
        buffer, overwritten by the next call: clone it to keep it)."""
        for k, f in enumerate(self.frames):
            hip.film_l2_loss([(f["film"], f["target"], f["film_gradient"])], 1.0 / (3.0 * f["pixels"]),
                             losses=self.losses[k:k + 1])
        return self.losses

    def adjoint(self):
        """Phase 3: texture gradients of the summed loss, replayed from the tapes."""
        for p in self.parameters:
            p["gradient"].zero_()
        self.last_adjoint_report = hip.trace_xpu_wavefront_adjoint(
            self.workspace, [(f["scene"], f["frame"], f["tape"], f["film_gradient"], f["gradients"])
                             for f in self.frames], self.options, synchronize=False)

    def update(self):
        """Phase 4: one projected Adam step of every fitted texture."""
        self.step_count += 1
        hip.adam_step([(p["value"], p["gradient"], p["first"], p["second"], p["lr"], p["lower"], p["upper"])
                       for p in self.parameters], self.step_count, self.betas[0], self.betas[1], self.epsilon)

    def step(self):
        """forward, loss, adjoint, update; returns the per-frame losses of this step's render
        (a CUDA float64 tensor, valid once the stream reaches it)."""
        if self.use_graph:
            if self.graph is None:
                self.graph = hip.FitGraph(
                    self.workspace,
                    [(f["scene"], f["frame"], f["film"], f["tape"], f["target"], f["film_gradient"], f["gradients"],
                      1.0 / (3.0 * f["pixels"])) for f in self.frames],
                    [(p["value"], p["gradient"], p["first"], p["second"], p["lr"], p["lower"], p["upper"])
                     for p in self.parameters],
                    self.losses, self.options, step=self.step_count + 1, betas=self.betas, epsilon=self.epsilon,
                    max_steps=self.max_graph_steps)
            self.graph.launch()
            self.step_count += 1
            return self.losses.clone()
        self.forward()
        losses = self.loss()
        self.adjoint()
        self.update()
        return losses.clone()

    def graph_reports(self):
        """(forward report, adjoint report) of the last graph replay (waits for the stream)."""
        forward, adjoint, _ = self.graph.report()
        return forward, adjoint
