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
    def __init__(self, scene, frame, options, textures, workspace=None):
        """scene: an XpuScene whose texture slots k were created from textures[k] (device=(ptr, ...));
        textures: the float32 CUDA tensors bound to those slots, in slot order."""
        self.scene, self.frame, self.options = scene, frame, options
        self.textures = list(textures)
        cam = frame.camera
        self.width, self.height = cam.width, cam.height
        self.path_count = cam.width * cam.height * frame.samples_per_pixel
        self.workspace = workspace if workspace is not None else hip.XpuWavefrontTransportWorkspace(self.path_count)
        self.film = hip.Film(cam.width, cam.height)
        self.tape = hip.Tape(self.path_count)
        self.last_forward_report = None
        self.last_adjoint_report = None

    def render(self):
        """Forward render with a tape; returns the (height, width, 3) film mean."""
        for t in (self.film.float_sum, self.film.compensated, self.film.counts):
            t.zero_()
        r = hip.trace_xpu_wavefront_transport(self.workspace, self.scene, self.frame, self.options, film=self.film,
                                              tape=self.tape, synchronize=False)
        self.last_forward_report = r["report"]
        counts = self.film.counts.to(torch.float32).clamp_min(1.0).view(self.height, self.width, 1)
        return self.film.float_sum.view(self.height, self.width, 3) / counts

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
