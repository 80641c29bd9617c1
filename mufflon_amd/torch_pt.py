"""torch.autograd binding of the HIP path tracer's adjoint.

`render(tracer, albedo_textures, iterations)` renders the Radiance image of
`iterations` iterations from a reset with the given albedo textures and
returns it as a torch tensor; its backward pass runs the path-replay adjoint
(mfh_pt_adjoint) and returns dLoss/d(texture) for every texture. PyTorch is
used only at this layer: the primal and adjoint run in the HIP library.
"""
import numpy as np
import torch

from mufflon_amd import hip


class Tracer:
    """A path tracer plus the device texture ids that are differentiable."""

    def __init__(self, ctx, pt, texture_ids, texture_shapes, width, height):
        self.ctx = ctx
        self.pt = pt
        self.texture_ids = list(texture_ids)
        self.texture_shapes = list(texture_shapes)  # (h, w, channels)
        self.width = width
        self.height = height

    def primal(self, textures, iterations):
        for tid, tex in zip(self.texture_ids, textures):
            self.ctx.texture_update(tid, tex)
        self.ctx.pt_reset(self.pt)
        self.ctx.pt_iterate(self.pt, iterations)
        return self.ctx.pt_target(self.pt, hip.TARGET_RADIANCE)

    def adjoint(self, delta, iterations):
        self.ctx.pt_adjoint(self.pt, iterations, delta)
        return [self.ctx.pt_texture_gradient(self.pt, tid, shape)
                for tid, shape in zip(self.texture_ids, self.texture_shapes)]


class _Render(torch.autograd.Function):
    @staticmethod
    def forward(ctx, tracer, iterations, *textures):
        arrays = [t.detach().cpu().numpy().astype(np.float32) for t in textures]
        img = tracer.primal(arrays, iterations)
        ctx.tracer = tracer
        ctx.iterations = iterations
        ctx.devices = [t.device for t in textures]
        out = torch.from_numpy(img.copy())
        return out.to(textures[0].device) if textures else out

    @staticmethod
    def backward(ctx, grad_image):
        delta = grad_image.detach().cpu().numpy().astype(np.float32)
        grads = ctx.tracer.adjoint(delta, ctx.iterations)
        return (None, None) + tuple(torch.from_numpy(g).to(d) for g, d in zip(grads, ctx.devices))


def render(tracer, textures, iterations):
    """Radiance image (h, w, 3) of `iterations` iterations from a reset,
    differentiable with respect to the albedo textures."""
    return _Render.apply(tracer, iterations, *textures)
