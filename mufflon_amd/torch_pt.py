"""torch.autograd binding of the HIP path tracer's adjoint.

`render(tracer, textures, iterations, emission_scales=())` renders the
Radiance image of `iterations` iterations from a reset with the given
textures (albedo or roughness) and emission scales and returns it as a torch
tensor; its backward pass runs the path-replay adjoint and returns
dLoss/d(texture) for every texture and dLoss/d(scale) for every emission
scale. PyTorch is used only at this layer: the primal and adjoint run in the
HIP library.

Two tracers bind it:
* `CoreTracer` over the core C API (mufflon_amd.core, the calls the
  reference's Python binding makes, bindings.py:373-386): a loaded scenario
  of a CoreInterface, its textures by handle and its emissive materials by
  material slot (render_set_texture_data, render_set_emission_scale,
  render_iterate, mufflon_get_target_image, render_adjoint,
  render_get_texture_gradient, render_get_emission_gradient);
* `Tracer` over the device layer (mufflon_amd.hip, mfh_*), for callers that
  build scenes there.
"""
import numpy as np
import torch

from mufflon_amd import hip


class CoreTracer:
    """A scenario loaded in a core instance (CoreInterface) with the
    "Pathtracer" renderer enabled. textures: texture handles
    (world_add_texture / world_get_texture) with their (h, w, channels)
    shapes in the stored format; emission_slots: material slots of EMISSIVE
    materials whose radiance scale is differentiable."""

    def __init__(self, core, width, height, textures=(), texture_shapes=(), emission_slots=()):
        self.core = core
        self.width = width
        self.height = height
        self.textures = list(textures)
        self.texture_shapes = [tuple(s) for s in texture_shapes]
        self.emission_slots = list(emission_slots)

    def primal(self, textures, scales, iterations):
        c = self.core
        for h, tex in zip(self.textures, textures):
            c.render_set_texture_data(h, tex)
        for slot, s in zip(self.emission_slots, scales):
            c.render_set_emission_scale(slot, s)
        if not c.render_reset():
            raise RuntimeError(c.core_get_dll_error())
        for _ in range(iterations):
            ok, _ = c.render_iterate()
            if not ok:
                raise RuntimeError(c.core_get_dll_error())
        return c.mufflon_get_target_image("Radiance", False, self.width, self.height)

    def adjoint(self, delta, iterations):
        c = self.core
        if c.render_get_current_iteration() != iterations:
            raise RuntimeError("the renderer was iterated after the primal pass")
        c.render_adjoint(delta)
        grads = [c.render_get_texture_gradient(h, s) for h, s in zip(self.textures, self.texture_shapes)]
        egrads = [c.render_get_emission_gradient(slot) for slot in self.emission_slots]
        return grads, egrads


class Tracer:
    """A device-layer path tracer plus its differentiable device texture ids
    and emissive material indices."""

    def __init__(self, ctx, pt, texture_ids, texture_shapes, width, height, emission_materials=(),
                 materials=None):
        self.ctx = ctx
        self.pt = pt
        self.texture_ids = list(texture_ids)
        self.texture_shapes = list(texture_shapes)  # (h, w, channels)
        self.width = width
        self.height = height
        self.emission_materials = list(emission_materials)
        self.materials = None if materials is None else materials.copy()  # records of the emission materials

    def primal(self, textures, scales, iterations):
        for tid, tex in zip(self.texture_ids, textures):
            self.ctx.texture_update(tid, tex)
        for mi, s in zip(self.emission_materials, scales):
            rec = self.materials[mi].copy()
            rec["scale"] = np.asarray(s, np.float32)
            self.ctx.pt_set_material(self.pt, mi, rec)
        self.ctx.pt_reset(self.pt)
        self.ctx.pt_iterate(self.pt, iterations)
        return self.ctx.pt_target(self.pt, hip.TARGET_RADIANCE)

    def adjoint(self, delta, iterations):
        self.ctx.pt_adjoint(self.pt, iterations, delta)
        grads = [self.ctx.pt_texture_gradient(self.pt, tid, shape)
                 for tid, shape in zip(self.texture_ids, self.texture_shapes)]
        egrads = [self.ctx.pt_emission_gradient(self.pt, mi) for mi in self.emission_materials]
        return grads, egrads


class _Render(torch.autograd.Function):
    @staticmethod
    def forward(ctx, tracer, iterations, num_textures, *params):
        arrays = [t.detach().cpu().numpy().astype(np.float32) for t in params]
        img = tracer.primal(arrays[:num_textures], arrays[num_textures:], iterations)
        ctx.tracer = tracer
        ctx.iterations = iterations
        ctx.num_textures = num_textures
        ctx.devices = [t.device for t in params]
        out = torch.from_numpy(np.ascontiguousarray(img))
        return out.to(params[0].device) if params else out

    @staticmethod
    def backward(ctx, grad_image):
        delta = grad_image.detach().cpu().numpy().astype(np.float32)
        grads, egrads = ctx.tracer.adjoint(delta, ctx.iterations)
        out = [torch.from_numpy(np.asarray(g, np.float32)).to(d) for g, d in zip(grads + egrads, ctx.devices)]
        return (None, None, None) + tuple(out)


def render(tracer, textures, iterations, emission_scales=()):
    """Radiance image (h, w, 3) of `iterations` iterations from a reset,
    differentiable with respect to the textures and the emission scales
    (3 floats each)."""
    return _Render.apply(tracer, iterations, len(textures), *textures, *emission_scales)
