"""Differentiable rendering through the C ABI: a torch.autograd.Function whose
forward is rt_trace_rays (the accumulated radiance of one frame) and whose backward
is rt_trace_rays_adjoint (path replay), with parameter textures as the
differentiable inputs.

The parameter textures are RGBA32F slots of the scene's texture array (base colour
RGB, metallic-roughness G = roughness, B = metallic) whose device memory is the
parameter tensor itself, so an optimiser's in-place update is what the next trace
reads."""
import torch

from . import layouts as L
from . import rt


class TraceFunction(torch.autograd.Function):
    """accum = trace(params).  Inputs after the context arguments are the parameter
    tensors, in the order of `slots`; their gradients come from the adjoint."""

    @staticmethod
    def forward(fctx, rctx, tlas, buffers, ubo, width, height, slots, record, *params):
        for k, p in zip(slots, params):
            if buffers.textures[k].data_ptr() != p.data_ptr():
                buffers.set_texture(k, p)
        accum = rt.new_accumulation(width, height)
        if record:  # megakernel forward that records its paths for the adjoint
            rctx.set_option(rt.OPTION_RECORD_PATHS, 1)
            try:
                rt.trace_rays(rctx, tlas, buffers, ubo, accum, width, height, mode=rt.MODE_MEGAKERNEL)
            finally:
                rctx.set_option(rt.OPTION_RECORD_PATHS, 0)
        else:
            rt.trace_rays(rctx, tlas, buffers, ubo, accum, width, height, mode=rt.MODE_WAVEFRONT)
        fctx.state = (rctx, tlas, buffers, ubo, width, height, slots)
        fctx.save_for_backward(*params)
        return accum

    @staticmethod
    def backward(fctx, grad_accum):
        rctx, tlas, buffers, ubo, width, height, slots = fctx.state
        params = fctx.saved_tensors
        grads = {k: torch.zeros_like(p) for k, p in zip(slots, params)}
        rt.trace_rays_adjoint(rctx, tlas, buffers, ubo, grad_accum.contiguous(), width, height, grads)
        return (None,) * 8 + tuple(grads[k] for k in slots)


class DifferentiableScene:
    """Device scene whose parameter texture slots are trainable tensors.

    slots: texture indices (RGBA32F) to optimise; their initial values come from
    the scene's textures.  record_paths: the forward traces in the megakernel mode
    and records its paths (RT_OPTION_RECORD_PATHS), so the backward replays them
    instead of tracing again (88 bytes per pixel, sample and bounce)."""

    def __init__(self, ctx, scene, slots, record_paths=False):
        self.ctx = ctx
        self.record_paths = bool(record_paths)
        self.scene = scene
        self.buffers = rt.Buffers(scene)
        self.blases, self.tlas = rt.create_as(ctx, scene, self.buffers)
        self.slots = list(slots)
        self.params = []
        for k in self.slots:
            if scene.textures[k].format != L.FMT_RGBA32F:
                raise ValueError(f"texture {k} is not RGBA32F")
            p = self.buffers.textures[k].view(torch.float32).reshape(scene.textures[k].data.shape).clone()
            p.requires_grad_(True)
            self.buffers.set_texture(k, p.detach())
            self.params.append(p)

    def render(self, ubo=None):
        sc = self.scene
        ubo = sc.ubo if ubo is None else ubo
        return TraceFunction.apply(self.ctx, self.tlas, self.buffers, ubo, sc.width, sc.height, self.slots,
                                   self.record_paths, *self.params)
