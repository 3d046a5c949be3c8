"""Differentiable rendering through the C ABI: a torch.autograd.Function whose
forward is rt_trace_rays (the accumulated radiance of one frame) and whose backward
is rt_trace_rays_adjoint (path replay), with parameter textures and the point lights'
intensities as the differentiable inputs.

The parameter textures are RGBA32F slots of the scene's texture array (base colour
RGB, metallic-roughness G = roughness, B = metallic) whose device memory is the
parameter tensor itself, so an optimiser's in-place update is what the next trace
reads.  The light intensities are copied into the device light buffer (the
`intensity` field of each 48-byte Light record) before every forward trace."""
import torch

from . import layouts as L
from . import rt

# float offset of `intensity` in a Light record (lib/PunctualLight.glsl Light)
_INTENSITY = L.LIGHT.fields["intensity"][1] // 4
_LIGHT_FLOATS = L.LIGHT.itemsize // 4


def _set_intensity(buffers, intensity):
    buffers.plights.view(torch.float32).view(-1, _LIGHT_FLOATS)[:, _INTENSITY].copy_(intensity)


class TraceFunction(torch.autograd.Function):
    """accum = trace(intensity, *textures).  `intensity` (float32, one per point light)
    or None; the parameter textures follow, in the order of `slots`.  `mode`: the
    forward's rt.MODE_* (both give the same image; the backward replays the paths)."""

    @staticmethod
    def forward(fctx, rctx, tlas, buffers, ubo, width, height, slots, mode, intensity, *params):
        for k, p in zip(slots, params):
            if buffers.textures[k].data_ptr() != p.data_ptr():
                buffers.set_texture(k, p)
        held = None
        if intensity is not None:
            held = intensity.detach().clone()
            _set_intensity(buffers, held)
        accum = rt.new_accumulation(width, height)
        rt.trace_rays(rctx, tlas, buffers, ubo, accum, width, height, mode=mode)
        fctx.state = (rctx, tlas, buffers, ubo.copy(), width, height, slots, held)
        fctx.save_for_backward(*params)
        return accum

    @staticmethod
    def backward(fctx, grad_accum):
        rctx, tlas, buffers, ubo, width, height, slots, held = fctx.state
        params = fctx.saved_tensors
        # the bindings of this forward (another forward may have run since)
        for k, p in zip(slots, params):
            if buffers.textures[k].data_ptr() != p.data_ptr():
                buffers.set_texture(k, p)
        gi = None
        if held is not None:
            _set_intensity(buffers, held)
            gi = torch.zeros(len(buffers.scene.plights), dtype=torch.float32, device=grad_accum.device)
        grads = {k: torch.zeros_like(p) for k, p in zip(slots, params)}
        rt.trace_rays_adjoint(rctx, tlas, buffers, ubo, grad_accum.contiguous(), width, height, grads,
                              grad_point_intensity=gi)
        return (None,) * 8 + (gi,) + tuple(grads[k] for k in slots)


class DifferentiableScene:
    """Device scene whose parameter texture slots (and, optionally, point-light
    intensities) are trainable tensors.

    slots: texture indices (RGBA32F) to optimise; their initial values come from
    the scene's textures.  lights=True adds `intensity`, a float32 tensor with one
    entry per point light, initialised from the scene.  mode: the forward trace's
    rt.MODE_* (the megakernel is the faster one on small frames)."""

    def __init__(self, ctx, scene, slots, lights=False, mode=rt.MODE_WAVEFRONT):
        self.ctx = ctx
        self.mode = mode
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
        self.intensity = None
        if lights:
            if len(scene.plights) == 0:
                raise ValueError("the scene has no point lights")
            self.intensity = torch.tensor(scene.plights["intensity"], dtype=torch.float32,
                                          device=self.buffers.plights.device, requires_grad=True)

    def render(self, ubo=None):
        sc = self.scene
        ubo = sc.ubo if ubo is None else ubo
        return TraceFunction.apply(self.ctx, self.tlas, self.buffers, ubo, sc.width, sc.height, self.slots,
                                   self.mode, self.intensity, *self.params)
