"""3D texture sampling and the differentiable emission-absorption march.

Volume mirrors a single-component FP32 3D texture as glTexImage3D /
lsglTexImage3DPointer create it (gles/gles_texture.cc:198-357) and sampled
by texture3D -> FilterTexture3DFloat (render/render_texture.cc:946-1058).
march() is the forward pass; VolumeRender is a torch.autograd.Function whose
backward calls the hand-written adjoint kernel (srf_volume_march_bwd).
"""
from __future__ import annotations

import ctypes as C

import torch

from ._lib import check, load
from .accel import _ptr, _stream, default_context

REPEAT, CLAMP_TO_EDGE = 0, 1


class Volume:
    def __init__(self, voxels, dims, wrap=(REPEAT, REPEAT, REPEAT), ctx=None):
        """voxels: float32 GPU tensor of dims[0]*dims[1]*dims[2] values, x fastest.
        The tensor is referenced (like lsglTexImage3DPointer), not copied."""
        self.ctx = ctx or default_context()
        self.lib = load()
        self.voxels = voxels
        self.dims = tuple(int(x) for x in dims)
        self.wrap = tuple(int(x) for x in wrap)
        h = C.c_void_p()
        check(self.lib.srf_volume_create(self.ctx.handle, _ptr(voxels), *self.dims, *self.wrap,
                                         C.byref(h)))
        self.handle = h

    def __del__(self):
        try:
            if self.handle:
                self.lib.srf_volume_destroy(self.handle)
                self.handle = None
        except Exception:
            pass

    def sample(self, uvw, stream=None):
        n = uvw.numel() // 3
        out = torch.empty(n, dtype=torch.float32, device=uvw.device)
        check(self.lib.srf_volume_sample(self.handle, _ptr(uvw), n, _ptr(out), _stream(stream)))
        return out

    def checkpoint_slots(self, step):
        slots = C.c_uint32()
        check(self.lib.srf_volume_checkpoint_slots(self.handle, float(step), C.byref(slots)))
        return slots.value

    def march(self, tf, step, rays, checkpoints=None, rgba=None, stream=None):
        """Forward march; with checkpoints ([n, checkpoint_slots(step)] float32)
        the per-segment transmittance is kept for march_backward."""
        n = rays.shape[0]
        if rgba is None:
            rgba = torch.empty((n, 4), dtype=torch.float32, device=rays.device)
        check(self.lib.srf_volume_march_fwd(self.handle, _ptr(tf), tf.shape[0], float(step),
                                            _ptr(rays), n, _ptr(rgba), _ptr(checkpoints),
                                            _stream(stream)))
        return rgba

    def march_backward(self, tf, step, rays, grad_rgba, grad_vox, grad_tf, checkpoints=None,
                       stream=None):
        """Accumulates dL/dvoxels and dL/dtf (the adjoint kernel)."""
        n = rays.shape[0]
        check(self.lib.srf_volume_march_bwd(self.handle, _ptr(tf), tf.shape[0], float(step),
                                            _ptr(rays), n, _ptr(checkpoints), _ptr(grad_rgba),
                                            _ptr(grad_vox), _ptr(grad_tf), _stream(stream)))


class VolumeRender(torch.autograd.Function):
    """rgba = march(voxels, tf); gradients w.r.t. voxels and tf by the adjoint
    kernel (no per-sample state kept between forward and backward)."""

    @staticmethod
    def forward(ctx, voxels, tf, dims, wrap, step, rays, keep_checkpoints=True):
        vol = Volume(voxels.detach().contiguous(), dims, wrap)
        tf_c = tf.detach().contiguous()
        ck = None
        if keep_checkpoints:
            ck = torch.empty((rays.shape[0], vol.checkpoint_slots(step)), dtype=torch.float32,
                             device=rays.device)
        rgba = vol.march(tf_c, step, rays, ck)
        if ck is None:
            ctx.save_for_backward(voxels, tf, rays)
        else:
            ctx.save_for_backward(voxels, tf, rays, ck)
        ctx.dims, ctx.wrap, ctx.step = dims, wrap, step
        return rgba

    @staticmethod
    def backward(ctx, grad_rgba):
        saved = ctx.saved_tensors
        voxels, tf, rays = saved[:3]
        ck = saved[3] if len(saved) > 3 else None
        vol = Volume(voxels.detach().contiguous(), ctx.dims, ctx.wrap)
        gv = torch.zeros_like(voxels)
        gt = torch.zeros_like(tf)
        vol.march_backward(tf.detach().contiguous(), ctx.step, rays,
                           grad_rgba.contiguous(), gv, gt, ck)
        return gv, gt, None, None, None, None, None


def render(voxels, tf, dims, step, rays, wrap=(CLAMP_TO_EDGE,) * 3, keep_checkpoints=True):
    return VolumeRender.apply(voxels, tf, tuple(dims), tuple(wrap), float(step), rays,
                              keep_checkpoints)
