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


BYTE, FLOAT, DOUBLE = 1, 8, 16  # LSGL_RENDER_TEXTURE3D_FORMAT_*


class Texture3D:
    """A dense 3D texture as Texture::Fetch samples it
    (gles/gles_common.h:344-403): byte, float or double texels with 1-4
    interleaved components (FilterTexture3D{Byte,Float,Double},
    render/render_texture.cc:832-1172) and optional per-axis remap tables
    (local::remap, gles/gles_common.h:263-269). image: GPU tensor of
    uint8 / float32 / float64, referenced, not copied."""

    _FORMATS = {torch.uint8: BYTE, torch.float32: FLOAT, torch.float64: DOUBLE}

    def __init__(self, image, dims, components=1, wrap=(REPEAT, REPEAT, REPEAT), ctx=None):
        self.ctx = ctx or default_context()
        self.lib = load()
        self.image = image
        self.remaps = [None, None, None]
        fmt = self._FORMATS[image.dtype]
        h = C.c_void_p()
        check(self.lib.srf_texture3d_create(self.ctx.handle, _ptr(image), fmt, int(components),
                                            *(int(x) for x in dims), *(int(x) for x in wrap),
                                            C.byref(h)))
        self.handle = h

    def __del__(self):
        try:
            if self.handle:
                self.lib.srf_volume_destroy(self.handle)
                self.handle = None
        except Exception:
            pass

    def set_remap(self, axis, table):
        """table: float32 GPU tensor (kept referenced) or None to clear."""
        self.remaps[axis] = table
        check(self.lib.srf_texture3d_remap(self.handle, int(axis), _ptr(table),
                                           0 if table is None else table.numel()))

    def fetch(self, uvw, stream=None):
        """RGBA [n, 4] at n texture coordinates (uvw [n, 3] float32 GPU)."""
        n = uvw.numel() // 3
        out = torch.empty((n, 4), dtype=torch.float32, device=uvw.device)
        check(self.lib.srf_texture3d_fetch(self.handle, _ptr(uvw), n, _ptr(out),
                                           _stream(stream)))
        return out


class SparseVolume:
    """SparseVolumeAccel (render/render_accel_volume.cc:35-139) over volume
    blocks: blocks int32 [nb, 9] (offset, extent, size per block), data a
    list of GPU tensors (uint8 / float32 / float64 voxels, referenced)."""

    _FORMATS = {torch.uint8: BYTE, torch.float32: FLOAT, torch.float64: DOUBLE}

    def __init__(self, blocks, data, components, global_dim, ctx=None):
        import numpy as np
        self.ctx = ctx or default_context()
        self.lib = load()
        self.data = list(data)

        class Block(C.Structure):
            _fields_ = [("offset", C.c_int32 * 3), ("extent", C.c_int32 * 3),
                        ("size", C.c_int32 * 3), ("id", C.c_int32), ("data", C.c_void_p)]

        bl = (Block * len(self.data))()
        for i, row in enumerate(np.asarray(blocks, np.int32).reshape(-1, 9)):
            bl[i].offset[:] = [int(x) for x in row[:3]]
            bl[i].extent[:] = [int(x) for x in row[3:6]]
            bl[i].size[:] = [int(x) for x in row[6:9]]
            bl[i].id = i
            bl[i].data = self.data[i].data_ptr()
        gd = (C.c_int * 3)(*[int(x) for x in global_dim])
        h = C.c_void_p()
        check(self.lib.srf_sparse_volume_create(self.ctx.handle, bl, len(self.data), gd,
                                                int(components),
                                                self._FORMATS[self.data[0].dtype], C.byref(h)))
        self.handle = h

    def __del__(self):
        try:
            if self.handle:
                self.lib.srf_sparse_volume_destroy(self.handle)
                self.handle = None
        except Exception:
            pass

    def sample(self, positions, stream=None):
        """values [n, 4] float64 at positions [n, 3] float64 (GPU)."""
        n = positions.numel() // 3
        out = torch.empty((n, 4), dtype=torch.float64, device=positions.device)
        check(self.lib.srf_sparse_volume_sample(self.handle, _ptr(positions), n, _ptr(out),
                                                _stream(stream)))
        return out

    def fetch(self, uvw, stream=None):
        """Texture::Fetch's sparse branch: RGBA [n, 4] float32 at uvw [n, 3]."""
        n = uvw.numel() // 3
        out = torch.empty((n, 4), dtype=torch.float32, device=uvw.device)
        check(self.lib.srf_sparse_volume_fetch(self.handle, _ptr(uvw), n, _ptr(out),
                                               _stream(stream)))
        return out


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
