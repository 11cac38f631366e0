"""Python host mirror of SURFACE's accelerator interface over libsrf.so.

The classes keep the reference's names and meaning -- TriangleAccel,
TetraAccel, ParticleAccel and SolidAccel with Build32 / Traverse /
BoundingBox (render/render_accel_triangle.h:69-105,
render/render_accel_tetra.h:82-125, render/render_accel_particle.h:68-110,
render/render_accel_solid.h:82-125) and LineAccel with Build / Traverse
(render/render_accel_line.h:73-103) -- with Traverse taking a batch of
rays. Geometry, rays and hits are torch tensors on the GPU (torch is only the
allocator and stream provider here); every computation runs in libsrf.so.
Failures raise SrfError where the reference returns false or asserts.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib
from ._lib import (AccelInfo, LineDesc, SolidDesc, SphereDesc, SrfError, TetDesc, TriDesc, check,
                   load)
from .workloads import HIT_DTYPE, RAY_DTYPE, TET_EXT_DTYPE

try:
    import torch
except ImportError:  # pragma: no cover
    torch = None

TRI_NODE_DTYPE = np.dtype([("bmin", "<f4", (2, 3)), ("bmax", "<f4", (2, 3)),
                           ("flag", "<i4"), ("axis", "<i4"), ("data", "<u4", (2,))])
TET_NODE_DTYPE = np.dtype([("bmin", "<f4", (3,)), ("bmax", "<f4", (3,)),
                           ("flag", "<i4"), ("axis", "<i4"), ("data", "<u4", (2,))])
SPH_NODE_DTYPE = np.dtype([("bmin", "<f4", (2, 3)), ("bmax", "<f4", (2, 3)),
                           ("data", "<u4", (2,)), ("flag", "<i2"), ("axis", "<i2")])


def _stream(stream=None):
    if stream is not None:
        return C.c_void_p(stream)
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def device_tensor(arr, device="cuda"):
    """numpy array (any dtype, C-contiguous) -> torch tensor on the GPU holding
    the same bytes; structured ray/hit arrays become [n, words] float32."""
    arr = np.ascontiguousarray(arr)
    if arr.dtype.fields is not None:
        words = arr.dtype.itemsize // 4
        arr = arr.view(np.float32).reshape(len(arr), words)
    elif arr.dtype == np.uint32:
        arr = arr.view(np.int32)
    return torch.from_numpy(arr).to(device)


def hits_to_numpy(t, dtype=HIT_DTYPE):
    a = t.detach().cpu().numpy()
    return np.ascontiguousarray(a).view(dtype).reshape(-1)


class Context:
    """srf_ctx: one per device; owns the build scratch space."""

    def __init__(self, device=0):
        self.lib = load()
        h = C.c_void_p()
        check(self.lib.srf_ctx_create(int(device), C.byref(h)))
        self.handle = h
        self.device = device

    def close(self):
        if self.handle:
            self.lib.srf_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(torch.cuda.current_device())
    return _default_ctx


class _Accel:
    kind = None
    node_dtype = None

    def __init__(self, ctx=None):
        self.ctx = ctx or default_context()
        self.lib = self.ctx.lib
        self.handle = None
        self._keep = ()

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass

    def release(self):
        if self.handle:
            self.lib.srf_accel_destroy(self.handle)
            self.handle = None

    def _build(self, fn, desc, stream):
        """First call allocates (srf_accel_build_*); later calls rebuild in
        place like a second TriangleAccel::Build32 on the same object."""
        if self.handle:
            check(self.lib.srf_accel_rebuild(self.ctx.handle, self.handle, C.byref(desc),
                                             _stream(stream)))
        else:
            h = C.c_void_p()
            check(fn(self.ctx.handle, C.byref(desc), _stream(stream), C.byref(h)))
            self.handle = h
        return True

    def _require(self):
        if not self.handle:
            raise SrfError(1, "accelerator has not been built")

    def get_info(self):
        self._require()
        info = AccelInfo()
        check(self.lib.srf_accel_get_info(self.handle, C.byref(info)))
        return {f: getattr(info, f) for f, _ in AccelInfo._fields_}

    def BoundingBox(self):
        """BoundingBox(double bmin[3], double bmax[3]) -> (bmin, bmax)."""
        self._require()
        bmin = (C.c_double * 3)()
        bmax = (C.c_double * 3)()
        check(self.lib.srf_accel_bounds(self.handle, bmin, bmax))
        return np.array(bmin[:]), np.array(bmax[:])

    def export(self):
        """(nodes in the reference layout, DFS preorder; sorted prim indices)."""
        info = self.get_info()
        nodes = np.zeros(info["num_nodes"], self.node_dtype)
        idx = np.zeros(info["num_prims"], np.uint32)
        check(self.lib.srf_accel_export(self.handle, nodes.ctypes.data_as(C.c_void_p),
                                        idx.ctypes.data_as(C.c_void_p)))
        return nodes, idx

    def Dump(self, filename):
        """The tree and indices in the reference's Dump file format
        (render/render_accel_triangle.cc:1692-1719)."""
        check(self.lib.srf_accel_dump(self.handle, os.fsencode(filename)))

    def Traverse(self, rays, hits=None, stream=None):
        """Closest hit for every ray. rays: [n, 8] float32 GPU tensor (srf_ray);
        returns hits [n, 4] float32 (t, u, v, prim_id bits)."""
        self._require()
        n = rays.shape[0]
        if hits is None:
            hits = torch.empty((n, 4), dtype=torch.float32, device=rays.device)
        check(self.lib.srf_trace_closest(self.handle, _ptr(rays), n, _ptr(hits),
                                         _stream(stream)))
        return hits

    def TraverseAny(self, rays, t_max=None, out=None, stream=None):
        """Occlusion flags (int32 0/1) for every ray."""
        self._require()
        n = rays.shape[0]
        if out is None:
            out = torch.empty(n, dtype=torch.int32, device=rays.device)
        check(self.lib.srf_trace_any(self.handle, _ptr(rays), _ptr(t_max), n, _ptr(out),
                                     _stream(stream)))
        return out

    def _build_host(self, kind, desc, keep):
        """srf_accel_build_host / srf_accel_rebuild_host: desc holds host
        (numpy) arrays, kept alive in `keep` for the call; synchronises."""
        if self.handle:
            check(self.lib.srf_accel_rebuild_host(self.ctx.handle, self.handle, C.byref(desc),
                                                  _stream(None)))
        else:
            h = C.c_void_p()
            check(self.lib.srf_accel_build_host(self.ctx.handle, kind, C.byref(desc),
                                                _stream(None), C.byref(h)))
            self.handle = h
        del keep
        return True

    def TraverseHost(self, rays, ext_dtype=None, isect=True):
        """Closest hits of host (numpy) rays through srf_trace_closest_host:
        returns (hits, ext or None, isect fields or None) as numpy arrays."""
        from .workloads import ISECT_DTYPE
        self._require()
        rays = np.ascontiguousarray(rays)
        n = len(rays)
        hits = np.zeros(n, HIT_DTYPE)
        ext = None if ext_dtype is None else np.zeros(n, ext_dtype)
        out = np.zeros(n, ISECT_DTYPE) if isect else None
        check(self.lib.srf_trace_closest_host(
            self.handle, rays.ctypes.data_as(C.c_void_p), n, hits.ctypes.data_as(C.c_void_p),
            None if ext is None else ext.ctypes.data_as(C.c_void_p),
            None if out is None else out.ctypes.data_as(C.c_void_p)))
        return hits, ext, out

    def force_wide_stack(self, on=True):
        """Testing switch: trace with the kernels meant for trees of 2^24 or
        more nodes (32-bit stack entries); results are identical."""
        self._require()
        check(self.lib.srf_debug_force_wide_stack(self.handle, 1 if on else 0))

    def TraverseAnySorted(self, rays, t_max=None, key_bits=0, mode="plain", out=None,
                          stream=None):
        """Occlusion flags with the rays traced in sorted order (key of
        key_bits bits; 0 = given order); mode "plain" or "persistent"."""
        self._require()
        n = rays.shape[0]
        if out is None:
            out = torch.empty(n, dtype=torch.int32, device=rays.device)
        m = {"plain": _lib.LAUNCH_PLAIN, "persistent": _lib.LAUNCH_PERSISTENT}[mode]
        check(self.lib.srf_trace_any_sorted(self.ctx.handle, self.handle, _ptr(rays), _ptr(t_max),
                                            n, int(key_bits), m, _ptr(out), _stream(stream)))
        return out

    def TraverseAnySortedNoAlloc(self, rays, t_max=None, key_bits=0, mode="plain", out=None,
                                 stream=None):
        """TraverseAnySorted through srf_trace_any_sorted_noalloc: the scratch
        is the library's static per-GPU array (at most SRF_NOALLOC_MAX_RAYS
        rays); out must be given by callers that cannot allocate."""
        self._require()
        n = rays.shape[0]
        if out is None:
            out = torch.empty(n, dtype=torch.int32, device=rays.device)
        m = {"plain": _lib.LAUNCH_PLAIN, "persistent": _lib.LAUNCH_PERSISTENT}[mode]
        check(self.lib.srf_trace_any_sorted_noalloc(self.handle, _ptr(rays), _ptr(t_max), n,
                                                    int(key_bits), m, _ptr(out),
                                                    _stream(stream)))
        return out

    def autotune_any(self, rays, t_max=None, out=None, stream=None):
        """Fastest (key_bits, persistent) for TraverseAnySorted on these rays."""
        self._require()
        n = rays.shape[0]
        if out is None:
            out = torch.empty(n, dtype=torch.int32, device=rays.device)
        cfg = _lib.AnyConfig()
        check(self.lib.srf_autotune_any(self.ctx.handle, self.handle, _ptr(rays), _ptr(t_max), n,
                                        _ptr(out), _stream(stream), C.byref(cfg)))
        return {"key_bits": cfg.key_bits, "persistent": bool(cfg.persistent), "ms": cfg.ms}


def ao_rays_spheres(primary, hits, positions, radius=None, constant_radius=0.0, per_hit=8,
                    eps=1e-3, dist=0.05, seed=1, stream=None, out=None):
    """Cosine-weighted ambient-occlusion rays from sphere hits
    (srf_ao_rays_spheres). Returns (rays [m, 8] float32, t_max [m],
    pixel [m] int32) with m the number written (hits x per_hit). With
    out = (rays, t_max, pixel, count) preallocated (n * per_hit entries and
    one int32 counter on the device) the call only enqueues the kernel and
    returns out; the count stays on the device."""
    lib = load()
    n = primary.shape[0]
    cap = n * per_hit
    if out is not None:
        rays, tmax, pixel, count = out
        if rays.shape[0] < cap or tmax.shape[0] < cap or pixel.shape[0] < cap:
            raise SrfError(1, "ao_rays_spheres: out buffers hold fewer than n * per_hit rays")
    else:
        rays = torch.empty((cap, 8), dtype=torch.float32, device=primary.device)
        tmax = torch.empty(cap, dtype=torch.float32, device=primary.device)
        pixel = torch.empty(cap, dtype=torch.int32, device=primary.device)
        count = torch.zeros(1, dtype=torch.int32, device=primary.device)
    check(lib.srf_ao_rays_spheres(_ptr(primary), _ptr(hits), n, _ptr(positions), _ptr(radius),
                                  float(constant_radius), int(per_hit), float(eps), float(dist),
                                  int(seed), _ptr(rays), _ptr(tmax), _ptr(pixel), _ptr(count),
                                  _stream(stream)))
    if out is not None:
        return out
    m = int(count.item())
    return rays[:m], tmax[:m], pixel[:m]


def _isect_out(n, device):
    return torch.empty((n, 20), dtype=torch.float32, device=device)


def build_intersection_triangles(vertices, faces, rays, hits, out=None, stream=None):
    """BuildIntersection of triangle hits (srf_build_intersection_triangles):
    [n, 20] float32 rows in the srf_isect layout (position, f0, geometric,
    f1, normal, f2, tangent, f3, binormal, pad)."""
    lib = load()
    n = rays.shape[0]
    out = _isect_out(n, rays.device) if out is None else out
    d = TriDesc(vertices.data_ptr(), vertices.numel() // 3, faces.data_ptr(), faces.numel() // 3)
    check(lib.srf_build_intersection_triangles(C.byref(d), _ptr(rays), _ptr(hits), n, _ptr(out),
                                               _stream(stream)))
    return out


def build_intersection_tetras(vertices, indices, rays, hits, ext, out=None, stream=None):
    """BuildIntersection of tetra hits (entry faces from the srf_tet_hit_ext
    rows of TraverseTetra)."""
    lib = load()
    n = rays.shape[0]
    out = _isect_out(n, rays.device) if out is None else out
    d = TetDesc(vertices.data_ptr(), vertices.numel() // 3, indices.data_ptr(),
                indices.numel() // 4)
    check(lib.srf_build_intersection_tetras(C.byref(d), _ptr(rays), _ptr(hits), _ptr(ext), n,
                                            _ptr(out), _stream(stream)))
    return out


def build_intersection_spheres(positions, rays, hits, out=None, stream=None):
    """BuildIntersection of sphere hits (position, normal, tangent,
    binormal; u and v are in the hits)."""
    lib = load()
    n = rays.shape[0]
    out = _isect_out(n, rays.device) if out is None else out
    d = SphereDesc(positions.data_ptr(), None, 0.0, positions.numel() // 3)
    check(lib.srf_build_intersection_spheres(C.byref(d), _ptr(rays), _ptr(hits), n, _ptr(out),
                                             _stream(stream)))
    return out


class TriangleAccel(_Accel):
    """TriangleAccel::Build32 / Traverse / BoundingBox on the GPU."""
    kind = _lib.PRIM_TRIANGLES
    node_dtype = TRI_NODE_DTYPE

    def Build32(self, vertices, faces, stream=None):
        d = TriDesc(vertices.data_ptr(), vertices.numel() // 3, faces.data_ptr(),
                    faces.numel() // 3)
        return self._build(self.lib.srf_accel_build_triangles, d, stream)

    def Build32Host(self, vertices, faces):
        """Build32 from host (numpy) arrays; synchronises."""
        V = np.ascontiguousarray(vertices, np.float32)
        F = np.ascontiguousarray(faces, np.uint32)
        d = TriDesc(V.ctypes.data, V.size // 3, F.ctypes.data, F.size // 3)
        return self._build_host(_lib.PRIM_TRIANGLES, d, (V, F))


def build_triangles_batch(meshes, accels=None, ctx=None, stream=None):
    """Builds many triangle meshes with one launch per build stage
    (srf_accel_build_triangles_batch); meshes is a list of (vertices, faces)
    GPU tensors. With `accels` (from an earlier call) they are rebuilt in
    place. Returns one TriangleAccel per mesh, each the tree Build32 of that
    mesh alone gives."""
    ctx = ctx or default_context()
    lib = ctx.lib
    n = len(meshes)
    descs = (TriDesc * n)()
    for i, (V, F) in enumerate(meshes):
        descs[i] = TriDesc(V.data_ptr(), V.numel() // 3, F.data_ptr(), F.numel() // 3)
    if accels is None:
        out = (C.c_void_p * n)()
        check(lib.srf_accel_build_triangles_batch(ctx.handle, n, descs, _stream(stream), out))
        accels = []
        for i in range(n):
            a = TriangleAccel(ctx)
            a.handle = C.c_void_p(out[i])
            accels.append(a)
        return accels
    hs = (C.c_void_p * n)(*[a.handle.value for a in accels])
    check(lib.srf_accel_rebuild_triangles_batch(ctx.handle, n, hs, descs, _stream(stream)))
    return accels


class TetraAccel(_Accel):
    """TetraAccel::Build32 / Traverse (+ InterpolateTetra) / BoundingBox."""
    kind = _lib.PRIM_TETRAHEDRONS
    node_dtype = TET_NODE_DTYPE

    def Build32(self, vertices, indices, stream=None):
        d = TetDesc(vertices.data_ptr(), vertices.numel() // 3, indices.data_ptr(),
                    indices.numel() // 4)
        return self._build(self.lib.srf_accel_build_tetras, d, stream)

    def Build32Host(self, vertices, indices):
        """Build32 from host (numpy) arrays; synchronises."""
        V = np.ascontiguousarray(vertices, np.float32)
        F = np.ascontiguousarray(indices, np.uint32)
        d = TetDesc(V.ctypes.data, V.size // 3, F.ctypes.data, F.size // 4)
        return self._build_host(_lib.PRIM_TETRAHEDRONS, d, (V, F))

    def TraverseTetra(self, rays, scalars=None, hits=None, ext=None, stream=None):
        """Closest hit + InterpolateTetra weights and the interpolated scalar;
        ext is [n, 8] float32 (srf_tet_hit_ext)."""
        self._require()
        n = rays.shape[0]
        if hits is None:
            hits = torch.empty((n, 4), dtype=torch.float32, device=rays.device)
        if ext is None:
            ext = torch.empty((n, 8), dtype=torch.float32, device=rays.device)
        check(self.lib.srf_trace_tetra(self.handle, _ptr(rays), n, _ptr(scalars), _ptr(hits),
                                       _ptr(ext), _stream(stream)))
        return hits, ext


class SolidAccel(_Accel):
    """SolidAccel::Build32 / Traverse / BoundingBox over pyramids (5
    vertices per solid), prisms (6) or hexahedra (8)."""
    kind = _lib.PRIM_HEXAHEDRONS
    node_dtype = TET_NODE_DTYPE

    def Build32(self, vertices, indices, vertices_per_solid, stream=None):
        k = int(vertices_per_solid)
        d = SolidDesc(vertices.data_ptr(), vertices.numel() // 3, indices.data_ptr(),
                      indices.numel() // k if k else 0, k)
        self.kind = _lib.PRIM_SOLID.get(k, _lib.PRIM_HEXAHEDRONS)
        return self._build(self.lib.srf_accel_build_solids, d, stream)

    def TraverseSolid(self, rays, scalars=None, hits=None, ext=None, stream=None):
        """Closest crossing + interpolate() weights, position, normal and the
        varying sum of the per-index-position scalars; ext is [n, 16] float32
        (srf_solid_hit_ext)."""
        self._require()
        n = rays.shape[0]
        if hits is None:
            hits = torch.empty((n, 4), dtype=torch.float32, device=rays.device)
        if ext is None:
            ext = torch.empty((n, 16), dtype=torch.float32, device=rays.device)
        check(self.lib.srf_trace_solids(self.handle, _ptr(rays), n, _ptr(scalars), _ptr(hits),
                                        _ptr(ext), _stream(stream)))
        return hits, ext


class LineAccel(_Accel):
    """LineAccel::Build / Traverse / BoundingBox (capped cylinders along line
    segments), built on the LBVH."""
    kind = _lib.PRIM_LINES
    node_dtype = TET_NODE_DTYPE

    def Build(self, positions, segments, radius=None, constant_radius=0.0, cap=True,
              stream=None):
        d = LineDesc(positions.data_ptr(), positions.numel() // 3, segments.data_ptr(),
                     segments.numel() // 2, radius.data_ptr() if radius is not None else None,
                     float(constant_radius), 1 if cap else 0)
        return self._build(self.lib.srf_accel_build_lines, d, stream)

    def TraverseLines(self, rays, hits=None, ext=None, stream=None):
        """Closest hit + position, normal and cap flag; ext is [n, 8] float32
        (srf_line_hit_ext)."""
        self._require()
        n = rays.shape[0]
        if hits is None:
            hits = torch.empty((n, 4), dtype=torch.float32, device=rays.device)
        if ext is None:
            ext = torch.empty((n, 8), dtype=torch.float32, device=rays.device)
        check(self.lib.srf_trace_lines(self.handle, _ptr(rays), n, _ptr(hits), _ptr(ext),
                                       _stream(stream)))
        return hits, ext


class ParticleAccel(_Accel):
    """ParticleAccel::Build32 / Traverse / BoundingBox (spheres)."""
    kind = _lib.PRIM_POINTS
    node_dtype = SPH_NODE_DTYPE

    def Build32(self, positions, radius=None, constant_radius=0.0, stream=None):
        d = SphereDesc(positions.data_ptr(), radius.data_ptr() if radius is not None else None,
                       float(constant_radius), positions.numel() // 3)
        return self._build(self.lib.srf_accel_build_spheres, d, stream)

    def Build32Host(self, positions, radius=None, constant_radius=0.0):
        """Build32 from host (numpy) arrays; synchronises."""
        P = np.ascontiguousarray(positions, np.float32)
        R = None if radius is None else np.ascontiguousarray(radius, np.float32)
        d = SphereDesc(P.ctypes.data, None if R is None else R.ctypes.data,
                       float(constant_radius), P.size // 3)
        return self._build_host(_lib.PRIM_POINTS, d, (P, R))

    def TraverseSelf(self, rays, prev, hits=None, stream=None):
        """Traverse with per-ray prev_hit_t / prev_hit_normal ([n, 6] float32
        tensor holding srf_sphere_prev records)."""
        self._require()
        n = rays.shape[0]
        if hits is None:
            hits = torch.empty((n, 4), dtype=torch.float32, device=rays.device)
        check(self.lib.srf_trace_spheres(self.handle, _ptr(rays), _ptr(prev), n, _ptr(hits),
                                         _stream(stream)))
        return hits


# ---- LBVH stage entry points (parity of the intermediate products) ----

def morton30(kind, desc_tensors, bmin, bmax, ctx=None, stream=None):
    ctx = ctx or default_context()
    if kind == _lib.PRIM_TRIANGLES:
        V, F = desc_tensors
        d = TriDesc(V.data_ptr(), V.numel() // 3, F.data_ptr(), F.numel() // 3)
        n = F.numel() // 3
    elif kind == _lib.PRIM_TETRAHEDRONS:
        V, F = desc_tensors
        d = TetDesc(V.data_ptr(), V.numel() // 3, F.data_ptr(), F.numel() // 4)
        n = F.numel() // 4
    elif kind in (_lib.PRIM_PYRAMIDS, _lib.PRIM_PRISMS, _lib.PRIM_HEXAHEDRONS):
        V, F = desc_tensors
        k = {_lib.PRIM_PYRAMIDS: 5, _lib.PRIM_PRISMS: 6, _lib.PRIM_HEXAHEDRONS: 8}[kind]
        d = SolidDesc(V.data_ptr(), V.numel() // 3, F.data_ptr(), F.numel() // k, k)
        n = F.numel() // k
    else:
        P, R = desc_tensors
        d = SphereDesc(P.data_ptr(), R.data_ptr() if R is not None else None, 0.0, P.numel() // 3)
        n = P.numel() // 3
    codes = torch.empty(n, dtype=torch.int32, device="cuda")
    mn = (C.c_float * 3)(*[float(x) for x in bmin])
    mx = (C.c_float * 3)(*[float(x) for x in bmax])
    check(ctx.lib.srf_lbvh_morton30(ctx.handle, kind, C.byref(d), mn, mx, _ptr(codes),
                                    _stream(stream)))
    return codes


def sort30(codes, ctx=None, stream=None):
    ctx = ctx or default_context()
    n = codes.numel()
    sc = torch.empty(n, dtype=torch.int32, device=codes.device)
    si = torch.empty(n, dtype=torch.int32, device=codes.device)
    check(ctx.lib.srf_lbvh_sort30(ctx.handle, _ptr(codes), n, _ptr(sc), _ptr(si), _stream(stream)))
    return sc, si


def radix_tree30(sorted_codes, ctx=None, stream=None):
    ctx = ctx or default_context()
    n = sorted_codes.numel()
    out = torch.empty((max(n - 1, 0), 3), dtype=torch.int32, device=sorted_codes.device)
    check(ctx.lib.srf_lbvh_radix_tree30(ctx.handle, _ptr(sorted_codes), n, _ptr(out),
                                        _stream(stream)))
    return out


__all__ = ["Context", "TriangleAccel", "TetraAccel", "ParticleAccel", "SrfError",
           "device_tensor", "hits_to_numpy", "morton30", "sort30", "radix_tree30",
           "RAY_DTYPE", "HIT_DTYPE", "TET_EXT_DTYPE", "default_context"]


def rigid_inverses(T):
    """invT and invT33 of object-to-world matrices T [n, 4, 4] (row-vector
    convention) computed with numpy; element_inverses gives the reference's
    own (RenderElement::Update's Cramer-rule inverse), bit for bit."""
    T = np.asarray(T, np.float64)
    invT = np.linalg.inv(T)
    T33 = T.copy()
    T33[:, 3, :3] = 0.0
    return invT, np.linalg.inv(T33)


def element_inverses(T):
    """invT and invT33 of object-to-world matrices T [n, 4, 4] as
    RenderElement::Update computes them (srf_instance_update; host only)."""
    T = np.asarray(T, np.float64).reshape(-1, 4, 4)
    lib = _lib.load()
    invT = np.zeros_like(T)
    invT33 = np.zeros_like(T)
    for i in range(len(T)):
        inst = _lib.Instance()
        inst.T[:] = list(T[i].reshape(16))
        check(lib.srf_instance_update(C.byref(inst)))
        invT[i] = np.array(inst.invT[:]).reshape(4, 4)
        invT33[i] = np.array(inst.invT33[:]).reshape(4, 4)
    return invT, invT33


class Scene:
    """RenderGraph over accelerators of any kind (triangles, tetrahedra,
    spheres, solids, lines): Build (top-level BVH over the elements' world
    boxes) at construction, Trace for batches of rays."""

    def __init__(self, accels, T, invT=None, invT33=None, ctx=None):
        self.ctx = ctx or default_context()
        self.lib = self.ctx.lib
        self.accels = list(accels)  # keeps the elements alive
        T = np.asarray(T, np.float64).reshape(-1, 4, 4)
        if invT is None or invT33 is None:
            invT, invT33 = element_inverses(T)
        n = len(self.accels)
        arr = (_lib.Instance * n)()
        for i, a in enumerate(self.accels):
            a._require()
            arr[i].accel = a.handle.value
            arr[i].T[:] = list(T[i].reshape(16))
            arr[i].invT[:] = list(np.asarray(invT[i], np.float64).reshape(16))
            arr[i].invT33[:] = list(np.asarray(invT33[i], np.float64).reshape(16))
        h = C.c_void_p()
        check(self.lib.srf_scene_create(self.ctx.handle, n, arr, C.byref(h)))
        self.handle = h

    def __del__(self):
        try:
            if self.handle:
                self.lib.srf_scene_destroy(self.handle)
                self.handle = None
        except Exception:
            pass

    def BoundingBox(self):
        bmin = (C.c_double * 3)()
        bmax = (C.c_double * 3)()
        check(self.lib.srf_scene_bounds(self.handle, bmin, bmax))
        return np.array(bmin[:]), np.array(bmax[:])

    def export_tlas(self):
        cnt = C.c_uint32()
        check(self.lib.srf_scene_export(self.handle, None, C.byref(cnt)))
        arr = (_lib.TlasNode * cnt.value)()
        check(self.lib.srf_scene_export(self.handle, arr, C.byref(cnt)))
        return np.frombuffer(bytes(arr), dtype=np.dtype([
            ("bmin", "<f8", (3,)), ("bmax", "<f8", (3,)), ("child", "<i4", (2,)),
            ("axis", "<i4"), ("is_leaf", "<i4"), ("node_id", "<u4"), ("pad", "<u4")])).copy()

    def Trace(self, rays, hits=None, instance_ids=None, mode="tuned", stream=None):
        """Closest hit of every ray over all elements: hits [n, 4] (world t,
        u, v, prim id bits) and instance_ids [n] int32 (-1 on a miss).
        mode: "tuned" (the autotuned launch), "plain" or "persistent"."""
        n = rays.shape[0]
        if hits is None:
            hits = torch.empty((n, 4), dtype=torch.float32, device=rays.device)
        if instance_ids is None:
            instance_ids = torch.empty(n, dtype=torch.int32, device=rays.device)
        m = {"tuned": _lib.LAUNCH_TUNED, "plain": _lib.LAUNCH_PLAIN,
             "persistent": _lib.LAUNCH_PERSISTENT}[mode]
        check(self.lib.srf_scene_trace(self.handle, _ptr(rays), n, _ptr(hits),
                                       _ptr(instance_ids), m, _stream(stream)))
        return hits, instance_ids

    def autotune(self, rays, hits=None, instance_ids=None, stream=None):
        """Picks the fastest launch configuration on these rays; returns it."""
        n = rays.shape[0]
        if hits is None:
            hits = torch.empty((n, 4), dtype=torch.float32, device=rays.device)
        if instance_ids is None:
            instance_ids = torch.empty(n, dtype=torch.int32, device=rays.device)
        cfg = _lib.LaunchConfig()
        check(self.lib.srf_scene_autotune(self.handle, _ptr(rays), n, _ptr(hits),
                                          _ptr(instance_ids), _stream(stream), C.byref(cfg)))
        return {"block_size": cfg.block_size, "persistent": bool(cfg.persistent),
                "blocks_per_cu": cfg.blocks_per_cu, "ms": cfg.ms}
