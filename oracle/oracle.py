"""ctypes front-end of the parity checkers (test infrastructure only).

`Oracle` wraps liboracle.so, the C restatement of the reference algorithms;
`Reference` wraps oracle/_ref/libsrfref_{serial,omp}.so, the reference's own
render/ sources compiled by oracle/Makefile. Only tests/, the smoke check and
the bench's cpu_baseline leg import this module.
"""
from __future__ import annotations

import ctypes as C
import os
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

TRI_NODE = np.dtype([("bmin", "<f4", (2, 3)), ("bmax", "<f4", (2, 3)),
                     ("flag", "<i4"), ("axis", "<i4"), ("data", "<u4", (2,))])
TET_NODE = np.dtype([("bmin", "<f4", (3,)), ("bmax", "<f4", (3,)),
                     ("flag", "<i4"), ("axis", "<i4"), ("data", "<u4", (2,))])
SPH_NODE = np.dtype([("bmin", "<f4", (2, 3)), ("bmax", "<f4", (2, 3)),
                     ("data", "<u4", (2,)), ("flag", "<i2"), ("axis", "<i2")])
TLAS_NODE = np.dtype([("bmin", "<f8", (3,)), ("bmax", "<f8", (3,)), ("child", "<i4", (2,)),
                      ("axis", "<i4"), ("is_leaf", "<i4"), ("node_id", "<u4"), ("pad", "<u4")])
assert TRI_NODE.itemsize == 64 and TET_NODE.itemsize == 40 and SPH_NODE.itemsize == 60
assert TLAS_NODE.itemsize == 72


class TriElement(C.Structure):
    _fields_ = [("nodes", C.c_void_p), ("indices", C.c_void_p), ("V", C.c_void_p),
                ("F", C.c_void_p), ("eps_scale", C.c_double), ("T", C.c_double * 16),
                ("invT", C.c_double * 16), ("invT33", C.c_double * 16)]

class Element(C.Structure):
    """or_element: one element of a scene mixing primitive kinds."""
    _fields_ = [("kind", C.c_int32), ("nv", C.c_int32), ("cap", C.c_int32),
                ("rconst", C.c_float), ("nodes", C.c_void_p), ("indices", C.c_void_p),
                ("V", C.c_void_p), ("F", C.c_void_p), ("R", C.c_void_p),
                ("eps_scale", C.c_double), ("T", C.c_double * 16), ("invT", C.c_double * 16),
                ("invT33", C.c_double * 16)]


ELEMENT_KINDS = {"tri": (0, TRI_NODE), "tet": (1, TET_NODE), "sph": (2, SPH_NODE),
                 "sol": (3, TET_NODE), "lin": (4, TET_NODE)}

_p = C.c_void_p
_i64 = C.c_int64


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


class Oracle:
    """The C restatement (oracle/srf_oracle.c)."""

    def __init__(self, path=None):
        path = path or os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
        L = self.lib = C.CDLL(path)
        for name in ("or_scene_bounds_tri", "or_scene_bounds_tet", "or_morton_tri",
                     "or_morton_tet", "or_morton_pts", "or_sort_keys",
                     "or_radix_tree30", "or_trace_tri", "or_trace_tet",
                     "or_trace_sph", "or_tex3d_sample", "or_volume_fwd_f32",
                     "or_volume_fwd_f64", "or_volume_bwd_f64",
                     "or_scene_bounds_sph", "or_transform_box", "or_scene_trace",
                     "or_scene_bounds_sol", "or_morton_sol", "or_trace_sol",
                     "or_interpolate_solid", "or_scene_bounds_lin", "or_morton_lin",
                     "or_trace_lin", "or_isect_tri", "or_isect_tet", "or_isect_sph",
                     "or_scene_trace_mixed"):
            getattr(L, name).restype = None
        L.or_tlas_build.restype = _i64
        for name in ("or_build_tri", "or_build_tet", "or_build_sph", "or_build_sol",
                     "or_build_lin"):
            getattr(L, name).restype = _i64
        L.or_morton30_point.restype = C.c_uint32
        L.or_solid_isect.restype = C.c_int
        L.or_line_isect.restype = C.c_int

    # ---- stages ----
    def scene_bounds(self, kind, V, F=None, R=None, rconst=0.0):
        bmin = np.zeros(3, np.float32)
        bmax = np.zeros(3, np.float32)
        V = _c(V, np.float32)
        if kind == "tri":
            F = _c(F, np.uint32)
            self.lib.or_scene_bounds_tri(_ptr(V), _ptr(F), _i64(len(F)), _ptr(bmin), _ptr(bmax))
        elif kind == "tet":
            F = _c(F, np.uint32)
            self.lib.or_scene_bounds_tet(_ptr(V), _ptr(F), _i64(len(F)), _ptr(bmin), _ptr(bmax))
        else:
            R = None if R is None else _c(R, np.float32)
            self.lib.or_scene_bounds_sph(_ptr(V), _ptr(R), C.c_float(rconst),
                                         _i64(len(V)), _ptr(bmin), _ptr(bmax))
        return bmin, bmax

    def morton(self, kind, V, F, bmin, bmax):
        V = _c(V, np.float32)
        bmin = _c(bmin, np.float32)
        bmax = _c(bmax, np.float32)
        if kind == "pts":
            codes = np.zeros(len(V), np.uint32)
            self.lib.or_morton_pts(_ptr(V), _i64(len(V)), _ptr(bmin), _ptr(bmax), _ptr(codes))
            return codes
        F = _c(F, np.uint32)
        codes = np.zeros(len(F), np.uint32)
        fn = self.lib.or_morton_tri if kind == "tri" else self.lib.or_morton_tet
        fn(_ptr(V), _ptr(F), _i64(len(F)), _ptr(bmin), _ptr(bmax), _ptr(codes))
        return codes

    def morton_point(self, p, bmin, inv):
        p = _c(p, np.float32)
        bmin = _c(bmin, np.float32)
        return int(self.lib.or_morton30_point(_ptr(p), _ptr(bmin), C.c_float(inv[0]),
                                              C.c_float(inv[1]), C.c_float(inv[2])))

    def sort_keys(self, codes):
        codes = _c(codes, np.uint32)
        sc = np.zeros_like(codes)
        si = np.zeros_like(codes)
        self.lib.or_sort_keys(_ptr(codes), _i64(len(codes)), _ptr(sc), _ptr(si))
        return sc, si

    def radix_tree(self, sorted_codes):
        sc = _c(sorted_codes, np.uint32)
        out = np.zeros((max(len(sc) - 1, 0), 3), np.uint32)
        self.lib.or_radix_tree30(_ptr(sc), _i64(len(sc)), _ptr(out))
        return out

    # ---- builds ----
    def build_tri(self, V, F):
        V = _c(V, np.float32)
        F = _c(F, np.uint32)
        n = len(F)
        nodes = np.zeros(2 * n, TRI_NODE)
        idx = np.zeros(n, np.uint32)
        bounds = np.zeros(6, np.float64)
        es = C.c_double()
        m = self.lib.or_build_tri(_ptr(V), _ptr(F), _i64(n), _ptr(nodes), _ptr(idx),
                                  _ptr(bounds), C.byref(es))
        return dict(nodes=nodes[:m].copy(), indices=idx, bounds=bounds,
                    eps_scale=es.value)

    def build_tet(self, V, F):
        V = _c(V, np.float32)
        F = _c(F, np.uint32)
        n = len(F)
        nodes = np.zeros(2 * n, TET_NODE)
        idx = np.zeros(n, np.uint32)
        bounds = np.zeros(6, np.float64)
        m = self.lib.or_build_tet(_ptr(V), _ptr(F), _i64(n), _ptr(nodes), _ptr(idx), _ptr(bounds))
        return dict(nodes=nodes[:m].copy(), indices=idx, bounds=bounds)

    def build_sol(self, V, F, nv):
        """SolidAccel::Build32 over solids of nv vertices (F: nv per solid)."""
        V = _c(V, np.float32)
        F = _c(F, np.uint32).reshape(-1, nv)
        n = len(F)
        nodes = np.zeros(2 * n, TET_NODE)
        idx = np.zeros(n, np.uint32)
        bounds = np.zeros(6, np.float64)
        m = self.lib.or_build_sol(_ptr(V), _ptr(F), _i64(n), C.c_int(nv), _ptr(nodes), _ptr(idx),
                                  _ptr(bounds))
        return dict(nodes=nodes[:m].copy(), indices=idx, bounds=bounds, nv=nv)

    def scene_bounds_sol(self, V, F, nv):
        bmin = np.zeros(3, np.float32)
        bmax = np.zeros(3, np.float32)
        F = _c(F, np.uint32).reshape(-1, nv)
        self.lib.or_scene_bounds_sol(_ptr(_c(V, np.float32)), _ptr(F), _i64(len(F)), C.c_int(nv),
                                     _ptr(bmin), _ptr(bmax))
        return bmin, bmax

    def morton_sol(self, V, F, nv, bmin, bmax):
        F = _c(F, np.uint32).reshape(-1, nv)
        codes = np.zeros(len(F), np.uint32)
        self.lib.or_morton_sol(_ptr(_c(V, np.float32)), _ptr(F), _i64(len(F)), C.c_int(nv),
                               _ptr(_c(bmin, np.float32)), _ptr(_c(bmax, np.float32)),
                               _ptr(codes))
        return codes

    def solid_isect(self, nv, o, d, verts):
        """One ray against one solid: (status, t[2], pos[2,3], nrm[2,3])."""
        t = np.zeros(2, np.float32)
        pos = np.zeros(6, np.float32)
        nrm = np.zeros(6, np.float32)
        st = self.lib.or_solid_isect(C.c_int(nv), _ptr(_c(o, np.float64)), _ptr(_c(d, np.float64)),
                                     _ptr(_c(verts, np.float64)), _ptr(t), _ptr(pos), _ptr(nrm))
        return st, t, pos.reshape(2, 3), nrm.reshape(2, 3)

    def interpolate_solid(self, nv, p, verts):
        d = np.zeros(8, np.float32)
        self.lib.or_interpolate_solid(C.c_int(nv), _ptr(_c(p, np.float32)),
                                      _ptr(_c(verts, np.float64)), _ptr(d))
        return d

    def build_lin(self, P, segs, R=None, rconst=0.0):
        """The LBVH over line segments (segs: 2 vertex ids per line; radius
        per vertex R, or rconst)."""
        P = _c(P, np.float32)
        segs = _c(segs, np.uint32).reshape(-1, 2)
        R = None if R is None else _c(R, np.float32)
        n = len(segs)
        nodes = np.zeros(2 * n, TET_NODE)
        idx = np.zeros(n, np.uint32)
        bounds = np.zeros(6, np.float64)
        m = self.lib.or_build_lin(_ptr(P), _ptr(segs), _ptr(R), C.c_float(rconst), _i64(n),
                                  _ptr(nodes), _ptr(idx), _ptr(bounds))
        return dict(nodes=nodes[:m].copy(), indices=idx, bounds=bounds)

    def line_isect(self, p0, p1, rr, o, d, cap=True, tmax=np.finfo(np.float32).max):
        out = np.zeros(3, np.float32)
        hc = C.c_int()
        h = self.lib.or_line_isect(_ptr(_c(p0, np.float32)), _ptr(_c(p1, np.float32)),
                                   C.c_float(rr), _ptr(_c(o, np.float32)), _ptr(_c(d, np.float32)),
                                   C.c_int(1 if cap else 0), C.c_float(tmax), _ptr(out),
                                   C.byref(hc))
        return h, out, hc.value

    def trace_lin(self, tree, P, segs, R, rconst, rays, cap=True, threads=1, stats=False):
        from surface_amd.workloads import HIT_DTYPE, LIN_EXT_DTYPE
        hits = np.zeros(len(rays), HIT_DTYPE)
        ext = np.zeros(len(rays), LIN_EXT_DTYPE)
        st = np.zeros((len(rays), 2), np.uint32) if stats else None
        Rc = None if R is None else _c(R, np.float32)
        self.lib.or_trace_lin(_ptr(tree["nodes"]), _ptr(tree["indices"]), _ptr(_c(P, np.float32)),
                              _ptr(_c(segs, np.uint32)), _ptr(Rc), C.c_float(rconst),
                              C.c_int(1 if cap else 0), _ptr(rays), _i64(len(rays)), _ptr(hits),
                              _ptr(ext), _ptr(st), C.c_int(threads))
        return (hits, ext, st) if stats else (hits, ext)

    def build_sph(self, P, R=None, rconst=0.0):
        P = _c(P, np.float32)
        R = None if R is None else _c(R, np.float32)
        n = len(P)
        nodes = np.zeros(2 * n, SPH_NODE)
        idx = np.zeros(n, np.uint32)
        bounds = np.zeros(6, np.float64)
        m = self.lib.or_build_sph(_ptr(P), _ptr(R), C.c_float(rconst), _i64(n), _ptr(nodes),
                                  _ptr(idx), _ptr(bounds))
        return dict(nodes=nodes[:m].copy(), indices=idx, bounds=bounds)

    # ---- traversal ----
    def trace_tri(self, tree, V, F, rays, threads=1, stats=False):
        from surface_amd.workloads import HIT_DTYPE
        hits = np.zeros(len(rays), HIT_DTYPE)
        st = np.zeros((len(rays), 2), np.uint32) if stats else None
        self.lib.or_trace_tri(_ptr(tree["nodes"]), _ptr(tree["indices"]), _ptr(_c(V, np.float32)),
                              _ptr(_c(F, np.uint32)), C.c_double(tree["eps_scale"]),
                              _ptr(rays), _i64(len(rays)), _ptr(hits), _ptr(st), C.c_int(threads))
        return (hits, st) if stats else hits

    def trace_tet(self, tree, V, F, rays, scalars=None, threads=1, stats=False):
        from surface_amd.workloads import HIT_DTYPE, TET_EXT_DTYPE
        hits = np.zeros(len(rays), HIT_DTYPE)
        ext = np.zeros(len(rays), TET_EXT_DTYPE)
        st = np.zeros((len(rays), 2), np.uint32) if stats else None
        sc = None if scalars is None else _c(scalars, np.float32)
        self.lib.or_trace_tet(_ptr(tree["nodes"]), _ptr(tree["indices"]), _ptr(_c(V, np.float32)),
                              _ptr(_c(F, np.uint32)), _ptr(sc), _ptr(rays), _i64(len(rays)),
                              _ptr(hits), _ptr(ext), _ptr(st), C.c_int(threads))
        return (hits, ext, st) if stats else (hits, ext)

    def trace_sol(self, tree, V, F, rays, scalars=None, threads=1, stats=False):
        from surface_amd.workloads import HIT_DTYPE, SOL_EXT_DTYPE
        nv = tree["nv"]
        hits = np.zeros(len(rays), HIT_DTYPE)
        ext = np.zeros(len(rays), SOL_EXT_DTYPE)
        st = np.zeros((len(rays), 2), np.uint32) if stats else None
        sc = None if scalars is None else _c(scalars, np.float32)
        self.lib.or_trace_sol(_ptr(tree["nodes"]), _ptr(tree["indices"]), _ptr(_c(V, np.float32)),
                              _ptr(_c(F, np.uint32)), C.c_int(nv), _ptr(sc), _ptr(rays),
                              _i64(len(rays)), _ptr(hits), _ptr(ext), _ptr(st), C.c_int(threads))
        return (hits, ext, st) if stats else (hits, ext)

    def trace_sph(self, tree, P, R, rconst, rays, prev_t=None, prev_n=None, threads=1,
                  stats=False):
        from surface_amd.workloads import HIT_DTYPE
        hits = np.zeros(len(rays), HIT_DTYPE)
        st = np.zeros((len(rays), 2), np.uint32) if stats else None
        Rr = None if R is None else _c(R, np.float32)
        pt = None if prev_t is None else _c(prev_t, np.float64)
        pn = None if prev_n is None else _c(prev_n, np.float32)
        self.lib.or_trace_sph(_ptr(tree["nodes"]), _ptr(tree["indices"]), _ptr(_c(P, np.float32)),
                              _ptr(Rr), C.c_float(rconst), _ptr(rays), _ptr(pt), _ptr(pn),
                              _i64(len(rays)), _ptr(hits), _ptr(st), C.c_int(threads))
        return (hits, st) if stats else hits

    # ---- volume ----
    def tex3d(self, vox, dims, clamp, uvw):
        vox = _c(vox, np.float32)
        uvw = _c(uvw, np.float32).reshape(-1, 3)
        out = np.zeros(len(uvw), np.float32)
        d = _c(dims, np.int32)
        cl = _c(clamp, np.int32)
        self.lib.or_tex3d_sample(_ptr(vox), _ptr(d), _ptr(cl), _ptr(uvw), _i64(len(uvw)), _ptr(out))
        return out

    def tex3d_fetch(self, image, comps, dims, clamp, uvw, remap=(None, None, None)):
        """Texture::Fetch: image uint8 / float32 / float64 with comps
        interleaved components; remap: per-axis float32 tables or None.
        Returns rgba [n, 4] float32."""
        fmt = {np.dtype(np.uint8): 1, np.dtype(np.float32): 8, np.dtype(np.float64): 16}[
            np.asarray(image).dtype]
        img = np.ascontiguousarray(image)
        uvw = _c(uvw, np.float32).reshape(-1, 3)
        out = np.zeros((len(uvw), 4), np.float32)
        tabs = [None if t is None else _c(t, np.float32) for t in remap]
        ptrs = (C.c_void_p * 3)(*[None if t is None else t.ctypes.data for t in tabs])
        ns = np.array([0 if t is None else len(t) for t in tabs], np.int32)
        self.lib.or_tex3d_fetch.restype = None
        self.lib.or_tex3d_fetch(_ptr(img), C.c_int(fmt), C.c_int(comps), _ptr(_c(dims, np.int32)),
                                _ptr(_c(clamp, np.int32)), ptrs, _ptr(ns), _ptr(uvw),
                                _i64(len(uvw)), _ptr(out))
        return out

    def sparse_sample(self, blocks, data, components, pos=None, uvw=None, global_dim=(0, 0, 0)):
        """SparseVolumeAccel::Sample over blocks (int32 [nb, 9]: offset,
        extent, size) with voxel arrays data[b] (uint8 / float32 / float64);
        uvw instead of pos samples Texture::Fetch's sparse branch."""
        fmt = {np.dtype(np.uint8): 1, np.dtype(np.float32): 8, np.dtype(np.float64): 16}[
            np.asarray(data[0]).dtype]
        blocks = _c(blocks, np.int32).reshape(-1, 9)
        arrs = [np.ascontiguousarray(d) for d in data]
        ptrs = (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
        n = len(pos) if pos is not None else len(uvw)
        out = np.zeros((n, 4), np.float64)
        pd = None if pos is None else _c(pos, np.float64).reshape(-1, 3)
        ud = None if uvw is None else _c(uvw, np.float32).reshape(-1, 3)
        self.lib.or_sparse_sample.restype = None
        self.lib.or_sparse_sample(_ptr(blocks), ptrs, _i64(len(blocks)), C.c_int(components),
                                  C.c_int(fmt), _ptr(_c(global_dim, np.int32)),
                                  None if pd is None else _ptr(pd),
                                  None if ud is None else _ptr(ud), _i64(n), _ptr(out))
        return out

    @staticmethod
    def _vparams(dims, clamp, tf_size, step):
        class P(C.Structure):
            _fields_ = [("dims", C.c_int * 3), ("clamp", C.c_int * 3),
                        ("tf_size", C.c_int), ("step", C.c_float)]
        p = P()
        p.dims[:] = [int(x) for x in dims]
        p.clamp[:] = [int(x) for x in clamp]
        p.tf_size = int(tf_size)
        p.step = float(step)
        return p

    def volume_fwd(self, vox, dims, clamp, tf, step, rays, threads=1):
        p = self._vparams(dims, clamp, len(tf), step)
        rgba = np.zeros((len(rays), 4), np.float32)
        ns = np.zeros(len(rays), np.uint32)
        self.lib.or_volume_fwd_f32(C.byref(p), _ptr(_c(vox, np.float32)), _ptr(_c(tf, np.float32)),
                                   _ptr(rays), _i64(len(rays)), _ptr(rgba), _ptr(ns), C.c_int(threads))
        return rgba, ns

    def volume_fwd64(self, vox, dims, clamp, tf, step, rays, threads=1):
        p = self._vparams(dims, clamp, len(tf), step)
        rgba = np.zeros((len(rays), 4), np.float64)
        self.lib.or_volume_fwd_f64(C.byref(p), _ptr(_c(vox, np.float64)), _ptr(_c(tf, np.float64)),
                                   _ptr(rays), _i64(len(rays)), _ptr(rgba), C.c_int(threads))
        return rgba

    def volume_bwd64(self, vox, dims, clamp, tf, step, rays, grad_rgba):
        p = self._vparams(dims, clamp, len(tf), step)
        gv = np.zeros(len(vox), np.float64)
        gt = np.zeros(np.shape(tf), np.float64)
        self.lib.or_volume_bwd_f64(C.byref(p), _ptr(_c(vox, np.float64)), _ptr(_c(tf, np.float64)),
                                   _ptr(rays), _i64(len(rays)), _ptr(_c(grad_rgba, np.float64)),
                                   _ptr(gv), _ptr(gt))
        return gv, gt


    def isect(self, kind, V, F, rays, hits, ext=None):
        """BuildIntersection fields (ISECT_DTYPE) of traced hits: kind "tri"
        (V, F faces), "tet" (V, F, ext with the entry faces) or "sph" (V =
        centres, F unused)."""
        from surface_amd.workloads import ISECT_DTYPE
        rays = np.ascontiguousarray(rays)
        hits = np.ascontiguousarray(hits)
        out = np.zeros(len(rays), ISECT_DTYPE)
        Vc = _c(V, np.float32)
        if kind == "tri":
            self.lib.or_isect_tri(_ptr(Vc), _ptr(_c(F, np.uint32)), _ptr(rays), _ptr(hits),
                                  _i64(len(rays)), _ptr(out))
        elif kind == "tet":
            self.lib.or_isect_tet(_ptr(Vc), _ptr(_c(F, np.uint32)), _ptr(rays), _ptr(hits),
                                  _ptr(np.ascontiguousarray(ext)), _i64(len(rays)), _ptr(out))
        else:
            self.lib.or_isect_sph(_ptr(Vc), _ptr(rays), _ptr(hits), _i64(len(rays)), _ptr(out))
        return out

    # ---- two-level scene ----
    def transform_box(self, bmin, bmax, T):
        out = np.zeros(6, np.float64)
        self.lib.or_transform_box(_ptr(_c(bmin, np.float64)), _ptr(_c(bmax, np.float64)),
                                  _ptr(_c(T, np.float64).reshape(16)), _ptr(out))
        return out

    def tlas_build(self, boxes):
        boxes = _c(boxes, np.float64).reshape(-1, 6)
        nodes = np.zeros(2 * len(boxes), TLAS_NODE)
        m = self.lib.or_tlas_build(_ptr(boxes), _i64(len(boxes)), _ptr(nodes))
        return nodes[:m].copy()

    def scene_trace(self, tlas, elements, rays, threads=1):
        """elements: list of dicts with the element's tree (build_tri), V, F,
        T, invT, invT33 (4x4 float64, row-vector convention)."""
        from surface_amd.workloads import HIT_DTYPE
        keep = []
        arr = (TriElement * len(elements))()
        for k, e in enumerate(elements):
            nodes = _c(e["tree"]["nodes"], TRI_NODE)
            idx = _c(e["tree"]["indices"], np.uint32)
            V = _c(e["V"], np.float32)
            F = _c(e["F"], np.uint32)
            keep += [nodes, idx, V, F]
            arr[k].nodes, arr[k].indices = _ptr(nodes).value, _ptr(idx).value
            arr[k].V, arr[k].F = _ptr(V).value, _ptr(F).value
            arr[k].eps_scale = e["tree"]["eps_scale"]
            for name in ("T", "invT", "invT33"):
                getattr(arr[k], name)[:] = list(np.asarray(e[name], np.float64).reshape(16))
        tl = _c(tlas, TLAS_NODE)
        hits = np.zeros(len(rays), HIT_DTYPE)
        inst = np.zeros(len(rays), np.uint32)
        self.lib.or_scene_trace(_ptr(tl), arr, _ptr(rays), _i64(len(rays)), _ptr(hits),
                                _ptr(inst), C.c_int(threads))
        return hits, inst


    def build_element(self, e):
        """The tree of one element of a mixed scene (a dict as
        surface_amd.workloads.mixed_scene makes them)."""
        k = e["kind"]
        if k == "tri":
            return self.build_tri(e["V"], e["F"])
        if k == "tet":
            return self.build_tet(e["V"], e["F"])
        if k == "sph":
            return self.build_sph(e["V"], e.get("R"), e.get("rconst", 0.0))
        if k == "sol":
            return self.build_sol(e["V"], e["F"], e["nv"])
        return self.build_lin(e["V"], e["F"], e.get("R"), e.get("rconst", 0.0))

    def scene_trace_mixed(self, tlas, elements, rays, threads=1):
        """elements: list of dicts with kind ("tri", "tet", "sph", "sol" or
        "lin"), the element's tree (from the matching build_*), its arrays
        (V, F; spheres: V = centres, R or rconst; solids: nv; lines: R or
        rconst, cap) and T, invT, invT33 (4x4 float64)."""
        from surface_amd.workloads import HIT_DTYPE
        keep = []
        arr = (Element * len(elements))()
        for k, e in enumerate(elements):
            code, ndt = ELEMENT_KINDS[e["kind"]]
            nodes = _c(e["tree"]["nodes"], ndt)
            idx = _c(e["tree"]["indices"], np.uint32)
            V = _c(e["V"], np.float32)
            F = None if e.get("F") is None else _c(e["F"], np.uint32)
            R = None if e.get("R") is None else _c(e["R"], np.float32)
            keep += [nodes, idx, V, F, R]
            arr[k].kind, arr[k].nv, arr[k].cap = code, int(e.get("nv", 0)), int(e.get("cap", 1))
            arr[k].rconst = float(e.get("rconst", 0.0))
            arr[k].nodes, arr[k].indices = _ptr(nodes).value, _ptr(idx).value
            arr[k].V = _ptr(V).value
            arr[k].F = None if F is None else _ptr(F).value
            arr[k].R = None if R is None else _ptr(R).value
            arr[k].eps_scale = e["tree"].get("eps_scale", 0.0)
            for name in ("T", "invT", "invT33"):
                getattr(arr[k], name)[:] = list(np.asarray(e[name], np.float64).reshape(16))
        tl = _c(tlas, TLAS_NODE)
        rays = np.ascontiguousarray(rays)
        hits = np.zeros(len(rays), HIT_DTYPE)
        inst = np.zeros(len(rays), np.uint32)
        self.lib.or_scene_trace_mixed(_ptr(tl), arr, _ptr(rays), _i64(len(rays)), _ptr(hits),
                                      _ptr(inst), C.c_int(threads))
        return hits, inst


def ref_available(variant="serial"):
    return os.path.exists(os.path.join(HERE, "_ref", f"libsrfref_{variant}.so"))


class Reference:
    """The reference's own render/ code (oracle/_ref), via ref_driver.cc."""

    def __init__(self, variant="serial"):
        path = os.path.join(HERE, "_ref", f"libsrfref_{variant}.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle ref`")
        L = self.lib = C.CDLL(path)
        for name in ("ref_build_tri", "ref_build_tet", "ref_build_sph", "ref_build_sol",
                     "ref_build_lin"):
            getattr(L, name).restype = _p
        for name in ("ref_morton_tri", "ref_morton_tet", "ref_morton_pts", "ref_sort_keys",
                     "ref_radix_tree30", "ref_bounds", "ref_free", "ref_trace",
                     "ref_tex3d_sample"):
            getattr(L, name).restype = None
        L.ref_dump.restype = C.c_int
        L.ref_has_openmp.restype = C.c_int
        for name in ("ref_free", "ref_bounds", "ref_dump", "ref_trace"):
            getattr(L, name).argtypes = None

    def morton(self, kind, V, F, bmin, bmax, simd=True):
        V = _c(V, np.float32)
        bmin = _c(bmin, np.float32)
        bmax = _c(bmax, np.float32)
        if kind == "pts":
            codes = np.zeros(len(V), np.uint32)
            self.lib.ref_morton_pts(_ptr(V), _i64(len(V)), _ptr(bmin), _ptr(bmax), _ptr(codes),
                                    C.c_int(1 if simd else 0))
            return codes
        F = _c(F, np.uint32)
        codes = np.zeros(len(F), np.uint32)
        fn = self.lib.ref_morton_tri if kind == "tri" else self.lib.ref_morton_tet
        fn(_ptr(V), _ptr(F), _i64(len(F)), _ptr(bmin), _ptr(bmax), _ptr(codes))
        return codes

    def sort_keys(self, codes):
        codes = _c(codes, np.uint32)
        sc = np.zeros_like(codes)
        si = np.zeros_like(codes)
        self.lib.ref_sort_keys(_ptr(codes), _i64(len(codes)), _ptr(sc), _ptr(si))
        return sc, si

    def radix_tree(self, sorted_codes):
        sc = _c(sorted_codes, np.uint32)
        out = np.zeros((max(len(sc) - 1, 0), 3), np.uint32)
        self.lib.ref_radix_tree30(_ptr(sc), _i64(len(sc)), _ptr(out))
        return out

    def _read_dump(self, h, dtype):
        fd, path = tempfile.mkstemp(suffix=".bvh")
        os.close(fd)
        try:
            if self.lib.ref_dump(_p(h), path.encode()) != 0:
                raise RuntimeError("reference Dump failed")
            raw = np.fromfile(path, np.uint8)
        finally:
            os.unlink(path)
        nn = int(raw[:8].view(np.uint64)[0])
        off = 8 + nn * dtype.itemsize
        nodes = raw[8:off].view(dtype).copy()
        ni = int(raw[off:off + 8].view(np.uint64)[0])
        idx = raw[off + 8:off + 8 + 4 * ni].view(np.uint32).copy()
        return nodes, idx

    def _bounds(self, h):
        bmin = np.zeros(3, np.float64)
        bmax = np.zeros(3, np.float64)
        self.lib.ref_bounds(_p(h), _ptr(bmin), _ptr(bmax))
        return np.concatenate([bmin, bmax])

    def build(self, kind, V, F=None, R=None, rconst=0.0, keep=False, nv=0, cap=True):
        """Run Build32; return nodes (reference layout), indices and bounds.
        With keep=True the handle stays alive for trace() (call free()).
        kind "sol" builds solids of nv vertices (SolidAccel::Build32); kind
        "lin" runs LineAccel::Build (F = 2 vertex ids per line)."""
        V = _c(V, np.float32)
        keepalive = [V]
        if kind == "lin":
            F = _c(F, np.uint32).reshape(-1, 2)
            R = None if R is None else _c(R, np.float32)
            keepalive += [F, R]
            h = self.lib.ref_build_lin(_ptr(V), _ptr(F), _ptr(R), C.c_float(rconst),
                                       _i64(len(F)), C.c_int(1 if cap else 0))
            dtype = TET_NODE
        elif kind == "sol":
            F = _c(F, np.uint32).reshape(-1, nv)
            keepalive.append(F)
            h = self.lib.ref_build_sol(_ptr(V), _ptr(F), _i64(len(F)), C.c_int(nv))
            dtype = TET_NODE
        elif kind == "tri":
            F = _c(F, np.uint32)
            keepalive.append(F)
            h = self.lib.ref_build_tri(_ptr(V), _i64(len(V)), _ptr(F), _i64(len(F)))
            dtype = TRI_NODE
        elif kind == "tet":
            F = _c(F, np.uint32)
            keepalive.append(F)
            h = self.lib.ref_build_tet(_ptr(V), _ptr(F), _i64(len(F)))
            dtype = TET_NODE
        else:
            R = None if R is None else _c(R, np.float32)
            keepalive.append(R)
            h = self.lib.ref_build_sph(_ptr(V), _ptr(R), C.c_float(rconst), _i64(len(V)))
            dtype = SPH_NODE
        nodes, idx = self._read_dump(h, dtype)
        out = dict(nodes=nodes, indices=idx, bounds=self._bounds(h), kind=kind, nv=nv)
        if keep:
            out["handle"] = h
            out["_keep"] = keepalive
        else:
            self.lib.ref_free(_p(h))
        return out

    def trace(self, tree, rays, width=0, scalars=None, threads=0):
        from surface_amd.workloads import HIT_DTYPE, LIN_EXT_DTYPE, SOL_EXT_DTYPE, TET_EXT_DTYPE
        hits = np.zeros(len(rays), HIT_DTYPE)
        ext = np.zeros(len(rays), {"sol": SOL_EXT_DTYPE, "lin": LIN_EXT_DTYPE}.get(
            tree.get("kind"), TET_EXT_DTYPE))
        sc = None if scalars is None else _c(scalars, np.float32)
        self.lib.ref_trace(_p(tree["handle"]), _ptr(rays), _i64(len(rays)), _i64(width),
                           _ptr(hits), _ptr(ext), _ptr(sc), C.c_int(threads))
        return hits, ext

    def trace_isect(self, tree, rays):
        """The Intersection fields Traverse's BuildIntersection fills
        (ISECT_DTYPE; triangles, tetras, spheres; serial)."""
        from surface_amd.workloads import ISECT_DTYPE
        out = np.zeros(len(rays), ISECT_DTYPE)
        self.lib.ref_trace_isect.restype = None
        self.lib.ref_trace_isect(_p(tree["handle"]), _ptr(np.ascontiguousarray(rays)),
                                 _i64(len(rays)), _ptr(out))
        return out

    def free(self, tree):
        if tree.get("handle"):
            self.lib.ref_free(_p(tree["handle"]))
            tree["handle"] = None

    def scene_build(self, trees, T):
        """RenderGraph over accelerators of any kind (trees from build(...,
        keep=True); each becomes the RenderElement glDrawElements records)
        with object-to-world matrices T [n, 4, 4]; returns the reference's
        invT, invT33, world boxes, top-level BVH nodes and the graph handle."""
        n = len(trees)
        T = _c(T, np.float64).reshape(n, 16)
        hs = (C.c_void_p * n)(*[t["handle"] for t in trees])
        invT = np.zeros((n, 16), np.float64)
        invT33 = np.zeros((n, 16), np.float64)
        xbox = np.zeros((n, 6), np.float64)
        cap = 2 * n + 2
        tbox = np.zeros((cap, 6), np.float64)
        tmeta = np.zeros((cap, 5), np.int32)
        cnt = C.c_int()
        self.lib.ref_scene_build.restype = _p
        h = self.lib.ref_scene_build(C.c_int(n), hs, _ptr(T), _ptr(invT), _ptr(invT33),
                                     _ptr(xbox), _ptr(tbox), _ptr(tmeta), C.byref(cnt),
                                     C.c_int(cap))
        m = cnt.value
        tl = np.zeros(m, TLAS_NODE)
        tl["bmin"], tl["bmax"] = tbox[:m, :3], tbox[:m, 3:]
        tl["child"] = tmeta[:m, :2]
        tl["axis"], tl["is_leaf"], tl["node_id"] = tmeta[:m, 2], tmeta[:m, 3], tmeta[:m, 4]
        return dict(handle=h, invT=invT.reshape(n, 4, 4), invT33=invT33.reshape(n, 4, 4),
                    xbox=xbox, tlas=tl)

    def scene_trace(self, scene, rays, width=0, threads=0):
        from surface_amd.workloads import HIT_DTYPE
        hits = np.zeros(len(rays), HIT_DTYPE)
        inst = np.zeros(len(rays), np.uint32)
        self.lib.ref_scene_trace.restype = None
        self.lib.ref_scene_trace(_p(scene["handle"]), _ptr(rays), _i64(len(rays)), _i64(width),
                                 _ptr(hits), _ptr(inst), C.c_int(threads))
        return hits, inst

    def scene_free(self, scene):
        if scene.get("handle"):
            self.lib.ref_scene_free.restype = None
            self.lib.ref_scene_free(_p(scene["handle"]))
            scene["handle"] = None

    def tex3d(self, vox, dims, clamp, uvw):
        vox = _c(vox, np.float32)
        uvw = _c(uvw, np.float32).reshape(-1, 3)
        out = np.zeros(len(uvw), np.float32)
        self.lib.ref_tex3d_sample(_ptr(vox), _ptr(_c(dims, np.int32)), _ptr(_c(clamp, np.int32)),
                                  _ptr(uvw), _i64(len(uvw)), _ptr(out))
        return out


def _ref_tex3d_fetch(self, image, comps, dims, clamp, uvw):
    """FilterTexture3D{Byte,Float,Double} of the reference (no remap)."""
    fmt = {np.dtype(np.uint8): 1, np.dtype(np.float32): 8, np.dtype(np.float64): 16}[
        np.asarray(image).dtype]
    img = np.ascontiguousarray(image)
    uvw = _c(uvw, np.float32).reshape(-1, 3)
    out = np.zeros((len(uvw), 4), np.float32)
    self.lib.ref_tex3d_fetch.restype = None
    self.lib.ref_tex3d_fetch(_ptr(img), C.c_int(fmt), C.c_int(comps), _ptr(_c(dims, np.int32)),
                             _ptr(_c(clamp, np.int32)), _ptr(uvw), _i64(len(uvw)), _ptr(out))
    return out


Reference.tex3d_fetch = _ref_tex3d_fetch


def _ref_sparse_sample(self, blocks, data, components, pos):
    """SparseVolumeAccel::Build + Sample of the reference."""
    fmt = {np.dtype(np.uint8): 1, np.dtype(np.float32): 8, np.dtype(np.float64): 16}[
        np.asarray(data[0]).dtype]
    blocks = _c(blocks, np.int32).reshape(-1, 9)
    arrs = [np.ascontiguousarray(d) for d in data]
    ptrs = (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
    pos = _c(pos, np.float64).reshape(-1, 3)
    out = np.zeros((len(pos), 4), np.float64)
    self.lib.ref_sparse_sample.restype = None
    self.lib.ref_sparse_sample(_ptr(blocks), ptrs, _i64(len(blocks)), C.c_int(components),
                               C.c_int(fmt), _ptr(pos), _i64(len(pos)), _ptr(out))
    return out


Reference.sparse_sample = _ref_sparse_sample


def canonical_tri(nodes):
    """Zero the fields the reference leaves uninitialised (leaf axis and the
    unused second box slot of a leaf) so trees compare byte for byte."""
    n = nodes.copy()
    leaf = n["flag"] == 1
    n["axis"][leaf] = 0
    n["bmin"][leaf, 1] = 0
    n["bmax"][leaf, 1] = 0
    return n


def canonical_tet(nodes):
    n = nodes.copy()
    n["axis"][n["flag"] == 1] = 0
    return n


canonical_sph = canonical_tri


def canonical_tlas(nodes):
    """Zero what BVHTree leaves uninitialised: a branch's node id, a leaf's
    axis, and the padding."""
    n = nodes.copy()
    n["node_id"][n["is_leaf"] == 0] = 0
    n["axis"][n["is_leaf"] == 1] = 0
    n["pad"] = 0
    return n
