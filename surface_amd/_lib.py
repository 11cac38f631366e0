"""ctypes declarations of include/srf.h (libsrf.so).

torch is imported first when it is installed: torch ships its own
libamdhip64.so.7, and loading it before libsrf.so makes both use that one
HIP runtime (the SONAMEs are equal, so the dynamic linker reuses it).
"""
from __future__ import annotations

import ctypes as C
import os

try:  # noqa: SIM105 - see module docstring
    import torch  # noqa: F401
except ImportError:  # pragma: no cover
    torch = None

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsrf.so")

SRF_OK = 0
STATUS_NAMES = {0: "SRF_OK", 1: "SRF_ERR_INVALID_ARG", 2: "SRF_ERR_OUT_OF_MEMORY",
                3: "SRF_ERR_HIP", 4: "SRF_ERR_UNSUPPORTED", 5: "SRF_ERR_INTERNAL"}
PRIM_POINTS, PRIM_TRIANGLES, PRIM_TETRAHEDRONS = 0x0000, 0x0004, 0x0010
PRIM_PYRAMIDS, PRIM_PRISMS, PRIM_HEXAHEDRONS = 0x0011, 0x0012, 0x0013
PRIM_LINES = 0x0001
PRIM_SOLID = {5: PRIM_PYRAMIDS, 6: PRIM_PRISMS, 8: PRIM_HEXAHEDRONS}


class SrfError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class TriDesc(C.Structure):
    _fields_ = [("vertices", C.c_void_p), ("num_vertices", C.c_uint64),
                ("faces", C.c_void_p), ("num_faces", C.c_uint64)]


class TetDesc(C.Structure):
    _fields_ = [("vertices", C.c_void_p), ("num_vertices", C.c_uint64),
                ("indices", C.c_void_p), ("num_tetras", C.c_uint64)]


class SolidDesc(C.Structure):
    _fields_ = [("vertices", C.c_void_p), ("num_vertices", C.c_uint64),
                ("indices", C.c_void_p), ("num_solids", C.c_uint64),
                ("vertices_per_solid", C.c_uint32)]


class LineDesc(C.Structure):
    _fields_ = [("positions", C.c_void_p), ("num_vertices", C.c_uint64),
                ("segments", C.c_void_p), ("num_lines", C.c_uint64),
                ("radius", C.c_void_p), ("constant_radius", C.c_float), ("cap", C.c_uint32)]


class SphereDesc(C.Structure):
    _fields_ = [("positions", C.c_void_p), ("radius", C.c_void_p),
                ("constant_radius", C.c_float), ("num_particles", C.c_uint64)]


class AccelInfo(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("num_prims", C.c_uint64), ("num_nodes", C.c_uint64),
                ("num_leaves", C.c_uint64), ("max_depth", C.c_uint32),
                ("eps_scale", C.c_double), ("node_bytes", C.c_uint64)]


class Instance(C.Structure):
    """srf_instance: accelerator + T, invT, invT33 (row-major 4x4 doubles)."""
    _fields_ = [("accel", C.c_void_p), ("T", C.c_double * 16), ("invT", C.c_double * 16),
                ("invT33", C.c_double * 16)]


class LaunchConfig(C.Structure):
    """srf_launch_config."""
    _fields_ = [("block_size", C.c_uint32), ("persistent", C.c_uint32),
                ("blocks_per_cu", C.c_uint32), ("ms", C.c_float)]


LAUNCH_TUNED, LAUNCH_PLAIN, LAUNCH_PERSISTENT = 0, 1, 2


class AnyConfig(C.Structure):
    """srf_any_config."""
    _fields_ = [("key_bits", C.c_uint32), ("persistent", C.c_uint32), ("ms", C.c_float)]


class TlasNode(C.Structure):
    """srf_tlas_node (72 bytes)."""
    _fields_ = [("bmin", C.c_double * 3), ("bmax", C.c_double * 3), ("child", C.c_int32 * 2),
                ("axis", C.c_int32), ("is_leaf", C.c_int32), ("node_id", C.c_uint32),
                ("pad", C.c_uint32)]


_vp, _u64, _i = C.c_void_p, C.c_uint64, C.c_int
SIGNATURES = {
    "srf_last_error_string": (C.c_char_p, []),
    "srf_abi_version": (C.c_int, []),
    "srf_ctx_create": (_i, [_i, C.POINTER(_vp)]),
    "srf_ctx_destroy": (_i, [_vp]),
    "srf_accel_build_triangles": (_i, [_vp, C.POINTER(TriDesc), _vp, C.POINTER(_vp)]),
    "srf_accel_build_tetras": (_i, [_vp, C.POINTER(TetDesc), _vp, C.POINTER(_vp)]),
    "srf_accel_build_spheres": (_i, [_vp, C.POINTER(SphereDesc), _vp, C.POINTER(_vp)]),
    "srf_accel_build_solids": (_i, [_vp, C.POINTER(SolidDesc), _vp, C.POINTER(_vp)]),
    "srf_accel_build_lines": (_i, [_vp, C.POINTER(LineDesc), _vp, C.POINTER(_vp)]),
    "srf_accel_build_triangles_batch": (_i, [_vp, C.c_uint32, _vp, _vp, _vp]),
    "srf_accel_rebuild_triangles_batch": (_i, [_vp, C.c_uint32, _vp, _vp, _vp]),
    "srf_accel_rebuild": (_i, [_vp, _vp, _vp, _vp]),
    "srf_accel_bounds": (_i, [_vp, _vp, _vp]),
    "srf_accel_get_info": (_i, [_vp, C.POINTER(AccelInfo)]),
    "srf_accel_export": (_i, [_vp, _vp, _vp]),
    "srf_accel_dump": (_i, [_vp, C.c_char_p]),
    "srf_accel_destroy": (_i, [_vp]),
    "srf_trace_closest": (_i, [_vp, _vp, _u64, _vp, _vp]),
    "srf_trace_tetra": (_i, [_vp, _vp, _u64, _vp, _vp, _vp, _vp]),
    "srf_trace_solids": (_i, [_vp, _vp, _u64, _vp, _vp, _vp, _vp]),
    "srf_trace_lines": (_i, [_vp, _vp, _u64, _vp, _vp, _vp]),
    "srf_trace_spheres": (_i, [_vp, _vp, _vp, _u64, _vp, _vp]),
    "srf_trace_any": (_i, [_vp, _vp, _vp, _u64, _vp, _vp]),
    "srf_debug_force_wide_stack": (_i, [_vp, _i]),
    "srf_build_intersection_triangles": (_i, [C.POINTER(TriDesc), _vp, _vp, _u64, _vp, _vp]),
    "srf_build_intersection_tetras": (_i, [C.POINTER(TetDesc), _vp, _vp, _vp, _u64, _vp, _vp]),
    "srf_build_intersection_spheres": (_i, [C.POINTER(SphereDesc), _vp, _vp, _u64, _vp, _vp]),
    "srf_accel_build_host": (_i, [_vp, _i, _vp, _vp, C.POINTER(_vp)]),
    "srf_accel_rebuild_host": (_i, [_vp, _vp, _vp, _vp]),
    "srf_trace_closest_host": (_i, [_vp, _vp, _u64, _vp, _vp, _vp]),
    "srf_lbvh_morton30": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _vp]),
    "srf_lbvh_sort30": (_i, [_vp, _vp, _u64, _vp, _vp, _vp]),
    "srf_lbvh_radix_tree30": (_i, [_vp, _vp, _u64, _vp, _vp]),
    "srf_scene_create": (_i, [_vp, C.c_uint32, _vp, C.POINTER(_vp)]),
    "srf_scene_destroy": (_i, [_vp]),
    "srf_scene_bounds": (_i, [_vp, _vp, _vp]),
    "srf_scene_export": (_i, [_vp, _vp, C.POINTER(C.c_uint32)]),
    "srf_scene_trace": (_i, [_vp, _vp, _u64, _vp, _vp, _i, _vp]),
    "srf_instance_update": (_i, [_vp]),
    "srf_scene_autotune": (_i, [_vp, _vp, _u64, _vp, _vp, _vp, _vp]),
    "srf_trace_any_sorted": (_i, [_vp, _vp, _vp, _vp, _u64, C.c_uint32, _i, _vp, _vp]),
    "srf_trace_any_sorted_noalloc": (_i, [_vp, _vp, _vp, _u64, C.c_uint32, _i, _vp, _vp]),
    "srf_autotune_any": (_i, [_vp, _vp, _vp, _vp, _u64, _vp, _vp, _vp]),
    "srf_ao_rays_spheres": (_i, [_vp, _vp, _u64, _vp, _vp, C.c_float, C.c_uint32, C.c_float,
                                 C.c_float, C.c_uint64, _vp, _vp, _vp, _vp, _vp]),
    "srf_volume_create": (_i, [_vp, _vp, _i, _i, _i, _i, _i, _i, C.POINTER(_vp)]),
    "srf_volume_destroy": (_i, [_vp]),
    "srf_texture3d_create": (_i, [_vp, _vp, _i, _i, _i, _i, _i, _i, _i, _i, C.POINTER(_vp)]),
    "srf_texture3d_remap": (_i, [_vp, _i, _vp, _i]),
    "srf_texture3d_fetch": (_i, [_vp, _vp, _u64, _vp, _vp]),
    "srf_sparse_volume_create": (_i, [_vp, _vp, C.c_uint32, _vp, _i, _i, C.POINTER(_vp)]),
    "srf_sparse_volume_destroy": (_i, [_vp]),
    "srf_sparse_volume_sample": (_i, [_vp, _vp, _u64, _vp, _vp]),
    "srf_sparse_volume_fetch": (_i, [_vp, _vp, _u64, _vp, _vp]),
    "srf_volume_sample": (_i, [_vp, _vp, _u64, _vp, _vp]),
    "srf_volume_march_fwd": (_i, [_vp, _vp, _i, C.c_float, _vp, _u64, _vp, _vp, _vp]),
    "srf_volume_checkpoint_slots": (_i, [_vp, C.c_float, C.POINTER(C.c_uint32)]),
    "srf_volume_march_bwd": (_i, [_vp, _vp, _i, C.c_float, _vp, _u64, _vp, _vp, _vp, _vp,
                                  _vp]),
}

_lib = None


def load(path=None):
    """Load libsrf.so (raises if it has not been built: there is no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise FileNotFoundError(
            f"{path} is missing; build it with `python -m surface_amd.build` "
            "(the HIP path has no CPU fallback)")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status):
    if status != SRF_OK:
        msg = load().srf_last_error_string()
        raise SrfError(status, msg.decode() if msg else "")
    return status
