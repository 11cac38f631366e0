"""The C-ABI library loads and exports every entry point include/srf.h
declares (no device calls: this runs without a GPU)."""
from __future__ import annotations

import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "srf.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(srf_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("srf_accel_build_triangles", "srf_accel_build_tetras", "srf_accel_build_spheres",
                 "srf_trace_closest", "srf_trace_any", "srf_accel_bounds",
                 "srf_volume_march_fwd", "srf_volume_march_bwd", "srf_lbvh_sort30"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from surface_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libsrf.so is not built: run `python -m surface_amd.build`")
    lib = _lib.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(_lib.SIGNATURES) == set(declared_functions())
    assert lib.srf_abi_version() == 1
    assert lib.srf_last_error_string() is not None


def test_wrapper_fails_loudly_without_library(tmp_path):
    from surface_amd import _lib
    saved = _lib._lib
    try:
        _lib._lib = None
        with pytest.raises(FileNotFoundError):
            _lib.load(str(tmp_path / "missing.so"))
    finally:
        _lib._lib = saved


def test_cpp_host_classes_compile_against_the_header(tmp_path):
    """include/srf_accel.hpp (the C++ mirror of TriangleAccel/TetraAccel/
    ParticleAccel) is plain C++17 over srf.h: it compiles with the host
    compiler alone, with no HIP headers."""
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    src = tmp_path / "use.cpp"
    src.write_text('#include "srf_accel.hpp"\n'
                   "int main() { srf::Context c(0); srf::TriangleAccel t(c);"
                   " srf::TetraAccel k(c); srf::ParticleAccel p(c); return c.ok() ? 0 : 1; }\n")
    inc = os.path.join(ROOT, "include")
    r = subprocess.run([gxx, "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", "-I", inc,
                        str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_instance_update_equals_reference_element_matrices():
    """srf_instance_update (host code of the library, no device call) gives
    RenderElement::Update's invT and invT33 bit for bit: against the matrices
    the reference's own RenderGraph computed for rigid, scaled and sheared
    transforms, and against the closed-form inverse of a translation."""
    import numpy as np

    from oracle.oracle import Reference, ref_available
    from surface_amd import _lib
    from surface_amd import workloads as W
    lib = _lib.load()

    def update(T):
        inst = _lib.Instance()
        inst.T[:] = list(np.asarray(T, np.float64).reshape(16))
        assert lib.srf_instance_update(C.byref(inst)) == 0
        return np.array(inst.invT[:]).reshape(4, 4), np.array(inst.invT33[:]).reshape(4, 4)

    M = np.eye(4)
    M[3, :3] = [2.0, -3.0, 0.5]
    invT, invT33 = update(M)
    want = np.eye(4)
    want[3, :3] = [-2.0, 3.0, -0.5]
    assert np.array_equal(invT[:, :3], want[:, :3]) and np.array_equal(invT33, np.eye(4))
    if not ref_available():
        pytest.skip("reference build not present")
    ref = Reference("serial")
    meshes, T = W.batch_scene(count=8, lo=1100, hi=1500, seed=W.BASE_SEED + 40)
    rng = np.random.default_rng(3)
    for i in range(0, 8, 2):  # scale and shear every other element
        S = np.eye(4)
        S[:3, :3] = np.diag(rng.uniform(0.3, 3.0, 3)) + np.triu(rng.uniform(-0.4, 0.4, (3, 3)), 1)
        T[i] = S @ T[i]
    trees = [ref.build("tri", V, F, keep=True) for V, F in meshes]
    sc = ref.scene_build(trees, T)
    for i in range(len(T)):
        invT, invT33 = update(T[i])
        assert invT.tobytes() == sc["invT"][i].tobytes(), i
        assert invT33.tobytes() == sc["invT33"][i].tobytes(), i
    ref.scene_free(sc)
    for t in trees:
        ref.free(t)
