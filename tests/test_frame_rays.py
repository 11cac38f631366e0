"""Eye rays of the frame loop (include/srf_frame.hpp, host code) against the
reference's own camera: Camera::buildFrame (render/render_camera.cc:17-63)
with RaytraceEngine::OnData's per-pixel expression
(gles/gles_raytrace_engine.cc:730-812), run from the reference build
(oracle/_ref). Origins, directions and the skipped-pixel flags must match bit
for bit for square and non-square frames, odd sizes and several poses."""
from __future__ import annotations

import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle.oracle import HERE as ORACLE_DIR, ref_available

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CAMERAS = [
    ((0.0, -6.0, 1.0), (0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 30.0, 64, 64),
    ((1.8, -1.2, 1.5), (0.5, 0.5, 0.5), (0.0, 0.0, 1.0), 35.0, 97, 61),
    ((-3.0, 2.5, 0.25), (0.1, -0.2, 0.3), (0.0, 1.0, 0.0), 60.0, 33, 128),
    ((100.0, 40.0, -20.0), (0.0, 0.0, 0.0), (0.3, 0.9, 0.1), 5.0, 50, 50),
]


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    gxx = shutil.which("g++")
    if not gxx:
        pytest.skip("g++ not available")
    out = tmp_path_factory.mktemp("eye") / "eye_rays"
    subprocess.run([gxx, "-O2", "-ffp-contract=off", "-std=c++17", "-D__HIP_PLATFORM_AMD__",
                    "-I/opt/rocm/include", "-I" + os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "eye_rays.cpp"), "-o", str(out)],
                   check=True)
    return str(out)


@pytest.mark.parametrize("cam", range(len(CAMERAS)))
def test_eye_rays_match_reference_camera(exe, cam, tmp_path):
    if not ref_available("serial"):
        pytest.skip("oracle/_ref not built (the reference checkout was absent at build time)")
    eye, look, up, fov, w, h = CAMERAS[cam]
    path = str(tmp_path / "rays.bin")
    args = [str(v) for v in (*eye, *look, *up, fov, w, h)]
    subprocess.run([exe, *args, path], check=True)
    raw = np.fromfile(path, np.uint8)
    n = w * h
    got = raw[:24 * n].view(np.float32).reshape(n, 6)
    got_valid = raw[24 * n:]
    L = C.CDLL(os.path.join(ORACLE_DIR, "_ref", "libsrfref_serial.so"))
    f3 = lambda v: (C.c_float * 3)(*[float(np.float32(x)) for x in v])
    org = np.zeros((n, 3), np.float32)
    dirs = np.zeros((n, 3), np.float32)
    valid = np.zeros(n, np.uint8)
    L.ref_eye_rays(f3(eye), f3(look), f3(up), C.c_float(fov), C.c_int(w), C.c_int(h),
                   org.ctypes.data_as(C.c_void_p), dirs.ctypes.data_as(C.c_void_p),
                   valid.ctypes.data_as(C.c_void_p))
    assert np.array_equal(got_valid, valid)
    assert valid.mean() > 0.99
    assert got[:, :3].tobytes() == org.tobytes()
    assert got[:, 3:].tobytes() == dirs.tobytes()
