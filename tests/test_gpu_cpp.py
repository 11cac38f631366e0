"""The C++ drop-in classes (include/srf_accel.hpp) on the GPU, through the
C ABI, checked bit for bit against the oracle by tests/cpp/test_accel_cpp."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(__file__), "cpp", "test_accel_cpp")


def test_cpp_host_classes_bit_exact():
    assert os.path.exists(EXE), "tests/cpp/test_accel_cpp not built (run __graft_entry__.build())"
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    assert "PASS" in r.stdout


FRAME = os.path.join(os.path.dirname(__file__), "cpp", "test_frame_cpp")


def test_cpp_frame_loop_and_ray_queues():
    assert os.path.exists(FRAME), "tests/cpp/test_frame_cpp not built (run __graft_entry__.build())"
    r = subprocess.run([FRAME], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    assert "PASS" in r.stdout


GL = os.path.join(os.path.dirname(__file__), "cpp", "test_gl_cpp")


def test_cpp_gl_draw_path_over_mixed_elements():
    assert os.path.exists(GL), "tests/cpp/test_gl_cpp not built (run __graft_entry__.build())"
    r = subprocess.run([GL], capture_output=True, text=True, timeout=180)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    assert "PASS" in r.stdout
