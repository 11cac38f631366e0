"""srf_trace_any_sorted_noalloc (the sorted occlusion entry point that
allocates nothing; its scratch is a static per-GPU device array) against
srf_trace_any_sorted, flag for flag, on two calls in a row, the second with
fewer rays, for triangles and spheres, sorted and persistent launches; and
its size limit, which is rejected before any GPU work."""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np
import pytest

from surface_amd import workloads as W

MAX_RAYS = 1048576  # SRF_NOALLOC_MAX_RAYS


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch():
    import torch as T
    assert T.cuda.is_available(), "no GPU visible"
    return T


@pytest.fixture(scope="module")
def S(torch):
    from surface_amd import accel
    return accel


def _scene(S, kind):
    if kind == "tri":
        V, F = W.heightfield(256)
        a = S.TriangleAccel()
        a.Build32(S.device_tensor(V), S.device_tensor(F))
        big = W.heightfield_rays(1024, 1024)
        small = W.heightfield_rays(300, 200)
        tm = 3.0e38
    else:
        P, R = W.plummer_spheres(300_000, seed=W.BASE_SEED + 60, mu=-5.0)
        a = S.ParticleAccel()
        a.Build32(S.device_tensor(P), S.device_tensor(R), 0.0)
        big = W.sphere_rays(1024, 1024)
        small = W.sphere_rays(256, 130)
        tm = 6.0
    rng = np.random.default_rng(7)
    big = big[rng.permutation(len(big))]
    small = small[rng.permutation(len(small))]
    return a, big, small, tm


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["tri", "sph"])
def test_noalloc_equals_sorted_any_hit(S, torch, kind):
    a, big, small, tm = _scene(S, kind)
    assert len(big) == MAX_RAYS
    for bits, mode in ((30, "persistent"), (15, "plain"), (0, "persistent")):
        for rays in (big, small):
            say(f"noalloc {kind} n={len(rays)} bits={bits} mode={mode}")
            rd = S.device_tensor(rays)
            td = torch.full((len(rays),), tm, dtype=torch.float32, device="cuda")
            want = a.TraverseAnySorted(rd, td, key_bits=bits, mode=mode).cpu().numpy()
            out = torch.full((len(rays),), 7, dtype=torch.int32, device="cuda")
            got = a.TraverseAnySortedNoAlloc(rd, td, key_bits=bits, mode=mode, out=out)
            got = got.cpu().numpy()
            assert 0.02 < want.mean() < 0.98
            assert np.array_equal(got, want), int((got != want).sum())
            plain = a.TraverseAny(rd, td).cpu().numpy()
            assert np.array_equal(got, plain)


@pytest.mark.gpu
def test_noalloc_rejects_more_rays_than_its_scratch_holds(S, torch):
    from surface_amd import _lib
    a, _, small, _ = _scene(S, "tri")
    lib = _lib.load()
    st = lib.srf_trace_any_sorted_noalloc(a.handle, C.c_void_p(1), None, MAX_RAYS + 1, 30,
                                          _lib.LAUNCH_PERSISTENT, C.c_void_p(1), None)
    assert st == 1  # SRF_ERR_INVALID_ARG, before any device work
    assert b"at most 1048576" in lib.srf_last_error_string()
    # tetras are not sorted-occlusion accelerators
    V, F, _ = W.tet_grid(12)
    t = S.TetraAccel()
    t.Build32(S.device_tensor(V), S.device_tensor(F))
    rd = S.device_tensor(small[:1000])
    st = lib.srf_trace_any_sorted_noalloc(t.handle, C.c_void_p(rd.data_ptr()), None, 1000, 0,
                                          _lib.LAUNCH_PLAIN, C.c_void_p(1), None)
    assert st == 4  # SRF_ERR_UNSUPPORTED
