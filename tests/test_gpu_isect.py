"""BuildIntersection on the GPU (srf_build_intersection_{triangles,tetras,
spheres}) against the oracle's restatement, bit for bit: position, geometric
and shading normal, tangent, binormal and f0..f3 of every hit, zeros for
misses. The oracle's restatement is pinned against the reference's own
Traverse + BuildIntersection in tests/test_oracle.py."""
from __future__ import annotations

import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle.oracle import Oracle  # noqa: E402
from surface_amd import workloads as W  # noqa: E402


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


@pytest.fixture(scope="module")
def orc():
    return Oracle()


def _isect(S, t):
    return S.hits_to_numpy(t, W.ISECT_DTYPE)


def test_triangle_intersection_fields(S, orc):
    say("isect triangles")
    V, F = W.heightfield(128)
    rays = W.heightfield_rays(128, 128)
    rays["flags"][::3] = 0
    Vd, Fd, rd = S.device_tensor(V), S.device_tensor(F), S.device_tensor(rays)
    a = S.TriangleAccel()
    a.Build32(Vd, Fd)
    hd = a.Traverse(rd)
    got = _isect(S, S.build_intersection_triangles(Vd, Fd, rd, hd))
    hits = S.hits_to_numpy(hd)
    want = orc.isect("tri", V, F, rays, hits)
    assert (hits["prim_id"] != W.NO_PRIM).mean() > 0.5
    assert got.tobytes() == want.tobytes()


def test_tetra_intersection_fields(S, orc):
    say("isect tetras")
    V, F, Sc = W.tet_grid(16)
    # camera rays and rays from inside the mesh in random directions, so that
    # every entry face (subface 0..3) occurs
    rng = np.random.default_rng(12)
    inside = np.zeros(4000, W.RAY_DTYPE)
    inside["org"] = rng.uniform(1.0, 15.0, (4000, 3)).astype(np.float32)
    d = rng.normal(size=(4000, 3))
    inside["dir"] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    inside["flags"] = 1
    inside["prev_prim_id"] = W.NO_PRIM
    rays = np.concatenate([W.tet_rays(16, 96, 96), inside])
    Vd, Fd, rd = S.device_tensor(V), S.device_tensor(F), S.device_tensor(rays)
    a = S.TetraAccel()
    a.Build32(Vd, Fd)
    hd, ed = a.TraverseTetra(rd, S.device_tensor(Sc))
    got = _isect(S, S.build_intersection_tetras(Vd, Fd, rd, hd, ed))
    hits = S.hits_to_numpy(hd)
    ext = S.hits_to_numpy(ed, W.TET_EXT_DTYPE)
    want = orc.isect("tet", V, F, rays, hits, ext)
    assert (hits["prim_id"] != W.NO_PRIM).mean() > 0.5
    assert set(np.unique(ext["subface"][hits["prim_id"] != W.NO_PRIM])) == {0, 1, 2, 3}
    assert got.tobytes() == want.tobytes()


@pytest.mark.parametrize("case", ["plummer", "poles"])
def test_sphere_intersection_fields(S, orc, case):
    say(f"isect spheres {case}")
    if case == "plummer":
        P, R = W.plummer_spheres(50_000, mu=-4.0)
        rays = W.sphere_rays(128, 128)
    else:
        g = np.arange(40, dtype=np.float32)
        x, z = np.meshgrid(g, g, indexing="ij")
        P = np.stack([x.ravel(), np.zeros(x.size, np.float32), z.ravel()], 1).astype(np.float32)
        R = np.full(len(P), 0.25, np.float32)
        rays = np.zeros(2 * len(P), W.RAY_DTYPE)
        rays["org"][:len(P)] = P + np.array([0, 5, 0], np.float32)
        rays["dir"][:len(P)] = (0, -1, 0)
        rays["org"][len(P):] = P + np.array([0, -5, 0], np.float32)
        rays["dir"][len(P):] = (0, 1, 0)
        rays["flags"] = 1
        rays["prev_prim_id"] = W.NO_PRIM
    Pd, Rd, rd = S.device_tensor(P), S.device_tensor(R), S.device_tensor(rays)
    a = S.ParticleAccel()
    a.Build32(Pd, Rd, 0.0)
    hd = a.Traverse(rd)
    got = _isect(S, S.build_intersection_spheres(Pd, rd, hd))
    hits = S.hits_to_numpy(hd)
    want = orc.isect("sph", P, None, rays, hits)
    assert (hits["prim_id"] != W.NO_PRIM).mean() > 0.3
    assert got.tobytes() == want.tobytes()
