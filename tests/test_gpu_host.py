"""The host-memory side of the boundary: srf_accel_build_host /
srf_accel_rebuild_host (Build32 with host arrays, as AccelBuilder hands the
reference its client arrays) and srf_trace_closest_host (Traverse on host
rays, per ray or batched, from several host threads). Trees equal the
device builds' and the oracle's; hits, tetra extras and BuildIntersection
fields equal the oracle's, bit for bit."""
from __future__ import annotations

import sys
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle.oracle import Oracle, canonical_sph, canonical_tet, canonical_tri  # noqa: E402
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


def test_triangles_host_build_and_trace(S, orc):
    say("host triangles")
    V, F = W.heightfield(128)
    rays = W.heightfield_rays(96, 96)
    t = orc.build_tri(V, F)
    want = orc.trace_tri(t, V, F, rays)
    a = S.TriangleAccel()
    a.Build32Host(V, F)
    nodes, idx = a.export()
    assert canonical_tri(nodes).tobytes() == canonical_tri(t["nodes"]).tobytes()
    assert np.array_equal(idx, t["indices"])
    hits, _, isect = a.TraverseHost(rays)
    assert hits.tobytes() == want.tobytes()
    assert isect.tobytes() == orc.isect("tri", V, F, rays, want).tobytes()
    # per-ray calls from 4 host threads
    got = np.zeros(len(rays), W.HIT_DTYPE)

    def work(w):
        for i in range(w, len(rays), 4 * 5):
            h, _, _ = a.TraverseHost(rays[i:i + 1], isect=False)
            got[i] = h[0]

    th = [threading.Thread(target=work, args=(w,)) for w in range(4)]
    for x in th:
        x.start()
    for x in th:
        x.join()
    sel = np.zeros(len(rays), bool)
    for w in range(4):
        sel[w::20] = True
    assert got[sel].tobytes() == want[sel].tobytes()
    # rebuild in place from other host arrays (a larger mesh)
    V2, F2 = W.heightfield(160)
    a.Build32Host(V2, F2)
    t2 = orc.build_tri(V2, F2)
    nodes, idx = a.export()
    assert canonical_tri(nodes).tobytes() == canonical_tri(t2["nodes"]).tobytes()
    hits, _, isect = a.TraverseHost(rays)
    want2 = orc.trace_tri(t2, V2, F2, rays)
    assert hits.tobytes() == want2.tobytes()
    assert isect.tobytes() == orc.isect("tri", V2, F2, rays, want2).tobytes()


def test_tetras_host_build_and_trace(S, orc):
    say("host tetras")
    V, F, Sc = W.tet_grid(12)
    rays = W.tet_rays(12, 64, 64)
    t = orc.build_tet(V, F)
    want_h, want_e = orc.trace_tet(t, V, F, rays, None)
    a = S.TetraAccel()
    a.Build32Host(V, F)
    nodes, idx = a.export()
    assert canonical_tet(nodes).tobytes() == canonical_tet(t["nodes"]).tobytes()
    hits, ext, isect = a.TraverseHost(rays, ext_dtype=W.TET_EXT_DTYPE)
    assert hits.tobytes() == want_h.tobytes()
    hit = want_h["prim_id"] != W.NO_PRIM
    for f in ("d", "subface", "t_leave"):
        assert np.array_equal(ext[f][hit], want_e[f][hit]), f
    assert isect.tobytes() == orc.isect("tet", V, F, rays, want_h, want_e).tobytes()


def test_spheres_host_build_and_trace(S, orc):
    say("host spheres")
    P, R = W.plummer_spheres(30_000, mu=-4.0)
    rays = W.sphere_rays(64, 64)
    t = orc.build_sph(P, R)
    want = orc.trace_sph(t, P, R, 0.0, rays)
    a = S.ParticleAccel()
    a.Build32Host(P, R)
    nodes, idx = a.export()
    assert canonical_sph(nodes).tobytes() == canonical_sph(t["nodes"]).tobytes()
    hits, _, isect = a.TraverseHost(rays)
    assert np.array_equal(hits["prim_id"], want["prim_id"])
    assert np.array_equal(hits["t"], want["t"])
    assert isect.tobytes() == orc.isect("sph", P, None, rays, want).tobytes()
