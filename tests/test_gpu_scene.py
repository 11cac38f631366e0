"""Two-level scenes on the GPU (srf_scene_*): the top-level BVH and every
ray's nearest element hit bit-exact against the reference's RenderGraph
(fixture) and against the oracle on larger scenes, for the plain and the
persistent (work-counter) launch, and for scenes that mix triangle, tetra,
particle, solid and line elements."""
from __future__ import annotations

import hashlib
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle.oracle import Oracle, canonical_tlas  # noqa: E402
from surface_amd import workloads as W  # noqa: E402

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def S():
    import torch
    assert torch.cuda.is_available(), "no GPU visible"
    from surface_amd import accel
    return accel


@pytest.fixture(scope="module")
def orc():
    return Oracle()


def _hits(S, h, ids):
    hits = S.hits_to_numpy(h)
    return hits, ids.cpu().numpy().view(np.uint32)


def test_scene_matches_reference_fixture(S):
    g = np.load(os.path.join(GOLD, "scene_small.npz"), allow_pickle=False)
    meshes, T = W.batch_scene(count=12, lo=1100, hi=6000, seed=W.BASE_SEED + 21)
    h = hashlib.sha256()
    for m in meshes:
        for a in m:
            h.update(np.ascontiguousarray(a).tobytes())
    assert h.hexdigest() == str(g["meshes_sha"])
    say("scene fixture: 12 elements, 8192 rays")
    accs = S.build_triangles_batch([(S.device_tensor(V), S.device_tensor(F)) for V, F in meshes])
    sc = S.Scene(accs, g["T"], g["invT"], g["invT33"])
    tl = sc.export_tlas()
    assert canonical_tlas(tl).tobytes() == g["tlas"].tobytes()
    bmin, bmax = sc.BoundingBox()
    assert np.array_equal(np.concatenate([bmin, bmax]),
                          np.concatenate([g["tlas"]["bmin"][0], g["tlas"]["bmax"][0]]))
    rays = S.device_tensor(g["rays"])
    for persistent in (False, True):
        hits, inst = _hits(S, *sc.Trace(rays, mode=("persistent" if persistent else "plain")))
        assert hits.tobytes() == g["hits"].tobytes(), persistent
        assert np.array_equal(inst, g["inst"]), persistent


@pytest.mark.parametrize("count", [1, 2, 64])
def test_scene_matches_oracle(S, orc, count):
    meshes, T = W.batch_scene(count=count, lo=1100, hi=9000, seed=W.BASE_SEED + 30 + count)
    say(f"scene of {count} elements vs oracle")
    invT, invT33 = S.rigid_inverses(T)
    dm = [(S.device_tensor(V), S.device_tensor(F)) for V, F in meshes]
    accs = S.build_triangles_batch(dm) if count > 1 else [S.TriangleAccel()]
    if count == 1:
        accs[0].Build32(*dm[0])
    sc = S.Scene(accs, T, invT, invT33)
    trees = [orc.build_tri(V, F) for V, F in meshes]
    xb = np.array([orc.transform_box(t["bounds"][:3], t["bounds"][3:], T[i])
                   for i, t in enumerate(trees)])
    tl = orc.tlas_build(xb)
    assert canonical_tlas(sc.export_tlas()).tobytes() == canonical_tlas(tl).tobytes()
    els = [dict(tree=trees[i], V=meshes[i][0], F=meshes[i][1], T=T[i], invT=invT[i],
                invT33=invT33[i]) for i in range(count)]
    for k, cam in enumerate(W.scene_cameras(T, count=2, width=128, height=128)):
        want_h, want_i = orc.scene_trace(tl, els, cam, threads=8)
        rays = S.device_tensor(cam)
        if k == 1:
            # every launch shape the autotuner tries must give the same hits
            cfg = sc.autotune(rays)
            say(f"autotuned: {cfg}")
            got_h, got_i = _hits(S, *sc.Trace(rays, mode="tuned"))
        else:
            got_h, got_i = _hits(S, *sc.Trace(rays, mode="persistent"))
        assert got_h.tobytes() == want_h.tobytes()
        assert np.array_equal(got_i, want_i)
        assert (want_i != W.NO_PRIM).mean() > 0.05


def _gpu_element(S, e):
    k = e["kind"]
    if k == "tri":
        a = S.TriangleAccel()
        a.Build32(S.device_tensor(e["V"]), S.device_tensor(e["F"]))
    elif k == "tet":
        a = S.TetraAccel()
        a.Build32(S.device_tensor(e["V"]), S.device_tensor(e["F"]))
    elif k == "sph":
        a = S.ParticleAccel()
        a.Build32(S.device_tensor(e["V"]), S.device_tensor(e["R"]), e.get("rconst", 0.0))
    elif k == "sol":
        a = S.SolidAccel()
        a.Build32(S.device_tensor(e["V"]), S.device_tensor(e["F"]), e["nv"])
    else:
        a = S.LineAccel()
        a.Build(S.device_tensor(e["V"]), S.device_tensor(e["F"]), None, e["rconst"],
                cap=bool(e.get("cap", 1)))
    return a


@pytest.mark.parametrize("copies", [1, 2])
def test_mixed_scene_matches_oracle(S, orc, copies):
    """RenderGraph::Trace over triangle, tetra, particle, hexahedron and line
    elements (the element passes): every ray's world hit and element equal
    the oracle's mixed walk, whose walk equals the reference's RenderGraph
    (tests/test_oracle.py); the launch modes agree."""
    elems, T = W.mixed_scene(copies=copies)
    say(f"mixed scene of {len(elems)} elements vs oracle")
    invT, invT33 = S.rigid_inverses(T)
    accs = [_gpu_element(S, e) for e in elems]
    sc = S.Scene(accs, T, invT, invT33)
    trees = [orc.build_element(e) for e in elems]
    xb = np.array([orc.transform_box(t["bounds"][:3], t["bounds"][3:], T[i])
                   for i, t in enumerate(trees)])
    tl = orc.tlas_build(xb)
    assert canonical_tlas(sc.export_tlas()).tobytes() == canonical_tlas(tl).tobytes()
    els = [dict(e, tree=trees[i], T=T[i], invT=invT[i], invT33=invT33[i])
           for i, e in enumerate(elems)]
    kinds = set()
    for k, cam in enumerate(W.scene_cameras(T, count=2, width=128, height=128)):
        want_h, want_i = orc.scene_trace_mixed(tl, els, cam, threads=8)
        rays = S.device_tensor(cam)
        for mode in ("plain", "persistent", "tuned"):
            got_h, got_i = _hits(S, *sc.Trace(rays, mode=mode))
            assert got_h.tobytes() == want_h.tobytes(), mode
            assert np.array_equal(got_i, want_i), mode
        kinds |= {elems[j]["kind"] for j in np.unique(want_i[want_i != W.NO_PRIM])}
    assert kinds == {"tri", "tet", "sph", "sol", "lin"}


def test_scene_trace_in_hip_graph_replays_match_eager(S):
    """Persistent scene traces captured in a hipGraph (the work counter is
    cleared by a kernel, not a memset): every replay equals the eager run."""
    import torch
    meshes, T = W.batch_scene(count=8, lo=1100, hi=20000, seed=W.BASE_SEED + 40)
    accs = S.build_triangles_batch([(S.device_tensor(V), S.device_tensor(F)) for V, F in meshes])
    sc = S.Scene(accs, T)
    cams = [S.device_tensor(c) for c in W.scene_cameras(T, count=3, width=128, height=128)]
    eager = [sc.Trace(r, mode="persistent") for r in cams]
    want = [(h.clone(), i.clone()) for h, i in eager]
    outs = [(torch.full_like(h, -1.0), torch.full_like(i, -7)) for h, i in eager]
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        for r, (h, i) in zip(cams, outs):
            sc.Trace(r, h, i, mode="persistent")
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        for r, (h, i) in zip(cams, outs):
            sc.Trace(r, h, i, mode="persistent")
    for rep in range(4):
        for h, i in outs:
            h.fill_(-1.0)
            i.fill_(-7)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for (h, i), (wh, wi) in zip(outs, want):
            assert torch.equal(h.view(torch.int32), wh.view(torch.int32)), rep
            assert torch.equal(i, wi), rep
        say(f"graph replay {rep} equals the eager trace")
