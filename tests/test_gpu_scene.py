"""Two-level scenes on the GPU (srf_scene_*): the top-level BVH and every
ray's nearest element hit bit-exact against the reference's RenderGraph
(fixture) and against the oracle on larger scenes, for the plain and the
persistent (work-counter) launch."""
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
        hits, inst = _hits(S, *sc.Trace(rays, persistent=persistent))
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
    for cam in W.scene_cameras(T, count=2, width=128, height=128):
        want_h, want_i = orc.scene_trace(tl, els, cam, threads=8)
        got_h, got_i = _hits(S, *sc.Trace(S.device_tensor(cam), persistent=True))
        assert got_h.tobytes() == want_h.tobytes()
        assert np.array_equal(got_i, want_i)
        assert (want_i != W.NO_PRIM).mean() > 0.05


def test_scene_rejects_non_triangle_elements(S):
    from surface_amd._lib import SrfError
    V, F, _ = W.tet_grid(6)  # 1,296 tets
    tet = S.TetraAccel()
    tet.Build32(S.device_tensor(V), S.device_tensor(F))
    with pytest.raises(SrfError):
        S.Scene([tet], np.eye(4)[None])
