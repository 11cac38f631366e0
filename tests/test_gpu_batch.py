"""Batched (segmented) LBVH build on the GPU: many meshes, one launch per
stage, each mesh's tree and hits bit-exact against the oracle's Build32 /
Traverse of that mesh alone, also after an in-place rebuild of the batch."""
from __future__ import annotations

import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle.oracle import Oracle, canonical_tri  # noqa: E402
from surface_amd import workloads as W  # noqa: E402


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


def _local_rays(V, n=48):
    """A pinhole camera looking at the mesh's box from outside it."""
    lo, hi = V.min(0).astype(np.float64), V.max(0).astype(np.float64)
    c = 0.5 * (lo + hi)
    ext = float(np.max(hi - lo))
    eye = c + ext * np.array([1.3, -0.9, 1.1])
    return W.pinhole_rays(tuple(eye), tuple(c), (0, 0, 1), 45.0, n, n)


def _check(S, orc, acc, V, F):
    want = orc.build_tri(V, F)
    nodes, idx = acc.export()
    assert canonical_tri(nodes).tobytes() == canonical_tri(want["nodes"]).tobytes()
    assert np.array_equal(idx, want["indices"])
    bmin, bmax = acc.BoundingBox()
    assert np.array_equal(np.concatenate([bmin, bmax]), want["bounds"])
    rays = _local_rays(V)
    got = S.hits_to_numpy(acc.Traverse(S.device_tensor(rays)))
    assert got.tobytes() == orc.trace_tri(want, V, F, rays).tobytes()
    return (got["prim_id"] != W.NO_PRIM).mean()


def test_batch_build_matches_single_builds(S, orc):
    meshes, _ = W.batch_scene(7, 1500, 60000, seed=W.BASE_SEED + 11)
    say("batch of 7 meshes: " + ", ".join(str(len(F)) for _, F in meshes))
    dm = [(S.device_tensor(V), S.device_tensor(F)) for V, F in meshes]
    accs = S.build_triangles_batch(dm)
    assert len(accs) == len(meshes)
    for k, (acc, (V, F)) in enumerate(zip(accs, meshes)):
        say(f"mesh {k}: {len(F)} triangles")
        frac = _check(S, orc, acc, V, F)
        assert frac > 0.05
    # in place, same sizes
    S.build_triangles_batch(dm, accels=accs)
    for acc, (V, F) in zip(accs, meshes):
        _check(S, orc, acc, V, F)


def test_batch_rebuild_with_other_meshes(S, orc):
    a, _ = W.batch_scene(4, 1200, 20000, seed=W.BASE_SEED + 12)
    b, _ = W.batch_scene(4, 1200, 90000, seed=W.BASE_SEED + 13)
    accs = S.build_triangles_batch([(S.device_tensor(V), S.device_tensor(F)) for V, F in a])
    say("rebuild batch: " + ", ".join(f"{len(x[1])}->{len(y[1])}" for x, y in zip(a, b)))
    S.build_triangles_batch([(S.device_tensor(V), S.device_tensor(F)) for V, F in b],
                            accels=accs)
    for acc, (V, F) in zip(accs, b):
        _check(S, orc, acc, V, F)


@pytest.mark.parametrize("mix", ["single_tile_only", "mixed"])
def test_batch_of_single_tile_meshes(S, orc, mix):
    """Meshes of at most one sort tile (3,840 primitives) are sorted by one
    workgroup each (all four digit passes in LDS) and own no onesweep tiles;
    alone or next to larger meshes their trees must stay Build32's."""
    hi = 3500 if mix == "single_tile_only" else 30000
    meshes, _ = W.batch_scene(9, 1100, hi, seed=W.BASE_SEED + 14)
    say(f"batch ({mix}): " + ", ".join(str(len(F)) for _, F in meshes))
    if mix == "single_tile_only":
        assert max(len(F) for _, F in meshes) <= 3840
    else:
        assert min(len(F) for _, F in meshes) <= 3840 < max(len(F) for _, F in meshes)
    dm = [(S.device_tensor(V), S.device_tensor(F)) for V, F in meshes]
    accs = S.build_triangles_batch(dm)
    for acc, (V, F) in zip(accs, meshes):
        _check(S, orc, acc, V, F)


def test_batch_rejects_small_mesh(S):
    from surface_amd._lib import SrfError
    good = W.heightfield(40)
    small = W.heightfield(20)  # 722 triangles: below Build32's 1024 floor
    dm = [(S.device_tensor(V), S.device_tensor(F)) for V, F in (good, small)]
    with pytest.raises(SrfError):
        S.build_triangles_batch(dm)
