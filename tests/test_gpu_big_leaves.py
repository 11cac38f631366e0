"""Big sphere leaves (depth-capped leaves of clustered particles) on the GPU.

Build32 leaves clusters whose centres share one 30-bit Morton cell in single
leaves of thousands of spheres; the reference scans such a leaf in order. The
HIP path searches it through a secondary tree and must return the same
sphere, t, u, v as that scan (oracle: oracle/srf_oracle.c trace_sph, the
restatement of ParticleAccel::Traverse, render/render_accel_particle.cc:
3053-3218), including ties between duplicated spheres (the first in leaf
order wins), rays that start inside a cluster, axis-parallel rays, secondary
rays that skip their own sphere, and any-hit queries.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle.oracle import Oracle, canonical_sph  # noqa: E402
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
def scene(S, torch):
    say("clustered spheres: build")
    P, R, centres = W.clustered_spheres()
    orc = Oracle()
    tree = orc.build_sph(P, R)
    a = S.ParticleAccel()
    a.Build32(S.device_tensor(P), S.device_tensor(R), 0.0)
    torch.cuda.synchronize()
    return dict(P=P, R=R, centres=centres, orc=orc, tree=tree, a=a)


def _rays(org, dirs):
    rays = np.zeros(len(dirs), W.RAY_DTYPE)
    rays["org"] = org
    rays["dir"] = dirs
    rays["flags"] = 1
    rays["prev_prim_id"] = W.NO_PRIM
    return rays


def _cases(centres):
    rng = np.random.default_rng(W.BASE_SEED + 60)
    out = {}
    c = centres[0]
    # a camera close to a cluster, one far away (large padding relative to
    # the cluster), and rays from inside each cluster in random directions
    out["near"] = W.pinhole_rays(c + np.array([0.0, -0.02, 0.004]), c, (0, 0, 1), 12.0, 96, 96)
    out["far"] = W.pinhole_rays(c + np.array([0.4, -3.0, 0.3]), c, (0, 0, 1), 0.12, 64, 64)
    n = 4096
    org = centres[rng.integers(0, len(centres), n)] + rng.uniform(-1e-3, 1e-3, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d[::5, 0] = 0.0
    d[1::5, 1] = -0.0
    out["inside"] = _rays(org.astype(np.float32), d.astype(np.float32))
    return out


def test_workload_has_big_leaves(scene):
    n = scene["tree"]["nodes"]
    counts = n["data"][n["flag"] == 1][:, 0]
    assert (counts > 64).sum() >= 6
    assert counts.max() > 2000


def test_tree_unchanged(scene):
    nodes, idx = scene["a"].export()
    assert canonical_sph(nodes).tobytes() == canonical_sph(scene["tree"]["nodes"]).tobytes()
    assert np.array_equal(idx, scene["tree"]["indices"])


@pytest.mark.parametrize("case", ["near", "far", "inside"])
def test_closest_hits_match_in_order_scan(S, scene, case):
    say(f"big leaves: closest {case}")
    rays = _cases(scene["centres"])[case]
    want = scene["orc"].trace_sph(scene["tree"], scene["P"], scene["R"], 0.0, rays, threads=8)
    got = S.hits_to_numpy(scene["a"].Traverse(S.device_tensor(rays)))
    hit = want["prim_id"] != W.NO_PRIM
    assert hit.mean() > 0.2
    assert np.array_equal(got["prim_id"], want["prim_id"])
    assert np.array_equal(got["t"], want["t"])
    assert np.allclose(got["u"], want["u"], rtol=1e-5, atol=1e-6)
    assert np.allclose(got["v"], want["v"], rtol=1e-5, atol=1e-6)


def test_duplicates_tie_to_the_first_in_leaf_order(S, scene):
    # rays aimed straight at duplicated centres: both copies report the same
    # t, the scan keeps the one earlier in the leaf
    P, R = scene["P"], scene["R"]
    key = np.ascontiguousarray(np.concatenate([P, R[:, None]], 1)).view(np.dtype((np.void, 16)))
    _, first, cnt = np.unique(key.ravel(), return_index=True, return_counts=True)
    dup = first[cnt > 1][:512]
    assert len(dup) > 100
    org = P[dup] + np.array([0.0, -0.05, 0.0], np.float32)
    d = P[dup] - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = _rays(org.astype(np.float32), d.astype(np.float32))
    want = scene["orc"].trace_sph(scene["tree"], P, R, 0.0, rays, threads=8)
    got = S.hits_to_numpy(scene["a"].Traverse(S.device_tensor(rays)))
    assert np.array_equal(got["prim_id"], want["prim_id"])
    assert np.array_equal(got["t"], want["t"])


def test_secondary_rays_skip_their_own_sphere(S, scene, torch):
    say("big leaves: secondary rays")
    P, R = scene["P"], scene["R"]
    rays = _cases(scene["centres"])["near"]
    h0 = scene["orc"].trace_sph(scene["tree"], P, R, 0.0, rays, threads=8)
    hit = h0["prim_id"] != W.NO_PRIM
    sec = rays.copy()
    sec["org"] = rays["org"] + rays["dir"] * np.where(hit, h0["t"], 0)[:, None]
    sec["prev_prim_id"] = h0["prim_id"]
    prev_t = np.where(hit, h0["t"], 0).astype(np.float64)
    prev_n = (sec["org"] - np.where(hit[:, None], P[np.where(hit, h0["prim_id"], 0)], 0))
    prev_n = prev_n.astype(np.float32)
    want = scene["orc"].trace_sph(scene["tree"], P, R, 0.0, sec, prev_t, prev_n, threads=8)
    prev = np.zeros((len(sec), 6), np.float32)
    prev.view(np.float64)[:, 0] = prev_t
    prev[:, 2:5] = prev_n
    got = S.hits_to_numpy(scene["a"].TraverseSelf(S.device_tensor(sec),
                                                  torch.from_numpy(prev).cuda()))
    assert np.array_equal(got["prim_id"], want["prim_id"])
    assert np.array_equal(got["t"], want["t"])


@pytest.mark.parametrize("case", ["near", "inside"])
def test_any_hit_matches_closest_distance(S, scene, torch, case):
    say(f"big leaves: any-hit {case}")
    rays = _cases(scene["centres"])[case]
    want_h = scene["orc"].trace_sph(scene["tree"], scene["P"], scene["R"], 0.0, rays, threads=8)
    t = want_h["t"]
    # half the rays stop just short of the closest hit, half just past it
    tmax = np.where(np.arange(len(rays)) % 2 == 0, np.nextafter(t, np.float32(0)),
                    np.nextafter(t, np.float32(np.inf))).astype(np.float32)
    tmax[want_h["prim_id"] == W.NO_PRIM] = np.float32(1e30)
    want = ((want_h["prim_id"] != W.NO_PRIM) & (t < tmax)).astype(np.int32)
    rd, td = S.device_tensor(rays), torch.from_numpy(tmax).cuda()
    plain = scene["a"].TraverseAny(rd, td).cpu().numpy()
    assert np.array_equal(plain, want)
    for bits, mode in ((0, "persistent"), (30, "plain"), (15, "persistent")):
        got = scene["a"].TraverseAnySorted(rd, td, bits, mode).cpu().numpy()
        assert np.array_equal(got, want), (bits, mode)
