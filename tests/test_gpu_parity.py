"""GPU parity of the HIP path (libsrf.so) against the oracle and the
reference's own outputs. Every call goes through the C ABI (ctypes).

Bar: bit-exact Morton codes, sort permutation, Karras nodes, trees (node
bytes in the reference layout, depth-first preorder), permutations, bounds,
per-ray hits (prim id, t, u, v) and tetra weights; sphere u, v (libm
atan2/acos) to 1e-5 relative; volume RGBA to 1e-4 absolute; adjoint
gradients to finite differences at 1e-3 relative.
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle.oracle import Oracle, canonical_sph, canonical_tet, canonical_tri  # noqa: E402
from surface_amd import workloads as W  # noqa: E402

GOLD = os.path.join(os.path.dirname(__file__), "golden")


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


def gold(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def dev(S, a):
    return S.device_tensor(a)


def np_u32(t):
    return t.cpu().numpy().view(np.uint32)


# ------------------------------------------------------------ LBVH stages

def test_library_is_the_hip_path(S):
    from surface_amd import _lib
    lib = _lib.load()
    assert os.path.basename(lib._name) == "libsrf.so"


@pytest.mark.parametrize("kind", ["tri", "tet", "pts"])
def test_morton_codes_bit_exact(S, orc, kind):
    say(f"morton {kind}")
    if kind == "tri":
        V, F = W.heightfield(200)
        bmin, bmax = orc.scene_bounds("tri", V, F)
        want = orc.morton("tri", V, F, bmin, bmax)
        got = S.morton30(0x4, (dev(S, V), dev(S, F)), bmin, bmax)
    elif kind == "tet":
        V, F, _ = W.tet_grid(20)
        bmin, bmax = orc.scene_bounds("tet", V, F)
        want = orc.morton("tet", V, F, bmin, bmax)
        got = S.morton30(0x10, (dev(S, V), dev(S, F)), bmin, bmax)
    else:
        P, R = W.plummer_spheres(100_000)
        bmin, bmax = orc.scene_bounds("sph", P, R=R)
        want = orc.morton("pts", P, None, bmin, bmax)
        got = S.morton30(0x0, (dev(S, P), dev(S, R)), bmin, bmax)
    assert np.array_equal(np_u32(got), want)


@pytest.mark.parametrize("n", [1, 5, 3839, 3840, 3841, 100_000, 2_000_003])
def test_sort_permutation_bit_exact(S, orc, n):
    say(f"sort n={n}")
    rng = np.random.default_rng(n)
    codes = rng.integers(0, 1 << 30, n).astype(np.uint32)
    codes[rng.integers(0, n, n // 2)] = codes[0]  # heavy duplication
    sc, si = S.sort30(dev(S, codes))
    want_c, want_i = orc.sort_keys(codes)
    assert np.array_equal(np_u32(si), want_i)
    assert np.array_equal(np_u32(sc), want_c)


def test_radix_tree_bit_exact(S, orc):
    say("karras")
    g = gold("karras.npz")
    for key in ("dup", "same"):
        got = np_u32(S.radix_tree30(dev(S, g[f"codes_{key}"])))
        assert np.array_equal(got, g[f"info_{key}"])
    rng = np.random.default_rng(3)
    codes = np.sort(rng.integers(0, 1 << 30, 300_000).astype(np.uint32))
    assert np.array_equal(np_u32(S.radix_tree30(dev(S, codes))), orc.radix_tree(codes))


# ----------------------------------------------------- builds and traces

def _build(S, kind, arrays, **kw):
    if kind == "tri":
        a = S.TriangleAccel()
        a.Build32(dev(S, arrays[0]), dev(S, arrays[1]))
    elif kind == "tet":
        a = S.TetraAccel()
        a.Build32(dev(S, arrays[0]), dev(S, arrays[1]))
    else:
        a = S.ParticleAccel()
        R = arrays[1]
        a.Build32(dev(S, arrays[0]), dev(S, R) if R is not None else None,
                  kw.get("rconst", 0.0))
    return a


def test_triangle_fixture_tree_and_hits(S):
    say("tri fixture")
    g = gold("tri_small.npz")
    a = _build(S, "tri", (g["V"], g["F"]))
    nodes, idx = a.export()
    assert canonical_tri(nodes).tobytes() == g["nodes"].tobytes()
    assert np.array_equal(idx, g["indices"])
    bmin, bmax = a.BoundingBox()
    assert np.array_equal(np.concatenate([bmin, bmax]), g["bounds"])
    hits = S.hits_to_numpy(a.Traverse(dev(S, g["rays"])))
    assert hits.tobytes() == g["hits"].tobytes()


def test_tetra_fixture_tree_hits_and_weights(S):
    say("tet fixture")
    g = gold("tet_small.npz")
    a = _build(S, "tet", (g["V"], g["F"]))
    nodes, idx = a.export()
    assert canonical_tet(nodes).tobytes() == g["nodes"].tobytes()
    assert np.array_equal(idx, g["indices"])
    bmin, bmax = a.BoundingBox()
    assert np.array_equal(np.concatenate([bmin, bmax]), g["bounds"])
    h, e = a.TraverseTetra(dev(S, g["rays"]), dev(S, g["S"]))
    hits = S.hits_to_numpy(h)
    ext = S.hits_to_numpy(e, W.TET_EXT_DTYPE)
    assert hits.tobytes() == g["hits"].tobytes()
    assert np.array_equal(ext["d"], g["ext_d"])
    assert np.array_equal(ext["scalar"], g["ext_scalar"])
    assert np.array_equal(ext["subface"], g["ext_subface"])


def test_sphere_fixture_tree_and_hits(S):
    say("sphere fixture")
    g = gold("sph_small.npz")
    a = _build(S, "sph", (g["P"], g["R"]))
    nodes, idx = a.export()
    assert canonical_sph(nodes).tobytes() == g["nodes"].tobytes()
    assert np.array_equal(idx, g["indices"])
    hits = S.hits_to_numpy(a.Traverse(dev(S, g["rays"])))
    assert np.array_equal(hits["prim_id"], g["hits"]["prim_id"])
    assert np.array_equal(hits["t"], g["hits"]["t"])
    hit = hits["prim_id"] != W.NO_PRIM
    assert np.allclose(hits["u"][hit], g["hits"]["u"][hit], rtol=1e-5, atol=1e-6)
    assert np.allclose(hits["v"][hit], g["hits"]["v"][hit], rtol=1e-5, atol=1e-6)


def test_full_heightfield_matches_reference_digests(S):
    say("heightfield 512: 522,242 triangles, 1024^2 rays")
    d = json.load(open(os.path.join(GOLD, "full_digests.json")))["heightfield512"]
    V, F = W.heightfield(512)
    a = _build(S, "tri", (V, F))
    nodes, idx = a.export()
    assert len(nodes) == d["num_nodes"]
    assert sha(canonical_tri(nodes)) == d["nodes"]
    assert sha(idx) == d["indices"]
    bmin, bmax = a.BoundingBox()
    assert list(np.concatenate([bmin, bmax])) == d["bounds"]
    hits = S.hits_to_numpy(a.Traverse(dev(S, W.heightfield_rays(1024, 1024))))
    assert sha(hits) == d["hits"]


def test_full_tetgrid_matches_reference_digests(S):
    say("tet grid 64: 1,572,864 tets, 1024^2 rays")
    d = json.load(open(os.path.join(GOLD, "full_digests.json")))["tetgrid64"]
    V, F, Sc = W.tet_grid(64)
    a = _build(S, "tet", (V, F))
    nodes, idx = a.export()
    assert len(nodes) == d["num_nodes"]
    assert sha(canonical_tet(nodes)) == d["nodes"]
    assert sha(idx) == d["indices"]
    h, e = a.TraverseTetra(dev(S, W.tet_rays(64, 1024, 1024)), dev(S, Sc))
    hits = S.hits_to_numpy(h)
    ext = S.hits_to_numpy(e, W.TET_EXT_DTYPE)
    assert sha(hits) == d["hits"]
    assert sha(ext["d"], ext["scalar"], ext["subface"]) == d["ext"]


def _soup(rng, n):
    c = rng.uniform(-1, 1, (n, 3))
    V = (c[:, None, :] + rng.normal(0, 0.02, (n, 3, 3))).reshape(-1, 3).astype(np.float32)
    return V, np.arange(3 * n, dtype=np.uint32).reshape(n, 3)


@pytest.mark.parametrize("case", ["soup", "duplicates", "identical", "min1024", "flags"])
def test_triangle_edge_cases_match_oracle(S, orc, case):
    say(f"triangle edge case {case}")
    rng = np.random.default_rng(11)
    if case == "duplicates":
        V, F = _soup(rng, 300)
        F = np.tile(F, (30, 1))
    elif case == "identical":
        V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0.2]], np.float32)
        F = np.zeros((5000, 3), np.uint32)
    elif case == "min1024":
        V, F = _soup(rng, 1024)
    else:
        V, F = _soup(rng, 40_000)
    t = orc.build_tri(V, F)
    a = _build(S, "tri", (V, F))
    nodes, idx = a.export()
    assert canonical_tri(nodes).tobytes() == canonical_tri(t["nodes"]).tobytes()
    assert np.array_equal(idx, t["indices"])
    rays = W.pinhole_rays((0.2, -3.0, 0.5), (0, 0, 0), (0, 0, 1), 50.0, 64, 64)
    if case == "flags":
        rays["flags"][::3] = 0
        rays["prev_prim_id"][::5] = np.arange(0, len(rays), 5) % len(F)
        rays["dir"][::7, 0] = 0.0  # infinite inverse direction
        rays["dir"][1::7, 1] = -0.0
    want = orc.trace_tri(t, V, F, rays)
    got = S.hits_to_numpy(a.Traverse(dev(S, rays)))
    assert got.tobytes() == want.tobytes()


def _tet_edge_rays(case, V, F, cells, rng):
    """Rays for the tetra edge cases: exact ties on an unjittered lattice,
    origins inside the mesh, grazing directions and self-skipping rays."""
    rays = []

    def add(org, d):
        r = np.zeros(len(d), W.RAY_DTYPE)
        r["org"] = org
        r["dir"] = d
        r["flags"] = 1
        r["prev_prim_id"] = W.NO_PRIM
        rays.append(r)

    c = float(cells)
    if case == "lattice":
        # along lattice lines (every tetra around an edge is crossed at the
        # same distance) and aimed exactly at lattice vertices
        j, k = np.meshgrid(np.arange(cells + 1), np.arange(cells + 1), indexing="ij")
        org = np.stack([np.full(j.size, -3.0), j.ravel(), k.ravel()], 1)
        add(org, np.tile([1.0, 0.0, 0.0], (j.size, 1)))
        org = np.stack([j.ravel(), np.full(j.size, c + 2.5), k.ravel() + 0.5], 1)
        add(org, np.tile([0.0, -1.0, 0.0], (j.size, 1)))
        eye = np.array([1.7 * c, -0.55 * c, 1.25 * c])
        tgt = V[rng.choice(len(V), 1500, replace=False)].astype(np.float64)
        tgt = np.concatenate([tgt, 0.5 * (tgt[:750] + tgt[750:])])  # edge and face midpoints
        d = tgt - eye
        add(np.tile(eye, (len(d), 1)), d / np.linalg.norm(d, axis=1, keepdims=True))
    elif case == "inside":
        org = rng.uniform(1.0, c - 1.0, (3000, 3))
        d = rng.normal(size=(3000, 3))
        add(org, d / np.linalg.norm(d, axis=1, keepdims=True))
    elif case == "grazing":
        # nearly parallel to the lattice planes
        org = np.stack([np.full(2000, -2.0), rng.uniform(0, c, 2000), rng.uniform(0, c, 2000)], 1)
        d = np.stack([np.ones(2000), rng.normal(0, 1e-4, 2000), rng.normal(0, 1e-4, 2000)], 1)
        add(org, d / np.linalg.norm(d, axis=1, keepdims=True))
    out = np.concatenate(rays)
    out["dir"] = out["dir"].astype(np.float32)
    return out


@pytest.mark.parametrize("case", ["lattice", "inside", "grazing", "prev"])
def test_tetra_edge_cases_match_oracle(S, orc, case):
    """Closest hits, weights, entry faces and exit distances equal the
    oracle's (TetraAccel::Traverse, render/render_accel_tetra.cc:2203-2289)
    where the nearest-first pass cannot vouch for its hit: ties between
    tetras around a lattice edge or vertex, rays starting inside the mesh,
    grazing rays and rays that skip the tetra they start on."""
    say(f"tetra edge case {case}")
    cells = 12
    rng = np.random.default_rng(23)
    V, F, Sc = W.tet_grid(cells, jitter=0.0 if case == "lattice" else 0.2)
    t = orc.build_tet(V, F)
    a = _build(S, "tet", (V, F))
    if case == "prev":
        rays = W.tet_rays(cells, 48, 48)
        first, _ = orc.trace_tet(t, V, F, rays, Sc)
        rays["prev_prim_id"] = first["prim_id"]
    else:
        rays = _tet_edge_rays(case, V, F, cells, rng)
    want_h, want_e = orc.trace_tet(t, V, F, rays, Sc)
    h, e = a.TraverseTetra(dev(S, rays), dev(S, Sc))
    hits = S.hits_to_numpy(h)
    ext = S.hits_to_numpy(e, W.TET_EXT_DTYPE)
    hit = want_h["prim_id"] != W.NO_PRIM
    assert hit.mean() > 0.3
    assert hits.tobytes() == want_h.tobytes()
    # InterpolateTetra divides by zero where an unjittered lattice point
    # lies on a degenerate sub-volume: those weights are NaN on both sides
    for f in ("d", "scalar", "t_leave"):
        assert np.array_equal(ext[f][hit], want_e[f][hit], equal_nan=True), f
    assert np.array_equal(ext["subface"][hit], want_e["subface"][hit])


def _branch_levels(nodes):
    """Branch nodes on the longest root-to-leaf path of a DFS-ordered tree."""
    best, stack = 0, [(0, 1)]
    while stack:
        i, d = stack.pop()
        if nodes[i]["flag"] == 1:
            continue
        best = max(best, d)
        stack += [(int(nodes[i]["data"][0]), d + 1), (int(nodes[i]["data"][1]), d + 1)]
    return best


def test_depth_capped_leaves(S, orc):
    say("depth-capped caterpillar tree")
    P, R = W.deep_chain_points()
    t = orc.build_sph(P, R)
    assert t["nodes"]["data"][t["nodes"]["flag"] == 1][:, 0].max() > 17
    a = _build(S, "sph", (P, R))
    nodes, idx = a.export()
    assert canonical_sph(nodes).tobytes() == canonical_sph(t["nodes"]).tobytes()
    assert np.array_equal(idx, t["indices"])
    # the caterpillar reaches the cap: 24 branch levels (root, depths 0..22),
    # a traversal stack of kTraceStack = 24 entries filled to the last one
    assert _branch_levels(t["nodes"]) == 24
    assert a.get_info()["max_depth"] == 24
    rays = W.pinhole_rays((300.0, -900.0, 200.0), (8.0, 8.0, 8.0), (0, 0, 1), 3.0, 64, 64)
    want = orc.trace_sph(t, P, R, 0.0, rays)
    got = S.hits_to_numpy(a.Traverse(dev(S, rays)))
    assert np.array_equal(got["prim_id"], want["prim_id"])
    assert np.array_equal(got["t"], want["t"])


def test_fewer_than_1024_primitives_is_reported(S):
    V, F = _soup(np.random.default_rng(1), 1000)
    with pytest.raises(S.SrfError) as e:
        _build(S, "tri", (V, F))
    assert e.value.status == 4


def test_sphere_self_intersection_and_constant_radius(S, orc, torch):
    say("sphere self-intersection")
    P, R = W.plummer_spheres(50_000, seed=21)
    t = orc.build_sph(P, None, 0.01)
    a = _build(S, "sph", (P, None), rconst=0.01)
    nodes, idx = a.export()
    assert canonical_sph(nodes).tobytes() == canonical_sph(t["nodes"]).tobytes()
    rays = W.sphere_rays(96, 96)
    h0 = orc.trace_sph(t, P, None, 0.01, rays)
    hit = h0["prim_id"] != W.NO_PRIM
    # secondary rays leaving the first hit point, told to skip that sphere
    sec = rays.copy()
    sec["org"] = rays["org"] + rays["dir"] * np.where(hit, h0["t"], 0)[:, None]
    sec["prev_prim_id"] = h0["prim_id"]
    prev_t = np.where(hit, h0["t"], 0).astype(np.float64)
    prev_n = (sec["org"] - np.where(hit[:, None], P[np.where(hit, h0["prim_id"], 0)], 0)).astype(np.float32)
    want = orc.trace_sph(t, P, None, 0.01, sec, prev_t, prev_n)
    prev = np.zeros((len(sec), 6), np.float32)
    prev.view(np.float64)[:, 0] = prev_t
    prev[:, 2:5] = prev_n
    got = S.hits_to_numpy(a.TraverseSelf(dev(S, sec), torch.from_numpy(prev).cuda()))
    assert np.array_equal(got["prim_id"], want["prim_id"])
    assert np.array_equal(got["t"], want["t"])


def test_any_hit_agrees_with_closest_hit(S, torch):
    say("any hit")
    V, F = W.heightfield(128)
    a = _build(S, "tri", (V, F))
    rays = W.heightfield_rays(256, 256)
    hits = S.hits_to_numpy(a.Traverse(dev(S, rays)))
    tmax = np.where(np.arange(len(rays)) % 2 == 0, np.float32(3.0e38), hits["t"] * 0.999)
    tmax = tmax.astype(np.float32)
    occ = a.TraverseAny(dev(S, rays), torch.from_numpy(tmax).cuda()).cpu().numpy()
    want = (hits["prim_id"] != W.NO_PRIM) & (hits["t"] < tmax)
    assert np.array_equal(occ.astype(bool), want)


# ------------------------------------------------------------------ volume

def test_texture_sample_bit_exact(S, torch):
    say("tex3d")
    from surface_amd.volume import Volume
    g = gold("tex3d.npz")
    vox = torch.from_numpy(g["vox"]).cuda()
    uvw = torch.from_numpy(g["uvw"]).cuda()
    for name, clamp in (("repeat", (0, 0, 0)), ("clamp", (1, 1, 1)), ("mixed", (1, 0, 1))):
        v = Volume(vox, tuple(int(x) for x in g["dims"]), clamp)
        out = v.sample(uvw).cpu().numpy()
        assert out.tobytes() == g[f"out_{name}"].tobytes(), name


def _vol_case(dim=24):
    vox = W.gaussian_volume(dim, n_gauss=8)
    tf = W.transfer_function(32, max_alpha=0.2)
    # opaque bins (a sample there has 1 - a = 0, and T = 0 behind it) take the
    # adjoint's re-march fallback; near-opaque ones take the division
    tf[20:22, 3] = 1.0
    tf[22:24, 3] = 0.9
    rays = W.volume_rays(40, 40)
    return vox, tf, rays


def test_volume_forward_matches_oracle(S, orc, torch):
    say("volume forward")
    from surface_amd.volume import Volume
    vox, tf, rays = _vol_case()
    dims = (24, 24, 24)
    step = 0.5 / 23
    want, ns = orc.volume_fwd(vox, dims, (1, 1, 1), tf, step, rays)
    v = Volume(torch.from_numpy(vox).cuda(), dims, (1, 1, 1))
    got = v.march(torch.from_numpy(tf).cuda(), step, dev(S, rays)).cpu().numpy()
    assert ns.max() > 20
    assert np.abs(got - want).max() <= 1e-4


@pytest.mark.parametrize("keep", [True, False], ids=["checkpointed", "replayed"])
def test_volume_adjoint_matches_f64_adjoint_and_fd(S, orc, torch, keep):
    say(f"volume adjoint keep_checkpoints={keep}")
    from surface_amd.volume import render
    vox, tf, rays = _vol_case()
    dims = (24, 24, 24)
    step = 0.5 / 23
    rng = np.random.default_rng(5)
    g = rng.normal(size=(len(rays), 4)).astype(np.float32)
    gv64, gt64 = orc.volume_bwd64(vox.astype(np.float64), dims, (1, 1, 1), tf.astype(np.float64),
                                  step, rays, g.astype(np.float64))
    vt = torch.from_numpy(vox).cuda().requires_grad_(True)
    tt = torch.from_numpy(tf).cuda().requires_grad_(True)
    out = render(vt, tt, dims, step, dev(S, rays), keep_checkpoints=keep)
    (out * torch.from_numpy(g).cuda()).sum().backward()
    gv = vt.grad.cpu().numpy()
    gt = tt.grad.cpu().numpy()
    sv = np.abs(gv64).max()
    st = np.abs(gt64).max()
    assert np.abs(gv - gv64).max() <= 1e-3 * sv
    assert np.abs(gt - gt64).max() <= 1e-3 * st

    def loss(v64, t64):
        return float(np.sum(g.astype(np.float64) *
                            orc.volume_fwd64(v64, dims, (1, 1, 1), t64, step, rays)))

    h = 1e-6
    v64, t64 = vox.astype(np.float64), tf.astype(np.float64)
    for k in np.argsort(-np.abs(gv64))[:8]:
        vp, vm = v64.copy(), v64.copy()
        vp[k] += h
        vm[k] -= h
        fd = (loss(vp, t64) - loss(vm, t64)) / (2 * h)
        assert abs(gv[k] - fd) <= 1e-3 * abs(fd) + 1e-4 * sv
    for k in np.argsort(-np.abs(gt64.ravel()))[:8]:
        i, c = divmod(int(k), 4)
        tp, tm = t64.copy(), t64.copy()
        tp[i, c] += h
        tm[i, c] -= h
        fd = (loss(v64, tp) - loss(v64, tm)) / (2 * h)
        assert abs(gt[i, c] - fd) <= 1e-3 * abs(fd) + 1e-4 * st
