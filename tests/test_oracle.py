"""Pins of the CPU oracle (oracle/srf_oracle.c) -- runs without a GPU.

Each check is chosen so a plausible slip in the restatement (a dropped term,
a swapped operand, a wrong tie rule or index) fails it:
  * the reference's only known-answer vector (test/cctest/test-prefix-tree.cc:27-41)
    and an independent bit-by-bit interleave;
  * the reference's Karras invariants (test-prefix-tree.cc:418-510,622-625)
    plus a brute-force top-down radix tree, incl. all-equal keys (:732-826);
  * stable-sort equality with numpy's stable argsort (stability :127-129);
  * byte equality with fixtures the reference itself produced
    (tests/golden/gen_golden.py) and, where oracle/_ref is built, with the
    reference run live on fresh inputs;
  * brute-force closest hit over every primitive, analytic sphere/tetra
    cases, exact trilinear cases, and central finite differences of the
    volume adjoint.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

from oracle.oracle import (Oracle, Reference, canonical_sph, canonical_tet, canonical_tri,
                           ref_available)
from surface_amd import workloads as W

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def orc():
    return Oracle()


def gold(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


# ------------------------------------------------------------------ Morton

def test_morton_known_answer(orc):
    # test/cctest/test-prefix-tree.cc:27-41
    code = orc.morton_point(np.array([0.5, 0.5, 0.5], np.float32), np.zeros(3, np.float32),
                            (1024.0, 1024.0, 1024.0))
    assert code == (1 << 29) | (1 << 28) | (1 << 27)


def _interleave_bits(ix, iy, iz):
    code = np.zeros_like(ix, dtype=np.uint64)
    for b in range(10):
        code |= ((ix >> b) & 1).astype(np.uint64) << np.uint64(3 * b + 2)
        code |= ((iy >> b) & 1).astype(np.uint64) << np.uint64(3 * b + 1)
        code |= ((iz >> b) & 1).astype(np.uint64) << np.uint64(3 * b + 0)
    return code.astype(np.uint32)


def test_morton_matches_bit_interleave(orc):
    rng = np.random.default_rng(1)
    V = rng.uniform(-3, 5, (3000, 3)).astype(np.float32)
    F = rng.integers(0, 3000, (5000, 3)).astype(np.uint32)
    bmin, bmax = orc.scene_bounds("tri", V, F)
    codes = orc.morton("tri", V, F, bmin, bmax)
    inv = np.float32(1024) / (bmax - bmin)
    c = np.float32(1.0 / 3.0) * ((V[F[:, 0]] + V[F[:, 1]]) + V[F[:, 2]])
    q = ((c - bmin) * inv).astype(np.int64).astype(np.uint32) & 0x3FF
    assert np.array_equal(codes, _interleave_bits(q[:, 0], q[:, 1], q[:, 2]))


def test_morton_special_cases(orc):
    bmin = np.zeros(3, np.float32)
    inv = (1024.0, 1024.0, 1024.0)
    assert orc.morton_point(np.zeros(3, np.float32), bmin, inv) == 0
    # quantised 1024 on an axis wraps to 0 (PartBy2_30 masks 10 bits)
    assert orc.morton_point(np.array([1.0, 0, 0], np.float32), bmin, inv) == 0
    top = orc.morton_point(np.array([1023.5 / 1024, 1023.5 / 1024, 1023.5 / 1024], np.float32),
                           bmin, inv)
    assert top == (1 << 30) - 1


# -------------------------------------------------------------------- sort

def test_sort_is_stable_sort(orc):
    rng = np.random.default_rng(2)
    for n in (1, 7, 1024, 50_000):
        codes = rng.integers(0, 1 << 30, n).astype(np.uint32)
        codes[: n // 3] = codes[0]  # long runs of equal keys
        rng.shuffle(codes)
        sc, si = orc.sort_keys(codes)
        order = np.argsort(codes, kind="stable")
        assert np.array_equal(si, order) and np.array_equal(sc, codes[order])


# ------------------------------------------------------------------ Karras

def _delta(c, i, j):
    n = len(c)
    if j < 0 or j >= n:
        return -1
    a, b = int(c[i]), int(c[j])
    off = 0
    if a == b:
        a, b, off = i, j, 32
    x = a ^ b
    return (32 - x.bit_length()) + off


def _topdown(codes):
    """Binary radix tree top down: the split of [l, r] is where the adjacent
    common prefix is shortest (unique once equal codes are told apart by
    position). Returns {node index: (child, left_leaf, right_leaf)}."""
    n = len(codes)
    adj = np.array([_delta(codes, k, k + 1) for k in range(n - 1)])
    out = {}
    todo = [(0, n - 1, 0)]
    while todo:
        l, r, idx = todo.pop()
        g = l + int(np.argmin(adj[l:r]))
        assert np.sum(adj[l:r] == adj[g]) == 1
        out[idx] = (g, int(l == g), int(g + 1 == r))
        if g > l:
            todo.append((l, g, g))
        if g + 1 < r:
            todo.append((g + 1, r, g + 1))
    return out


@pytest.mark.parametrize("kind", ["random", "dups", "same"])
def test_karras_matches_topdown_bruteforce(orc, kind):
    rng = np.random.default_rng(3)
    n = 3000
    if kind == "random":
        codes = np.sort(rng.integers(0, 1 << 30, n).astype(np.uint32))
    elif kind == "dups":
        codes = np.sort(rng.integers(0, 40, n).astype(np.uint32))
    else:
        codes = np.full(n, 777, np.uint32)
    info = orc.radix_tree(codes)
    ref = _topdown(codes)
    assert len(ref) == n - 1
    for i in range(n - 1):
        assert tuple(int(x) for x in info[i]) == ref[i], i


def _check_karras_invariants(info, n):
    # CountLeafs32 / VisitAndCountLeafs32: test/cctest/test-prefix-tree.cc:418-510
    seen_leaf = np.zeros(n, np.int64)
    seen_node = np.zeros(n - 1, np.int64)
    for i in range(n - 1):
        c, ll, rl = (int(x) for x in info[i])
        (seen_leaf if ll else seen_node)[c] += 1
        (seen_leaf if rl else seen_node)[c + 1] += 1
    assert np.all(seen_leaf == 1)
    assert seen_node[0] == 0 and np.all(seen_node[1:] == 1)
    depth, stack, reached = 0, [(0, 0)], np.zeros(n, np.int64)
    while stack:
        i, d = stack.pop()
        depth = max(depth, d)
        c, ll, rl = (int(x) for x in info[i])
        for child, leaf in ((c, ll), (c + 1, rl)):
            if leaf:
                reached[child] += 1
            else:
                stack.append((child, d + 1))
    assert np.all(reached == 1)
    return depth


def test_karras_invariants_and_fixtures(orc):
    g = gold("karras.npz")
    for key in ("dup", "same"):
        codes = g[f"codes_{key}"]
        info = orc.radix_tree(codes)
        assert np.array_equal(info, g[f"info_{key}"])
        _check_karras_invariants(info, len(codes))
    rng = np.random.default_rng(4)
    codes = np.sort(rng.integers(0, 1 << 30, 20000).astype(np.uint32))
    assert _check_karras_invariants(orc.radix_tree(codes), 20000) < 60


# -------------------------------------------------------- builds vs fixtures

def _tree_invariants(nodes, indices, n, max_leaf, two_box):
    assert sorted(indices.tolist()) == list(range(n))
    covered = []
    stack = [(0, 0)]
    while stack:
        i, depth = stack.pop()
        nd = nodes[i]
        if nd["flag"] == 1:
            cnt, first = (int(x) for x in nd["data"])
            covered.append((first, cnt))
            assert cnt <= max_leaf + 1 or depth > 23
        else:
            assert depth <= 23
            stack.append((int(nd["data"][1]), depth + 1))
            stack.append((int(nd["data"][0]), depth + 1))
    # DFS visits leaves left to right and they tile [0, n)
    pos = 0
    for first, cnt in covered:
        assert first == pos
        pos += cnt
    assert pos == n
    assert len(nodes) == 2 * len(covered) - 1


def test_triangle_build_matches_reference_fixture(orc):
    g = gold("tri_small.npz")
    bmin, bmax = orc.scene_bounds("tri", g["V"], g["F"])
    assert np.array_equal(orc.morton("tri", g["V"], g["F"], bmin, bmax), g["codes"])
    sc, si = orc.sort_keys(g["codes"])
    assert np.array_equal(sc, g["sorted_codes"]) and np.array_equal(si, g["sorted_index"])
    assert np.array_equal(orc.radix_tree(sc), g["nodeinfo"])
    t = orc.build_tri(g["V"], g["F"])
    assert canonical_tri(t["nodes"]).tobytes() == g["nodes"].tobytes()
    assert np.array_equal(t["indices"], g["indices"])
    assert np.array_equal(t["bounds"], g["bounds"])
    _tree_invariants(t["nodes"], t["indices"], len(g["F"]), 16, True)
    hits = orc.trace_tri(t, g["V"], g["F"], g["rays"])
    assert hits.tobytes() == g["hits"].tobytes()


def test_tetra_build_and_trace_match_reference_fixture(orc):
    g = gold("tet_small.npz")
    t = orc.build_tet(g["V"], g["F"])
    assert canonical_tet(t["nodes"]).tobytes() == g["nodes"].tobytes()
    assert np.array_equal(t["indices"], g["indices"])
    assert np.array_equal(t["bounds"], g["bounds"])
    _tree_invariants(t["nodes"], t["indices"], len(g["F"]), 4, False)
    hits, ext = orc.trace_tet(t, g["V"], g["F"], g["rays"], g["S"])
    assert hits.tobytes() == g["hits"].tobytes()
    assert np.array_equal(ext["d"], g["ext_d"])
    assert np.array_equal(ext["scalar"], g["ext_scalar"])
    assert np.array_equal(ext["subface"], g["ext_subface"])


def test_sphere_build_and_trace_match_reference_fixture(orc):
    g = gold("sph_small.npz")
    t = orc.build_sph(g["P"], g["R"])
    assert canonical_sph(t["nodes"]).tobytes() == g["nodes"].tobytes()
    assert np.array_equal(t["indices"], g["indices"])
    _tree_invariants(t["nodes"], t["indices"], len(g["P"]), 16, True)
    hits = orc.trace_sph(t, g["P"], g["R"], 0.0, g["rays"])
    assert hits.tobytes() == g["hits"].tobytes()


def test_texture_matches_reference_fixture(orc):
    g = gold("tex3d.npz")
    for name, clamp in (("repeat", (0, 0, 0)), ("clamp", (1, 1, 1)), ("mixed", (1, 0, 1))):
        out = orc.tex3d(g["vox"], g["dims"], clamp, g["uvw"])
        assert out.tobytes() == g[f"out_{name}"].tobytes()


# ------------------------------------------------------ live reference runs

needs_ref = pytest.mark.skipif(not ref_available("serial"),
                               reason="oracle/_ref not built (reference checkout absent)")


def _soup(rng, n, spread=1.0):
    c = rng.uniform(-spread, spread, (n, 3))
    V = (c[:, None, :] + rng.normal(0, 0.02, (n, 3, 3))).reshape(-1, 3).astype(np.float32)
    return V, np.arange(3 * n, dtype=np.uint32).reshape(n, 3)


@needs_ref
@pytest.mark.parametrize("case", ["soup", "duplicates", "identical", "min1024"])
def test_triangle_oracle_equals_live_reference(orc, case):
    ref = Reference("serial")
    rng = np.random.default_rng(5)
    if case == "soup":
        V, F = _soup(rng, 20000)
    elif case == "duplicates":
        V, F = _soup(rng, 300)
        F = np.tile(F, (30, 1))  # each triangle 30 times -> equal Morton codes
    elif case == "identical":
        V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0.2]], np.float32)
        F = np.zeros((5000, 3), np.uint32)
        F[:] = [0, 1, 2]
    else:
        V, F = _soup(rng, 1024)
    t = orc.build_tri(V, F)
    r = ref.build("tri", V, F, keep=True)
    assert canonical_tri(t["nodes"]).tobytes() == canonical_tri(r["nodes"]).tobytes()
    assert np.array_equal(t["indices"], r["indices"])
    rays = W.pinhole_rays((0.2, -3.0, 0.5), (0, 0, 0), (0, 0, 1), 50.0, 48, 48)
    rays["flags"][::3] = 0  # some single-sided rays
    rays["prev_prim_id"][::5] = np.arange(0, len(rays), 5) % len(F)
    h1 = orc.trace_tri(t, V, F, rays)
    h2, _ = ref.trace(r, rays)
    ref.free(r)
    assert h1.tobytes() == h2.tobytes()


@needs_ref
def test_tetra_and_sphere_oracle_equal_live_reference(orc):
    ref = Reference("serial")
    V, F, S = W.tet_grid(7, seed=11)
    t = orc.build_tet(V, F)
    r = ref.build("tet", V, F, keep=True)
    assert canonical_tet(t["nodes"]).tobytes() == canonical_tet(r["nodes"]).tobytes()
    rays = W.tet_rays(7, 40, 40)
    h1, e1 = orc.trace_tet(t, V, F, rays, S)
    h2, e2 = ref.trace(r, rays, scalars=S)
    ref.free(r)
    assert h1.tobytes() == h2.tobytes()
    assert np.array_equal(e1["d"], e2["d"]) and np.array_equal(e1["scalar"], e2["scalar"])
    P, R = W.plummer_spheres(30000, seed=12)
    t = orc.build_sph(P, None, 0.02)
    r = ref.build("sph", P, R=None, rconst=0.02, keep=True)
    assert canonical_sph(t["nodes"]).tobytes() == canonical_sph(r["nodes"]).tobytes()
    rays = W.sphere_rays(40, 40)
    h1 = orc.trace_sph(t, P, None, 0.02, rays)
    h2, _ = ref.trace(r, rays)
    ref.free(r)
    assert h1.tobytes() == h2.tobytes()


@needs_ref
def test_depth_capped_leaves_equal_live_reference(orc):
    # a 30-deep caterpillar: leaves below depth 22 are cut (render/render_accel_particle.cc:578)
    ref = Reference("serial")
    P, R = W.deep_chain_points()
    t = orc.build_sph(P, R)
    r = ref.build("sph", P, R=R)
    assert canonical_sph(t["nodes"]).tobytes() == canonical_sph(r["nodes"]).tobytes()
    assert t["nodes"]["data"][t["nodes"]["flag"] == 1][:, 0].max() > 17


@needs_ref
def test_texture_oracle_equals_live_reference(orc):
    ref = Reference("serial")
    rng = np.random.default_rng(6)
    dims = np.array([17, 5, 33], np.int32)
    vox = rng.normal(size=int(np.prod(dims))).astype(np.float32)
    uvw = rng.uniform(-3, 3, (20000, 3)).astype(np.float32)
    for clamp in ((0, 0, 0), (1, 1, 1), (0, 1, 0)):
        a = orc.tex3d(vox, dims, clamp, uvw)
        b = ref.tex3d(vox, dims, np.array(clamp, np.int32), uvw)
        assert a.tobytes() == b.tobytes()


# -------------------------------------------------- brute force / analytic

def _mt_all(V, F, ray, eps_scale):
    """Moller-Trumbore of one ray against every triangle, float32 numpy."""
    f = np.float32
    o, d = ray["org"].astype(f), ray["dir"].astype(f)
    p0, p1, p2 = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    e1, e2 = p1 - p0, p2 - p0
    p = np.stack([d[1] * e2[:, 2] - d[2] * e2[:, 1], d[2] * e2[:, 0] - d[0] * e2[:, 2],
                  d[0] * e2[:, 1] - d[1] * e2[:, 0]], -1)
    det = (e1[:, 0] * p[:, 0] + e1[:, 1] * p[:, 1]) + e1[:, 2] * p[:, 2]
    keps = f(np.finfo(f).eps) * f(eps_scale)
    ok = np.abs(det) >= keps
    with np.errstate(all="ignore"):
        inv = f(1.0) / det
        s = o - p0
        q = np.stack([s[:, 1] * e1[:, 2] - s[:, 2] * e1[:, 1], s[:, 2] * e1[:, 0] - s[:, 0] * e1[:, 2],
                      s[:, 0] * e1[:, 1] - s[:, 1] * e1[:, 0]], -1)
        u = ((s[:, 0] * p[:, 0] + s[:, 1] * p[:, 1]) + s[:, 2] * p[:, 2]) * inv
        v = ((q[:, 0] * d[0] + q[:, 1] * d[1]) + q[:, 2] * d[2]) * inv
        t = ((e2[:, 0] * q[:, 0] + e2[:, 1] * q[:, 1]) + e2[:, 2] * q[:, 2]) * inv
    ok &= (u >= 0) & (u <= 1) & (v >= 0) & (u + v <= 1) & (t >= 0)
    return np.where(ok, t, np.inf)


def test_triangle_trace_equals_bruteforce_closest_hit(orc):
    V, F = W.heightfield(24, seed=99)
    t = orc.build_tri(V, F)
    rays = W.heightfield_rays(40, 40)
    hits = orc.trace_tri(t, V, F, rays)
    nhit = 0
    for k, ray in enumerate(rays):
        tt = _mt_all(V, F, ray, t["eps_scale"])
        best = tt.min()
        if np.isinf(best):
            assert hits["prim_id"][k] == W.NO_PRIM
            continue
        nhit += 1
        assert hits["t"][k] == np.float32(best)
        if np.sum(tt == best) == 1:
            assert hits["prim_id"][k] == np.argmin(tt)
    assert nhit > 1000


def test_sphere_trace_analytic(orc):
    P = np.zeros((2000, 3), np.float32)
    P[:, 0] = np.arange(2000, dtype=np.float32) * 3.0
    t = orc.build_sph(P, None, 1.0)
    rays = np.zeros(50, W.RAY_DTYPE)
    rays["org"][:, 0] = np.arange(50) * 3.0
    rays["org"][:, 1] = -10.0
    rays["dir"][:, 1] = 1.0
    rays["prev_prim_id"] = W.NO_PRIM
    hits = orc.trace_sph(t, P, None, 1.0, rays)
    assert np.array_equal(hits["prim_id"], np.arange(50))
    assert np.allclose(hits["t"], 9.0, rtol=0, atol=1e-6)
    assert np.allclose(hits["u"], 0.5, atol=1e-6)  # normal (0,-1,0): atan2(0, 0)+pi / 2pi
    assert np.allclose(hits["v"], 1.0, atol=1e-6)  # acos(-1) / pi


def test_tetra_trace_analytic(orc):
    # 2000 disjoint unit right tetras along x; a ray along +z through (x+0.2, 0.2).
    # RayTetraPluecker's enter/leave labels assume det(v1-v0, v2-v0, v3-v0) < 0.
    n = 2000
    base = np.array([[0, 0, 0], [0, 1, 0], [1, 0, 0], [0, 0, 1]], np.float32)
    V = (base[None] + np.stack([np.arange(n) * 3.0, np.zeros(n), np.zeros(n)], -1)[:, None]
         ).reshape(-1, 3).astype(np.float32)
    F = np.arange(4 * n, dtype=np.uint32).reshape(n, 4)
    t = orc.build_tet(V, F)
    rays = np.zeros(20, W.RAY_DTYPE)
    rays["org"] = np.stack([np.arange(20) * 3.0 + 0.2, np.full(20, 0.2), np.full(20, -5.0)], -1)
    rays["dir"][:, 2] = 1.0
    rays["prev_prim_id"] = W.NO_PRIM
    hits, ext = orc.trace_tet(t, V, F, rays, np.ones(len(V), np.float32))
    assert np.array_equal(hits["prim_id"], np.arange(20))
    assert np.allclose(hits["t"], 5.0, atol=1e-5)  # enters through the z = 0 face
    assert np.allclose(ext["t_leave"], 5.6, atol=1e-5)  # leaves through x + y + z = 1
    assert np.allclose(ext["d"].sum(1), 1.0, atol=1e-5)  # weights of a partition of unity
    assert np.allclose(ext["scalar"], 1.0, atol=1e-5)


def test_trilinear_exact_cases(orc):
    dims = (6, 5, 4)
    n = int(np.prod(dims))
    const = np.full(n, 2.5, np.float32)
    rng = np.random.default_rng(7)
    uvw = rng.uniform(0, 1, (500, 3)).astype(np.float32)
    # Lerp (1 - t) a + t b of equal values is the value to within rounding
    assert np.allclose(orc.tex3d(const, dims, (1, 1, 1), uvw), 2.5, rtol=0, atol=1e-6)
    z, y, x = np.meshgrid(np.arange(dims[2]), np.arange(dims[1]), np.arange(dims[0]), indexing="ij")
    lin = (0.5 * x + 0.25 * y - 0.125 * z).astype(np.float32).ravel()
    want = 0.5 * uvw[:, 0] * 5 + 0.25 * uvw[:, 1] * 4 - 0.125 * uvw[:, 2] * 3
    assert np.allclose(orc.tex3d(lin, dims, (1, 1, 1), uvw), want, atol=1e-5)
    corners = np.array([[i / 5, j / 4, k / 3] for i in (0, 5) for j in (0, 4) for k in (0, 3)],
                       np.float32)
    got = orc.tex3d(lin, dims, (1, 1, 1), corners)
    idx = [(k * 5 + j) * 6 + i for i in (0, 5) for j in (0, 4) for k in (0, 3)]
    assert np.array_equal(got, lin[idx])


# ------------------------------------------------------------------ volume

def _small_volume():
    dims = (12, 10, 9)
    rng = np.random.default_rng(8)
    vox = rng.uniform(0, 1, int(np.prod(dims)))
    tf = np.clip(rng.uniform(0, 1, (16, 4)), 0, 1)
    tf[:, 3] *= 0.3
    rays = W.pinhole_rays((1.7, -1.1, 1.4), (0.5, 0.5, 0.5), (0, 0, 1), 35.0, 12, 12)
    return dims, vox, tf, rays


def test_volume_f32_and_f64_forwards_agree(orc):
    dims, vox, tf, rays = _small_volume()
    step = 0.5 / 11
    a, ns = orc.volume_fwd(vox.astype(np.float32), dims, (1, 1, 1), tf.astype(np.float32), step, rays)
    b = orc.volume_fwd64(vox.astype(np.float32).astype(np.float64), dims, (1, 1, 1),
                         tf.astype(np.float32).astype(np.float64), step, rays)
    assert ns.max() > 10
    assert np.abs(a - b).max() < 1e-5


def test_volume_adjoint_matches_central_differences(orc):
    dims, vox, tf, rays = _small_volume()
    step = 0.5 / 11
    rng = np.random.default_rng(9)
    g = rng.normal(size=(len(rays), 4))
    gv, gt = orc.volume_bwd64(vox, dims, (1, 1, 1), tf, step, rays, g)

    def loss(v, t):
        return float(np.sum(g * orc.volume_fwd64(v, dims, (1, 1, 1), t, step, rays)))

    h = 1e-6
    for k in np.argsort(-np.abs(gv))[:12]:
        vp, vm = vox.copy(), vox.copy()
        vp[k] += h
        vm[k] -= h
        fd = (loss(vp, tf) - loss(vm, tf)) / (2 * h)
        assert abs(fd - gv[k]) <= 1e-6 * max(1.0, abs(fd)) + 1e-4 * abs(fd), (k, fd, gv[k])
    for k in range(tf.size):
        i, c = divmod(k, 4)
        tp, tm = tf.copy(), tf.copy()
        tp[i, c] += h
        tm[i, c] -= h
        fd = (loss(vox, tp) - loss(vox, tm)) / (2 * h)
        assert abs(fd - gt[i, c]) <= 1e-6 * max(1.0, abs(fd)) + 1e-4 * abs(fd), (i, c, fd)


def _ramp_case(tf_size=256):
    """A volume whose value is its x texture coordinate (x-linear ramp), a TF
    whose alpha and red are linear in the value, and +x rays: the march has a
    closed form computed here independently of the oracle's code."""
    dims = (12, 7, 5)
    x = np.arange(dims[0]) / (dims[0] - 1)
    vox = np.broadcast_to(x[None, None, :], dims[::-1]).astype(np.float32).ravel()
    s = np.arange(tf_size) / (tf_size - 1)
    tf = np.stack([s, np.ones_like(s), np.full_like(s, 0.5), 0.02 * s], 1).astype(np.float32)
    rays = np.zeros(4, W.RAY_DTYPE)
    rays["org"] = [(-0.5, 0.3, 0.6), (-0.25, 0.8, 0.1), (-0.5, 1.5, 0.5), (0.5, 0.5, 2.0)]
    rays["dir"] = [(1, 0, 0), (1, 0, 0), (1, 0, 0), (1, 0, 0)]
    rays["prev_prim_id"] = W.NO_PRIM
    return dims, vox, tf, rays


def test_volume_forward_closed_form(orc):
    dims, vox, tf, rays = _ramp_case()
    step = 0.5 / (dims[0] - 1)
    want = np.zeros((len(rays), 4))
    nwant = np.zeros(len(rays), int)
    for i, r in enumerate(rays):
        o = r["org"].astype(np.float64)
        if not (0 <= o[1] <= 1 and 0 <= o[2] <= 1):
            continue  # misses the unit cube
        tn, tf_ = -o[0], 1.0 - o[0]
        C, T = np.zeros(3), 1.0
        k = 0
        while tn + (k + 0.5) * step < tf_:
            sx = o[0] + tn + (k + 0.5) * step  # sample's x = its value
            a = 0.02 * sx
            C += T * a * np.array([sx, 1.0, 0.5])
            T *= 1 - a
            k += 1
        want[i, :3], want[i, 3], nwant[i] = C, 1 - T, k
    got, ns = orc.volume_fwd(vox, dims, (1, 1, 1), tf, step, rays)
    got64 = orc.volume_fwd64(vox.astype(np.float64), dims, (1, 1, 1), tf.astype(np.float64),
                             step, rays)
    assert list(ns) == list(nwant) and nwant[0] == 22
    assert np.abs(got - want).max() < 2e-6
    # f64 path: float sample geometry and float32-rounded voxels/TF entries
    assert np.abs(got64 - want).max() < 1e-7


# ------------------------------------------------------------------ scenes

def _scene_elements(orc, meshes, T, invT, invT33):
    trees = [orc.build_tri(V, F) for V, F in meshes]
    els = [dict(tree=trees[i], V=meshes[i][0], F=meshes[i][1], T=T[i], invT=invT[i],
                invT33=invT33[i]) for i in range(len(meshes))]
    return trees, els


def test_scene_oracle_matches_reference_fixture(orc):
    from oracle.oracle import canonical_tlas
    import hashlib
    g = np.load(os.path.join(GOLD, "scene_small.npz"), allow_pickle=False)
    meshes, T = W.batch_scene(count=12, lo=1100, hi=6000, seed=W.BASE_SEED + 21)
    h = hashlib.sha256()
    for m in meshes:
        for a in m:
            h.update(np.ascontiguousarray(a).tobytes())
    assert h.hexdigest() == str(g["meshes_sha"]), "scene generator drifted from the fixture"
    assert np.array_equal(T, g["T"])
    trees, els = _scene_elements(orc, meshes, T, g["invT"], g["invT33"])
    xb = np.array([orc.transform_box(t["bounds"][:3], t["bounds"][3:], T[i])
                   for i, t in enumerate(trees)])
    assert np.array_equal(xb, g["xbox"])
    tl = orc.tlas_build(xb)
    assert canonical_tlas(tl).tobytes() == g["tlas"].tobytes()
    hits, inst = orc.scene_trace(tl, els, g["rays"])
    assert hits.tobytes() == g["hits"].tobytes()
    assert np.array_equal(inst, g["inst"])


def test_scene_transform_box_closed_forms(orc):
    bmin, bmax = np.array([-1.0, 0.5, 2.0]), np.array([3.0, 1.5, 2.25])
    I = np.eye(4)
    assert np.array_equal(orc.transform_box(bmin, bmax, I), np.concatenate([bmin, bmax]))
    M = np.eye(4)
    M[3, :3] = [10.0, -4.0, 0.5]  # translation row (row-vector convention)
    assert np.array_equal(orc.transform_box(bmin, bmax, M),
                          np.concatenate([bmin + M[3, :3], bmax + M[3, :3]]))
    R = np.zeros((4, 4))  # p' = p R: x -> y, y -> -x, z -> z (90 degrees about z)
    R[0, 1], R[1, 0], R[2, 2], R[3, 3] = 1.0, -1.0, 1.0, 1.0
    out = orc.transform_box(bmin, bmax, R)
    assert np.array_equal(out, [-1.5, -1.0, 2.0, -0.5, 3.0, 2.25])


def test_tlas_invariants(orc):
    rng = np.random.default_rng(5)
    for n in (1, 2, 3, 17, 64):
        lo = rng.uniform(-5, 5, (n, 3))
        boxes = np.concatenate([lo, lo + rng.uniform(0.1, 2, (n, 3))], 1)
        if n == 17:
            boxes[5:9] = boxes[4]  # identical boxes force the object-median fallback
        tl = orc.tlas_build(boxes)
        leaves = tl[tl["is_leaf"] == 1]
        assert sorted(leaves["node_id"].tolist()) == list(range(n))
        assert len(tl) == 2 * n - 1
        # each leaf's box contains its element's box; children inside parents
        for k, nd in enumerate(tl):
            if nd["is_leaf"]:
                b = boxes[nd["node_id"]]
                assert np.all(nd["bmin"] <= b[:3]) and np.all(nd["bmax"] >= b[3:])
            else:
                for c in nd["child"]:
                    if c > 0 and k > 0:
                        assert np.all(tl[c]["bmin"] >= nd["bmin"]) and np.all(tl[c]["bmax"] <= nd["bmax"])


def test_scene_trace_equals_bruteforce_over_elements(orc):
    """The ordered, culled visit of elements gives the nearest world hit of
    all elements, each traced on its own."""
    meshes, T = W.batch_scene(count=10, lo=1100, hi=5000, seed=W.BASE_SEED + 22)
    invT = np.linalg.inv(T)
    invT33 = np.array([np.linalg.inv(np.block([[M[:3, :3], np.zeros((3, 1))],
                                               [np.zeros((1, 3)), np.ones((1, 1))]]))
                       for M in T])
    trees, els = _scene_elements(orc, meshes, T, invT, invT33)
    xb = np.array([orc.transform_box(t["bounds"][:3], t["bounds"][3:], T[i])
                   for i, t in enumerate(trees)])
    tl = orc.tlas_build(xb)
    rays = W.scene_cameras(T, count=1, width=48, height=48)[0]
    hits, inst = orc.scene_trace(tl, els, rays)
    best = np.full(len(rays), np.inf)
    for i, e in enumerate(els):
        lr = rays.copy()
        o = rays["org"].astype(np.float64)
        d = rays["dir"].astype(np.float64)
        lo = o @ e["invT"][:3, :3] + e["invT"][3, :3]
        ld = d @ e["invT33"][:3, :3] + e["invT33"][3, :3]
        lr["org"], lr["dir"] = lo.astype(np.float32), ld.astype(np.float32)
        lr["prev_prim_id"] = W.NO_PRIM
        h = orc.trace_tri(e["tree"], e["V"], e["F"], lr)
        ok = h["prim_id"] != W.NO_PRIM
        lp = (lo + h["t"].astype(np.float64)[:, None] * ld).astype(np.float32).astype(np.float64)
        wp = (lp @ e["T"][:3, :3] + e["T"][3, :3]).astype(np.float32)
        with np.errstate(over="ignore", invalid="ignore"):  # misses carry t = FLT_MAX
            tw = np.sqrt(np.sum((wp - rays["org"]) ** 2, axis=1, dtype=np.float32))
        best = np.where(ok & (tw < best), tw, best)
    hit = inst != W.NO_PRIM
    assert hit.mean() > 0.2
    assert np.array_equal(np.isfinite(best), hit)
    assert np.allclose(hits["t"][hit], best[hit], rtol=1e-6, atol=0)


@pytest.mark.skipif(not ref_available(), reason="reference build not present")
def test_scene_oracle_equals_live_reference(orc):
    from oracle.oracle import Reference, canonical_tlas
    ref = Reference("serial")
    meshes, T = W.batch_scene(count=64, lo=1100, hi=4000, seed=W.BASE_SEED + 23)
    rtrees = [ref.build("tri", V, F, keep=True) for V, F in meshes]
    sc = ref.scene_build(rtrees, T)
    trees, els = _scene_elements(orc, meshes, T, sc["invT"], sc["invT33"])
    xb = np.array([orc.transform_box(t["bounds"][:3], t["bounds"][3:], T[i])
                   for i, t in enumerate(trees)])
    assert np.array_equal(xb, sc["xbox"])
    tl = orc.tlas_build(xb)
    assert canonical_tlas(tl).tobytes() == canonical_tlas(sc["tlas"]).tobytes()
    rays = W.scene_cameras(T, count=2, width=96, height=96)
    for r in rays:
        h1, i1 = ref.scene_trace(sc, r)
        h2, i2 = orc.scene_trace(tl, els, r, threads=4)
        assert h1.tobytes() == h2.tobytes() and np.array_equal(i1, i2)
    ref.scene_free(sc)
    for t in rtrees:
        ref.free(t)


def _mixed_trees(orc, elems):
    return [orc.build_element(e) for e in elems]


def _mixed_elements(elems, trees, T, invT, invT33):
    return [dict(e, tree=trees[i], T=T[i], invT=invT[i], invT33=invT33[i])
            for i, e in enumerate(elems)]


@pytest.mark.skipif(not ref_available(), reason="reference build not present")
def test_mixed_scene_oracle_equals_live_reference(orc):
    """RenderGraph::Trace over triangle, tetra, particle and hexahedron
    elements (the kinds whose trees the oracle reproduces; the reference's
    line tree is its own spatial-median one): the oracle's mixed walk equals
    the reference's, hit for hit."""
    from oracle.oracle import Reference, canonical_tlas
    ref = Reference("serial")
    elems, T = W.mixed_scene(kinds=("tri", "tet", "sph", "sol"), copies=2)
    rtrees = []
    for e in elems:
        rtrees.append(ref.build(e["kind"], e["V"], e.get("F"), R=e.get("R"),
                                rconst=e.get("rconst", 0.0), keep=True, nv=e.get("nv", 0)))
    sc = ref.scene_build(rtrees, T)
    trees = _mixed_trees(orc, elems)
    for t, r in zip(trees, rtrees):
        assert np.array_equal(t["bounds"], r["bounds"])
    els = _mixed_elements(elems, trees, T, sc["invT"], sc["invT33"])
    xb = np.array([orc.transform_box(t["bounds"][:3], t["bounds"][3:], T[i])
                   for i, t in enumerate(trees)])
    assert np.array_equal(xb, sc["xbox"])
    tl = orc.tlas_build(xb)
    assert canonical_tlas(tl).tobytes() == canonical_tlas(sc["tlas"]).tobytes()
    hit_kinds = set()
    is_sol = np.array([e["kind"] == "sol" for e in elems] + [False])
    for r in W.scene_cameras(T, count=3, width=96, height=96):
        h1, i1 = ref.scene_trace(sc, r)
        h2, i2 = orc.scene_trace_mixed(tl, els, r, threads=4)
        assert np.array_equal(i1, i2)
        assert h1["t"].tobytes() == h2["t"].tobytes()
        assert np.array_equal(h1["prim_id"], h2["prim_id"])
        # SolidAccel::Traverse never writes u and v (the reference returns
        # whatever its Intersection held); every other kind's are exact
        sol = is_sol[np.minimum(i2, len(elems))]
        assert h1[~sol].tobytes() == h2[~sol].tobytes()
        hit_kinds |= {elems[k]["kind"] for k in np.unique(i2[i2 != W.NO_PRIM])}
    assert hit_kinds == {"tri", "tet", "sph", "sol"}
    ref.scene_free(sc)
    for t in rtrees:
        ref.free(t)


def test_mixed_scene_trace_equals_bruteforce_over_elements(orc):
    """Every kind, lines included: the ordered, culled walk gives the
    nearest world hit over all elements, each traced on its own."""
    elems, T = W.mixed_scene()
    invT = np.linalg.inv(T)
    invT33 = np.array([np.linalg.inv(np.block([[M[:3, :3], np.zeros((3, 1))],
                                               [np.zeros((1, 3)), np.ones((1, 1))]]))
                       for M in T])
    trees = _mixed_trees(orc, elems)
    els = _mixed_elements(elems, trees, T, invT, invT33)
    xb = np.array([orc.transform_box(t["bounds"][:3], t["bounds"][3:], T[i])
                   for i, t in enumerate(trees)])
    tl = orc.tlas_build(xb)
    rays = W.scene_cameras(T, count=1, width=64, height=64)[0]
    hits, inst = orc.scene_trace_mixed(tl, els, rays)
    best = np.full(len(rays), np.inf)
    o = rays["org"].astype(np.float64)
    d = rays["dir"].astype(np.float64)
    for e in els:
        lr = rays.copy()
        lo = o @ e["invT"][:3, :3] + e["invT"][3, :3]
        ld = d @ e["invT33"][:3, :3] + e["invT33"][3, :3]
        lr["org"], lr["dir"] = lo.astype(np.float32), ld.astype(np.float32)
        lr["prev_prim_id"] = W.NO_PRIM
        k = e["kind"]
        if k == "tri":
            h = orc.trace_tri(e["tree"], e["V"], e["F"], lr)
        elif k == "tet":
            h = orc.trace_tet(e["tree"], e["V"], e["F"], lr)[0]
        elif k == "sph":
            h = orc.trace_sph(e["tree"], e["V"], e["R"], e["rconst"], lr)
        elif k == "sol":
            h = orc.trace_sol(e["tree"], e["V"], e["F"], lr)[0]
        else:
            h = orc.trace_lin(e["tree"], e["V"], e["F"], e.get("R"), e["rconst"], lr)[0]
        ok = h["prim_id"] != W.NO_PRIM
        lp = (lo + h["t"].astype(np.float64)[:, None] * ld).astype(np.float32).astype(np.float64)
        wp = (lp @ e["T"][:3, :3] + e["T"][3, :3]).astype(np.float32)
        with np.errstate(over="ignore", invalid="ignore"):
            tw = np.sqrt(np.sum((wp - rays["org"]) ** 2, axis=1, dtype=np.float32))
        best = np.where(ok & (tw < best), tw, best)
    hit = inst != W.NO_PRIM
    assert hit.mean() > 0.1
    assert {elems[k]["kind"] for k in np.unique(inst[hit])} == {"tri", "tet", "sph", "sol", "lin"}
    assert np.array_equal(np.isfinite(best), hit)
    assert np.allclose(hits["t"][hit], best[hit], rtol=1e-6, atol=0)


# ------------------------------------------------------------------ solids

SOL_NV = (8, 6, 5)


def _eq_nan(a, b):
    return np.array_equal(a, b, equal_nan=True)


@pytest.mark.parametrize("nv", SOL_NV)
def test_solid_build_and_trace_match_reference_fixture(orc, nv):
    g = gold("sol_small.npz")
    p = f"n{nv}_"
    V, F, S = g[p + "V"], g[p + "F"], g[p + "S"]
    t = orc.build_sol(V, F, nv)
    assert canonical_tet(t["nodes"]).tobytes() == g[p + "nodes"].tobytes()
    assert np.array_equal(t["indices"], g[p + "indices"])
    assert np.array_equal(t["bounds"], g[p + "bounds"])
    _tree_invariants(t["nodes"], t["indices"], len(F), 4, False)
    hits, ext = orc.trace_sol(t, V, F, g[p + "rays"], S)
    assert np.array_equal(hits["prim_id"], g[p + "prim_id"])
    assert np.array_equal(hits["t"], g[p + "t"])
    assert (hits["prim_id"] != W.NO_PRIM).mean() > 0.5
    assert _eq_nan(ext["d"], g[p + "d"]) and _eq_nan(ext["scalar"], g[p + "scalar"])
    assert np.array_equal(ext["pos"], g[p + "pos"]) and np.array_equal(ext["nrm"], g[p + "nrm"])


@needs_ref
@pytest.mark.parametrize("nv", SOL_NV)
def test_solid_oracle_equals_live_reference(orc, nv):
    # fresh jitter, self-skipping rays, axis-parallel directions, inside origins
    ref = Reference("serial")
    V, F, S = W.solid_grid({8: 12, 6: 10, 5: 7}[nv], nv, seed=W.BASE_SEED + 77, jitter=0.3)
    t = orc.build_sol(V, F, nv)
    r = ref.build("sol", V, F, keep=True, nv=nv)
    assert canonical_tet(t["nodes"]).tobytes() == canonical_tet(r["nodes"]).tobytes()
    assert np.array_equal(t["indices"], r["indices"])
    cells = {8: 12, 6: 10, 5: 7}[nv]
    rays = W.solid_rays(cells, 40, 40)
    rng = np.random.default_rng(nv)
    rays["prev_prim_id"][::7] = rng.integers(0, len(F), len(rays[::7]))
    inner = rays[:800].copy()
    inner["org"] = rng.uniform(0.2, cells - 0.2, (800, 3)).astype(np.float32)
    d = rng.normal(size=(800, 3))
    d[::4, 1] = 0.0
    d[1::4, 0] = 0.0
    d[2::8, 2] = 0.0
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    inner["dir"] = d.astype(np.float32)
    rays = np.concatenate([rays, inner])
    h1, e1 = orc.trace_sol(t, V, F, rays, S, threads=4)
    h2, e2 = ref.trace(r, rays, scalars=S)
    ref.free(r)
    assert np.array_equal(h1["prim_id"], h2["prim_id"]) and np.array_equal(h1["t"], h2["t"])
    hit = h2["prim_id"] != W.NO_PRIM
    assert hit.mean() > 0.5
    assert _eq_nan(e1["d"][hit], e2["d"][hit]) and _eq_nan(e1["scalar"][hit], e2["scalar"][hit])
    assert np.array_equal(e1["pos"], e2["pos"]) and np.array_equal(e1["nrm"], e2["nrm"])


def _unit_solids():
    c = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1],
                  [0, 1, 1]], np.float64)
    return {8: c, 6: c[[0, 1, 2, 4, 5, 6]],
            5: np.concatenate([c[[0, 1, 2, 3]], [[0.5, 0.5, 1.0]]])}


@pytest.mark.parametrize("nv", SOL_NV)
def test_solid_crossings_analytic(orc, nv):
    # a vertical ray through each unit solid crosses its base at z = 0 and
    # its top (face or apex side) at a height the geometry fixes
    verts = _unit_solids()[nv]
    x, y = (0.5, 0.5) if nv != 6 else (0.6, 0.3)
    st, t, pos, nrm = orc.solid_isect(nv, [x, y, -1.0], [0.0, 0.0, 1.0], verts)
    assert st == 1
    top = 1.0 if nv != 5 else 1.0 - 2 * max(abs(x - 0.5), abs(y - 0.5))
    assert sorted(t.tolist()) == [1.0, 1.0 + top]
    for k in range(2):
        assert np.array_equal(pos[k][:2], np.float32([x, y]))
        assert np.abs(np.linalg.norm(nrm[k]) - 1) < 1e-6
    # the base point's normal is vertical, the entry before the exit
    base = int(np.argmin(t))
    assert pos[base][2] == 0.0 and abs(abs(nrm[base][2]) - 1) < 1e-6
    # a parallel ray beside the solid misses it; one starting inside crosses
    # its base behind the origin
    assert orc.solid_isect(nv, [2.5, 0.5, -1.0], [0.0, 0.0, 1.0], verts)[0] == 0
    st, t, _, _ = orc.solid_isect(nv, [x, y, 0.25], [0.0, 0.0, 1.0], verts)
    assert st == 1 and min(t) == np.float32(-0.25)
    # no axis with |dir| > 0.1: the reference leaves t unset
    assert orc.solid_isect(nv, [x, y, -1.0], [0.0, 0.0, 0.05], verts)[0] == -1


def test_solid_weights_analytic(orc):
    # at the centre of the unit cube's base the four base corners share the
    # weight and the top corners, whose rays through the point end on it, get 0
    c = _unit_solids()[8]
    d = orc.interpolate_solid(8, np.float32([0.5, 0.5, 0.0]), c)
    assert np.allclose(d[:4], 0.25, atol=1e-6) and np.all(d[4:] == 0)
    # at a corner of the cube that corner takes all the weight
    d = orc.interpolate_solid(8, np.float32([1.0, 1.0, 1.0]) - np.float32(1e-6), c)
    assert d[6] > 0.99 and abs(d.sum() - 1) < 1e-5
    # the prism's top-face centroid (2/3, 1/3, 1): the bottom corners get 0,
    # the corners (0,0,1) and (1,1,1), mirror images across the centroid's
    # line x + y = 1, tie, and the nearer right-angle corner weighs more
    p = _unit_solids()[6]
    d = orc.interpolate_solid(6, np.float32([2 / 3, 1 / 3, 1.0]), p)
    assert np.allclose(d[:3], 0, atol=1e-6) and abs(d[3:6].sum() - 1) < 1e-5
    assert abs(d[3] - d[5]) < 1e-6 and d[4] > d[3] + 0.1


@pytest.mark.parametrize("nv", SOL_NV)
def test_solid_trace_equals_bruteforce(orc, nv):
    # the traversal reports the smallest non-negative crossing over all solids
    V, F, _ = W.solid_grid({8: 11, 6: 9, 5: 6}[nv], nv, seed=W.BASE_SEED + 91)
    t = orc.build_sol(V, F, nv)
    rays = W.solid_rays({8: 11, 6: 9, 5: 6}[nv], 12, 12)
    hits, _ = orc.trace_sol(t, V, F, rays)
    Fr = F.reshape(-1, nv)
    for i in range(0, len(rays), 3):
        o = rays["org"][i].astype(np.float64)
        d = rays["dir"][i].astype(np.float64)
        best = np.float32(np.finfo(np.float32).max)
        for q in range(len(Fr)):
            st, ts, _, _ = orc.solid_isect(nv, o, d, V[Fr[q]].astype(np.float64))
            if st == 1:
                for tk in ts:
                    if 0 <= tk < best:
                        best = tk
        assert hits["t"][i] == best
        if best < np.finfo(np.float32).max:
            q = hits["prim_id"][i]
            st, ts, _, _ = orc.solid_isect(nv, o, d, V[Fr[q]].astype(np.float64))
            assert st == 1 and best in ts


# ------------------------------------------------------------------ lines

def _line_bruteforce(orc, P, S, R, rconst, o, d, cap=True):
    """Closest CylinderIsect crossing over every line (running t as
    TestLeafNode keeps it): (t, line) or (FLT_MAX, -1)."""
    best, bi = np.float32(np.finfo(np.float32).max), -1
    for li in range(len(S)):
        a, b = S[li]
        rr = max(R[a], R[b]) if R is not None else rconst
        h, out, _ = orc.line_isect(P[a], P[b], rr, o, d, cap, best)
        if h:
            best, bi = out[0], li
    return best, bi


@pytest.mark.parametrize("case", ["const", "nocap", "vary"])
def test_line_trace_matches_reference_fixture(orc, case):
    # constant radius: hits, u, v, position and normal bit-exact with the
    # reference's LineAccel; per-vertex radii: equal except where the
    # reference's boxes (p0 +- r0, p1 +- r1) hide a crossing of its max(r0, r1)
    # cylinder, and there ours is the brute-force closest
    g = gold("lin_small.npz")
    P, S, rays = g["P"], g["S"], g["rays"]
    R = g["R"] if case == "vary" else None
    rconst = 0.0 if case == "vary" else 0.004
    t = orc.build_lin(P, S, R, rconst)
    _tree_invariants(t["nodes"], t["indices"], len(S), 4, False)
    hits, ext = orc.trace_lin(t, P, S, R, rconst, rays, cap=case != "nocap")
    want = g[f"{case}_hits"]
    same = hits.tobytes() == want.tobytes()
    if case != "vary":
        assert same
        hit = want["prim_id"] != W.NO_PRIM
        assert hit.mean() > 0.1
        assert np.array_equal(ext["pos"][hit], g[f"{case}_pos"][hit])
        assert np.array_equal(ext["nrm"][hit], g[f"{case}_nrm"][hit])
        return
    diff = np.nonzero((hits["t"] != want["t"]) | (hits["prim_id"] != want["prim_id"]))[0]
    assert len(diff) < 0.01 * len(rays)
    for i in diff:
        bt, bi = _line_bruteforce(orc, P, S, R, 0.0, rays["org"][i], rays["dir"][i])
        assert hits["t"][i] == bt and hits["prim_id"][i] == bi and bt < want["t"][i]


@needs_ref
def test_line_oracle_equals_live_reference(orc):
    ref = Reference("serial")
    P, S, _ = W.polylines(800, 48, seed=W.BASE_SEED + 41, step=0.015)
    t = orc.build_lin(P, S, None, 0.003)
    r = ref.build("lin", P, S, R=None, rconst=0.003, keep=True)
    rays = W.line_rays(96, 96)
    # plus rays from inside the polyline cloud, some axis-parallel. Only unit
    # directions: CylinderIsect measures a cap crossing along the normalised
    # direction but the side and the box culling along the given one, so for
    # other lengths the reference's answer depends on its tree's visit order
    # (and short directions make its float quadratic report crossings no
    # cylinder has)
    extra = rays[:2000].copy()
    rng = np.random.default_rng(7)
    extra["org"] = rng.uniform(-1.2, 1.2, (2000, 3)).astype(np.float32)
    dd = rng.normal(size=(2000, 3))
    dd[::6, 2] = 0.0
    dd /= np.linalg.norm(dd, axis=1, keepdims=True)
    extra["dir"] = dd.astype(np.float32)
    rays = np.concatenate([rays, extra])
    h1, e1 = orc.trace_lin(t, P, S, None, 0.003, rays, threads=4)
    h2, e2 = ref.trace(r, rays, threads=4)
    ref.free(r)
    assert (h2["prim_id"] != W.NO_PRIM).mean() > 0.1
    assert h1.tobytes() == h2.tobytes()
    hit = h2["prim_id"] != W.NO_PRIM
    assert np.array_equal(e1["pos"][hit], e2["pos"][hit])
    assert np.array_equal(e1["nrm"][hit], e2["nrm"][hit])


def test_line_isect_analytic(orc):
    # unit-radius cylinder along z from 0 to 4; rays along +x at height 2
    # meet the side at x = -1 (t = 4 from x = -5), rays along +z from below
    # meet the lower cap at t = 1 (u = distance from the axis, v = 0)
    p0, p1 = np.float32([0, 0, 0]), np.float32([0, 0, 4])
    h, out, cap = orc.line_isect(p0, p1, 1.0, [-5, 0, 2], [1, 0, 0])
    assert h and not cap and out[0] == 4.0 and out[1] == 0.0 and out[2] == 0.5
    h, out, cap = orc.line_isect(p0, p1, 1.0, [0.5, 0, -1], [0, 0, 1])
    assert h and cap and out[0] == 1.0 and out[1] == 0.5 and out[2] == 0.0
    # without caps the same ray starts below the side and never meets it
    h, out, cap = orc.line_isect(p0, p1, 1.0, [0.5, 0, -1], [0, 0, 1], cap=False)
    assert not h
    # from above, the upper cap (v = 1); a miss beside the cylinder; a hit
    # beyond the running distance is rejected
    h, out, cap = orc.line_isect(p0, p1, 1.0, [0, 0.25, 9], [0, 0, -1])
    assert h and cap and out[0] == 5.0 and out[2] == 1.0
    assert not orc.line_isect(p0, p1, 1.0, [-5, 2, 2], [1, 0, 0])[0]
    assert not orc.line_isect(p0, p1, 1.0, [-5, 0, 2], [1, 0, 0], tmax=3.5)[0]


def test_line_trace_equals_bruteforce(orc):
    P, S, R = W.polylines(120, 20, seed=W.BASE_SEED + 42, radius="vary")
    t = orc.build_lin(P, S, R, 0.0)
    rays = W.line_rays(24, 24)
    hits, _ = orc.trace_lin(t, P, S, R, 0.0, rays)
    assert (hits["prim_id"] != W.NO_PRIM).mean() > 0.02
    for i in range(0, len(rays), 2):
        bt, bi = _line_bruteforce(orc, P, S, R, 0.0, rays["org"][i], rays["dir"][i])
        assert hits["t"][i] == bt
        if bi >= 0:
            assert hits["prim_id"][i] == bi or _line_bruteforce(
                orc, P, S[[hits["prim_id"][i]]], R, 0.0, rays["org"][i], rays["dir"][i])[0] == bt


# ------------------------------------------------------- BuildIntersection

def _pole_spheres():
    """Spheres on a grid with rays straight down and straight up onto their
    poles: the normals are (0, +-1, 0), the branch of BuildIntersection where
    n.x^2 + n.z^2 == 0."""
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
    return P, R, rays


@needs_ref
@pytest.mark.parametrize("kind", ["tri", "tet", "sph", "sph_poles"])
def test_build_intersection_oracle_equals_live_reference(orc, kind):
    """position, geometric / shading normal, tangent, binormal and f0..f3 of
    the oracle's BuildIntersection equal the Intersection the reference's
    Traverse fills (render/render_accel_{triangle,tetra,particle}.cc), bit for
    bit, including the pole branch of the sphere's tangent frame."""
    ref = Reference("serial")
    if kind == "tri":
        V, F = W.heightfield(64)
        rays = W.heightfield_rays(64, 64)
        h = orc.trace_tri(orc.build_tri(V, F), V, F, rays)
        got = orc.isect("tri", V, F, rays, h)
        r = ref.build("tri", V, F, keep=True)
    elif kind == "tet":
        V, F, Sc = W.tet_grid(8)
        rays = W.tet_rays(8, 64, 64)
        h, e = orc.trace_tet(orc.build_tet(V, F), V, F, rays, Sc)
        got = orc.isect("tet", V, F, rays, h, e)
        r = ref.build("tet", V, F, keep=True)
    else:
        if kind == "sph":
            P, R = W.plummer_spheres(20000, mu=-4.0)
            rays = W.sphere_rays(64, 64)
        else:
            P, R, rays = _pole_spheres()
        h = orc.trace_sph(orc.build_sph(P, R), P, R, 0.0, rays)
        got = orc.isect("sph", P, None, rays, h)
        r = ref.build("sph", P, R=R, keep=True)
        F = None
    want = ref.trace_isect(r, rays)
    ref.free(r)
    hit = h["prim_id"] != W.NO_PRIM
    assert hit.mean() > 0.3
    assert got.tobytes() == want.tobytes()
    if kind == "sph_poles":
        assert np.all(np.abs(got["normal"][hit][:, 1]) == 1.0)
        assert np.all(np.abs(got["binormal"][hit][:, 0]) == 1.0)


def test_build_intersection_oracle_closed_forms(orc):
    """Without the reference: the triangle normal is the unit normal of the
    plane (an axis-aligned quad gives exactly +-z), the position lies on the
    ray at t, and a sphere's normal has unit length and points from the
    centre to the hit point."""
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]], np.float32)
    F = np.array([[0, 1, 2], [1, 3, 2]], np.uint32)
    rays = np.zeros(2, W.RAY_DTYPE)
    rays["org"] = [[0.25, 0.25, 2.0], [0.75, 0.75, 2.0]]
    rays["dir"] = [[0, 0, -1], [0, 0, -1]]
    hits = np.zeros(2, W.HIT_DTYPE)
    hits["t"] = 2.0
    hits["prim_id"] = [0, 1]
    out = orc.isect("tri", V, F, rays, hits)
    assert np.array_equal(out["normal"], np.array([[0, 0, 1], [0, 0, 1]], np.float32))
    assert np.array_equal(out["position"][:, 2], [0, 0])
    assert list(out["f0"]) == [0, 1] and list(out["f2"]) == [2, 2]
    P, R = W.plummer_spheres(2000, mu=-3.0)
    rays = W.sphere_rays(32, 32)
    h = orc.trace_sph(orc.build_sph(P, R), P, R, 0.0, rays)
    o = orc.isect("sph", P, None, rays, h)
    hit = h["prim_id"] != W.NO_PRIM
    n = o["normal"][hit].astype(np.float64)
    assert np.allclose(np.linalg.norm(n, axis=1), 1.0, atol=1e-6)
    c = o["position"][hit] - P[h["prim_id"][hit]]
    assert np.allclose(n, c / np.linalg.norm(c, axis=1, keepdims=True), atol=1e-5)
    # tangent = 2 pi (-n.x, n.z, 0), the u-derivative direction; the
    # binormal's z component is pi * sqrt(n.x^2 + n.z^2)
    assert np.allclose(o["tangent"][hit][:, 0], -2 * np.pi * n[:, 0], atol=1e-5)
    assert np.allclose(o["tangent"][hit][:, 1], 2 * np.pi * n[:, 2], atol=1e-5)
    assert np.allclose(o["binormal"][hit][:, 2], np.pi * np.hypot(n[:, 0], n[:, 2]), atol=1e-5)


# ------------------------------------------------- byte / double textures

def _tex_inputs(rng, comps, fmt):
    dims = np.array([7, 5, 6], np.int32)
    n = int(np.prod(dims)) * comps
    if fmt == "byte":
        img = rng.integers(0, 256, n).astype(np.uint8)
    elif fmt == "float":
        img = rng.uniform(-1, 1, n).astype(np.float32)
    else:
        img = rng.uniform(-1, 1, n).astype(np.float64)
    uvw = rng.uniform(-1.5, 2.5, (3000, 3)).astype(np.float32)
    edge = np.array([[0, 0, 0], [1, 1, 1], [0.5, 0.5, 0.5], [-0.0, 1.0, 0.25],
                     [-0.01, 1.01, 0.999], [1e-8, -1e-8, 2.0]], np.float32)
    return dims, img, np.concatenate([uvw, edge])


@needs_ref
@pytest.mark.parametrize("fmt", ["byte", "float", "double"])
@pytest.mark.parametrize("comps", [1, 2, 3, 4])
def test_texture_fetch_oracle_equals_live_reference(orc, fmt, comps):
    """FilterTexture3DByte (/255.5f), FilterTexture3DFloat and
    FilterTexture3DDouble with 1-4 interleaved components, REPEAT, CLAMP and
    mixed wraps: the oracle's fetch equals the reference's, bit for bit."""
    ref = Reference("serial")
    rng = np.random.default_rng(100 + comps)
    dims, img, uvw = _tex_inputs(rng, comps, fmt)
    for clamp in ((0, 0, 0), (1, 1, 1), (1, 0, 1)):
        want = ref.tex3d_fetch(img, comps, dims, clamp, uvw)
        got = orc.tex3d_fetch(img, comps, dims, clamp, uvw)
        assert got.tobytes() == want.tobytes(), clamp


def test_texture_fetch_oracle_closed_forms(orc):
    """Byte texels are value / 255.5 at the corners; a constant field is
    returned exactly; the remap is a table lookup at (int)(x * n) clamped to
    the table, so fetching with a remap equals fetching at the looked-up
    coordinates (the lookup done here in numpy float32)."""
    dims = np.array([4, 3, 5], np.int32)
    img = np.arange(int(np.prod(dims)), dtype=np.uint8) * 3
    corners = np.array([[0, 0, 0], [1, 1, 1], [1, 0, 0]], np.float32)
    out = orc.tex3d_fetch(img, 1, dims, (1, 1, 1), corners)
    for k, c in enumerate(corners):
        idx = int(c[2] * 4) * 12 + int(c[1] * 2) * 4 + int(c[0] * 3)
        assert out[k, 0] == np.float32(np.float32(img[idx]) / np.float32(255.5))
    const = np.full(int(np.prod(dims)) * 2, 0.375, np.float64)
    out = orc.tex3d_fetch(const, 2, dims, (0, 0, 0), np.random.default_rng(1).random((50, 3)))
    assert np.all(out[:, :2] == np.float32(0.375)) and np.all(out[:, 2:] == 0)
    rng = np.random.default_rng(3)
    img = rng.uniform(0, 1, int(np.prod(dims)) * 4).astype(np.float32)
    tabs = [np.sort(rng.uniform(0, 1, m)).astype(np.float32) for m in (17, 5, 64)]
    uvw = rng.uniform(-0.2, 1.2, (2000, 3)).astype(np.float32)
    looked = uvw.copy()
    for k in range(3):
        n = len(tabs[k])
        idx = (uvw[:, k] * np.float32(n)).astype(np.float32).astype(np.int64)
        looked[:, k] = tabs[k][np.clip(idx, 0, n - 1)]
    a = orc.tex3d_fetch(img, 4, dims, (0, 1, 0), uvw, remap=tabs)
    b = orc.tex3d_fetch(img, 4, dims, (0, 1, 0), looked)
    assert a.tobytes() == b.tobytes()


# ------------------------------------------------------ SparseVolumeAccel

@needs_ref
@pytest.mark.parametrize("dtype", [np.uint8, np.float32, np.float64])
@pytest.mark.parametrize("comps,overlap", [(1, False), (3, True), (4, False), (5, True)])
def test_sparse_volume_oracle_equals_live_reference(orc, dtype, comps, overlap):
    """SparseVolumeAccel::Build + Sample (render/render_accel_volume.cc:35-139)
    -- the block BVHTree, Locate's first hit (overlapping blocks included),
    the nearest voxel with the reference's scaleX on every axis, byte / 255,
    float, double through float, alpha 1 or 0 -- bit for bit."""
    ref = Reference("serial")
    blocks, data, gdim = W.sparse_blocks(components=comps, dtype=dtype, overlap=overlap,
                                         seed=W.BASE_SEED + 13 + comps)
    rng = np.random.default_rng(comps)
    pos = rng.uniform(-4, gdim + 4, (5000, 3))
    # block corners and faces exactly
    pos[:len(blocks)] = blocks[:, :3]
    pos[len(blocks):2 * len(blocks)] = blocks[:, :3] + blocks[:, 3:6]
    want = ref.sparse_sample(blocks, data, comps, pos)
    got = orc.sparse_sample(blocks, data, comps, pos=pos)
    assert got.tobytes() == want.tobytes()
    assert 0.3 < (got[:, 3] != 0).mean() < 1.0


def test_sparse_volume_oracle_closed_forms(orc):
    """One block of extent 8 with 8^3 cells: the sample at a cell's corner is
    that cell's voxel; outside every block the result is (0, 0, 0, 0); with
    fewer cells than extent the reference scales y and z by size[0] / extent[0]
    too (so a block of 8 x 2 x 2 cells reads y = 1.5 as cell 1, not 0)."""
    vox = np.arange(8 * 8 * 8, dtype=np.float32)
    blocks = np.array([[0, 0, 0, 8, 8, 8, 8, 8, 8]], np.int32)
    pos = np.array([[3, 5, 6], [7.9, 0.1, 2.5], [-1, 0, 0], [0, 0, 8.5]], np.float64)
    out = orc.sparse_sample(blocks, [vox], 1, pos=pos)
    assert out[0, 0] == 6 * 64 + 5 * 8 + 3 and out[0, 3] == 1.0
    assert out[1, 0] == 2 * 64 + 0 * 8 + 7
    assert np.all(out[2] == 0) and np.all(out[3] == 0)
    vox = np.arange(8 * 2 * 2, dtype=np.float32)
    blocks = np.array([[0, 0, 0, 8, 8, 8, 8, 2, 2]], np.int32)
    out = orc.sparse_sample(blocks, [vox], 1, pos=np.array([[0.5, 1.5, 0.5]]))
    assert out[0, 0] == 1 * 8 + 0  # y = 1.5 * (8 / 8) -> cell 1
