"""Occlusion (srf_trace_any, srf_trace_any_sorted) of triangles and tetras
against the oracle's closest hit, bit for bit.

The reference answers shadow() with a full closest-hit Traverse
(gles/gles_glsl_runtime.cc:179-217): a ray is occluded iff Traverse reports
a hit with t < t_max. The expected flags are therefore the oracle's closest
hit (TriangleAccel::Traverse, render/render_accel_triangle.cc:2511-2669;
TetraAccel::Traverse, render/render_accel_tetra.cc:2200-2289) compared with
t_max, where t_max is put one float step below the closest distance, on it,
and one float step above it, so that a kernel that accepts t == t_max or
misses the closest primitive fails.

Each kernel variant runs: the wavefront traversal (rays of a wavefront share
their direction signs: a tiled pinhole camera), the per-lane traversal
(every wavefront mixes direction signs), the sorted and persistent launches,
and the kernels for trees of 2^24 or more nodes (forced on a small tree).
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle.oracle import Oracle  # noqa: E402
from surface_amd import workloads as W  # noqa: E402

F32_MAX = np.float32(np.finfo(np.float32).max)


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


def bracket_tmax(want_h):
    """t_max per ray: one float below the closest t, equal to it, one float
    above it (cycling over the rays); misses get FLT_MAX."""
    t = want_h["t"]
    k = np.arange(len(t)) % 3
    with np.errstate(over="ignore"):
        up = np.nextafter(t, np.float32(np.inf))
    tmax = np.where(k == 0, np.nextafter(t, np.float32(0)), np.where(k == 1, t, up))
    tmax = tmax.astype(np.float32)
    miss = want_h["prim_id"] == W.NO_PRIM
    tmax[miss] = F32_MAX
    # every 7th ray keeps the default search distance
    tmax[::7] = F32_MAX
    return tmax


def expected(want_h, tmax):
    return ((want_h["prim_id"] != W.NO_PRIM) & (want_h["t"] < tmax)).astype(np.int32)


def sign_groups(rays):
    """Fraction of 64-ray groups whose direction signs all agree (the
    kernels take their wavefront traversal for those)."""
    s = (rays["dir"] < 0).astype(np.int32) @ np.array([1, 2, 4])
    n = len(s) // 64 * 64
    g = s[:n].reshape(-1, 64)
    return float((g == g[:, :1]).all(1).mean())


def mix_signs(rays, rng, lo, hi):
    """Every fourth ray replaced by one with a random direction from a random
    origin in [lo, hi]^3, so no wavefront has uniform direction signs."""
    out = rays.copy()
    m = np.arange(len(out)) % 4 == 3
    k = int(m.sum())
    out["org"][m] = rng.uniform(lo, hi, (k, 3)).astype(np.float32)
    d = rng.normal(size=(k, 3))
    d[np.arange(k) % 2 == 0] *= -1
    out["dir"][m] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    # make sure each group holds both signs on x
    out["dir"][3::8, 0] = np.abs(out["dir"][3::8, 0]) + np.float32(0.01)
    out["dir"][7::8, 0] = -np.abs(out["dir"][7::8, 0]) - np.float32(0.01)
    return out


# ---------------------------------------------------------------- triangles

@pytest.fixture(scope="module")
def tri(S, orc):
    V, F = W.heightfield(160)
    tree = orc.build_tri(V, F)
    a = S.TriangleAccel()
    a.Build32(S.device_tensor(V), S.device_tensor(F))
    return {"V": V, "F": F, "tree": tree, "a": a}


def _tri_rays(case, rng):
    rays = W.heightfield_rays(192, 192)
    if case == "wave":
        return rays
    if case == "mixed":
        return mix_signs(rays, rng, 0.0, 1.0)
    # flags: single-sided rays, rays told to skip a triangle, zero direction
    # components (infinite inverse direction), mixed signs
    rays = mix_signs(rays, rng, 0.0, 1.0)
    rays["flags"][::3] = 0
    rays["prev_prim_id"][::5] = np.arange(0, len(rays), 5) % 50_000
    rays["dir"][::11, 0] = 0.0
    rays["dir"][1::11, 1] = -0.0
    return rays


@pytest.mark.parametrize("wide", [False, True])
@pytest.mark.parametrize("case", ["wave", "mixed", "flags"])
def test_triangle_occlusion_matches_oracle(S, orc, tri, torch, case, wide):
    say(f"triangle occlusion {case} wide={wide}")
    rng = np.random.default_rng(31)
    rays = _tri_rays(case, rng)
    frac = sign_groups(rays)
    if case == "wave":
        assert frac > 0.95, frac
    else:
        assert frac == 0.0, frac
    want_h = orc.trace_tri(tri["tree"], tri["V"], tri["F"], rays, threads=8)
    tmax = bracket_tmax(want_h)
    want = expected(want_h, tmax)
    assert 0.2 < want.mean() < 0.9
    a = tri["a"]
    a.force_wide_stack(wide)
    try:
        rd, td = S.device_tensor(rays), torch.from_numpy(tmax).cuda()
        got = a.TraverseAny(rd, td).cpu().numpy()
        assert np.array_equal(got, want), int((got != want).sum())
        # default search distance: occluded iff there is any hit
        got = a.TraverseAny(rd).cpu().numpy()
        assert np.array_equal(got, (want_h["prim_id"] != W.NO_PRIM).astype(np.int32))
        if not wide:
            for bits, mode in ((0, "plain"), (0, "persistent"), (9, "plain"),
                               (21, "persistent"), (30, "plain")):
                got = a.TraverseAnySorted(rd, td, key_bits=bits, mode=mode).cpu().numpy()
                assert np.array_equal(got, want), (bits, mode, int((got != want).sum()))
    finally:
        a.force_wide_stack(False)


@pytest.mark.parametrize("wide", [False, True])
def test_triangle_closest_hit_wide_stack_matches_oracle(S, orc, tri, wide):
    say(f"triangle closest hit wide={wide}")
    rng = np.random.default_rng(32)
    rays = np.concatenate([W.heightfield_rays(128, 128), _tri_rays("flags", rng)[:8192]])
    want = orc.trace_tri(tri["tree"], tri["V"], tri["F"], rays, threads=8)
    a = tri["a"]
    a.force_wide_stack(wide)
    try:
        got = S.hits_to_numpy(a.Traverse(S.device_tensor(rays)))
    finally:
        a.force_wide_stack(False)
    assert got.tobytes() == want.tobytes()


# ------------------------------------------------------------------- tetras

@pytest.fixture(scope="module")
def tet(S, orc):
    cells = 20
    V, F, Sc = W.tet_grid(cells)
    tree = orc.build_tet(V, F)
    a = S.TetraAccel()
    a.Build32(S.device_tensor(V), S.device_tensor(F))
    return {"V": V, "F": F, "Sc": Sc, "tree": tree, "a": a, "cells": cells}


def _tet_rays(case, tet, rng, orc):
    cells = tet["cells"]
    c = float(cells)
    if case == "wave":
        return W.tet_rays(cells, 160, 160)
    if case == "mixed":
        return mix_signs(W.tet_rays(cells, 160, 160), rng, 0.0, c)
    if case == "inside":
        org = rng.uniform(1.0, c - 1.0, (6000, 3))
        d = rng.normal(size=(6000, 3))
    elif case == "grazing":
        org = np.stack([np.full(6000, -2.0), rng.uniform(0, c, 6000), rng.uniform(0, c, 6000)], 1)
        d = np.stack([np.ones(6000), rng.normal(0, 1e-4, 6000), rng.normal(0, 1e-4, 6000)], 1)
    elif case == "prev":
        rays = W.tet_rays(cells, 96, 96)
        first, _ = orc.trace_tet(tet["tree"], tet["V"], tet["F"], rays, tet["Sc"], threads=8)
        rays["prev_prim_id"] = first["prim_id"]
        return rays
    r = np.zeros(len(d), W.RAY_DTYPE)
    r["org"] = org
    r["dir"] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    r["flags"] = 1
    r["prev_prim_id"] = W.NO_PRIM
    return r


@pytest.mark.parametrize("wide", [False, True])
@pytest.mark.parametrize("case", ["wave", "mixed", "inside", "grazing", "prev"])
def test_tetra_occlusion_matches_oracle(S, orc, tet, torch, case, wide):
    say(f"tetra occlusion {case} wide={wide}")
    rng = np.random.default_rng(41)
    rays = _tet_rays(case, tet, rng, orc)
    if case == "wave":
        assert sign_groups(rays) > 0.95
    elif case == "mixed":
        assert sign_groups(rays) == 0.0
    want_h, _ = orc.trace_tet(tet["tree"], tet["V"], tet["F"], rays, tet["Sc"], threads=8)
    tmax = bracket_tmax(want_h)
    want = expected(want_h, tmax)
    assert 0.1 < want.mean() < 0.95, want.mean()
    a = tet["a"]
    a.force_wide_stack(wide)
    try:
        rd, td = S.device_tensor(rays), torch.from_numpy(tmax).cuda()
        got = a.TraverseAny(rd, td).cpu().numpy()
        assert np.array_equal(got, want), int((got != want).sum())
        got = a.TraverseAny(rd).cpu().numpy()
        assert np.array_equal(got, (want_h["prim_id"] != W.NO_PRIM).astype(np.int32))
    finally:
        a.force_wide_stack(False)


@pytest.mark.parametrize("case", ["wave", "mixed", "inside", "prev"])
def test_tetra_closest_hit_wide_stack_matches_oracle(S, orc, tet, case):
    say(f"tetra closest hit wide {case}")
    rng = np.random.default_rng(42)
    rays = _tet_rays(case, tet, rng, orc)
    want_h, want_e = orc.trace_tet(tet["tree"], tet["V"], tet["F"], rays, tet["Sc"], threads=8)
    a = tet["a"]
    a.force_wide_stack(True)
    try:
        h, e = a.TraverseTetra(S.device_tensor(rays), S.device_tensor(tet["Sc"]))
    finally:
        a.force_wide_stack(False)
    hits = S.hits_to_numpy(h)
    ext = S.hits_to_numpy(e, W.TET_EXT_DTYPE)
    assert hits.tobytes() == want_h.tobytes()
    hit = want_h["prim_id"] != W.NO_PRIM
    for f in ("d", "scalar", "subface", "t_leave"):
        assert np.array_equal(ext[f][hit], want_e[f][hit]), f
