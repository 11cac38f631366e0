"""Benchmark of the MI355X acceleration-structure path (one GPU).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-extras]

Headline workload (a "step"): the batched incoherent particle configuration
-- 16,777,216 Plummer-distributed spheres with lognormal radii (mu = -6,
sigma = 0.5) -- built into an LBVH (ParticleAccel::Build32 semantics, with
secondary trees for the depth-capped leaves), traced with 2048x2048 pinhole
primary rays (closest hit), 8 cosine-weighted ambient-occlusion rays
generated per hit, and those traced for occlusion with the autotuned ray
sort (persistent launch). Inputs are resident in HBM before timing starts.
`value` is the step's ray throughput: (primary + AO rays) / step time.

Every configuration also gets its own line in `configs` with GPU times, a
`roofline` object for its dominant kernel and a `cpu_baseline` measured on
this host in the same run:
  heightfield  522,242 triangles, 1024^2 closest-hit rays (k_trace_tri_wave)
  heightfield_1m  1,048,352 triangles, 1024^2 rays (the 1 Grays/s target)
  tetgrid      1,572,864 tetras, 1024^2 rays + InterpolateTetra (k_trace_tet)
  scene        64 meshes (15.3 M triangles) in one segmented build, 16 x 1024^2
               rays through the two-level trace (k_scene_tri)
  volume       256^3 FP32 voxels, 256-entry TF, 1024^2 rays (k_vol_fwd / k_vol_bwd)
  particles    the headline (k_trace_sph primary, k_any_ordered AO)

roofline: algorithmic bytes of the dominant kernel per launch divided by its
launch time (HIP events on the launch stream), against the 8 TB/s HBM peak.
Trace bytes follow the hot-path formula (40 B per ray + node and primitive
bytes) with the node and primitive loads the kernel itself issues on these
rays, counted once in an instrumented build and frozen below; `traffic` is
the HBM bytes per launch from rocprofv3 FETCH_SIZE / WRITE_SIZE (scaled by
the same counters on a plain copy), frozen below from the committed profiles.
cpu_baseline: the reference's own OpenMP code (oracle/_ref, built from the
reference sources) or, for the volume march that the reference leaves to
user GLSL, the oracle's OpenMP restatement; 3 warm-ups, median of 10 timed
repetitions, on a bounded sample of the same workload, on this host's cores.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import statistics
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "closest-hit Mrays/s + LBVH build Mprims/s; volume fwd+bwd ms vs plain fwd"
HBM_PEAK_GBS = 8000.0

# ---------------------------------------------------------------------------
# roofline.achieved follows the hot-path table's algorithmic bytes: per ray
# 40 B (ray in, hit out) + node bytes x nodes + primitive bytes x primitives,
# with the REFERENCE traversal's per-ray node and primitive counts on the
# configuration's rays (the oracle's stats, which count what
# {Triangle,Tetra,Particle}Accel::Traverse visits), frozen here; build
# bytes per primitive likewise. A wavefront traversal loads each node and
# primitive once for up to 64 rays, so these rates can exceed the HBM peak:
# `traffic` (rocprofv3 FETCH_SIZE + WRITE_SIZE per launch, each scaled by
# what the counter reads for a plain 1 GiB device copy: fetch x 2.000, write
# x 1.000) is the HBM traffic the kernel really causes, frozen from the
# committed profiles, and `own_bytes` (where counted) the kernel's own loads.
NODE_PRIM_BYTES = {"tri": (64, 48), "tet": (40, 80), "sph": (56, 16)}
BUILD_BYTES = {"tri": 288, "tet": 336, "sph": 192}
REF_COUNTS = {
    # oracle stats over all 1024^2 rays
    "heightfield512": (64.3038, 94.4435),
    "heightfield725": (79.4602, 128.2012),
    "tetgrid64": (180.0010, 60.3673),
    # 65,536 of the 2048^2 primary rays (every 64th row); the reference scans
    # a depth-capped leaf's spheres one by one (13,788 tests per ray)
    "plummer16M_primary": (123.38, 13787.78),
    # every element whose world box the ray enters traced to its end: an
    # upper bound of RenderGraph::Trace, which skips elements beyond its hit
    # (every 64th row of the 16 cameras)
    "scene64_upper": (103.0319, 47.7509),
}
# the kernel's own per-ray node and primitive loads (an instrumented build):
# k_trace_tet pass A 60.53 nodes and 28.24 tetras per wavefront of 64 rays,
# pass B 0.0588 nodes and 0.0187 tetras per ray
OWN_LOADS = {"k_trace_tet": (0.9458 + 0.0588, 0.4413 + 0.0187)}
# HBM bytes per launch (FETCH_SIZE x 2 + WRITE_SIZE, KiB -> bytes) from the
# counter passes over `bench.py --probe` summarised in
# profiles/counters_bench_probe.md; a configuration missing here reports null
TRAFFIC = {
    "heightfield512": 212_830_143,
    "heightfield725": 452_779_556,
    "tetgrid64": 131_817_300,
    "plummer16M_primary": 1_425_300_605,  # both launches
    "scene64_upper": 23_603_015_759,      # persistent launch, 128-lane blocks
    "k_vol_fwd": 275_953_791,             # the march without checkpoints
    "k_vol_bwd": 1_376_455_261,
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def timed_events(torch, fn, reps):
    """Average ms of fn() over reps runs, by HIP events on the current stream
    (the stream every call here launches on)."""
    s = torch.cuda.Event(enable_timing=True)
    e = torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(reps):
        fn()
    e.record()
    e.synchronize()
    return s.elapsed_time(e) / reps


def cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for ln in out.splitlines():
            if ln.startswith("Model name:"):
                return ln.split(":", 1)[1].strip()
    except Exception:
        pass
    return platform.processor() or "unknown"


def cpu_threads():
    """The host threads the CPU leg uses: this process's CPU share."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        return os.cpu_count() or 1


def median_s(fn, warmup=3, reps=10):
    """BASELINE protocol: 3 warm-ups, median of 10 timed runs (seconds)."""
    for _ in range(warmup):
        fn()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return statistics.median(ts)


def roofline(kernel, ms, n_rays, kind, counts, launches=1):
    """Roofline object of a traversal kernel: `counts` names the frozen
    reference traversal counts, `kind` the node/primitive sizes."""
    nb, pb = NODE_PRIM_BYTES[kind]
    nodes, prims = REF_COUNTS[counts]
    bpr = 40.0 + nb * nodes + pb * prims
    gbs = bpr * n_rays / (ms * 1e-3) / 1e9
    out = {"bound": "hbm", "peak": HBM_PEAK_GBS, "unit": "GB/s", "kernel": kernel,
           "launches": launches, "kernel_ms": round(ms, 4),
           "achieved": round(gbs, 2), "frac": round(gbs / HBM_PEAK_GBS, 4),
           "bytes_per_launch": round(bpr * n_rays),
           "reference_loads_per_ray": {"nodes": nodes, "prims": prims}}
    own = OWN_LOADS.get(kernel)
    if own is not None:
        ob = 40.0 + nb * own[0] + pb * own[1]
        out["own_bytes"] = round(ob * n_rays)
        out["own_frac"] = round(ob * n_rays / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    tr = TRAFFIC.get(counts)
    out["traffic"] = round(tr) if tr is not None else None
    if tr is not None:
        out["traffic_frac"] = round(tr / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    return out


def build_roofline(kind, prims, ms):
    """The LBVH build against HBM: algorithmic bytes per primitive (Morton,
    sort, Karras, leaves and refit) x primitives / build time."""
    bpp = BUILD_BYTES[kind]
    gbs = bpp * prims / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "peak": HBM_PEAK_GBS, "unit": "GB/s", "bytes_per_prim": bpp,
            "bytes_per_build": bpp * prims, "achieved": round(gbs, 2),
            "frac": round(gbs / HBM_PEAK_GBS, 4)}


def reference_or_none():
    try:
        from oracle.oracle import Reference, ref_available
    except ImportError:
        return None
    if not ref_available("omp"):
        return None
    return Reference("omp")


def row_sample(rays, width, every):
    """Every `every`-th chunk of `width` consecutive rays (an image row of the
    8x8-tiled ray order holds width / 8 tiles' rows)."""
    return np.ascontiguousarray(rays.reshape(-1, width)[::every].reshape(-1))


# ------------------------------------------------------------------- configs

def cfg_triangles(torch, S, W, ref, grid, do_cpu, kernel):
    n_side = 1024
    V, F = W.heightfield(grid)
    rays_np = W.heightfield_rays(n_side, n_side)
    n = len(rays_np)
    tri = S.TriangleAccel()
    Vd, Fd = S.device_tensor(V), S.device_tensor(F)
    rh = S.device_tensor(rays_np)
    hh = torch.empty((n, 4), dtype=torch.float32, device="cuda")
    tri.Build32(Vd, Fd)
    tri.Traverse(rh, hh)
    tb = timed_events(torch, lambda: tri.Build32(Vd, Fd), 10)
    tt = timed_events(torch, lambda: tri.Traverse(rh, hh), 10)
    hit = float((hh[:, 3].view(torch.int32) != -1).float().mean().item())
    out = {"workload": f"heightfield {grid}x{grid} grid = {len(F):,} FP32 triangles; LBVH build "
                       f"+ {n_side}x{n_side} pinhole closest-hit rays (prim id, t, u, v)",
           "prims": int(len(F)), "rays": n, "build_ms": round(tb, 4),
           "build_mprims_per_s": round(len(F) / tb / 1e3, 1), "trace_ms": round(tt, 4),
           "trace_mrays_per_s": round(n / tt / 1e3, 1), "hit_fraction": round(hit, 4),
           "roofline": roofline(kernel, tt, n, "tri", f"heightfield{grid}"),
           "build_roofline": build_roofline("tri", len(F), tb)}
    if grid == 725:
        # Four cameras' frames in one launch: a single frame's launch ends
        # with its slowest wavefronts (grazing rays near the horizon) running
        # nearly alone; several frames fill that tail with other work.
        rc = S.device_tensor(W.heightfield_cameras(4, n_side, n_side))
        hc = torch.empty((len(rc), 4), dtype=torch.float32, device="cuda")
        tri.Traverse(rc, hc)
        tc = timed_events(torch, lambda: tri.Traverse(rc, hc), 10)
        out["four_cameras"] = {"rays": int(len(rc)), "trace_ms": round(tc, 4),
                               "trace_mrays_per_s": round(len(rc) / tc / 1e3, 1),
                               "workload": "4 cameras turned 90 degrees apart about the grid "
                                           "centre, 4 x 1024x1024 rays in one Traverse"}
        del rc, hc
    del tri
    if do_cpu and ref is not None:
        th = cpu_threads()
        b = median_s(lambda: ref.free(ref.build("tri", V, F, keep=True)))
        tree = ref.build("tri", V, F, keep=True)
        sample = row_sample(rays_np, n_side, 8)
        t = median_s(lambda: ref.trace(tree, sample, width=n_side, threads=th))
        ref.free(tree)
        out["cpu_baseline"] = {
            "value": round(len(sample) / t / 1e6, 3), "unit": "Mrays/s", "cores": th,
            "kind": "reference", "cpu_model": cpu_model(),
            "build_mprims_per_s": round(len(F) / b / 1e6, 3),
            "sample": (f"TriangleAccel::Traverse over {len(sample)} of the {n} rays (every 8th "
                       f"row), OpenMP dynamic over rows; Build32 of all {len(F)} triangles; "
                       "3 warm-ups, median of 10")}
    return out


def cfg_tetgrid(torch, S, W, ref, do_cpu):
    V, F, Sc = W.tet_grid(64)
    rays_np = W.tet_rays(64, 1024, 1024)
    n = len(rays_np)
    Vd, Fd, Sd = S.device_tensor(V), S.device_tensor(F), S.device_tensor(Sc)
    rays = S.device_tensor(rays_np)
    hits = torch.empty((n, 4), dtype=torch.float32, device="cuda")
    ext = torch.empty((n, 8), dtype=torch.float32, device="cuda")
    acc = S.TetraAccel()
    acc.Build32(Vd, Fd)
    acc.TraverseTetra(rays, Sd, hits, ext)
    tb = timed_events(torch, lambda: acc.Build32(Vd, Fd), 10)
    tt = timed_events(torch, lambda: acc.TraverseTetra(rays, Sd, hits, ext), 10)

    def step():
        acc.Build32(Vd, Fd)
        acc.TraverseTetra(rays, Sd, hits, ext)

    ts = timed_events(torch, step, 10)
    info = acc.get_info()
    out = {"workload": "tetra grid 64^3 hexes x 6 = 1,572,864 tets; LBVH build + 1024x1024 "
                       "pinhole closest-hit rays with InterpolateTetra scalar",
           "prims": int(len(F)), "rays": n, "nodes": int(info["num_nodes"]),
           "build_ms": round(tb, 4), "build_mprims_per_s": round(len(F) / tb / 1e3, 1),
           "trace_ms": round(tt, 4), "trace_mrays_per_s": round(n / tt / 1e3, 1),
           "step_ms": round(ts, 4), "step_mrays_per_s": round(n / ts / 1e3, 1),
           "roofline": roofline("k_trace_tet", tt, n, "tet", "tetgrid64"),
           "build_roofline": build_roofline("tet", len(F), tb)}
    del acc
    if do_cpu and ref is not None:
        th = cpu_threads()
        b = median_s(lambda: ref.free(ref.build("tet", V, F, keep=True)))
        tree = ref.build("tet", V, F, keep=True)
        sample = row_sample(rays_np, 1024, 8)
        t = median_s(lambda: ref.trace(tree, sample, width=1024, scalars=Sc, threads=th))
        ref.free(tree)
        out["cpu_baseline"] = {
            "value": round(len(sample) / t / 1e6, 3), "unit": "Mrays/s", "cores": th,
            "kind": "reference", "cpu_model": cpu_model(),
            "build_mprims_per_s": round(len(F) / b / 1e6, 3),
            "sample": (f"TetraAccel::Traverse + InterpolateTetra over {len(sample)} of the {n} "
                       "rays (every 8th row), OpenMP dynamic over rows; Build32 of all "
                       f"{len(F)} tets; 3 warm-ups, median of 10")}
    return out


def cfg_scene(torch, S, W, ref, do_cpu):
    """64 meshes (log-uniform 10k..2M triangles) built by one segmented LBVH
    build, placed by rigid transforms under a top-level BVH, traced from 16
    orbit cameras of 1024x1024 rays: plain, persistent and autotuned
    launches over all 16M rays, the tuned launch captured in a hipGraph, and
    one launch per camera for comparison."""
    log("batched scene: 64 meshes, 16 cameras")
    meshes, T = W.batch_scene(64)
    ntris = int(sum(len(F) for _, F in meshes))
    dm = [(S.device_tensor(V), S.device_tensor(F)) for V, F in meshes]
    accs = S.build_triangles_batch(dm)
    torch.cuda.synchronize()
    batch_ms = timed_events(torch, lambda: S.build_triangles_batch(dm, accels=accs), 5)

    def loop_build():
        for a, (Vd, Fd) in zip(accs, dm):
            a.Build32(Vd, Fd)

    loop_build()
    loop_ms = timed_events(torch, loop_build, 3)
    S.build_triangles_batch(dm, accels=accs)
    t0 = time.perf_counter()
    scene = S.Scene(accs, T)
    scene_ms = (time.perf_counter() - t0) * 1e3
    cams = W.scene_cameras(T, count=16, width=1024, height=1024)
    rays = [S.device_tensor(c) for c in cams]
    nr = sum(len(c) for c in cams)
    allrays = torch.cat(rays)
    hits = torch.empty((nr, 4), dtype=torch.float32, device="cuda")
    ids = torch.empty(nr, dtype=torch.int32, device="cuda")
    out = {"workload": "64 triangle meshes, log-uniform 10k-2M triangles (perturbed "
                       "icospheres / heightfields), one segmented LBVH build; 16 cameras x "
                       "1024^2 rays through the two-level trace (top-level BVH + per-instance "
                       "transforms), persistent launch, autotuned, captured in a hipGraph",
           "meshes": len(meshes), "triangles": ntris, "rays": nr,
           "batch_build_ms": round(batch_ms, 3),
           "batch_build_mprims_per_s": round(ntris / batch_ms / 1e3, 1),
           "per_mesh_build_loop_ms": round(loop_ms, 3), "scene_create_ms": round(scene_ms, 3)}
    out["autotuned"] = scene.autotune(allrays, hits, ids)
    for mode in ("plain", "persistent", "tuned"):
        scene.Trace(allrays, hits, ids, mode=mode)
        ms = timed_events(torch, lambda: scene.Trace(allrays, hits, ids, mode=mode), 3)
        out[f"trace_{mode}_ms"] = round(ms, 3)
        out[f"trace_{mode}_mrays_per_s"] = round(nr / ms / 1e3, 1)
    # all 16 cameras in one tuned launch captured in a hipGraph and replayed
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        scene.Trace(allrays, hits, ids, mode="tuned")
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        scene.Trace(allrays, hits, ids, mode="tuned")
    graph.replay()
    torch.cuda.synchronize()
    ms = timed_events(torch, graph.replay, 3)
    out["trace_graph_16_cameras_ms"] = round(ms, 3)
    out["trace_graph_mrays_per_s"] = round(nr / ms / 1e3, 1)
    out["roofline"] = roofline("k_scene_tri", ms, nr, "tri", "scene64_upper")
    out["build_roofline"] = build_roofline("tri", ntris, batch_ms)
    # one camera per launch, for comparison
    ms = timed_events(torch, lambda: [scene.Trace(r, hits[:len(r)], ids[:len(r)], mode="tuned")
                                      for r in rays], 2)
    out["trace_per_camera_launches_ms"] = round(ms, 3)
    out["hit_fraction"] = round(float((ids != -1).float().mean().item()), 4)
    del graph, scene, accs, dm, allrays, rays, hits, ids
    torch.cuda.empty_cache()
    if do_cpu and ref is not None:
        th = cpu_threads()
        trees = []
        t0 = time.perf_counter()
        for V, F in meshes:
            trees.append(ref.build("tri", V, F, keep=True))
        build_s = time.perf_counter() - t0
        sc = ref.scene_build(trees, T)
        sample = np.concatenate([row_sample(c, 1024, 64) for c in cams])
        t = median_s(lambda: ref.scene_trace(sc, sample, width=1024, threads=th))
        ref.scene_free(sc)
        for tr in trees:
            ref.free(tr)
        out["cpu_baseline"] = {
            "value": round(len(sample) / t / 1e6, 3), "unit": "Mrays/s", "cores": th,
            "kind": "reference", "cpu_model": cpu_model(),
            "build_mprims_per_s": round(ntris / build_s / 1e6, 3),
            "sample": (f"RenderGraph::Trace (BVHTree + per-element TriangleAccel::Traverse) "
                       f"over {len(sample)} of the {nr} rays (every 64th row of each camera), "
                       "OpenMP dynamic over rows, 3 warm-ups, median of 10; the 64 Build32 "
                       "calls timed once")}
    return out


def cfg_volume(torch, S, W, orc, do_cpu):
    log("volume 256^3")
    from surface_amd.volume import Volume
    vox_np = W.gaussian_volume(256)
    tf_np = W.transfer_function(256)
    rays_np = W.volume_rays(1024, 1024)
    n = len(rays_np)
    vox = torch.from_numpy(vox_np).cuda()
    tf = torch.from_numpy(tf_np).cuda()
    vr = S.device_tensor(rays_np)
    vol = Volume(vox, (256, 256, 256), (1, 1, 1))
    step_len = 0.5 / 255
    gv = torch.zeros_like(vox)
    gt = torch.zeros_like(tf)
    grad = torch.ones((n, 4), dtype=torch.float32, device="cuda")
    ck = torch.empty((n, vol.checkpoint_slots(step_len)), dtype=torch.float32, device="cuda")
    rgba = vol.march(tf, step_len, vr)
    fwd = timed_events(torch, lambda: vol.march(tf, step_len, vr, rgba=rgba), 5)
    vol.march(tf, step_len, vr, ck, rgba)
    bwd = timed_events(torch, lambda: vol.march_backward(tf, step_len, vr, grad, gv, gt, ck), 3)

    def fwd_bwd():
        vol.march(tf, step_len, vr, ck, rgba)
        vol.march_backward(tf, step_len, vr, grad, gv, gt, ck)

    fwd_bwd()
    fb = timed_events(torch, fwd_bwd, 3)
    ns = samples_per_ray(rays_np, step_len)
    samples = float(ns.sum())
    fwd_bytes = 40 * n + 32 * samples
    bwd_bytes = 56 * n + 96 * samples
    out = {"workload": "volume 256^3 FP32 (64 anisotropic Gaussians), 256-entry RGBA TF, "
                       "1024x1024 rays, 0.5-voxel trilinear steps, emission-absorption; "
                       "forward and forward + adjoint (grads w.r.t. voxels and TF)",
           "rays": n, "samples": int(samples), "fwd_ms": round(fwd, 3), "bwd_ms": round(bwd, 3),
           "fwd_bwd_ms": round(fb, 3), "fwd_bwd_over_fwd": round(fb / fwd, 3),
           "fwd_mrays_per_s": round(n / fwd / 1e3, 1),
           "roofline": {"bound": "hbm", "kernel": "k_vol_fwd", "unit": "GB/s",
                        "peak": HBM_PEAK_GBS, "kernel_ms": round(fwd, 4),
                        "bytes_per_launch": round(fwd_bytes),
                        "achieved": round(fwd_bytes / (fwd * 1e-3) / 1e9, 2),
                        "frac": round(fwd_bytes / (fwd * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                        "traffic": TRAFFIC.get("k_vol_fwd"),
                        "traffic_frac": (round(TRAFFIC["k_vol_fwd"] / (fwd * 1e-3) / 1e9 /
                                               HBM_PEAK_GBS, 4)
                                         if TRAFFIC.get("k_vol_fwd") else None),
                        "bwd": {"kernel": "k_vol_bwd", "kernel_ms": round(bwd, 4),
                                "bytes_per_launch": round(bwd_bytes),
                                "achieved": round(bwd_bytes / (bwd * 1e-3) / 1e9, 2),
                                "frac": round(bwd_bytes / (bwd * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                                "traffic": TRAFFIC.get("k_vol_bwd"),
                                "traffic_frac": (round(TRAFFIC["k_vol_bwd"] / (bwd * 1e-3) /
                                                       1e9 / HBM_PEAK_GBS, 4)
                                                 if TRAFFIC.get("k_vol_bwd") else None)}}}
    del vol, vox, tf, vr, gv, gt, grad, ck, rgba
    torch.cuda.empty_cache()
    if do_cpu and orc is not None:
        th = cpu_threads()
        sample = row_sample(rays_np, 1024, 32)
        dims = (256, 256, 256)
        t = median_s(lambda: orc.volume_fwd(vox_np, dims, (1, 1, 1), tf_np, step_len, sample,
                                            threads=th))
        out["cpu_baseline"] = {
            "value": round(len(sample) / t / 1e6, 4), "unit": "Mrays/s", "cores": th,
            "kind": "port", "cpu_model": cpu_model(),
            "sample": (f"the oracle's C restatement of the march (FilterTexture3DFloat + TF + "
                       f"EA, OpenMP over rays) on {len(sample)} of the {n} rays (every 32nd "
                       "row), 3 warm-ups, median of 10; the reference leaves the march to "
                       "user GLSL and evaluates fragments serially")}
    return out


def samples_per_ray(rays, step):
    """Samples of the fixed-step march per ray: t_k = t0 + (k + 1/2) step
    below the exit of the unit cube (slab entry clamped at 0), in float32."""
    o = rays["org"].astype(np.float32)
    d = rays["dir"].astype(np.float32)
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = np.float32(1.0) / d
        a = (np.float32(0.0) - o) * inv
        b = (np.float32(1.0) - o) * inv
    tn = np.nanmax(np.minimum(a, b), axis=1)
    tf = np.nanmin(np.maximum(a, b), axis=1)
    tn = np.maximum(tn, np.float32(0.0))
    ns = np.where(tf > tn, np.ceil((tf - tn) / np.float32(step) - np.float32(0.5)), 0)
    return np.maximum(ns, 0)


class Particles:
    """The headline configuration: build + primary + AO generation + AO
    occlusion on 16,777,216 spheres and 2048^2 primary rays."""

    def __init__(self, torch, S, W):
        log("particles: 16M spheres, 2048^2 primary + 8 AO rays per hit")
        self.torch, self.S = torch, S
        P, R = W.plummer_spheres(1 << 24)
        self.P_np, self.R_np = P, R
        self.Pd, self.Rd = S.device_tensor(P), S.device_tensor(R)
        self.acc = S.ParticleAccel()
        self.acc.Build32(self.Pd, self.Rd, 0.0)
        self.prim_np = W.sphere_rays(2048, 2048)
        self.prim = S.device_tensor(self.prim_np)
        self.np = self.prim.shape[0]
        self.hits = torch.empty((self.np, 4), dtype=torch.float32, device="cuda")
        self.acc.Traverse(self.prim, self.hits)
        self.cap = 8 * self.np
        self.ao = torch.empty((self.cap, 8), dtype=torch.float32, device="cuda")
        self.tmax = torch.empty(self.cap, dtype=torch.float32, device="cuda")
        self.pix = torch.empty(self.cap, dtype=torch.int32, device="cuda")
        self.count = torch.zeros(1, dtype=torch.int32, device="cuda")
        self.occ = torch.empty(self.cap, dtype=torch.int32, device="cuda")
        self.gen()
        torch.cuda.synchronize()
        self.na = int(self.count.item())
        # the AO launch shape: tuned once on this frame's rays, before timing
        self.cfg = self.acc.autotune_any(self.ao[:self.na], self.tmax[:self.na],
                                         self.occ[:self.na])
        self.mode = "persistent" if self.cfg["persistent"] else "plain"

    def build(self):
        self.acc.Build32(self.Pd, self.Rd, 0.0)

    def primary(self):
        self.acc.Traverse(self.prim, self.hits)

    def gen(self):
        self.S.ao_rays_spheres(self.prim, self.hits, self.Pd, self.Rd, 0.0, 8, 1e-3, 0.05, 7,
                               out=(self.ao, self.tmax, self.pix, self.count))

    def ao_trace(self):
        # the ray count of the frame is fixed (the same hits every step)
        self.acc.TraverseAnySorted(self.ao[:self.na], self.tmax[:self.na], self.cfg["key_bits"],
                                   self.mode, self.occ[:self.na])

    def step(self):
        self.build()
        self.primary()
        self.gen()
        self.ao_trace()


def cfg_particles(torch, S, W, ref, pt, do_cpu):
    out = {"workload": "16,777,216 Plummer spheres (a = 1), lognormal radii (mu = -6, "
                       "sigma = 0.5); LBVH build, 2048x2048 pinhole primary closest hits, 8 "
                       "cosine-weighted AO rays per hit traced for occlusion (any hit, "
                       "autotuned ray sort, persistent launch)",
           "spheres": 1 << 24, "primary_rays": pt.np, "ao_rays": pt.na,
           "autotuned": pt.cfg}
    b = timed_events(torch, pt.build, 3)
    p = timed_events(torch, pt.primary, 3)
    g = timed_events(torch, pt.gen, 3)
    a = timed_events(torch, pt.ao_trace, 3)
    out.update({"build_ms": round(b, 3), "build_mprims_per_s": round((1 << 24) / b / 1e3, 1),
                "primary_ms": round(p, 3), "primary_mrays_per_s": round(pt.np / p / 1e3, 1),
                "ao_gen_ms": round(g, 3), "ao_ms": round(a, 3),
                "ao_mrays_per_s": round(pt.na / a / 1e3, 1),
                "primary_hit_fraction": round(pt.na / 8 / pt.np, 4),
                "ao_occluded_fraction": round(float(pt.occ[:pt.na].float().mean().item()), 4)})
    for name, bits, mode in (("unsorted_plain", 0, "plain"),
                             ("unsorted_persistent", 0, "persistent")):
        ms = timed_events(torch, lambda: pt.acc.TraverseAnySorted(
            pt.ao[:pt.na], pt.tmax[:pt.na], bits, mode, pt.occ[:pt.na]), 3)
        out[f"ao_{name}_ms"] = round(ms, 3)
    out["roofline"] = roofline("k_trace_sph", p, pt.np, "sph", "plummer16M_primary", launches=2)
    out["build_roofline"] = build_roofline("sph", 1 << 24, b)
    if do_cpu and ref is not None:
        th = cpu_threads()
        bs = median_s(lambda: ref.free(ref.build("sph", pt.P_np, R=pt.R_np, keep=True)))
        tree = ref.build("sph", pt.P_np, R=pt.R_np, keep=True)
        sample = row_sample(pt.prim_np, 2048, 512)
        ps = median_s(lambda: ref.trace(tree, sample, width=2048, threads=th))
        ao_np = pt.ao[:pt.na].cpu().numpy().view(W.RAY_DTYPE).reshape(-1)
        ao_sample = np.ascontiguousarray(ao_np[::4096])
        as_ = median_s(lambda: ref.trace(tree, ao_sample, width=64, threads=th))
        ref.free(tree)
        prate = len(sample) / ps
        arate = len(ao_sample) / as_
        step_s = bs + pt.np / prate + pt.na / arate
        out["cpu_baseline"] = {
            "value": round((pt.np + pt.na) / step_s / 1e6, 4), "unit": "Mrays/s", "cores": th,
            "kind": "reference", "cpu_model": cpu_model(),
            "build_mprims_per_s": round((1 << 24) / bs / 1e6, 3),
            "primary_mrays_per_s": round(prate / 1e6, 4),
            "ao_mrays_per_s": round(arate / 1e6, 4),
            "sample": (f"ParticleAccel::Build32 of all 16,777,216 spheres; "
                       f"ParticleAccel::Traverse over {len(sample)} of the {pt.np} primary rays "
                       f"(every 512th row) and {len(ao_sample)} of the {pt.na} AO rays (every "
                       "4096th; the reference answers shadow() with a closest-hit Traverse), "
                       "OpenMP dynamic, 3 warm-ups, median of 10 each; value = rays of the "
                       "step over build time + rays at the sampled rates")}
    return out


def solids_lines(torch, S, W):
    """Hexahedra (64^3 jittered grid, 262,144 solids) and capped lines
    (16,384 random-walk polylines x 64 = 1,048,576 segments): build and
    1024^2 closest-hit rays each. The solid search visits every node whose
    box the ray enters, as SolidAccel::Traverse does."""
    out = {}
    log("solids: 64^3 hexahedra")
    V, F, Sc = W.solid_grid(64, 8)
    Vd, Fd, Sd = S.device_tensor(V), S.device_tensor(F), S.device_tensor(Sc)
    rays = S.device_tensor(W.solid_rays(64, 1024, 1024))
    n = rays.shape[0]
    hits = torch.empty((n, 4), dtype=torch.float32, device="cuda")
    ext = torch.empty((n, 16), dtype=torch.float32, device="cuda")
    a = S.SolidAccel()
    a.Build32(Vd, Fd, 8)
    a.TraverseSolid(rays, Sd, hits, ext)
    tb = timed_events(torch, lambda: a.Build32(Vd, Fd, 8), 5)
    tt = timed_events(torch, lambda: a.TraverseSolid(rays, Sd, hits, ext), 3)
    hit = (hits[:, 3].view(torch.int32) != -1).float().mean().item()
    out["hexahedra_262144"] = {"build_ms": round(tb, 4),
                               "build_mprims_per_s": round(len(F) / tb / 1e3, 1),
                               "trace_ms": round(tt, 4),
                               "trace_mrays_per_s": round(n / tt / 1e3, 1),
                               "hit_fraction": round(hit, 4)}
    del a
    log("lines: 1M capped segments")
    P, segs, _ = W.polylines(16384, 64, seed=W.BASE_SEED + 46, step=0.01)
    Pd, Sgd = S.device_tensor(P), S.device_tensor(segs)
    lr = S.device_tensor(W.line_rays(1024, 1024))
    lh = torch.empty((n, 4), dtype=torch.float32, device="cuda")
    le = torch.empty((n, 8), dtype=torch.float32, device="cuda")
    b = S.LineAccel()
    b.Build(Pd, Sgd, None, 0.0015)
    b.TraverseLines(lr, lh, le)
    tb = timed_events(torch, lambda: b.Build(Pd, Sgd, None, 0.0015), 5)
    tt = timed_events(torch, lambda: b.TraverseLines(lr, lh, le), 3)
    hit = (lh[:, 3].view(torch.int32) != -1).float().mean().item()
    out["lines_1048576"] = {"build_ms": round(tb, 4),
                            "build_mprims_per_s": round(len(segs) / tb / 1e3, 1),
                            "trace_ms": round(tt, 4),
                            "trace_mrays_per_s": round(n / tt / 1e3, 1),
                            "hit_fraction": round(hit, 4)}
    return out


def probe(torch, S, W):
    """Each configuration's kernels a few times and a 1 GiB device copy (the
    scale of the HBM counters), for rocprofv3 counter passes; no JSON line."""
    pt = Particles(torch, S, W)
    for _ in range(2):
        pt.step()
    torch.cuda.synchronize()
    del pt
    torch.cuda.empty_cache()
    log("probe: particles done")
    for grid in (512, 725):
        V, F = W.heightfield(grid)
        tri = S.TriangleAccel()
        Vd, Fd = S.device_tensor(V), S.device_tensor(F)
        rh = S.device_tensor(W.heightfield_rays(1024, 1024))
        hh = torch.empty((len(rh), 4), dtype=torch.float32, device="cuda")
        for _ in range(2):
            tri.Build32(Vd, Fd)
            tri.Traverse(rh, hh)
        torch.cuda.synchronize()
    log("probe: heightfields done")
    V, F, Sc = W.tet_grid(64)
    acc = S.TetraAccel()
    Vd, Fd, Sd = S.device_tensor(V), S.device_tensor(F), S.device_tensor(Sc)
    rays = S.device_tensor(W.tet_rays(64, 1024, 1024))
    hits = torch.empty((len(rays), 4), dtype=torch.float32, device="cuda")
    ext = torch.empty((len(rays), 8), dtype=torch.float32, device="cuda")
    for _ in range(2):
        acc.Build32(Vd, Fd)
        acc.TraverseTetra(rays, Sd, hits, ext)
    torch.cuda.synchronize()
    log("probe: tetgrid done")
    from surface_amd.volume import Volume
    vox = torch.from_numpy(W.gaussian_volume(256)).cuda()
    tf = torch.from_numpy(W.transfer_function(256)).cuda()
    vr = S.device_tensor(W.volume_rays(1024, 1024))
    vol = Volume(vox, (256, 256, 256), (1, 1, 1))
    step_len = 0.5 / 255
    gv, gt = torch.zeros_like(vox), torch.zeros_like(tf)
    grad = torch.ones((len(vr), 4), dtype=torch.float32, device="cuda")
    ck = torch.empty((len(vr), vol.checkpoint_slots(step_len)), dtype=torch.float32,
                     device="cuda")
    rgba = vol.march(tf, step_len, vr)
    for _ in range(2):  # the march as the bench times it, without checkpoints
        vol.march(tf, step_len, vr, rgba=rgba)
    for _ in range(2):
        vol.march(tf, step_len, vr, ck, rgba)
        vol.march_backward(tf, step_len, vr, grad, gv, gt, ck)
    torch.cuda.synchronize()
    log("probe: volume done")
    meshes, T = W.batch_scene(64)
    dm = [(S.device_tensor(V), S.device_tensor(F)) for V, F in meshes]
    accs = S.build_triangles_batch(dm)
    scene = S.Scene(accs, T)
    allrays = torch.cat([S.device_tensor(c) for c in W.scene_cameras(T, 16, 1024, 1024)])
    sh = torch.empty((len(allrays), 4), dtype=torch.float32, device="cuda")
    ids = torch.empty(len(allrays), dtype=torch.int32, device="cuda")
    for _ in range(2):
        scene.Trace(allrays, sh, ids, mode="persistent")
    torch.cuda.synchronize()
    log("probe: scene done")
    src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    log(f"probe done: copy of {src.numel() * 4} bytes")
    del rgba


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline legs")
    ap.add_argument("--no-extras", action="store_true",
                    help="headline step only (no other configurations)")
    ap.add_argument("--probe", action="store_true",
                    help="each configuration's kernels twice plus a 1 GiB device copy (for "
                         "rocprofv3 counter passes); prints no JSON line")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); rays shard by "
              "screen tile across GPUs by design, but no multi-GPU path is built.",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from surface_amd import accel as S
    from surface_amd import workloads as W

    assert torch.cuda.is_available(), "no GPU visible"
    if args.probe:
        probe(torch, S, W)
        return
    pt = Particles(torch, S, W)
    log(f"warmup {args.warmup}")
    for _ in range(args.warmup):
        pt.step()
    torch.cuda.synchronize()
    log(f"timing {args.steps} steps")
    t0 = time.perf_counter()
    ms_step = timed_events(torch, pt.step, args.steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    rays_per_step = pt.np + pt.na
    value = rays_per_step / ms_step / 1e3

    do_cpu = not args.no_cpu
    ref = reference_or_none() if do_cpu else None
    orc = None
    if do_cpu:
        try:
            from oracle.oracle import Oracle
            orc = Oracle()
        except (ImportError, OSError):
            orc = None
    log("particles: per-phase times" + (" and CPU baseline" if ref else ""))
    parts = cfg_particles(torch, S, W, ref, pt, do_cpu)
    del pt
    torch.cuda.empty_cache()
    configs = {"particles": parts}
    extras = {}
    if not args.no_extras:
        log("heightfield 512")
        configs["heightfield"] = cfg_triangles(torch, S, W, ref, 512, do_cpu, "k_trace_tri_wave")
        log("heightfield 725 (1,048,352 triangles)")
        configs["heightfield_1m"] = cfg_triangles(torch, S, W, ref, 725, do_cpu,
                                                  "k_trace_tri_wave")
        log("tetra grid 64")
        configs["tetgrid"] = cfg_tetgrid(torch, S, W, ref, do_cpu)
        configs["volume"] = cfg_volume(torch, S, W, orc, do_cpu)
        configs["scene"] = cfg_scene(torch, S, W, ref, do_cpu)
        extras["solids_and_lines"] = solids_lines(torch, S, W)

    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": parts["workload"], "spheres": 1 << 24,
                   "primary_rays": parts["primary_rays"], "ao_rays": parts["ao_rays"],
                   "rays_per_step": rays_per_step},
        "build_ms": parts["build_ms"],
        "primary_ms": parts["primary_ms"],
        "ao_ms": parts["ao_ms"],
        "wall_s_timed": round(wall, 3),
        "roofline": parts["roofline"],
        "cpu_baseline": parts.get("cpu_baseline"),
        "configs": configs,
        "extras": extras,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
