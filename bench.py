"""Benchmark of the MI355X acceleration-structure path (one GPU).

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload (a "step"): the plain tetra configuration -- 64^3 jittered hexes split
into 1,572,864 tetrahedra with a per-vertex scalar -- built into an LBVH
(TetraAccel::Build32 semantics) and traced with 1024x1024 pinhole primary rays
(closest hit: entry t, InterpolateTetra weights, interpolated scalar). Inputs
are resident in HBM before timing starts. `value` is whole-step throughput
(rays / (build + trace time)); the JSON line also carries the trace-only and
build-only rates, the heightfield triangle case (522,242 triangles, 1024^2
rays) and the differentiable volume (256^3, 1024^2 rays, fwd and fwd+bwd ms).

roofline: the dominant kernel (the tetra traversal) against MI355X's 8 TB/s
HBM peak, with algorithmic bytes per ray = 40 B (ray in, hit out) + 40 B per
node loaded + 80 B per tetra loaded (64 B geometry + 16 B scalars), the node
and tetra counts being the kernel's own loads on these rays (a wavefront
loads each node and tetra once for its 64 rays; counted once and frozen
below; the reference's per-ray traversal would move about 100 times as many
bytes, reported beside them). cpu_baseline: the reference's own OpenMP code
(oracle/_ref/libsrfref_omp.so, built from the reference sources) on a bounded
sample of the same rays, on this host's cores.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "closest-hit Mrays/s + LBVH build Mprims/s; volume fwd+bwd ms vs plain fwd"
HBM_PEAK_GBS = 8000.0
# Work per ray on the bench rays. k_trace_tet's pass A runs once per
# wavefront and loads each node and tetra once for its 64 rays; it certifies
# almost every ray's hit, and the rest rerun the reference's traversal from
# just above it (pass B, per ray). Loads per ray, counted in an instrumented
# build of the kernel over tet_grid(64) x tet_rays(64, 1024, 1024): pass A
# 60.53 nodes and 28.24 tetras per wavefront = 0.9458 + 0.4413 per ray,
# pass B 0.0588 nodes (its pops reload the parent node) + 0.0187 tetras.
TET_NODES_PER_RAY = 0.9458 + 0.0588
TET_TESTS_PER_RAY = 0.4413 + 0.0187
# The reference's single traversal on the same rays (oracle, stats=True)
# loads 180.0010 nodes and 60.3673 tetras per ray.
TET_REF_NODES_PER_RAY = 180.0010
TET_REF_TESTS_PER_RAY = 60.3673
# The reference's triangle traversal (oracle counts on heightfield(512),
# heightfield_rays(1024, 1024)); k_trace_tri_wave visits exactly these nodes
# and triangles per ray but loads each once per wavefront, so the rate below
# is what a per-ray traversal would have to move, not HBM traffic.
TRI_NODES_PER_RAY = 64.3038
TRI_TESTS_PER_RAY = 94.4435
TET_BYTES_PER_RAY = 40 + 40 * TET_NODES_PER_RAY + 80 * TET_TESTS_PER_RAY
TET_REF_BYTES_PER_RAY = 40 + 40 * TET_REF_NODES_PER_RAY + 80 * TET_REF_TESTS_PER_RAY
# HBM bytes per k_trace_tet launch on the bench rays: rocprofv3 FETCH_SIZE and
# WRITE_SIZE (KiB/dispatch), each scaled by the same counter on a 1 GiB copy
# (fetch x 2.000, write x 1.000); profiles/traffic_k_trace_tet_tet64_k_trace_tri_heightfield512.md
TET_TRACE_HBM_BYTES = (2.000 * 32442.81 + 1.000 * 49152.0) * 1024
TRI_BYTES_PER_RAY = 40 + 64 * TRI_NODES_PER_RAY + 48 * TRI_TESTS_PER_RAY


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def timed_events(torch, fn, reps):
    """Average ms of fn() over reps runs, by HIP events on the current stream."""
    s = torch.cuda.Event(enable_timing=True)
    e = torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(reps):
        fn()
    e.record()
    e.synchronize()
    return s.elapsed_time(e) / reps


def cpu_baseline(V, F, S, rays):
    """Reference OpenMP path on a bounded sample: full Build32 of the tet mesh
    once, then Traverse over every 8th image row (OpenMP dynamic over rows)."""
    try:
        from oracle.oracle import Reference, ref_available
    except ImportError:
        return None
    if not ref_available("omp"):
        return None
    ref = Reference("omp")
    threads = os.cpu_count() or 1
    t0 = time.perf_counter()
    tree = ref.build("tet", V, F, keep=True)
    build_s = time.perf_counter() - t0
    width = 1024
    sample = np.ascontiguousarray(rays.reshape(-1, width)[::8].reshape(-1))
    t0 = time.perf_counter()
    ref.trace(tree, sample, width=width, scalars=S, threads=threads)
    trace_s = time.perf_counter() - t0
    ref.free(tree)
    mrays = len(sample) / trace_s / 1e6
    return {"value": round(mrays, 3), "unit": "Mrays/s", "cores": threads, "kind": "reference",
            "sample": (f"TetraAccel::Traverse over {len(sample)} of the 1024^2 rays (every 8th "
                       f"row), OpenMP dynamic over rows, {threads} threads; Build32 of all "
                       f"{len(F)} tets took {build_s:.2f} s ({len(F) / build_s / 1e6:.2f} Mprims/s)"),
            "build_mprims_per_s": round(len(F) / build_s / 1e6, 3)}


def traffic_probe(torch, S, W, build, trace):
    """Kernels for the FETCH_SIZE / WRITE_SIZE counter passes: the tetra build
    and trace (each once after a warm-up), the heightfield trace, and a plain
    1 GiB device-to-device copy whose counters give the scale factor."""
    build()
    trace()
    torch.cuda.synchronize()
    build()
    trace()
    Vh, Fh = W.heightfield(512)
    tri = S.TriangleAccel()
    tri.Build32(S.device_tensor(Vh), S.device_tensor(Fh))
    rh = S.device_tensor(W.heightfield_rays(1024, 1024))
    hh = torch.empty((len(rh), 4), dtype=torch.float32, device="cuda")
    tri.Traverse(rh, hh)
    src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    log(f"traffic probe done: copy of {src.numel() * 4} bytes")


def batched_scene(torch, S, W):
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
    out = {"meshes": len(meshes), "triangles": ntris, "rays": nr,
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
    # one camera per launch, for comparison
    ms = timed_events(torch, lambda: [scene.Trace(r, hits[:len(r)], ids[:len(r)], mode="tuned")
                                      for r in rays], 2)
    out["trace_per_camera_launches_ms"] = round(ms, 3)
    out["hit_fraction"] = round(float((ids != -1).float().mean().item()), 4)
    del graph, scene, accs
    return out


def particles_ao(torch, S, W):
    """16,777,216 Plummer spheres (lognormal radii): build, 2048x2048 primary
    closest hits, 8 cosine-weighted AO rays per hit, occlusion traced
    unsorted (plain and persistent) and with the autotuned ray sort."""
    log("particles: 16M spheres, 2048^2 primary + 8 AO rays per hit")
    P, R = W.plummer_spheres(1 << 24)
    Pd, Rd = S.device_tensor(P), S.device_tensor(R)
    del P, R
    acc = S.ParticleAccel()
    acc.Build32(Pd, Rd, 0.0)
    torch.cuda.synchronize()
    build_ms = timed_events(torch, lambda: acc.Build32(Pd, Rd, 0.0), 3)
    prim = S.device_tensor(W.sphere_rays(2048, 2048))
    np_ = prim.shape[0]
    hits = torch.empty((np_, 4), dtype=torch.float32, device="cuda")
    acc.Traverse(prim, hits)
    prim_ms = timed_events(torch, lambda: acc.Traverse(prim, hits), 3)
    rays, tmax, pix = S.ao_rays_spheres(prim, hits, Pd, Rd, 0.0, 8, 1e-3, 0.05, 7)
    gen_ms = timed_events(torch, lambda: S.ao_rays_spheres(prim, hits, Pd, Rd, 0.0, 8, 1e-3,
                                                           0.05, 7), 3)
    na = rays.shape[0]
    occ = torch.empty(na, dtype=torch.int32, device="cuda")
    out = {"spheres": 1 << 24, "primary_rays": np_, "ao_rays": na,
           "build_ms": round(build_ms, 3), "build_mprims_per_s": round((1 << 24) / build_ms / 1e3, 1),
           "primary_ms": round(prim_ms, 3), "primary_mrays_per_s": round(np_ / prim_ms / 1e3, 1),
           "primary_hit_fraction": round(na / 8 / np_, 4), "ao_gen_ms": round(gen_ms, 3)}
    for name, bits, mode in (("unsorted_plain", 0, "plain"), ("unsorted_persistent", 0, "persistent")):
        ms = timed_events(torch, lambda: acc.TraverseAnySorted(rays, tmax, bits, mode, occ), 3)
        out[f"ao_{name}_ms"] = round(ms, 3)
        out[f"ao_{name}_mrays_per_s"] = round(na / ms / 1e3, 1)
    cfg = acc.autotune_any(rays, tmax, occ)
    out["autotuned"] = cfg
    mode = "persistent" if cfg["persistent"] else "plain"
    ms = timed_events(torch, lambda: acc.TraverseAnySorted(rays, tmax, cfg["key_bits"], mode, occ), 3)
    out["ao_tuned_ms"] = round(ms, 3)
    out["ao_tuned_mrays_per_s"] = round(na / ms / 1e3, 1)
    out["ao_occluded_fraction"] = round(float(occ.float().mean().item()), 4)
    del acc, rays, tmax, pix, occ, hits, prim, Pd, Rd
    torch.cuda.empty_cache()
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extras", action="store_true", help="tetra step only")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="one launch of each traversal kernel plus a 1 GiB device copy "
                         "(the calibration for HBM-counter runs); prints no JSON line")
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
    log("generating the 64^3 tetra grid")
    V, F, Sc = W.tet_grid(64)
    rays_np = W.tet_rays(64, 1024, 1024)
    n_rays = len(rays_np)
    Vd, Fd, Sd = S.device_tensor(V), S.device_tensor(F), S.device_tensor(Sc)
    rays = S.device_tensor(rays_np)
    hits = torch.empty((n_rays, 4), dtype=torch.float32, device="cuda")
    ext = torch.empty((n_rays, 8), dtype=torch.float32, device="cuda")
    acc = S.TetraAccel()

    def build():
        acc.Build32(Vd, Fd)

    def trace():
        acc.TraverseTetra(rays, Sd, hits, ext)

    def step():
        build()
        trace()

    if args.traffic_probe:
        traffic_probe(torch, S, W, build, trace)
        return

    log(f"warmup {args.warmup}")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    log(f"timing {args.steps} steps")
    t0 = time.perf_counter()
    ms_step = timed_events(torch, step, args.steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_build = timed_events(torch, build, max(3, args.steps // 2))
    ms_trace = timed_events(torch, trace, max(3, args.steps // 2))
    info = acc.get_info()
    tet_gbs = TET_BYTES_PER_RAY * n_rays / (ms_trace * 1e-3) / 1e9

    extras = {}
    if not args.no_extras:
        log("heightfield triangles")
        Vh, Fh = W.heightfield(512)
        tri = S.TriangleAccel()
        Vhd, Fhd = S.device_tensor(Vh), S.device_tensor(Fh)
        rh = S.device_tensor(W.heightfield_rays(1024, 1024))
        hh = torch.empty((n_rays, 4), dtype=torch.float32, device="cuda")
        tri.Build32(Vhd, Fhd)
        tri.Traverse(rh, hh)
        tb = timed_events(torch, lambda: tri.Build32(Vhd, Fhd), 10)
        tt = timed_events(torch, lambda: tri.Traverse(rh, hh), 10)
        extras["heightfield_522242_tris"] = {
            "build_ms": round(tb, 4), "build_mprims_per_s": round(len(Fh) / tb / 1e3, 1),
            "trace_ms": round(tt, 4), "trace_mrays_per_s": round(n_rays / tt / 1e3, 1),
            "per_ray_traversal_equivalent_gbs": round(TRI_BYTES_PER_RAY * n_rays / (tt * 1e-3) / 1e9, 1)}
        del tri
        log("volume 256^3")
        from surface_amd.volume import Volume
        vox = torch.from_numpy(W.gaussian_volume(256)).cuda()
        tf = torch.from_numpy(W.transfer_function(256)).cuda()
        vr = S.device_tensor(W.volume_rays(1024, 1024))
        vol = Volume(vox, (256, 256, 256), (1, 1, 1))
        step_len = 0.5 / 255
        gv = torch.zeros_like(vox)
        gt = torch.zeros_like(tf)
        grad = torch.ones((n_rays, 4), dtype=torch.float32, device="cuda")
        ck = torch.empty((n_rays, vol.checkpoint_slots(step_len)), dtype=torch.float32,
                         device="cuda")
        rgba = vol.march(tf, step_len, vr)
        fwd = timed_events(torch, lambda: vol.march(tf, step_len, vr, rgba=rgba), 3)

        def fwd_bwd():
            vol.march(tf, step_len, vr, ck, rgba)
            vol.march_backward(tf, step_len, vr, grad, gv, gt, ck)

        fwd_bwd()
        fb = timed_events(torch, fwd_bwd, 2)
        extras["volume_256cubed_1M_rays"] = {"fwd_ms": round(fwd, 3), "fwd_bwd_ms": round(fb, 3),
                                             "fwd_bwd_over_fwd": round(fb / fwd, 3)}

    if not args.no_extras:
        extras["batched_scene_64_meshes_16_cameras"] = batched_scene(torch, S, W)
        extras["particles_16M_ao"] = particles_ao(torch, S, W)
        extras["solids_and_lines"] = solids_lines(torch, S, W)

    cpu = None if args.no_cpu else cpu_baseline(V, F, Sc, rays_np)

    value = n_rays / ms_step / 1e3  # Mrays/s over build + trace
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
        "config": {"workload": "tetra grid 64^3 hexes x 6 = 1,572,864 tets; LBVH build + "
                               "1024x1024 pinhole closest-hit rays with InterpolateTetra scalar",
                   "prims": int(len(F)), "rays": int(n_rays), "nodes": int(info["num_nodes"])},
        "trace_mrays_per_s": round(n_rays / ms_trace / 1e3, 3),
        "build_mprims_per_s": round(len(F) / ms_build / 1e3, 3),
        "build_ms": round(ms_build, 4),
        "trace_ms": round(ms_trace, 4),
        "wall_s_timed": round(wall, 3),
        "roofline": {"bound": "hbm", "achieved": round(tet_gbs, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(tet_gbs / HBM_PEAK_GBS, 4),
                     "traffic": round(TET_TRACE_HBM_BYTES), "kernel": "k_trace_tet",
                     "bytes_per_launch": round(TET_BYTES_PER_RAY * n_rays),
                     "reference_traversal_bytes_per_launch": round(TET_REF_BYTES_PER_RAY * n_rays)},
        "cpu_baseline": cpu,
        "extras": extras,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
