"""Write the golden fixtures of tests/golden/ from the reference itself.

Runs the reference's own render/ code (oracle/_ref/libsrfref_serial.so,
built from the reference sources by oracle/Makefile) on small deterministic
inputs and stores inputs + outputs as .npz fixtures; for the full-size
heightfield (522,242 triangles, 1024^2 rays) and tetra grid (1,572,864 tets,
1024^2 rays) it stores SHA-256 digests of the tree, the permutation and the
per-ray hits; --only full_extra adds the same for the 1M-triangle heightfield,
the 64-mesh scene and the 16M-sphere set. Run in the container that has the
reference checkout:

    make -C oracle && python tests/golden/gen_golden.py [--full] [--only scene|sol|lin|full_extra]
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle.oracle import (Reference, canonical_sph, canonical_tet, canonical_tlas,  # noqa: E402
                           canonical_tri)
from surface_amd import workloads as W  # noqa: E402


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def tri_case(ref, grid, res):
    V, F = W.heightfield(grid)
    bmin = V[F.reshape(-1)].min(0)
    bmax = V[F.reshape(-1)].max(0)
    codes = ref.morton("tri", V, F, bmin, bmax)
    sc, si = ref.sort_keys(codes)
    info = ref.radix_tree(sc)
    tree = ref.build("tri", V, F, keep=True)
    rays = W.heightfield_rays(res, res)
    hits, _ = ref.trace(tree, rays)
    ref.free(tree)
    return dict(V=V, F=F, bmin=bmin, bmax=bmax, codes=codes, sorted_codes=sc,
                sorted_index=si, nodeinfo=info, nodes=canonical_tri(tree["nodes"]),
                indices=tree["indices"], bounds=tree["bounds"], rays=rays, hits=hits)


def tet_case(ref, cells, res):
    V, F, S = W.tet_grid(cells)
    tree = ref.build("tet", V, F, keep=True)
    rays = W.tet_rays(cells, res, res)
    hits, ext = ref.trace(tree, rays, scalars=S)
    ref.free(tree)
    return dict(V=V, F=F, S=S, nodes=canonical_tet(tree["nodes"]), indices=tree["indices"],
                bounds=tree["bounds"], rays=rays, hits=hits, ext_d=ext["d"],
                ext_scalar=ext["scalar"], ext_subface=ext["subface"])


def sph_case(ref, n, res):
    P, R = W.plummer_spheres(n)
    tree = ref.build("sph", P, R=R, keep=True)
    rays = W.sphere_rays(res, res)
    hits, _ = ref.trace(tree, rays)
    ref.free(tree)
    codes_simd = ref.morton("pts", P, None, P.min(0) - 1e-3, P.max(0) + 1e-3, simd=True)
    codes_scalar = ref.morton("pts", P, None, P.min(0) - 1e-3, P.max(0) + 1e-3, simd=False)
    assert np.array_equal(codes_simd, codes_scalar)
    return dict(P=P, R=R, nodes=canonical_sph(tree["nodes"]), indices=tree["indices"],
                bounds=tree["bounds"], rays=rays, hits=hits)


def tex_case(ref):
    rng = np.random.Generator(np.random.PCG64(W.BASE_SEED + 9))
    dims = np.array([9, 7, 5], np.int32)
    vox = rng.uniform(-1, 1, int(np.prod(dims))).astype(np.float32)
    uvw = rng.uniform(-1.5, 2.5, (4000, 3)).astype(np.float32)
    edge = np.array([[0, 0, 0], [1, 1, 1], [0.5, 0.5, 0.5], [-0.0, 1.0, 0.25],
                     [-0.01, 1.01, 0.999], [1e-8, -1e-8, 2.0]], np.float32)
    uvw = np.concatenate([uvw, edge])
    out = {"dims": dims, "vox": vox, "uvw": uvw}
    for name, clamp in (("repeat", (0, 0, 0)), ("clamp", (1, 1, 1)), ("mixed", (1, 0, 1))):
        out[f"out_{name}"] = ref.tex3d(vox, dims, np.array(clamp, np.int32), uvw)
    return out


def karras_dup_case(ref):
    rng = np.random.Generator(np.random.PCG64(W.BASE_SEED + 10))
    codes = np.sort(rng.integers(0, 64, 5000).astype(np.uint32))
    same = np.full(3000, 12345, np.uint32)
    return {"codes_dup": codes, "info_dup": ref.radix_tree(codes),
            "codes_same": same, "info_same": ref.radix_tree(same)}


# solid grids of 1331 hexahedra, 1458 prisms and 1296 pyramids (Build32
# needs 1024 or more), seen from outside and from rays starting inside
SOL_CELLS = {8: 11, 6: 9, 5: 6}


def solid_rays_mixed(cells, res, n_inside, seed):
    rays = W.solid_rays(cells, res, res)
    rng = np.random.Generator(np.random.PCG64(seed))
    r2 = np.zeros(n_inside, W.RAY_DTYPE)
    r2["org"] = rng.uniform(0.3, cells - 0.3, (n_inside, 3)).astype(np.float32)
    d = rng.normal(size=(n_inside, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r2["dir"] = d.astype(np.float32)
    r2["flags"] = 1
    r2["prev_prim_id"] = W.NO_PRIM
    return np.concatenate([rays, r2])


def sol_case(ref):
    """SolidAccel::Build32 + Traverse for each solid kind. A solid hit's u/v
    are uninitialised in the reference and are not stored."""
    out = {}
    for nv, cells in SOL_CELLS.items():
        V, F, S = W.solid_grid(cells, nv)
        tree = ref.build("sol", V, F, keep=True, nv=nv)
        rays = solid_rays_mixed(cells, 32, 500, W.BASE_SEED + 30 + nv)
        hits, ext = ref.trace(tree, rays, scalars=S)
        ref.free(tree)
        p = f"n{nv}_"
        out.update({p + "V": V, p + "F": F, p + "S": S, p + "nodes": canonical_tet(tree["nodes"]),
                    p + "indices": tree["indices"], p + "bounds": tree["bounds"],
                    p + "rays": rays, p + "t": hits["t"], p + "prim_id": hits["prim_id"],
                    p + "d": ext["d"], p + "pos": ext["pos"], p + "nrm": ext["nrm"],
                    p + "scalar": ext["scalar"]})
    return out


def lin_case(ref):
    """LineAccel::Build + Traverse on random-walk polylines: constant radius
    with and without end caps, and per-vertex radii (where the reference's
    boxes can hide a closer crossing, see tests/test_oracle.py)."""
    rays = W.line_rays(48, 48)
    P, S, R = W.polylines(300, 30, radius="vary", seed=W.BASE_SEED + 40)
    out = {"rays": rays, "P": P, "S": S, "R": R}
    for name, radius, rconst, cap in (("const", None, 0.004, True), ("nocap", None, 0.004, False),
                                      ("vary", R, 0.0, True)):
        tree = ref.build("lin", P, S, R=radius, rconst=rconst, keep=True, cap=cap)
        hits, ext = ref.trace(tree, rays)
        ref.free(tree)
        out.update({f"{name}_hits": hits, f"{name}_pos": ext["pos"], f"{name}_nrm": ext["nrm"]})
    return out


SCENE_ARGS = dict(count=12, lo=1100, hi=6000, seed=W.BASE_SEED + 21)


def scene_case(ref):
    """RenderGraph of 12 triangle elements (batch_scene(**SCENE_ARGS)): the
    reference's inverse matrices, world boxes, top-level BVH, and the hits of
    two orbit cameras (64x64)."""
    meshes, T = W.batch_scene(**SCENE_ARGS)
    trees = [ref.build("tri", V, F, keep=True) for V, F in meshes]
    sc = ref.scene_build(trees, T)
    cams = W.scene_cameras(T, count=2, width=64, height=64)
    rays = np.concatenate(cams)
    hits, inst = ref.scene_trace(sc, rays)
    ref.scene_free(sc)
    for t in trees:
        ref.free(t)
    return dict(T=T, invT=sc["invT"], invT33=sc["invT33"], xbox=sc["xbox"],
                tlas=canonical_tlas(sc["tlas"]), rays=rays, hits=hits, inst=inst,
                meshes_sha=np.array(sha(*[a for m in meshes for a in m])))


def row_sample(rays, width, every):
    """Every `every`-th chunk of `width` consecutive rays."""
    return np.ascontiguousarray(rays.reshape(-1, width)[::every].reshape(-1))


def full_extra(ref, digests, t0):
    """Digests at full size for the configurations whose parity was pinned
    only at reduced size: the heightfield of 725^2 vertices (1,048,352
    triangles, 1024^2 rays), the 64-mesh scene (every mesh's tree and
    permutation; RenderGraph hits of every 16th ray row of the 16 cameras)
    and the 16,777,216 Plummer spheres (tree, permutation, primary hits of
    every 16th row of the 2048^2 rays, and occlusion of numpy AO rays from
    every 4th of those hits, 4 per hit: closest hit t < t_max)."""
    V, F = W.heightfield(725)
    tree = ref.build("tri", V, F, keep=True)
    hits, _ = ref.trace(tree, W.heightfield_rays(1024, 1024))
    ref.free(tree)
    digests["heightfield725"] = {
        "num_tris": int(len(F)), "num_nodes": int(len(tree["nodes"])),
        "nodes": sha(canonical_tri(tree["nodes"])), "indices": sha(tree["indices"]),
        "bounds": [float(x) for x in tree["bounds"]], "hits": sha(hits),
        "hit_fraction": float((hits["prim_id"] != W.NO_PRIM).mean())}
    print("heightfield725 done", time.time() - t0, flush=True)

    meshes, T = W.batch_scene(64)
    trees = [ref.build("tri", Vm, Fm, keep=True) for Vm, Fm in meshes]
    sc = ref.scene_build(trees, T)
    cams = W.scene_cameras(T, count=16, width=1024, height=1024)
    rays = np.concatenate([row_sample(c, 1024, 16) for c in cams])
    hits, inst = ref.scene_trace(sc, rays, width=1024)
    digests["scene64"] = {
        "meshes": len(meshes), "triangles": int(sum(len(Fm) for _, Fm in meshes)),
        "mesh_nodes": [sha(canonical_tri(t["nodes"])) for t in trees],
        "mesh_indices": [sha(t["indices"]) for t in trees],
        "tlas": sha(canonical_tlas(sc["tlas"])), "rays": int(len(rays)),
        "hits": sha(hits, inst), "hit_fraction": float((inst != W.NO_PRIM).mean())}
    ref.scene_free(sc)
    for t in trees:
        ref.free(t)
    print("scene64 done", time.time() - t0, flush=True)

    P, R = W.plummer_spheres(1 << 24)
    tree = ref.build("sph", P, R=R, keep=True)
    prim = row_sample(W.sphere_rays(2048, 2048), 2048, 16)
    hits, _ = ref.trace(tree, prim, width=2048)
    ao, tmax = W.ao_rays_from_hits(prim, hits, P, R, per_hit=4, every=4)
    aoh, _ = ref.trace(tree, ao, width=1024)
    occ = ((aoh["prim_id"] != W.NO_PRIM) & (aoh["t"] < tmax)).astype(np.int32)
    ref.free(tree)
    digests["plummer16M"] = {
        "num_spheres": int(len(P)), "num_nodes": int(len(tree["nodes"])),
        "nodes": sha(canonical_sph(tree["nodes"])), "indices": sha(tree["indices"]),
        "bounds": [float(x) for x in tree["bounds"]], "primary_rays": int(len(prim)),
        "primary_prim_t": sha(hits["prim_id"], hits["t"]),
        "hit_fraction": float((hits["prim_id"] != W.NO_PRIM).mean()),
        "ao_rays": int(len(ao)), "ao_rays_sha": sha(ao, tmax), "ao_occluded": sha(occ),
        "ao_occluded_fraction": float(occ.mean())}
    print("plummer16M done", time.time() - t0, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--full", action="store_true", help="also write full-size digests")
    ap.add_argument("--only", choices=["scene", "sol", "lin", "full_extra"],
                    help="write one fixture only (full_extra: the digests of the 1M-triangle "
                         "heightfield, the 64-mesh scene and the 16M-sphere set, merged into "
                         "full_digests.json)")
    args = ap.parse_args()
    ref = Reference("serial")
    t0 = time.time()
    if args.only == "full_extra":
        path = os.path.join(HERE, "full_digests.json")
        digests = json.load(open(path))
        full_extra(ref, digests, t0)
        with open(path, "w") as f:
            json.dump(digests, f, indent=1)
        return
    if args.only == "scene":
        np.savez_compressed(os.path.join(HERE, "scene_small.npz"), **scene_case(ref))
        return
    if args.only == "sol":
        np.savez_compressed(os.path.join(HERE, "sol_small.npz"), **sol_case(ref))
        return
    if args.only == "lin":
        np.savez_compressed(os.path.join(HERE, "lin_small.npz"), **lin_case(ref))
        return
    np.savez_compressed(os.path.join(HERE, "tri_small.npz"), **tri_case(ref, 48, 64))
    np.savez_compressed(os.path.join(HERE, "tet_small.npz"), **tet_case(ref, 6, 64))
    np.savez_compressed(os.path.join(HERE, "sph_small.npz"), **sph_case(ref, 6000, 64))
    np.savez_compressed(os.path.join(HERE, "tex3d.npz"), **tex_case(ref))
    np.savez_compressed(os.path.join(HERE, "karras.npz"), **karras_dup_case(ref))
    np.savez_compressed(os.path.join(HERE, "scene_small.npz"), **scene_case(ref))
    np.savez_compressed(os.path.join(HERE, "sol_small.npz"), **sol_case(ref))
    np.savez_compressed(os.path.join(HERE, "lin_small.npz"), **lin_case(ref))
    print(f"small fixtures written in {time.time() - t0:.1f}s", flush=True)
    if not args.full:
        return
    digests = {}
    V, F = W.heightfield(512)
    tree = ref.build("tri", V, F, keep=True)
    rays = W.heightfield_rays(1024, 1024)
    hits, _ = ref.trace(tree, rays)
    ref.free(tree)
    digests["heightfield512"] = {
        "num_tris": int(len(F)), "num_nodes": int(len(tree["nodes"])),
        "nodes": sha(canonical_tri(tree["nodes"])), "indices": sha(tree["indices"]),
        "bounds": [float(x) for x in tree["bounds"]], "hits": sha(hits),
        "hit_fraction": float((hits["prim_id"] != W.NO_PRIM).mean())}
    print("heightfield512 done", time.time() - t0, flush=True)
    V, F, S = W.tet_grid(64)
    tree = ref.build("tet", V, F, keep=True)
    rays = W.tet_rays(64, 1024, 1024)
    hits, ext = ref.trace(tree, rays, scalars=S)
    ref.free(tree)
    digests["tetgrid64"] = {
        "num_tets": int(len(F)), "num_nodes": int(len(tree["nodes"])),
        "nodes": sha(canonical_tet(tree["nodes"])), "indices": sha(tree["indices"]),
        "bounds": [float(x) for x in tree["bounds"]], "hits": sha(hits),
        "ext": sha(ext["d"], ext["scalar"], ext["subface"]),
        "hit_fraction": float((hits["prim_id"] != W.NO_PRIM).mean())}
    print("tetgrid64 done", time.time() - t0, flush=True)
    with open(os.path.join(HERE, "full_digests.json"), "w") as f:
        json.dump(digests, f, indent=1)


if __name__ == "__main__":
    main()
