"""srf_accel_dump against the reference's own Dump files.

{Triangle,Tetra,Particle,Solid}Accel::Dump (render/render_accel_triangle.cc:
1692-1719, render/render_accel_tetra.cc:1900-1928,
render/render_accel_particle.cc:1868-1896, render/render_accel_solid.cc:
2126-2154) write an unsigned long long node
count, the nodes, an unsigned long long index count and the indices. The
reference build (oracle/_ref, compiled from the reference's sources) dumps its
tree; the GPU build dumps the same meshes; the files must carry the same
counts and, once the fields the reference leaves uninitialised in leaves are
zeroed in both, the same bytes. The sphere case has depth-capped leaves with
secondary trees, whose links must not appear in the file.
"""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle.oracle import (SPH_NODE, TET_NODE, TRI_NODE, Reference, canonical_sph,  # noqa: E402
                           canonical_tet, canonical_tri, ref_available)
from surface_amd import workloads as W  # noqa: E402


def read_dump(path, dtype):
    raw = np.fromfile(path, np.uint8)
    nn = int(raw[:8].view(np.uint64)[0])
    off = 8 + nn * dtype.itemsize
    nodes = raw[8:off].view(dtype).copy()
    ni = int(raw[off:off + 8].view(np.uint64)[0])
    idx = raw[off + 8:off + 8 + 4 * ni].view(np.uint32).copy()
    assert len(raw) == off + 8 + 4 * ni
    return nodes, idx


@pytest.fixture(scope="module")
def S():
    import torch
    assert torch.cuda.is_available(), "no GPU visible"
    from surface_amd import accel
    return accel


CASES = {
    "tri": (lambda: W.heightfield(64), TRI_NODE, canonical_tri),
    "tet": (lambda: W.tet_grid(8)[:2], TET_NODE, canonical_tet),
    "sph": (lambda: W.clustered_spheres(n_bg=4000, clusters=3, per_cluster=500)[:2], SPH_NODE,
            canonical_sph),
    "sol": (lambda: W.solid_grid(12, 6)[:2], TET_NODE, canonical_tet),
}


@pytest.mark.parametrize("kind", ["tri", "tet", "sph", "sol"])
def test_dump_matches_reference_dump(S, kind, tmp_path):
    if not ref_available("serial"):
        pytest.skip("oracle/_ref not built (the reference checkout was absent at build time)")
    print(f"dump {kind}", flush=True)
    sys.stdout.flush()
    make, dtype, canon = CASES[kind]
    A, B = make()
    ref = Reference("serial")
    if kind == "sph":
        want = ref.build("sph", A, R=B)
        a = S.ParticleAccel()
        a.Build32(S.device_tensor(A), S.device_tensor(B), 0.0)
        counts = want["nodes"]["data"][want["nodes"]["flag"] == 1][:, 0]
        assert counts.max() > 64  # secondary trees exist
    elif kind == "sol":
        want = ref.build("sol", A, B, nv=6)
        a = S.SolidAccel()
        a.Build32(S.device_tensor(A), S.device_tensor(B), 6)
    else:
        want = ref.build(kind, A, B)
        a = S.TriangleAccel() if kind == "tri" else S.TetraAccel()
        a.Build32(S.device_tensor(A), S.device_tensor(B))
    path = os.path.join(tmp_path, f"{kind}.bvh")
    a.Dump(path)
    nodes, idx = read_dump(path, dtype)
    assert len(nodes) == len(want["nodes"]) and len(idx) == len(want["indices"])
    assert np.array_equal(idx, want["indices"])
    assert canon(nodes).tobytes() == canon(want["nodes"]).tobytes()
    # the fields the reference leaves uninitialised are written as zero here
    assert canon(nodes).tobytes() == nodes.tobytes()
