"""Texture::Fetch on the GPU (srf_texture3d_create / _remap / _fetch) against
the oracle, bit for bit: FilterTexture3DByte (texel / 255.5f),
FilterTexture3DFloat and FilterTexture3DDouble with 1-4 interleaved
components, REPEAT / CLAMP_TO_EDGE / mixed wraps, and the per-axis remap
tables of gles/gles_common.h:263-269. The oracle's fetch is pinned against
the reference's own filters in tests/test_oracle.py; the FP32 single-
component case is also checked against the reference-written fixture."""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle.oracle import Oracle  # noqa: E402

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
def orc():
    return Oracle()


def _image(rng, fmt, n):
    if fmt == "byte":
        return rng.integers(0, 256, n).astype(np.uint8)
    if fmt == "float":
        return rng.uniform(-1, 1, n).astype(np.float32)
    return rng.uniform(-1, 1, n).astype(np.float64)


@pytest.mark.parametrize("fmt", ["byte", "float", "double"])
@pytest.mark.parametrize("comps", [1, 2, 3, 4])
def test_texture_fetch_matches_oracle(torch, orc, fmt, comps):
    say(f"texture fetch {fmt} x{comps}")
    from surface_amd.volume import Texture3D
    rng = np.random.default_rng(7 * comps + len(fmt))
    dims = np.array([33, 17, 9], np.int32)
    img = _image(rng, fmt, int(np.prod(dims)) * comps)
    uvw = rng.uniform(-1.5, 2.5, (20000, 3)).astype(np.float32)
    uvw[:6] = [[0, 0, 0], [1, 1, 1], [0.5, 0.5, 0.5], [-0.0, 1.0, 0.25], [-0.01, 1.01, 0.999],
               [1e-8, -1e-8, 2.0]]
    ud = torch.from_numpy(uvw).cuda()
    for clamp in ((0, 0, 0), (1, 1, 1), (1, 0, 1)):
        tex = Texture3D(torch.from_numpy(img).cuda(), dims, comps, clamp)
        got = tex.fetch(ud).cpu().numpy()
        want = orc.tex3d_fetch(img, comps, dims, clamp, uvw)
        assert got.tobytes() == want.tobytes(), clamp
    # remap tables on every axis
    tabs = [np.sort(rng.uniform(0, 1, m)).astype(np.float32) for m in (40, 3, 129)]
    tex = Texture3D(torch.from_numpy(img).cuda(), dims, comps, (0, 1, 0))
    for k in range(3):
        tex.set_remap(k, torch.from_numpy(tabs[k]).cuda())
    got = tex.fetch(ud).cpu().numpy()
    want = orc.tex3d_fetch(img, comps, dims, (0, 1, 0), uvw, remap=tabs)
    assert got.tobytes() == want.tobytes()
    tex.set_remap(1, None)
    got = tex.fetch(ud).cpu().numpy()
    want = orc.tex3d_fetch(img, comps, dims, (0, 1, 0), uvw, remap=(tabs[0], None, tabs[2]))
    assert got.tobytes() == want.tobytes()


def test_texture_fetch_matches_reference_fixture(torch):
    from surface_amd.volume import Texture3D
    g = np.load(os.path.join(GOLD, "tex3d.npz"), allow_pickle=False)
    uvw = torch.from_numpy(g["uvw"]).cuda()
    for name, clamp in (("repeat", (0, 0, 0)), ("clamp", (1, 1, 1)), ("mixed", (1, 0, 1))):
        tex = Texture3D(torch.from_numpy(g["vox"]).cuda(), tuple(int(x) for x in g["dims"]), 1,
                        clamp)
        out = tex.fetch(uvw).cpu().numpy()
        assert out[:, 0].tobytes() == g[f"out_{name}"].tobytes(), name
        assert np.array_equal(out[:, 0], out[:, 3])


def test_march_rejects_non_float_textures(torch):
    from surface_amd import _lib
    from surface_amd.accel import SrfError
    from surface_amd.volume import Texture3D
    img = torch.zeros(8 * 8 * 8, dtype=torch.uint8, device="cuda")
    tex = Texture3D(img, (8, 8, 8), 1)
    uvw = torch.zeros((4, 3), dtype=torch.float32, device="cuda")
    with pytest.raises(SrfError):
        _lib.check(tex.lib.srf_volume_sample(tex.handle, uvw.data_ptr(), 4, uvw.data_ptr(), None))


@pytest.mark.parametrize("dtype", [np.uint8, np.float32, np.float64])
@pytest.mark.parametrize("comps,overlap", [(1, False), (3, True), (4, True)])
def test_sparse_volume_matches_oracle(torch, orc, dtype, comps, overlap):
    """SparseVolumeAccel::Sample and Texture::Fetch's sparse branch on the
    GPU against the oracle (pinned against the reference in
    tests/test_oracle.py), bit for bit."""
    say(f"sparse volume {np.dtype(dtype).name} x{comps} overlap={overlap}")
    from surface_amd import workloads as W
    from surface_amd.volume import SparseVolume
    blocks, data, gdim = W.sparse_blocks(grid=(6, 5, 4), components=comps, dtype=dtype,
                                         overlap=overlap, seed=W.BASE_SEED + 20 + comps)
    sv = SparseVolume(blocks, [torch.from_numpy(d).cuda() for d in data], comps, gdim)
    rng = np.random.default_rng(comps + 1)
    pos = rng.uniform(-4, gdim + 4, (40000, 3))
    pos[:len(blocks)] = blocks[:, :3]
    pos[len(blocks):2 * len(blocks)] = blocks[:, :3] + blocks[:, 3:6]
    got = sv.sample(torch.from_numpy(pos).cuda()).cpu().numpy()
    want = orc.sparse_sample(blocks, data, comps, pos=pos)
    assert got.tobytes() == want.tobytes()
    assert 0.3 < (got[:, 3] != 0).mean() < 1.0
    uvw = rng.uniform(-0.1, 1.1, (20000, 3)).astype(np.float32)
    got = sv.fetch(torch.from_numpy(uvw).cuda()).cpu().numpy()
    want = orc.sparse_sample(blocks, data, comps, uvw=uvw, global_dim=gdim).astype(np.float32)
    assert got.tobytes() == want.tobytes()
