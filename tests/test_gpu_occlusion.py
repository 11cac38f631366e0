"""Ray-sorted and persistent occlusion (srf_trace_any_sorted) and the
ambient-occlusion ray generator (srf_ao_rays_spheres) on the GPU.

Sorting only changes the order rays are traced in: every flag must equal
the unsorted any-hit's and the oracle's closest hit t < t_max, bit for bit.
The generated AO rays are checked against a numpy restatement of the
generator (PCG32 streams, cosine-weighted directions about the normal) to
float tolerance (the GPU's cosf/sinf/sqrtf need not match numpy's to the
last bit), and their occlusion against the oracle exactly."""
from __future__ import annotations

import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle.oracle import Oracle  # noqa: E402
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


def _pcg32_uniforms(seed, seq, count):
    """pcg32_srandom(seed, seq) then `count` pcg32_random() >> 8 * 2^-24."""
    M = (1 << 64) - 1
    state, inc = 0, ((seq << 1) | 1) & M

    def nxt():
        nonlocal state
        old = state
        state = (old * 6364136223846793005 + inc) & M
        xs = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        return ((xs >> rot) | (xs << ((-rot) & 31))) & 0xFFFFFFFF

    nxt()
    state = (state + seed) & M
    nxt()
    return [(nxt() >> 8) * (1.0 / 16777216.0) for _ in range(count)]


@pytest.mark.parametrize("kind", ["sph", "tri"])
def test_sorted_and_persistent_any_hit_equal_plain(S, kind):
    import torch
    if kind == "sph":
        P, R = W.plummer_spheres(200_000, seed=W.BASE_SEED + 50, mu=-5.0)
        acc = S.ParticleAccel()
        acc.Build32(S.device_tensor(P), S.device_tensor(R), 0.0)
        rays = W.sphere_rays(256, 256)
    else:
        V, F = W.heightfield(128)
        acc = S.TriangleAccel()
        acc.Build32(S.device_tensor(V), S.device_tensor(F))
        rays = W.heightfield_rays(256, 256)
    # shuffle so that sorting has something to do
    rays = rays[np.random.default_rng(3).permutation(len(rays))]
    rd = S.device_tensor(rays)
    # the sphere camera sits 6 units from the core: t_max 6 cuts the hits
    tmax = torch.full((len(rays),), 6.0, dtype=torch.float32, device="cuda")
    want = acc.TraverseAny(rd, tmax).cpu().numpy()
    assert 0.05 < want.mean() < 0.95
    for bits in (0, 9, 21, 30):
        for mode in ("plain", "persistent"):
            got = acc.TraverseAnySorted(rd, tmax, key_bits=bits, mode=mode).cpu().numpy()
            assert np.array_equal(got, want), (bits, mode)
    cfg = acc.autotune_any(rd, tmax)
    say(f"{kind}: autotuned {cfg}")
    got = acc.TraverseAnySorted(rd, tmax, key_bits=cfg["key_bits"],
                                mode="persistent" if cfg["persistent"] else "plain")
    assert np.array_equal(got.cpu().numpy(), want)


def test_ao_rays_match_restatement_and_oracle(S, orc):
    import torch
    P, R = W.plummer_spheres(100_000, seed=W.BASE_SEED + 51)
    acc = S.ParticleAccel()
    Pd, Rd = S.device_tensor(P), S.device_tensor(R)
    acc.Build32(Pd, Rd, 0.0)
    prim = W.sphere_rays(96, 96)
    pd = S.device_tensor(prim)
    hits_d = acc.Traverse(pd)
    hits = S.hits_to_numpy(hits_d)
    nh = int((hits["prim_id"] != W.NO_PRIM).sum())
    say(f"AO: {nh} primary hits")
    assert nh > 100
    per_hit, eps, dist, seed = 8, 1e-3, 0.05, 1234
    rays_d, tmax_d, pix_d = S.ao_rays_spheres(pd, hits_d, Pd, Rd, 0.0, per_hit, eps, dist, seed)
    rays = rays_d.cpu().numpy().view(W.RAY_DTYPE).reshape(-1)
    pix = pix_d.cpu().numpy()
    assert len(rays) == nh * per_hit
    assert np.all(tmax_d.cpu().numpy() == np.float32(dist))
    # every hit pixel appears exactly per_hit times, in a contiguous group
    assert np.array_equal(np.sort(np.unique(pix)), np.nonzero(hits["prim_id"] != W.NO_PRIM)[0])
    g = pix.reshape(-1, per_hit)
    assert np.all(g == g[:, :1])
    # restatement of the generator for a sample of groups
    for grp in np.random.default_rng(4).choice(len(g), 40, replace=False):
        i = int(g[grp, 0])
        c = hits["prim_id"][i]
        assert np.all(rays["prev_prim_id"][grp * per_hit:(grp + 1) * per_hit] == c)
        # hit point and normal in float32 as the generator forms them (the
        # normal of a small sphere is sensitive to the rounding of p)
        f32 = np.float32
        o = prim["org"][i] + hits["t"][i] * prim["dir"][i]
        nf = o - P[c]
        lf = np.sqrt(f32(nf[0] * nf[0] + nf[1] * nf[1]) + nf[2] * nf[2])
        n = (nf / lf).astype(np.float64)
        org = o.astype(np.float64) + eps * float(R[c]) * n
        sg = 1.0 if n[2] >= 0 else -1.0
        a = -1.0 / (sg + n[2])
        b = n[0] * n[1] * a
        b1 = np.array([1.0 + sg * n[0] * n[0] * a, sg * b, -sg * n[0]])
        b2 = np.array([b, sg + n[1] * n[1] * a, -n[1]])
        u = _pcg32_uniforms(seed, i, 2 * per_hit)
        for k in range(per_hit):
            r = rays[grp * per_hit + k]
            u1, u2 = u[2 * k], u[2 * k + 1]
            rr, phi = np.sqrt(u1), 2 * np.pi * u2
            d = rr * np.cos(phi) * b1 + rr * np.sin(phi) * b2 + np.sqrt(max(0.0, 1 - u1)) * n
            assert np.allclose(r["org"], org, rtol=0, atol=2e-6)
            assert np.allclose(r["dir"], d, rtol=0, atol=2e-5)
            assert np.dot(r["dir"], n) >= -1e-6
    # occlusion of the generated rays: GPU any-hit == oracle closest t < dist
    tree = orc.build_sph(P, R, 0.0)
    oh = orc.trace_sph(tree, P, R, 0.0, rays)
    want = ((oh["prim_id"] != W.NO_PRIM) & (oh["t"] < np.float32(dist))).astype(np.int32)
    got = acc.TraverseAny(rays_d, tmax_d).cpu().numpy()
    assert np.array_equal(got, want)
    assert 0.0 < want.mean() < 1.0
