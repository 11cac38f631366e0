"""The culling bound of the secondary sphere trees, checked on the CPU.

A big sphere leaf is searched through a secondary tree whose boxes are grown
by 1.25e-3 of the farthest distance from the origin to the tree's root boxes
(sub_pad in surface_amd/csrc/srf_trace.hip), which is at least the farthest
distance to any box below them. That is exact only if every t the float
sphere test reports lies no nearer than the entry into the ball of radius
r + 1.25e-3 D around the centre (D >= |o - c| + r, a lower bound of that farthest
distance). Here SphereIsect (render/render_accel_particle.cc:2423-2512) is
restated in numpy float32 / float64 with the kernel's evaluation order, and the
smallest padding each reported hit needs is measured over adversarial rays:
tiny spheres seen from far away, grazing rays, origins just outside the
surface, unnormalised and axis-parallel directions.
"""
from __future__ import annotations

import numpy as np

f32, f64 = np.float32, np.float64
PAD = 1.25e-3
# the rounding-error bound of the sphere test (sub_pad's comment): a reported
# root is never nearer than the entry into the ball of radius r + 9.2e-4 |o - c|
BOUND = 9.2e-4


def sphere_test(o, d, c, r):
    """Reported (hit, t) of the sphere test for rays o + t d and spheres
    (c, r), every operation in the kernel's precision and order."""
    with np.errstate(all="ignore"):
        oc = (o - c).astype(f32)
        dx, dy, dz = d[:, 0], d[:, 1], d[:, 2]
        ox, oy, oz = oc[:, 0], oc[:, 1], oc[:, 2]
        a = (dx * dx + dy * dy) + dz * dz
        b = (2.0 * ((dx * ox + dy * oy) + dz * oz).astype(f64)).astype(f32)
        cc = ((ox * ox + oy * oy) + oz * oz) - r * r
        disc = ((b * b).astype(f64) - (4.0 * a.astype(f64)) * cc.astype(f64)).astype(f32)
        dsq = np.sqrt(np.maximum(disc, f32(0)))
        q = np.where(b < 0, ((-b - dsq).astype(f64) / 2.0).astype(f32),
                     ((-b + dsq).astype(f64) / 2.0).astype(f32))
        t0, t1 = q / a, cc / q
        t00 = (-0.5 * (b / a).astype(f64)).astype(f32)
        t0 = np.where(disc == 0, t00, t0)
        t1 = np.where(disc == 0, t00, t1)
        lo, hi = np.minimum(t0, t1), np.maximum(t0, t1)
        hit = (disc >= 0) & (hi >= 0)
        t = np.where(lo < 0, hi, lo)
    return hit, t


def needed_pad(o, d, c, r, t):
    """Smallest k with t >= entry of the ray into ball(c, r + k D), D = |o-c| + r."""
    o, d, c, r, t = (x.astype(f64) for x in (o, d, c, r, t))
    dd = np.sqrt((d * d).sum(1))
    oc = c - o
    T = (oc * d).sum(1) / (dd * dd)
    dist = np.sqrt(np.maximum((oc * oc).sum(1) - (T * dd) ** 2, 0.0))
    D = np.sqrt((oc * oc).sum(1)) + r
    back = np.maximum(T - t, 0.0) * dd
    need = np.sqrt(dist * dist + back * back)
    return np.maximum(need - r, 0.0) / D


def adversarial(rng, n):
    c = rng.uniform(-1, 1, (n, 3))
    r = 10.0 ** rng.uniform(-6, 0, n)
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    dist_o = r * (1.0 + 10.0 ** rng.uniform(-6, 6, n))
    o = c + u * dist_o[:, None]
    # aim at a point near the surface (grazing) or near the centre
    w = rng.normal(size=(n, 3))
    w -= (w * u).sum(1, keepdims=True) * u
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    off = r * np.where(rng.uniform(size=n) < 0.7, 1.0 + rng.normal(scale=1e-3, size=n),
                       rng.uniform(0, 1, n))
    aim = c + w * off[:, None]
    d = (aim - o) * (10.0 ** rng.uniform(-3, 3, n))[:, None]
    # some axis-parallel components
    z = rng.uniform(size=n) < 0.05
    d[z, rng.integers(0, 3, z.sum())] = 0.0
    return o.astype(f32), d.astype(f32), c.astype(f32), r.astype(f32)


def test_reported_hits_lie_inside_the_padded_ball():
    rng = np.random.default_rng(20261202)
    worst, hits = 0.0, 0
    for _ in range(8):
        o, d, c, r = adversarial(rng, 250_000)
        hit, t = sphere_test(o, d, c, r)
        keep = hit & np.isfinite(t) & (np.abs(d).sum(1) > 0)
        k = needed_pad(o[keep], d[keep], c[keep], r[keep], t[keep])
        hits += int(keep.sum())
        worst = max(worst, float(k.max()))
    assert hits > 500_000
    # the measured worst case stays inside the analytical bound, which the
    # kernel's padding exceeds
    assert worst < BOUND, worst
    assert BOUND * 1.25 < PAD * 1.01


def test_restatement_reports_textbook_roots():
    # unit sphere at the origin, ray from -5 x along +x: entry at t = 4
    o = np.array([[-5.0, 0, 0]], f32)
    d = np.array([[1.0, 0, 0]], f32)
    c = np.zeros((1, 3), f32)
    r = np.ones(1, f32)
    hit, t = sphere_test(o, d, c, r)
    assert hit[0] and t[0] == f32(4.0)
    # origin inside: the exit distance
    hit, t = sphere_test(np.zeros((1, 3), f32), d, c, r)
    assert hit[0] and t[0] == f32(1.0)
    # a miss
    hit, _ = sphere_test(np.array([[-5.0, 2.0, 0]], f32), d, c, r)
    assert not hit[0]
