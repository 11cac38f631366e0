"""The early rejection in the sphere test, checked on the CPU.

sphere_isect (surface_amd/csrc/srf_trace.hip) rejects a sphere before
forming its exact roots when both estimated roots lie beyond
tcut * (1 + 2e-5): q is computed exactly as the exact path computes it, and
q / a and cc / q are estimated with hardware reciprocals (v_rcp_f32, within
1 ulp). Here SphereIsect (render/render_accel_particle.cc:2423-2512) is
restated in numpy float32 / float64 in the kernel's order, the reciprocals
are perturbed by one ulp either way, and no sphere whose reported t is at
most tcut may be rejected, for tcut at and around the reported t (and far
below it, where rejections must happen). The same check with q formed from a
square root one ulp off (an earlier form of the estimate) finds wrong
rejections: the reference's q formula cancels for b < 0, so q must be exact.
"""
from __future__ import annotations

import numpy as np

f32, f64 = np.float32, np.float64
TINY = np.finfo(f32).tiny


def _ulp(x, k):
    x = x.astype(f32)
    for _ in range(abs(k)):
        x = np.nextafter(x, f32(np.inf) if k > 0 else f32(-np.inf)).astype(f32)
    return x


def _cases(rng, n, mode):
    c = rng.normal(size=(n, 3)).astype(f32) * f32(10 ** rng.uniform(-3, 3))
    r = (np.abs(rng.normal(size=n)) * 10 ** rng.uniform(-4, 1, size=n)).astype(f32) + f32(1e-6)
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    if mode == "surface":
        o = c + u * (r[:, None] * (1 + rng.normal(size=(n, 1)) * 1e-6))
    elif mode == "inside":
        o = c + u * (r[:, None] * rng.uniform(0, 1, size=(n, 1)))
    else:
        o = c + u * (r[:, None] * 10 ** rng.uniform(0, 6, size=(n, 1)))
    o = o.astype(f32)
    d = (c - o + rng.normal(size=(n, 3)).astype(f32) * r[:, None] * f32(1.5)).astype(f32)
    return o, d, c, r


def _wrong_rejections(o, d, c, r, sqrt_ulps):
    with np.errstate(all="ignore"):
        oc = (o - c).astype(f32)
        dx, dy, dz = d[:, 0], d[:, 1], d[:, 2]
        a = ((dx * dx + dy * dy) + dz * dz).astype(f32)
        b = (2.0 * ((dx * oc[:, 0] + dy * oc[:, 1]) + dz * oc[:, 2]).astype(f64)).astype(f32)
        cc = (((oc[:, 0] * oc[:, 0] + oc[:, 1] * oc[:, 1]) + oc[:, 2] * oc[:, 2])
              - r * r).astype(f32)
        disc = ((b * b).astype(f64) - (4.0 * a.astype(f64)) * cc.astype(f64)).astype(f32)
        m = disc > 0
        dsq = np.sqrt(np.where(m, disc, f32(1))).astype(f32)
        x = np.where(b < 0, (-b - dsq).astype(f32), (-b + dsq).astype(f32))
        q = (x.astype(f64) / 2.0).astype(f32)
        t0, t1 = (q / a).astype(f32), (cc / q).astype(f32)
        lo, hi = np.minimum(t0, t1), np.maximum(t0, t1)
        rep = np.where(lo < 0, hi, lo)
        ok = m & (hi >= 0) & np.isfinite(rep)
        dse = _ulp(dsq, sqrt_ulps) if sqrt_ulps else dsq
        qe = (np.where(b < 0, (-b - dse).astype(f32), (-b + dse).astype(f32)) * f32(0.5)).astype(f32)
        wrong = rejected = 0
        for k in (-1, 1):
            ra = _ulp((f32(1) / a).astype(f32), k)
            rq = _ulp((f32(1) / np.where(qe == 0, f32(1), qe)).astype(f32), k)
            t0e, t1e = (qe * ra).astype(f32), (cc * rq).astype(f32)
            for tc in (rep, np.nextafter(rep, f32(np.inf)), np.nextafter(rep, f32(0)),
                       rep * f32(0.99999), rep * f32(1.00001), rep * f32(0.99997), rep * f32(0.5)):
                tc = tc.astype(f32)
                lim = (tc * f32(1.00002)).astype(f32)
                rej = (np.abs(qe) >= TINY) & (a >= TINY) & (t0e > lim) & (t1e > lim)
                wrong += int((rej & ok & (rep <= tc)).sum())
                rejected += int((rej & ok).sum())
    return wrong, rejected


def test_exact_q_estimate_never_rejects_a_kept_sphere():
    rng = np.random.default_rng(5)
    total_rej = 0
    for mode in ("surface", "inside", "far"):
        wrong, rej = _wrong_rejections(*_cases(rng, 200_000, mode), sqrt_ulps=0)
        assert wrong == 0, mode
        total_rej += rej
    assert total_rej > 10_000  # the filter does reject spheres beyond tcut


def test_estimate_from_an_inexact_square_root_is_caught():
    rng = np.random.default_rng(5)
    wrong = sum(_wrong_rejections(*_cases(rng, 200_000, mode), sqrt_ulps=1)[0]
                for mode in ("surface", "inside", "far"))
    assert wrong > 0
