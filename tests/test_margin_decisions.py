"""The margin-based fast decisions of the cull kernel, replayed on the CPU.

The kernel decides the frustum test (sphere-derived box) and the sub-pixel
test from cheap estimates and falls back to the reference arithmetic only
when an estimate is within its error margin. These tests replay the kernel's
estimates in float32 on inputs placed at the decision boundaries and check
that every decision the estimate makes agrees with the reference arithmetic
(indirect_dispatch.wgsl:88-98 and :127-137 in the canonical float semantics).
The hardware reciprocal is modelled by its documented bound: the correctly
rounded 1/w moved by up to one ulp either way.
"""
from __future__ import annotations

import numpy as np

F = np.float32
U = 2.0 ** -24


def f32(x):
    return np.asarray(x, dtype=np.float64).astype(F)


def fma32(a, b, c):
    # float32 fma: the exact product is representable in float64; one rounding
    # of a*b + c to float32 (float64 intermediate: within the margins here).
    return f32(a.astype(np.float64) * b.astype(np.float64) + c.astype(np.float64))


def ref_box_dist(n, d, c, r):
    """Reference distance of the p-vertex (float32, left to right, no contraction)."""
    sgn = np.where(n >= 0, F(1), F(-1))
    p = f32(c + sgn * r[:, None])                      # c + r or c - r, one rounding each
    prod = f32(n * p)
    return f32(f32(f32(prod[:, 0] + prod[:, 1]) + prod[:, 2]) + d)


def margin(planes, c, r):
    l1 = f32((np.abs(planes[:, 0]) + np.abs(planes[:, 1])) + np.abs(planes[:, 2]))
    l1max = float(np.max(np.abs(planes[:, :3]).sum(1)))
    dmax = float(np.max(np.abs(planes[:, 3])))
    kl = np.nextafter(F(l1max * 2.0 ** -20), F(np.inf))
    kd = np.nextafter(F(dmax * 2.0 ** -20 + 2.0 ** -140), F(np.inf))
    cmax = np.max(np.abs(c), axis=1)
    bb = f32(cmax + r)
    e = fma32(np.full_like(bb, kl), bb, np.full_like(bb, kd))
    return l1, e, cmax


def test_frustum_margin_never_contradicts_reference():
    g = np.random.default_rng(7)
    agree = decided = 0
    for _ in range(40):
        n = g.normal(size=(6, 3))
        n /= np.linalg.norm(n, axis=1, keepdims=True)
        d = g.uniform(-2000, 2000, 6)
        planes = f32(np.concatenate([n, d[:, None]], axis=1))
        m = 20000
        r = f32(np.exp(g.uniform(np.log(0.05), np.log(50.0), m)))
        c = f32(g.uniform(-1500, 1500, (m, 3)))
        # Put every candidate on (or within a few ulps of) one plane's boundary.
        k = g.integers(0, 6, m)
        nk, dk = planes[k, :3].astype(np.float64), planes[k, 3].astype(np.float64)
        l1k = np.abs(nk).sum(1)
        off = (nk * c).sum(1) + l1k * r + dk
        c = f32(c - nk * (off / (nk * nk).sum(1))[:, None] * g.uniform(0.999999, 1.000001, m)[:, None])
        l1, e, cmax = margin(planes, c, r)
        s = np.stack([fma32(np.full(m, planes[i, 0]), c[:, 0], fma32(np.full(m, planes[i, 1]), c[:, 1],
                       fma32(np.full(m, planes[i, 2]), c[:, 2], fma32(np.full(m, l1[i]), r, np.full(m, planes[i, 3])))))
                      for i in range(6)], axis=1)
        smin = s.min(axis=1)
        ref = np.stack([ref_box_dist(planes[i, :3][None, :].repeat(m, 0), planes[i, 3], c, r) for i in range(6)], 1)
        ref_visible = ~(ref < 0).any(axis=1)
        lane_ok = (r > cmax * F(2.0 ** -22)) & (f32(cmax + r) < F(2.0 ** 100))
        sure_vis = lane_ok & (smin >= e)
        sure_cull = lane_ok & (smin < -e)
        assert not np.any(sure_vis & ~ref_visible), "margin says visible, reference culls"
        assert not np.any(sure_cull & ref_visible), "margin says culled, reference keeps"
        decided += int((sure_vis | sure_cull).sum())
        agree += m
    # The boundary set is built to be hard; most lanes still decide fast.
    assert decided > 0.5 * agree


def test_subpixel_margin_never_contradicts_reference():
    g = np.random.default_rng(11)
    t = F(0.001)
    lo, hi = F(0.001 * (1 - 2.0 ** -20)), F(0.001 * (1 + 2.0 ** -20))
    m = 400_000
    w = f32(np.exp(g.uniform(np.log(1e-3), np.log(1e4), m)))
    # |a| / w within a few ulps of 0.001 for half the lanes, anywhere for the rest.
    a_near = f32(w.astype(np.float64) * 0.001 * (1 + g.uniform(-4e-7, 4e-7, m)))
    a_far = f32(w.astype(np.float64) * np.exp(g.uniform(np.log(1e-5), np.log(1e-1), m)))
    a = np.where(g.random(m) < 0.5, a_near, a_far)
    exact = np.abs(f32(a / w)) < t
    rc = f32(1.0 / w.astype(np.float64))
    for rcp in (rc, np.nextafter(rc, F(np.inf)), np.nextafter(rc, F(0))):
        q = f32(np.abs(a) * rcp)
        w_ok = (w >= F(2.0 ** -64)) & (w < F(2.0 ** 62))
        sure_tiny = w_ok & (q < lo)
        sure_not = w_ok & (q > hi)
        assert not np.any(sure_tiny & ~exact)
        assert not np.any(sure_not & exact)
    assert (sure_tiny | sure_not).mean() > 0.45  # the near half sits inside the margin by design
