"""The subdivision tree's boxes pinned from both sides by closed forms (TightBoundingBox,
src/kernel/kernel.cu:544-621; GetPoints 404-431; GetSecondDerivatives 375-402).

A node (bx, by) at level L (g = 8^L) samples the surface on the 9 x 9 grid
u = bx/g + tx/(8g), v = by/g + ty/(8g); its box is the samples' AABB widened by
K = (M1 + 2 M2 + M3) sqrt(3) / (8 (8g)^2), with M1, M2, M3 the largest |component| of S_uu, S_uv,
S_vv over the samples, then clamped to the parent's box.  Containment tests (tests/test_oracle.py)
only bound the boxes from one side: an oversized K or an extra ring of samples would pass them.
Here the expected boxes are computed in closed form from power-basis polynomial patches (values
and second derivatives are polynomials evaluated directly, not through the matrix form), so a
wrong K denominator, sample grid or derivative maximum fails at 1e-14.
"""
from math import comb

import numpy as np
import pytest

import oracle as O

LEVELS = 3


def _bezier_net(a):
    """Control points of the degree (n, m) Bezier patch equal to sum_kl a[k, l] u^k v^l (per
    coordinate): t^k = sum_i C(i,k)/C(n,k) B_i^n(t)."""
    n, m = a.shape[0] - 1, a.shape[1] - 1
    Tu = np.array([[comb(i, k) / comb(n, k) if k <= i else 0.0 for k in range(n + 1)] for i in range(n + 1)])
    Tv = np.array([[comb(j, l) / comb(m, l) if l <= j else 0.0 for l in range(m + 1)] for j in range(m + 1)])
    return np.einsum("ik,klc,jl->ijc", Tu, a, Tv)


def _poly(a, u, v, du, dv):
    """d^(du+dv)/du^du dv^dv of sum_kl a[k, l] u^k v^l at the points (u, v); returns (n, 3)."""
    out = np.zeros((u.size, a.shape[2]))
    for k in range(du, a.shape[0]):
        fk = np.prod(np.arange(k - du + 1, k + 1)) if du else 1.0
        for l in range(dv, a.shape[1]):
            fl = np.prod(np.arange(l - dv + 1, l + 1)) if dv else 1.0
            out += np.outer(fk * fl * u ** (k - du) * v ** (l - dv), a[k, l])
    return out


def _expected_boxes(a, levels):
    """Closed-form tree boxes (rows lb.xyz, rt.xyz) in the reference node numbering, with the
    geometric parent's clamp."""
    boxes = {}
    out = []
    g = 1
    for L in range(levels):
        t = np.arange(9)
        rows = []
        for bx in range(g):
            for by in range(g):
                u = (bx / g + t[:, None] * (1.0 / g / 8) + 0 * t[None, :]).ravel()
                v = (by / g + t[None, :] * (1.0 / g / 8) + 0 * t[:, None]).ravel()
                S = _poly(a, u, v, 0, 0)
                M1 = np.max(np.abs(_poly(a, u, v, 2, 0)))
                M2 = np.max(np.abs(_poly(a, u, v, 1, 1)))
                M3 = np.max(np.abs(_poly(a, u, v, 0, 2)))
                K = (M1 + 2 * M2 + M3) * np.sqrt(3.0) / (8.0 * 8 * g * 8 * g)
                lb, rt = S.min(axis=0) - K, S.max(axis=0) + K
                if L > 0:
                    plb, prt = boxes[(L - 1, bx // 8, by // 8)]
                    lb, rt = np.maximum(lb, plb), np.minimum(rt, prt)
                boxes[(L, bx, by)] = (lb, rt)
                rows.append(np.r_[lb, rt])
        out.append(np.array(rows))
        g *= 8
    return np.concatenate(out)


def _patch(a):
    n, m = a.shape[0] - 1, a.shape[1] - 1
    cp = _bezier_net(a)
    ku = np.r_[np.zeros(n + 1), np.ones(n + 1)]
    kv = np.r_[np.zeros(m + 1), np.ones(m + 1)]
    return O.OSurf(n, m, ku, kv, cp, np.ones((n + 1, m + 1)))


def test_bilinear_patch_boxes_are_corner_hulls_plus_uv_twist():
    """Bilinear patch S = A + B u + C v + D u v: S_uu = S_vv = 0 and S_uv = D everywhere, so
    K = 2 max|D_c| sqrt(3) / (8 (8g)^2), and each node's sample AABB is the hull of its four
    corners (a bilinear function takes its extremes there)."""
    rng = np.random.default_rng(3)
    a = np.zeros((2, 2, 3))
    a[0, 0], a[1, 0], a[0, 1], a[1, 1] = rng.normal(size=(4, 3))
    os_ = _patch(a)
    nodes, _, _ = os_.build_tree(LEVELS)
    exp = _expected_boxes(a, LEVELS)
    scale = np.max(np.abs(exp))
    assert np.max(np.abs(nodes[:, 4:10] - exp)) <= 1e-14 * scale
    # the twist term alone: K at level 2 from the box widths over the corner hull
    D = np.max(np.abs(a[1, 1]))
    K2 = 2 * D * np.sqrt(3) / (8 * 512 ** 2)
    leaf = nodes[O.tree_base(2):]
    u0, v0 = leaf[:, 0], leaf[:, 1]
    corners = [_poly(a, u0 + du / 64, v0 + dv / 64, 0, 0) for du in (0, 1) for dv in (0, 1)]
    hull_lo, hull_hi = np.min(corners, axis=0), np.max(corners, axis=0)
    assert np.allclose(leaf[:, 4:7], hull_lo - K2, rtol=0, atol=1e-14 * scale)
    assert np.allclose(leaf[:, 7:10], hull_hi + K2, rtol=0, atol=1e-14 * scale)


@pytest.mark.parametrize("n,m,seed", [(3, 3, 5), (2, 4, 6), (5, 2, 7)])
def test_polynomial_patch_boxes_closed_form(n, m, seed):
    """Non-rational Bezier patches of degrees (3,3), (2,4), (5,2) with random power-basis
    coefficients: every node's box equals the closed form (samples on the 9 x 9 grid, K from the
    exact second-derivative maxima over those samples, clamp to the geometric parent)."""
    rng = np.random.default_rng(seed)
    a = rng.normal(size=(n + 1, m + 1, 3)) / (1.0 + np.add.outer(np.arange(n + 1), np.arange(m + 1)))[..., None]
    os_ = _patch(a)
    nodes, _, _ = os_.build_tree(LEVELS)
    exp = _expected_boxes(a, LEVELS)
    scale = np.max(np.abs(exp))
    assert np.max(np.abs(nodes[:, 4:10] - exp)) <= 1e-14 * scale


def test_closed_form_detects_plausible_mistakes():
    """The comparison above has teeth: boxes with K's denominator (8g) instead of (8g)^2, or with
    a 10 x 10 sample grid, differ from the oracle's far beyond 1e-14."""
    rng = np.random.default_rng(5)
    a = rng.normal(size=(4, 4, 3)) / (1.0 + np.add.outer(np.arange(4), np.arange(4)))[..., None]
    nodes, _, _ = _patch(a).build_tree(2)
    good = _expected_boxes(a, 2)
    scale = np.max(np.abs(good))
    assert np.max(np.abs(nodes[:, 4:10] - good)) <= 1e-14 * scale

    def variant(kden=2, ppe=9):
        out, boxes, g = [], {}, 1
        for L in range(2):
            for bx in range(g):
                for by in range(g):
                    t = np.arange(ppe)
                    u = (bx / g + t[:, None] / (8.0 * g) + 0 * t[None, :]).ravel()
                    v = (by / g + t[None, :] / (8.0 * g) + 0 * t[:, None]).ravel()
                    S = _poly(a, u, v, 0, 0)
                    M = [np.max(np.abs(_poly(a, u, v, *d))) for d in ((2, 0), (1, 1), (0, 2))]
                    K = (M[0] + 2 * M[1] + M[2]) * np.sqrt(3.0) / (8.0 * (8.0 * g) ** kden)
                    lb, rt = S.min(axis=0) - K, S.max(axis=0) + K
                    if L > 0:
                        plb, prt = boxes[(L - 1, bx // 8, by // 8)]
                        lb, rt = np.maximum(lb, plb), np.minimum(rt, prt)
                    boxes[(L, bx, by)] = (lb, rt)
                    out.append(np.r_[lb, rt])
            g *= 8
        return np.array(out)

    assert np.allclose(variant(), good, rtol=0, atol=1e-14 * scale)
    for bad in (variant(kden=1), variant(ppe=10)):
        assert np.max(np.abs(nodes[:, 4:10] - bad)) > 1e-6 * scale


def test_tree_on_a_non_unit_knot_domain_is_the_reparametrised_tree():
    """The tree lies over the knot domain: reparametrising a surface affinely (knots a + W k)
    leaves the geometry unchanged, so every node's box and centre point, and the valid/divided
    flags, equal the [0,1] surface's, and its uv rectangle is the mapped one (S_uu scales by
    1/Wu^2, which the cell size Wu/(8g) in K undoes)."""
    from gnurbs import synth
    sd = synth.surface_G(11, 9, 3, 2, seed=41, jitter=0.3)
    base = O.OSurf.from_def(sd)
    a, W, b, V = -2.0, 10.0, 3.0, 0.25
    moved = O.OSurf(sd.p, sd.q, a + W * sd.ku, b + V * sd.kv, sd.cp, sd.w)
    n0, v0, d0 = base.build_tree(LEVELS)
    n1, v1, d1 = moved.build_tree(LEVELS)
    assert np.allclose(n1[:, 0], a + W * n0[:, 0], rtol=0, atol=1e-13)
    assert np.allclose(n1[:, 2], a + W * n0[:, 2], rtol=0, atol=1e-13)
    assert np.allclose(n1[:, 1], b + V * n0[:, 1], rtol=0, atol=1e-15)
    assert np.allclose(n1[:, 3], b + V * n0[:, 3], rtol=0, atol=1e-15)
    scale = np.max(np.abs(n0[:, 4:]))
    assert np.max(np.abs(n1[:, 4:] - n0[:, 4:])) <= 1e-12 * scale
    # the flags agree except where a box extent ratio sits at NeedDivide's threshold within rounding
    assert np.mean(v0 == v1) > 0.999 and np.mean(d0 == d1) > 0.999
    # every leaf's samples lie in its box (the containment the seeds rely on)
    leaf = n1[O.tree_base(LEVELS - 1):]
    rng = np.random.default_rng(2)
    for k in rng.choice(len(leaf), 200, replace=False):
        uu = rng.uniform(leaf[k, 0], leaf[k, 2], 50)
        vv = rng.uniform(leaf[k, 1], leaf[k, 3], 50)
        S = moved.eval(uu, vv, 0)
        assert np.all(S.T >= leaf[k, 4:7] - 1e-12) and np.all(S.T <= leaf[k, 7:10] + 1e-12)
