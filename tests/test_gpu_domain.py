"""Surfaces whose knots do not run from 0 to 1.  The reference lays its tree over [0,1]^2
(src/kernel/kernel.cu:383-387) and scales derivatives by (k_last - k_0)/h
(src/NURBS/Evaluate.cu:120), both right only on [0,1] knots; the library maps the tree onto the
knot domain and scales by 1/h.  On a surface with knots on [-2, 8] x [3, 3.25] (the config-2
generator's net, reparametrised): evaluation, the tree, greedy seeds, plain and mixed inversion
and the global projection against the oracle, and round trips to the constructed parameters."""
import numpy as np
import pytest

import oracle as O
import gnurbs as G
from gnurbs import synth
from helpers import blocks, field_close

pytestmark = pytest.mark.gpu

A, W, B, V = -2.0, 10.0, 3.0, 0.25


def _np(t):
    return t.detach().cpu().numpy()


@pytest.fixture(scope="module")
def moved():
    sd = synth.surface_G(32, 32, 3, 3, seed=1002)
    ku, kv = A + W * sd.ku, B + V * sd.kv
    gs = G.Surface(sd.p, sd.q, ku, kv, sd.cp, sd.w)
    os_ = O.OSurf(sd.p, sd.q, ku, kv, sd.cp, sd.w)
    gs.BuildTree(3)
    rng = np.random.default_rng(5)
    n = 6000
    x0, y0 = rng.random(n), rng.random(n)
    u0, v0 = A + W * x0, B + V * y0
    b = blocks(os_.eval(u0, v0, 1))
    P, _ = synth.offset_queries(b["S"], b["Su"], b["Sv"], 1e-2, 6)
    yield dict(gs=gs, os=os_, u0=u0, v0=v0, P=P, S=b["S"])
    gs.close()


def test_evaluation_on_the_domain(moved):
    gs, os_ = moved["gs"], moved["os"]
    u, v = moved["u0"], moved["v0"]
    got = blocks(_np(gs.Evaluate(u, v, 2)))
    ref = blocks(os_.eval(u, v, 2))
    floor = 1e-2 * np.max(np.abs(ref["S"]))
    for k in ("S", "Su", "Sv", "Suu", "Suv", "Svv"):
        ok, err = field_close(got[k], ref[k], 1e-12, floor)
        assert ok, (k, err)


def test_tree_covers_the_domain(moved):
    gs, os_ = moved["gs"], moved["os"]
    gn, gv, gd = gs.GetTree()
    on, ov, od = os_.build_tree(3)
    assert np.array_equal(gn[:, :4], on[:, :4])
    assert gn[0, 0] == A and gn[0, 2] == A + W and gn[0, 1] == B and gn[0, 3] == B + V
    scale = np.max(np.abs(on[:, 4:]))
    assert np.max(np.abs(gn[:, 4:] - on[:, 4:])) <= 1e-12 * scale
    assert np.array_equal(gv, ov) and np.array_equal(gd, od)


def test_seeds_and_inversion_on_the_domain(moved):
    gs, os_, P = moved["gs"], moved["os"], moved["P"]
    nodes, _, _ = gs.GetTree()
    leaf, su, sv = (_np(t) for t in gs.Seed(P[0], P[1], P[2], exact=False))
    rl, rsu, rsv = O.seed_greedy_leaf(nodes, 3, P[0], P[1], P[2])
    same = leaf == rl
    assert same.mean() > 0.999
    assert np.array_equal(su[same], rsu[same]) and np.array_equal(sv[same], rsv[same])
    # seeds lie in the domain and near the constructed parameters
    assert su.min() >= A and su.max() <= A + W and sv.min() >= B and sv.max() <= B + V
    assert np.median(np.abs(su - moved["u0"])) < W / 64 and np.median(np.abs(sv - moved["v0"])) < V / 64
    # plain FP64 against the oracle's same iteration from the same seeds
    u, v, st = (_np(t) for t in gs.NewInverse(P[0], P[1], P[2], G.default_opts(max_iter=20, tol_step=1e-10)))
    ru, rv, rst, _ = os_.gauss_newton(P[0], P[1], P[2], rsu, rsv, tol_step=1e-10, max_iter=20)
    both = same & (st == G.PT_CONVERGED) & (rst == O.CONV_STEP)
    assert both.mean() > 0.5
    assert np.max(np.maximum(np.abs(u - ru), np.abs(v - rv))[both]) <= 1e-9
    # mixed against the oracle's FP64 Newton solution
    mixed = G.default_opts(precision=G.GN_MIXED, newton=G.NEWTON_FP32 | G.NEWTON_FP64, tol_step32=1e-4)
    u, v, st = (_np(t) for t in gs.NewInverse(P[0], P[1], P[2], mixed))
    ru, rv, rst, _ = os_.gauss_newton(P[0], P[1], P[2], rsu, rsv, tol_step=1e-12, max_iter=40, newton=1)
    both = same & (st == G.PT_CONVERGED) & (rst == O.CONV_STEP)
    d = np.maximum(np.abs(u - ru), np.abs(v - rv))[both]
    sm = d < 1e-6
    assert both.mean() > 0.9 and sm.mean() > 0.98
    assert np.max(d[sm]) <= 1e-9
    # most reach the constructed foot point
    err = np.maximum(np.abs(u - moved["u0"]), np.abs(v - moved["v0"]))
    assert np.mean(err <= 1e-9) > 0.9


def test_global_projection_on_the_domain(moved):
    """GN_SEED_GLOBAL solves in leaf cells mapped onto the knot domain: on the near-surface
    queries it equals the exhaustive oracle (every leaf cell solved) in distance."""
    gs, os_, P = moved["gs"], moved["os"], moved["P"]
    sel = np.arange(400)
    Ps = P[:, sel]
    opts = G.default_opts(precision=G.GN_FP64, newton=G.NEWTON_FP64, seed_mode=G.SEED_GLOBAL, max_iter=20,
                          tol_step=1e-12)
    u, v, st = (_np(t) for t in gs.NewInverse(Ps[0], Ps[1], Ps[2], opts))
    nodes, _, _ = os_.build_tree(3)
    ou, ov, ost = os_.project_global(nodes, 3, Ps[0], Ps[1], Ps[2], max_iter=20, leaf_iter=8, tol_step=1e-12)
    assert u.min() >= A and u.max() <= A + W and v.min() >= B and v.max() <= B + V
    d = np.linalg.norm(os_.eval(u, v, 0) - Ps, axis=0)
    do = np.linalg.norm(os_.eval(ou, ov, 0) - Ps, axis=0)
    assert np.all(np.abs(d - do) <= 1e-12 + 1e-9 * do), np.max(np.abs(d - do))
