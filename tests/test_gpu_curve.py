"""GPU parity of the curve path (gn_curve_*) against the oracle's textbook de Boor curve.

Tolerances: FP64 evaluation 1e-12 relative to the largest component of each block; inversion
parameters within 1e-9 wherever both converge; seeds equal (up to float ties, which must then
be within 1e-6 in squared distance).
"""
import math

import numpy as np
import pytest

import gnurbs as G
import oracle as O
from gnurbs import synth

pytestmark = pytest.mark.gpu
R2 = math.sqrt(0.5)


def random_def(n, p, seed):
    rng = np.random.default_rng(seed)
    k = synth.clamped_knots(n, p, rng, jitter=0.3)
    return synth.CurveDef(p, k, rng.normal(size=(n, 3)), rng.uniform(0.5, 2.0, n))


def pair(cd):
    return G.Curve.from_def(cd), O.OCurve.from_def(cd)


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5, 7])
def test_eval_fp64_vs_oracle(p):
    cd = random_def(24, p, 100 + p)
    g, o = pair(cd)
    t = np.random.default_rng(p).uniform(0, 1, 20000)
    t[:4] = [0.0, 1.0, cd.k[p + 2], cd.k[p + 3]]  # end points and knots
    for order in (0, 1, 2):
        got = g.Evaluate(t, order).cpu().numpy()
        ref = o.eval(t, order)
        for b in range(order + 1):
            scale = np.max(np.abs(ref[3 * b:3 * b + 3])) + 1e-300
            err = np.max(np.abs(got[3 * b:3 * b + 3] - ref[3 * b:3 * b + 3])) / scale
            assert err < 1e-12, (order, b, err)


def test_eval_exact_quarter_circle():
    g = G.Curve(2, [0, 0, 0, 1, 1, 1], [[1, 0, 0], [1, 1, 0], [0, 1, 0]], [1, R2, 1])
    t = np.linspace(0, 1, 4097)
    C = g.Evaluate(t, 2).cpu().numpy()
    assert np.max(np.abs(np.hypot(C[0], C[1]) - 1)) < 4e-16
    assert np.max(np.abs(np.sum(C[0:2] * C[3:5], axis=0))) < 1e-14


def test_seeds_match_oracle():
    cd = synth.helix_curve(64, 3, 11)
    g, o = pair(cd)
    rng = np.random.default_rng(1)
    P = o.eval(rng.uniform(0, 1, 5000)) + rng.normal(0, 0.02, (3, 5000))
    t, st = g.Invert(*P, G.default_opts(max_iter=0))
    t = t.cpu().numpy()
    ref = o.seed(*P)
    same = t == ref
    assert same.mean() > 0.999
    if not same.all():  # near-ties of FP32 distances: both samples must be equally near
        Pf = P.astype(np.float32).astype(np.float64)
        a = o.eval(t[~same]).astype(np.float32).astype(np.float64)
        b = o.eval(ref[~same]).astype(np.float32).astype(np.float64)
        da = ((a - Pf[:, ~same]) ** 2).sum(0)
        db = ((b - Pf[:, ~same]) ** 2).sum(0)
        assert np.all(np.abs(da - db) <= 1e-6 * db + 1e-12)


@pytest.mark.parametrize("newton", [0, 1])
def test_invert_fp64_vs_oracle(newton):
    cd = synth.helix_curve(64, 3, 12)
    g, o = pair(cd)
    rng = np.random.default_rng(2)
    ts = rng.uniform(0, 1, 20000)
    P = o.eval(ts) + rng.normal(0, 0.01, (3, ts.size))
    seed = o.seed(*P)
    opts = G.default_opts(newton=G.NEWTON_FP64 if newton else 0, tol_quad=0.0)
    t, st = g.Invert(*P, opts, seeds=seed)
    t, st = t.cpu().numpy(), st.cpu().numpy()
    rt, rst, _ = o.invert(*P, seed, newton=newton)
    both = (st == O.CONV_STEP) & (rst == O.CONV_STEP)
    assert both.mean() > 0.97
    assert np.max(np.abs(t[both] - rt[both])) <= 1e-9
    assert (st == rst).mean() > 0.99


def test_non_finite_queries():
    """Queries with a NaN or infinite coordinate end GN_PT_NONFINITE (FP64 and mixed), as in the
    oracle, and the others keep their results."""
    cd = synth.helix_curve(64, 3, 14)
    g, o = pair(cd)
    rng = np.random.default_rng(4)
    ts = rng.uniform(0, 1, 2000)
    P = o.eval(ts) + rng.normal(0, 0.01, (3, ts.size))
    P[0, 3], P[1, 10], P[2, 20] = np.nan, np.inf, -np.inf
    bad = np.zeros(ts.size, dtype=bool)
    bad[[3, 10, 20]] = True
    seed = o.seed(*P)
    rt, rst, _ = o.invert(*P, seed, newton=1)
    assert np.all(rst[bad] == O.NONFINITE)
    for opts in (G.default_opts(newton=G.NEWTON_FP64), G.default_opts(precision=G.GN_MIXED, newton=3)):
        t, st = g.Invert(*P, opts, seeds=seed)
        t, st = t.cpu().numpy(), st.cpu().numpy()
        assert np.all(st[bad] == G.PT_NONFINITE), st[bad]
        both = (st == O.CONV_STEP) & (rst == O.CONV_STEP)
        assert both[~bad].mean() > 0.97 and np.max(np.abs(t[both] - rt[both])) <= 1e-9


@pytest.mark.parametrize("newton", [0, 3])
def test_mixed_round_trip(newton):
    cd = synth.helix_curve(64, 3, 13)
    g, o = pair(cd)
    rng = np.random.default_rng(3)
    ts = rng.uniform(0, 1, 100000)
    P = o.eval(ts)
    opts = G.default_opts(precision=G.GN_MIXED, newton=newton, tol_quad=1e-10 if newton else 0.0)
    t, st = g.Invert(*P, opts)
    err = np.abs(t.cpu().numpy() - ts)
    conv = st.cpu().numpy() == G.PT_CONVERGED
    same = conv & (err < 1e-6)  # converged to the sampled parameter, not another local minimum
    assert same.mean() > 0.99
    assert np.max(err[same]) <= 1e-9


def test_circle_foot_points():
    g = G.Curve(2, [0, 0, 0, 1, 1, 1], [[1, 0, 0], [1, 1, 0], [0, 1, 0]], [1, R2, 1])
    rng = np.random.default_rng(4)
    th = rng.uniform(0.05, math.pi / 2 - 0.05, 4000)
    r = rng.uniform(0.7, 1.3, th.size)
    P = np.stack([r * np.cos(th), r * np.sin(th), rng.normal(0, 0.2, th.size)])
    for prec in (G.GN_FP64, G.GN_MIXED):
        t, st = g.Invert(*P, G.default_opts(precision=prec, max_iter=50, fp64_tail=10))
        C = g.Evaluate(t).cpu().numpy()
        assert np.all(st.cpu().numpy() == G.PT_CONVERGED)
        assert np.max(np.abs(C[0] - np.cos(th))) < 1e-9 and np.max(np.abs(C[1] - np.sin(th))) < 1e-9


def test_curve_errors():
    g = G.Curve(2, [0, 0, 0, 1, 1, 1], [[1, 0, 0], [1, 1, 0], [0, 1, 0]], [1, R2, 1])
    P = np.zeros((3, 4))
    with pytest.raises(G.GnError):
        g.Invert(*P, G.default_opts(precision=G.GN_FP32))
    with pytest.raises(G.GnError):
        g.Invert(*P, G.default_opts(seed_mode=G.SEED_REF_FIRST_HIT))
    with pytest.raises(G.GnError):
        g.Evaluate(np.zeros(4), order=3)
    # empty input is a no-op
    t, st = g.Invert(np.zeros(0), np.zeros(0), np.zeros(0))
    assert t.numel() == 0


def test_long_curve_seeds_from_global_memory():
    # 400 spans: the seed data (48 KB) does not fit the LDS staging limit, samples stay in L2
    cd = random_def(403, 3, 808)
    g, o = pair(cd)
    rng = np.random.default_rng(5)
    P = o.eval(rng.uniform(0, 1, 3000)) + rng.normal(0, 0.05, (3, 3000))
    t, _ = g.Invert(*P, G.default_opts(max_iter=0))
    assert (t.cpu().numpy() == o.seed(*P)).mean() > 0.999
