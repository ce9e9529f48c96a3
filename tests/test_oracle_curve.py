"""Pins the oracle's curve functions (CPU only).

The reference has no curve code, so the curve oracle is pinned by what the mathematics fixes:
exact circle arcs (rational quadratic, w = cos(pi/4)), scipy's B-spline for non-rational
curves and their derivatives, finite differences for the rational quotient, a brute-force
nearest-sample check of the seeds and closed-form foot points on the circle.
"""
import math

import numpy as np
import pytest
from scipy.interpolate import BSpline

import oracle as O
from gnurbs import synth

R2 = math.sqrt(0.5)


def quarter_circle():
    return O.OCurve(2, [0, 0, 0, 1, 1, 1], [[1, 0, 0], [1, 1, 0], [0, 1, 0]], [1, R2, 1])


def full_circle():
    k = [0, 0, 0, .25, .25, .5, .5, .75, .75, 1, 1, 1]
    cp = [[1, 0, 0], [1, 1, 0], [0, 1, 0], [-1, 1, 0], [-1, 0, 0], [-1, -1, 0], [0, -1, 0], [1, -1, 0],
          [1, 0, 0]]
    w = [1, R2, 1, R2, 1, R2, 1, R2, 1]
    return O.OCurve(2, k, cp, w)


def spline_derivative(k, cp, p):
    """Knots, control points and degree of the derivative of a non-rational B-spline
    (The NURBS Book eq. 3.8; a zero-length knot interval contributes 0)."""
    den = k[p + 1:p + len(cp)] - k[1:len(cp)]
    d = np.where(den[:, None] > 0, p * np.diff(cp, axis=0) / np.where(den > 0, den, 1)[:, None], 0.0)
    return k[1:-1], d, p - 1


def random_curve(n, p, seed, rational=True, repeat=False):
    rng = np.random.default_rng(seed)
    k = synth.clamped_knots(n, p, rng, jitter=0.3)
    if repeat and n - p > 3:  # one interior knot of multiplicity 2
        k[p + 2] = k[p + 1]
    cp = rng.normal(size=(n, 3))
    w = rng.uniform(0.5, 2.0, n) if rational else np.ones(n)
    return O.OCurve(p, k, cp, w)


def test_quarter_circle_exact():
    c = quarter_circle()
    t = np.linspace(0, 1, 257)
    o = c.eval(t, order=2)
    C, C1 = o[0:3], o[3:6]
    assert np.max(np.abs(np.hypot(C[0], C[1]) - 1)) < 1e-15
    assert np.max(np.abs(C[2])) == 0
    assert np.allclose(C[:, 0], [1, 0, 0], atol=0) and np.allclose(C[:, -1], [0, 1, 0], atol=1e-16)
    assert np.allclose(c.eval([0.5])[:, 0], [R2, R2, 0], atol=1e-16)
    # tangent is orthogonal to the radius; C'(0) = 2 w1 (P1 - P0) / w0 = (0, sqrt 2, 0)
    assert np.max(np.abs(np.sum(C[:2] * C1[:2], axis=0))) < 1e-14
    assert np.allclose(o[3:6, 0], [0, 2 * R2, 0], atol=1e-15)


def test_full_circle_radius_and_periodicity():
    c = full_circle()
    t = np.linspace(0, 1, 1001)
    C = c.eval(t)
    assert np.max(np.abs(np.hypot(C[0], C[1]) - 1)) < 1e-15
    # the angle is monotone in t and covers the whole circle
    ang = np.unwrap(np.arctan2(C[1], C[0]))
    assert np.all(np.diff(ang) > 0)
    assert abs(ang[-1] - ang[0] - 2 * math.pi) < 1e-12


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("repeat", [False, True])
def test_nonrational_vs_scipy(p, repeat):
    c = random_curve(12, p, 10 + p, rational=False, repeat=repeat)
    t = np.random.default_rng(p).uniform(0, 1, 500)
    o = c.eval(t, order=2)
    assert np.allclose(o[0:3], BSpline(c.k, c.cp, p)(t).T, atol=1e-13, rtol=1e-12)
    k1, d1, p1 = spline_derivative(c.k, c.cp, p)
    assert np.allclose(o[3:6], BSpline(k1, d1, p1)(t).T, atol=1e-11, rtol=1e-11)
    if p >= 2:
        k2, d2, p2 = spline_derivative(k1, d1, p1)
        assert np.allclose(o[6:9], BSpline(k2, d2, p2)(t).T, atol=1e-9, rtol=1e-10)
    else:
        assert np.max(np.abs(o[6:9])) == 0


@pytest.mark.parametrize("p", [2, 3, 5])
def test_rational_finite_differences(p):
    c = random_curve(10, p, 40 + p)
    # stay away from knots, where the derivatives of order p are discontinuous
    t = np.random.default_rng(p).uniform(0.02, 0.98, 300)
    t = t[np.min(np.abs(t[:, None] - c.k[None, :]), axis=1) > 1e-3]
    h = 1e-6
    o = c.eval(t, order=2)
    op, om = c.eval(t + h, order=1), c.eval(t - h, order=1)
    fd1 = (op[0:3] - om[0:3]) / (2 * h)
    fd2 = (op[3:6] - om[3:6]) / (2 * h)
    assert np.max(np.abs(fd1 - o[3:6]) / (1 + np.abs(o[3:6]))) < 1e-6
    assert np.max(np.abs(fd2 - o[6:9]) / (1 + np.abs(o[6:9]))) < 1e-4


def test_seed_is_nearest_sample():
    c = random_curve(20, 3, 77)
    rng = np.random.default_rng(5)
    P = rng.normal(size=(3, 400))
    t = c.seed(*P)
    # the 8 sample parameters per domain span
    ts = []
    for m in range(c.p, c.n):
        lo, hi = c.k[m], c.k[m + 1]
        ts += [lo + (j + 0.5) * (hi - lo) / 8 for j in range(8)]
    ts = np.array(ts)
    S = c.eval(ts).astype(np.float32).astype(np.float64)
    d2 = ((S[:, None, :] - P.astype(np.float32).astype(np.float64)[:, :, None]) ** 2).sum(0)
    best = d2.min(axis=1)
    pos = np.searchsorted(ts, t)
    assert np.all(ts[pos] == t)
    got = d2[np.arange(P.shape[1]), pos]
    assert np.all(got <= best * (1 + 1e-6) + 1e-12)


@pytest.mark.parametrize("newton", [0, 1])
def test_round_trip_on_curve(newton):
    cd = synth.helix_curve(64, 3, 7)
    c = O.OCurve.from_def(cd)
    rng = np.random.default_rng(3)
    ts = rng.uniform(0, 1, 4000)
    P = c.eval(ts)
    t, st, it = c.invert(*P, c.seed(*P), newton=newton)
    ok = np.abs(t - ts) <= 1e-9
    assert ok.mean() > 0.99
    assert np.all(st[ok] <= 1)


@pytest.mark.parametrize("newton", [0, 1])
def test_circle_foot_point_closed_form(newton):
    c = quarter_circle()
    rng = np.random.default_rng(9)
    th = rng.uniform(0.05, math.pi / 2 - 0.05, 500)
    r = rng.uniform(0.7, 1.3, 500)
    P = np.stack([r * np.cos(th), r * np.sin(th), rng.normal(0, 0.2, 500)])
    t, st, _ = c.invert(*P, c.seed(*P), newton=newton, max_iter=50)
    C = c.eval(t)
    assert np.all(st == O.CONV_STEP)
    assert np.max(np.abs(C[0] - np.cos(th))) < 1e-9
    assert np.max(np.abs(C[1] - np.sin(th))) < 1e-9


def test_newton_takes_fewer_iterations_off_curve():
    cd = synth.helix_curve(64, 3, 8)
    c = O.OCurve.from_def(cd)
    rng = np.random.default_rng(4)
    ts = rng.uniform(0.05, 0.95, 2000)
    o = c.eval(ts, order=1)
    # offsets along a normal: any unit vector orthogonal to C'
    tan = o[3:6] / np.linalg.norm(o[3:6], axis=0)
    a = rng.normal(size=(3, ts.size))
    nrm = a - tan * np.sum(a * tan, axis=0)
    nrm /= np.linalg.norm(nrm, axis=0)
    P = o[0:3] + 0.01 * nrm
    seed = c.seed(*P)
    _, _, it_gn = c.invert(*P, seed, newton=0, max_iter=50)
    _, st_n, it_n = c.invert(*P, seed, newton=1, max_iter=50)
    assert np.mean(st_n == O.CONV_STEP) > 0.98
    assert it_n.mean() < it_gn.mean()


def test_singular_and_distance_rules():
    c = O.OCurve(3, synth.clamped_knots(5, 3), np.ones((5, 3)), np.ones(5))  # constant curve
    t, st, it = c.invert([2.0], [1.0], [1.0], [0.3])
    assert st[0] == O.SINGULAR and it[0] == 0 and t[0] == 0.3
    q = quarter_circle()
    P = q.eval([0.3])
    t, st, it = q.invert(*P, [0.3], tol_dist=1e-3)
    assert st[0] == O.CONV_DIST and it[0] == 0 and t[0] == 0.3
