"""Pins the CPU FP64 oracle (oracle/gnurbs_oracle.c) against checks that do not restate it:
brute force, an independent basis build, scipy's B-spline basis, closed forms, finite
differences and the reference's own known-answer points (src/kernel/kernel.cu:451-454)."""
import math
import os

import numpy as np
import pytest

import oracle as O
from gnurbs import synth
from helpers import blocks, field_close, read_nu, scipy_eval


# ---------------------------------------------------------------------------------------------
# a1 findSpan (src/NURBS/Evaluate.cu:504-533): brute-force linear scan
# ---------------------------------------------------------------------------------------------
def brute_span(p, cp, k, u):
    if u >= k[cp]:
        return cp - 1
    if u <= k[p]:
        return p
    for m in range(p, cp):
        if k[m] <= u < k[m + 1]:
            return m
    raise AssertionError


KNOT_CASES = [
    (3, [0, 0, 0, 0, 0.25, 0.5, 0.75, 1, 1, 1, 1]),
    (2, [0, 0, 0, 0.25, 0.25, 0.5, 0.5, 0.75, 0.75, 1, 1, 1]),  # repeated interior knots
    (1, [0, 0, 1, 1]),
    (5, [0] * 6 + [0.2, 0.2, 0.2, 0.29, 0.6] + [1] * 6),
    (3, [0, 0, 0, 0, 1, 1, 1, 1]),
    (2, [-1, -1, -1, 0.5, 2, 3, 3, 3]),  # not on [0, 1]
]


@pytest.mark.parametrize("p,k", KNOT_CASES)
def test_find_span_brute_force(p, k):
    k = np.array(k, dtype=float)
    cp = len(k) - p - 1
    rng = np.random.default_rng(7)
    us = list(rng.uniform(k[0] - 0.5, k[-1] + 0.5, 400)) + list(k) + [k[0], k[-1], -np.inf, np.inf]
    us += list(np.nextafter(k, np.inf)) + list(np.nextafter(k, -np.inf))
    for u in us:
        assert O.find_span(p, cp, k, u) == brute_span(p, cp, k, u), (p, u)


def test_find_span_nan_matches_reference_loop():
    # the reference's while-condition is false for NaN, so it returns the first midpoint
    p, k = KNOT_CASES[0]
    cp = len(k) - p - 1
    assert O.find_span(p, cp, np.array(k, float), float("nan")) == int((p + cp) / 2.0)


# ---------------------------------------------------------------------------------------------
# a2/a3 basis matrix (Nrc, src/NURBS/Evaluate.cu:90-108)
# ---------------------------------------------------------------------------------------------
@pytest.mark.parametrize("p,k", KNOT_CASES[:5])
def test_basis_nrc_vs_independent_cdb(p, k):
    k = np.array(k, dtype=float)
    cp = len(k) - p - 1
    for span in range(p, cp):
        if k[span] == k[span + 1]:
            continue
        A = O.basis_matrix(p, k, span, "nrc")
        B = O.basis_matrix(p, k, span, "cdb")
        assert np.max(np.abs(A - B)) < 1e-13
        # partition of unity: sum_s N_s(t) == 1  ->  row sums are (1, 0, 0, ...)
        e = np.zeros(p + 1)
        e[0] = 1
        assert np.allclose(A.sum(axis=1), e, atol=1e-13)


@pytest.mark.parametrize("p", [1, 2, 3, 5, 6])
def test_basis_bezier_is_bernstein(p):
    k = np.array([0.0] * (p + 1) + [1.0] * (p + 1))
    M = O.basis_matrix(p, k, p, "nrc")
    # N_s(t) = C(p,s) t^s (1-t)^(p-s) = sum_r C(p,s) C(p-s, r-s) (-1)^(r-s) t^r
    E = np.zeros((p + 1, p + 1))
    for s in range(p + 1):
        for r in range(s, p + 1):
            E[r, s] = math.comb(p, s) * math.comb(p - s, r - s) * (-1) ** (r - s)
    assert np.max(np.abs(M - E)) < 1e-12


@pytest.mark.parametrize("p", [2, 3, 5])
def test_basis_vs_scipy(p):
    rng = np.random.default_rng(p)
    n = 9
    interior = np.sort(rng.uniform(0.05, 0.95, n - p - 1))
    k = np.concatenate([np.zeros(p + 1), interior, np.ones(p + 1)])
    from scipy.interpolate import BSpline

    bs = BSpline(k, np.eye(n), p)
    for span in range(p, n):
        M = O.basis_matrix(p, k, span, "nrc")
        t = np.linspace(0, 0.999, 7)
        x = k[span] + t * (k[span + 1] - k[span])
        T = np.vander(t, p + 1, increasing=True)  # t^r
        mine = T @ M  # N_{span-p+s}(x)
        ref = bs(x)[:, span - p: span + 1]
        assert np.max(np.abs(mine - ref)) < 1e-12


# ---------------------------------------------------------------------------------------------
# a5-a9 span table + evaluation vs an independent scipy NURBS evaluation
# ---------------------------------------------------------------------------------------------
def _surf(nu, nv, p, q, seed, jitter=0.0):
    sd = synth.surface_G(nu, nv, p, q, seed=seed, jitter=jitter)
    return sd, O.OSurf.from_def(sd)


@pytest.mark.parametrize("nu,nv,p,q", [(8, 8, 3, 3), (12, 7, 2, 4), (9, 11, 5, 2), (6, 6, 5, 5), (4, 5, 1, 3)])
def test_eval_vs_scipy(nu, nv, p, q):
    sd, os_ = _surf(nu, nv, p, q, seed=11, jitter=0.3)
    u, v = synth.uniform_uv(300, seed=3)
    mine = blocks(os_.eval(u, v, order=2))
    ref = scipy_eval(p, q, sd.ku, sd.kv, sd.cp, sd.w, u, v, order=2)
    for key in ("S", "Su", "Sv", "Suu", "Suv", "Svv"):
        ok, err = field_close(mine[key], ref[key], 1e-11)
        assert ok, (key, err)


def test_span_table_nrc_equals_cdb():
    sd = synth.surface_G(10, 9, 3, 4, seed=5, jitter=0.3)
    a = O.OSurf.from_def(sd, use_cdb=False)
    b = O.OSurf.from_def(sd, use_cdb=True)
    assert np.array_equal(a.valid, b.valid)
    assert np.max(np.abs(a.MatrixN - b.MatrixN)) < 1e-12 * np.max(np.abs(a.MatrixN))


def test_eval_finite_differences():
    sd, os_ = _surf(8, 8, 3, 3, seed=1000)
    rng = np.random.default_rng(0)
    u = rng.uniform(0.05, 0.95, 50)
    v = rng.uniform(0.05, 0.95, 50)
    h = 1e-6
    b0 = blocks(os_.eval(u, v, 2))
    bu = [blocks(os_.eval(u + s * h, v, 1)) for s in (-1, 1)]
    bv = [blocks(os_.eval(u, v + s * h, 1)) for s in (-1, 1)]
    fd_su = (bu[1]["S"] - bu[0]["S"]) / (2 * h)
    fd_sv = (bv[1]["S"] - bv[0]["S"]) / (2 * h)
    fd_suu = (bu[1]["Su"] - bu[0]["Su"]) / (2 * h)
    fd_svv = (bv[1]["Sv"] - bv[0]["Sv"]) / (2 * h)
    fd_suv = (bv[1]["Su"] - bv[0]["Su"]) / (2 * h)
    for fd, key in ((fd_su, "Su"), (fd_sv, "Sv"), (fd_suu, "Suu"), (fd_svv, "Svv"), (fd_suv, "Suv")):
        scale = np.max(np.abs(b0[key]))
        assert np.max(np.abs(fd - b0[key])) < 1e-6 * max(scale, 1.0), key


def test_closed_form_bilinear():
    # p = q = 1, 2x2 net, unit weights: S(u, v) is the bilinear interpolant
    P = np.array([[[0, 0, 0], [0, 1, 0.5]], [[1, 0, 0.25], [1.5, 2, 1]]], dtype=float)
    w = np.ones((2, 2))
    k = np.array([0, 0, 1, 1.0])
    os_ = O.OSurf(1, 1, k, k, P, w)
    u, v = synth.uniform_uv(64, seed=9)
    b = blocks(os_.eval(u, v, 2))
    ref = ((1 - u) * (1 - v))[None] * P[0, 0][:, None] + ((1 - u) * v)[None] * P[0, 1][:, None] + \
          (u * (1 - v))[None] * P[1, 0][:, None] + (u * v)[None] * P[1, 1][:, None]
    assert np.max(np.abs(b["S"] - ref)) < 1e-14
    assert np.max(np.abs(b["Suu"])) < 1e-13 and np.max(np.abs(b["Svv"])) < 1e-13
    assert np.max(np.abs(b["Suv"] - (P[0, 0] - P[0, 1] - P[1, 0] + P[1, 1])[:, None])) < 1e-13


def test_closed_form_quarter_cylinder():
    # exact quarter circle (w = sqrt(2)/2 in the middle) swept linearly in v
    r2 = math.sqrt(2) / 2
    P = np.array([[[1, 0, 0], [1, 0, 2]], [[1, 1, 0], [1, 1, 2]], [[0, 1, 0], [0, 1, 2]]], dtype=float)
    w = np.array([[1, 1], [r2, r2], [1, 1]])
    ku = np.array([0, 0, 0, 1, 1, 1.0])
    kv = np.array([0, 0, 1, 1.0])
    os_ = O.OSurf(2, 1, ku, kv, P, w)
    u, v = synth.uniform_uv(200, seed=2)
    b = blocks(os_.eval(u, v, 2))
    radius = np.hypot(b["S"][0], b["S"][1])
    assert np.max(np.abs(radius - 1)) < 1e-14
    assert np.max(np.abs(b["S"][2] - 2 * v)) < 1e-14
    # the circle point is orthogonal to its tangent; curvature vector points to the axis
    assert np.max(np.abs(b["S"][0] * b["Su"][0] + b["S"][1] * b["Su"][1])) < 1e-13
    sp = np.hypot(b["Su"][0], b["Su"][1])
    kappa = np.abs(b["Su"][0] * b["Suu"][1] - b["Su"][1] * b["Suu"][0]) / sp ** 3
    assert np.max(np.abs(kappa - 1)) < 1e-12


@pytest.mark.parametrize("fname,expect", [
    ("center_bk.nu", (0.000000, 0.200049, 0.616000)),   # src/kernel/kernel.cu:453-454
    ("test_face2.nu", (0.700000, 0.455556, 0.055556)),  # src/kernel/kernel.cu:451-452
])
def test_reference_known_answers(golden_dir, fname, expect):
    p, q, ku, kv, cp, w = read_nu(os.path.join(golden_dir, fname))
    os_ = O.OSurf(p, q, ku, kv, cp, w)
    S = os_.eval([0.25], [0.25], 0)[:, 0]
    assert np.max(np.abs(S - np.array(expect))) < 5e-7


@pytest.mark.parametrize("fname", ["center_bk.nu", "part3_0.nu", "part3_2.nu", "test_face1.nu", "test_face2.nu"])
def test_models_vs_scipy(golden_dir, fname):
    p, q, ku, kv, cp, w = read_nu(os.path.join(golden_dir, fname))
    os_ = O.OSurf(p, q, ku, kv, cp, w)
    u, v = synth.uniform_uv(200, seed=4)
    mine = blocks(os_.eval(u, v, 2))
    ref = scipy_eval(p, q, ku, kv, cp, w, u, v, 2)
    floor = 1e-2 * np.max(np.abs(ref["S"]))
    for key in ("S", "Su", "Sv", "Suu", "Suv", "Svv"):
        ok, err = field_close(mine[key], ref[key], 1e-10, floor)
        assert ok, (fname, key, err)


# ---------------------------------------------------------------------------------------------
# a20 Gauss-Newton (src/kernel/kernel.cu:151-249)
# ---------------------------------------------------------------------------------------------
def test_gauss_newton_round_trip_on_surface():
    sd, os_ = _surf(8, 8, 3, 3, seed=1000)
    rng = np.random.default_rng(5)
    u0, v0 = rng.uniform(0.05, 0.95, 500), rng.uniform(0.05, 0.95, 500)
    S = os_.eval(u0, v0, 0)
    su = np.clip(u0 + rng.uniform(-0.02, 0.02, 500), 0, 1)
    sv = np.clip(v0 + rng.uniform(-0.02, 0.02, 500), 0, 1)
    u, v, st, it = os_.gauss_newton(S[0], S[1], S[2], su, sv, tol_step=1e-13, max_iter=30)
    assert np.all(st == O.CONV_STEP)
    assert np.max(np.abs(u - u0)) < 1e-12 and np.max(np.abs(v - v0)) < 1e-12


def test_gauss_newton_offset_points_orthogonal():
    sd, os_ = _surf(32, 32, 3, 3, seed=1002)
    rng = np.random.default_rng(6)
    u0, v0 = rng.uniform(0.05, 0.95, 300), rng.uniform(0.05, 0.95, 300)
    b = blocks(os_.eval(u0, v0, 1))
    # offsets well inside the radius of curvature: the foot point is (u0, v0) and both the plain
    # iteration and the exact-Hessian variant converge to it
    P, delta = synth.offset_queries(b["S"], b["Su"], b["Sv"], 1e-3, seed=1)
    for newton in (0, 1):
        u, v, st, it = os_.gauss_newton(P[0], P[1], P[2], u0 + 0.003, v0 - 0.003, tol_step=1e-13, max_iter=60,
                                        newton=newton)
        ok = st == O.CONV_STEP
        assert ok.mean() > 0.99
        assert np.max(np.abs(u[ok] - u0[ok])) < 1e-9 and np.max(np.abs(v[ok] - v0[ok])) < 1e-9
        bb = blocks(os_.eval(u[ok], v[ok], 1))
        r = bb["S"] - P[:, ok]
        assert np.max(np.abs(np.sum(r * bb["Su"], axis=0))) < 1e-12
        assert np.max(np.abs(np.sum(r * bb["Sv"], axis=0))) < 1e-12


def test_newton_variant_converges_faster_on_curved_offsets():
    # N(0, 1e-2) offsets on the bumpy 32x32 generator surface: the residual-curvature term turns
    # linear convergence into quadratic convergence
    sd, os_ = _surf(32, 32, 3, 3, seed=1002)
    rng = np.random.default_rng(6)
    u0, v0 = rng.uniform(0.05, 0.95, 400), rng.uniform(0.05, 0.95, 400)
    b = blocks(os_.eval(u0, v0, 1))
    P, delta = synth.offset_queries(b["S"], b["Su"], b["Sv"], 1e-2, seed=1)
    res = {}
    for newton in (0, 1):
        u, v, st, it = os_.gauss_newton(P[0], P[1], P[2], u0 + 0.003, v0 - 0.003, tol_step=1e-12, max_iter=100,
                                        newton=newton)
        err = np.maximum(np.abs(u - u0), np.abs(v - v0))
        res[newton] = (np.mean(err < 1e-9), np.median(it))
    assert res[1][0] >= res[0][0] and res[1][0] > 0.95
    assert res[1][1] < res[0][1]


def test_gauss_newton_reference_distance_rule_returns_current_point():
    # with the reference's rule, a seed already within tol stops before any update
    sd, os_ = _surf(8, 8, 3, 3, seed=1000)
    S = os_.eval([0.4], [0.6], 0)
    u, v, st, it = os_.gauss_newton(S[0], S[1], S[2], [0.4], [0.6], tol_dist=0.01, tol_step=0.0, max_iter=10)
    assert st[0] == O.CONV_DIST and it[0] == 0 and u[0] == 0.4 and v[0] == 0.6


def test_gauss_newton_singular_flag():
    # all control points equal: J == 0 -> singular (the reference would divide by zero)
    P = np.zeros((4, 4, 3))
    w = np.ones((4, 4))
    k = np.array([0, 0, 0, 0, 1, 1, 1, 1.0])
    os_ = O.OSurf(3, 3, k, k, P, w)
    u, v, st, it = os_.gauss_newton([1.0], [1.0], [1.0], [0.5], [0.5])
    assert st[0] == O.SINGULAR


# ---------------------------------------------------------------------------------------------
# a13-a19 tree and seeds (src/kernel/kernel.cu:375-469, 544-621, 707-733, 859-901)
# ---------------------------------------------------------------------------------------------
def test_tree_index_formula():
    assert [O.tree_base(L) for L in range(4)] == [0, 1, 65, 4161]
    assert O.tree_base(4) == 266305  # TreeArraySize, src/kernel/kernel.cuh:23


def test_need_divide_truth_table():
    assert O.need_divide([0, 0, 0], [1, 1, 0.4]) == 0     # flat in z
    assert O.need_divide([0, 0, 0], [1, 1, 0.6]) == 1
    assert O.need_divide([0, 0, 0], [0.3, 1, 1]) == 0     # flat in x
    assert O.need_divide([0, 0, 0], [1, 0.49, 1]) == 0    # flat in y
    assert O.need_divide([0, 0, 0], [1, 0.5, 0.9]) == 1   # 0.5 is not < 0.5*0.9... but 0.5 < 0.5*1 false


def test_tree_boxes_contain_dense_samples():
    sd, os_ = _surf(16, 16, 3, 3, seed=1002)
    nodes, valid, divided = os_.build_tree(3)
    assert valid[0] == 1
    rng = np.random.default_rng(1)
    base = O.tree_base(2)
    for t in rng.choice(np.arange(base, base + 4096), 200, replace=False):
        nd = nodes[t]
        uu = rng.uniform(nd[0], nd[2], 200)
        vv = rng.uniform(nd[1], nd[3], 200)
        S = os_.eval(uu, vv, 0)
        slack = 1e-12
        assert np.all(S >= nd[4:7, None] - slack) and np.all(S <= nd[7:10, None] + slack), t


def test_tree_uv_rects_and_parent_modes():
    sd, os_ = _surf(16, 16, 3, 3, seed=1002)
    a = os_.build_tree(2, 0)
    b = os_.build_tree(2, O.TREE_REF_PARENT)
    # levels 0 and 1 (reference level = 2): both parent rules coincide
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
    nodes = a[0]
    for idx in (0, 1, 9, 64):
        L = 0 if idx == 0 else 1
        g = 8 ** L
        loc = idx - O.tree_base(L)
        bx, by = divmod(loc, g)
        assert np.allclose(nodes[idx, :4], [bx / g, by / g, (bx + 1) / g, (by + 1) / g])


def test_seed_nearest_leaf_brute_force():
    sd, os_ = _surf(16, 16, 3, 3, seed=1002)
    nodes, valid, divided = os_.build_tree(2)
    rng = np.random.default_rng(2)
    P = rng.uniform(-0.2, 1.2, (3, 100))
    leaf, su, sv = O.seed_nearest_leaf(nodes, 2, P[0], P[1], P[2])
    C = nodes[1:65, 10:13]
    for i in range(100):
        d = np.sum((C - P[:, i]) ** 2, axis=1)
        assert leaf[i] == 1 + int(np.argmin(d))
        nd = nodes[leaf[i]]
        assert su[i] == (nd[0] + nd[2]) / 2 and sv[i] == (nd[1] + nd[3]) / 2


def test_first_hit_on_reference_model(golden_dir):
    # the reference's inversion chain on its own model: reference tree level 2 (levels 0, 1),
    # tolerance 0.01, 10 iterations (src/main.cu:283, 310); P is the hard-coded known answer.
    p, q, ku, kv, cp, w = read_nu(os.path.join(golden_dir, "center_bk.nu"))
    os_ = O.OSurf(p, q, ku, kv, cp, w)
    nodes, valid, divided = os_.build_tree(2)
    u, v, st, node = os_.invert_first_hit(nodes, valid, [0.0], [0.200049], [0.616], tol_dist=0.01, max_iter=10)
    assert st[0] == O.CONV_DIST
    S = os_.eval(u, v, 0)[:, 0]
    assert np.linalg.norm(S - [0.0, 0.200049, 0.616]) < 0.01


def test_seed_greedy_leaf_properties():
    # greedy descent + 3x3 refinement: a query exactly at a leaf centre gets that leaf, the result
    # is never farther than the plain descent leaf, and it agrees with the exact nearest leaf for
    # the large majority of near-surface queries
    sd, os_ = _surf(32, 32, 3, 3, seed=1002)
    nodes, valid, divided = os_.build_tree(3)
    base = O.tree_base(2)
    C = nodes[base:base + 4096, 10:13].T
    leaf, su, sv = O.seed_greedy_leaf(nodes, 3, C[0], C[1], C[2])
    assert np.mean(leaf == np.arange(base, base + 4096)) > 0.999
    u0, v0 = synth.uniform_uv(3000, 21)
    b = blocks(os_.eval(u0, v0, 1))
    P, _ = synth.offset_queries(b["S"], b["Su"], b["Sv"], 1e-2, 22)
    g_leaf, g_su, g_sv = O.seed_greedy_leaf(nodes, 3, P[0], P[1], P[2])
    e_leaf, _, _ = O.seed_nearest_leaf(nodes, 3, P[0], P[1], P[2])
    assert np.mean(g_leaf == e_leaf) > 0.95
    dg = np.sum((nodes[g_leaf, 10:13] - P.T) ** 2, axis=1)
    de = np.sum((nodes[e_leaf, 10:13] - P.T) ** 2, axis=1)
    assert np.all(dg >= de)  # the exact search is never beaten
    nd = nodes[g_leaf]
    assert np.array_equal(g_su, (nd[:, 0] + nd[:, 2]) / 2) and np.array_equal(g_sv, (nd[:, 1] + nd[:, 3]) / 2)
