"""GPU parity: the HIP kernels (through the C ABI of libgnurbs.so) against the CPU FP64 oracle.

Tolerances (from the project's correctness contract):
  span indices, tree validity/divided flags, seed leaf indices ..... bit-exact
  FP64 evaluation ................................................. 1e-12 of the field scale
  FP32 evaluation (per-point and MFMA grid) ....................... 1e-5 of the field scale
  inversion (u, v) ................................................ 1e-9 wherever both converge
"field scale" = max(|reference| over the field, 1e-2 * max|S|) (helpers.field_close).
"""
import os

import numpy as np
import pytest

import oracle as O
import gnurbs as G
from gnurbs import synth
from helpers import blocks, field_close, read_nu

pytestmark = pytest.mark.gpu

FIELDS = ("S", "Su", "Sv", "Suu", "Suv", "Svv", "Suuu", "Suuv", "Suvv", "Svvv")


def _np(t):
    return t.detach().cpu().numpy()


def _pair(sd):
    return G.Surface(sd.p, sd.q, sd.ku, sd.kv, sd.cp, sd.w), O.OSurf.from_def(sd)


SURFS = [
    ("bicubic8", lambda: synth.surface_G(8, 8, 3, 3, seed=1000)),
    ("bicubic64", lambda: synth.surface_G(64, 64, 3, 3, seed=1001)),
    ("p2q4_jitter", lambda: synth.surface_G(12, 9, 2, 4, seed=7, jitter=0.4)),
    ("p5q5_jitter", lambda: synth.surface_G(20, 17, 5, 5, seed=8, jitter=0.4)),
    ("p1q3", lambda: synth.surface_G(5, 7, 1, 3, seed=9)),
    ("p6q2", lambda: synth.surface_G(9, 8, 6, 2, seed=10)),
]


def _edge_uv(sd, n, seed):
    rng = np.random.default_rng(seed)
    u, v = rng.random(n), rng.random(n)
    ku, kv = np.unique(sd.ku), np.unique(sd.kv)
    m = min(len(ku), len(kv))
    u[:m], v[:m] = ku[:m], kv[:m]          # exactly on knots, including 0 and 1
    u[m:m + 4] = [-0.25, 1.25, 0.0, 1.0]   # outside the domain and the ends
    v[m:m + 4] = [1.0, 0.0, -0.5, 1.5]
    return u, v


def test_library_version():
    assert b"gfx950" in G.lib().gn_version()


@pytest.mark.parametrize("name,mk", SURFS)
def test_find_span_bit_exact(name, mk):
    sd = mk()
    gs, _ = _pair(sd)
    u, v = _edge_uv(sd, 4096, 1)
    u = np.concatenate([u, np.nextafter(sd.ku, 2.0), np.nextafter(sd.ku, -1.0)])
    v = np.concatenate([v, np.nextafter(sd.kv, 2.0), np.nextafter(sd.kv, -1.0)])
    for d, (x, k, deg, n) in enumerate(((u, sd.ku, sd.p, sd.nu), (v, sd.kv, sd.q, sd.nv))):
        got = _np(gs.findSpan(d, x))
        ref = np.array([O.find_span(deg, n, k, float(a)) for a in x], dtype=np.int32)
        assert np.array_equal(got, ref), name


def test_find_span_nan_and_empty():
    sd = synth.surface_G(8, 8, 3, 3, seed=1000)
    gs, _ = _pair(sd)
    got = _np(gs.findSpan(0, np.array([np.nan, np.inf, -np.inf])))
    assert list(got) == [O.find_span(3, 8, sd.ku, x) for x in (np.nan, np.inf, -np.inf)]
    assert gs.findSpan(0, np.zeros(0)).numel() == 0


@pytest.mark.parametrize("name,mk", SURFS)
def test_span_matrices(name, mk):
    sd = mk()
    gs, os_ = _pair(sd)
    scale = np.max(np.abs(os_.MatrixN))
    for i in range(sd.p, sd.nu):
        for j in range(sd.q, sd.nv):
            if not os_.valid[i * (sd.q + sd.nv) + j]:
                continue
            a = gs.GetMatrixN(i, j)
            b = os_.span_matrix(i, j)
            assert np.max(np.abs(a - b)) <= 1e-13 * scale, (name, i, j)


@pytest.mark.parametrize("order", [0, 1, 2, 3])
@pytest.mark.parametrize("name,mk", SURFS)
def test_eval_fp64(name, mk, order):
    sd = mk()
    gs, os_ = _pair(sd)
    u, v = _edge_uv(sd, 20000, 2)
    got = blocks(_np(gs.Evaluate(u, v, order)))
    ref = blocks(os_.eval(u, v, order))
    floor = 1e-2 * np.max(np.abs(ref["S"]))
    for key in FIELDS[: (1, 3, 6, 10)[order]]:
        ok, err = field_close(got[key], ref[key], 1e-12, floor)
        assert ok, (name, key, err)


def test_eval_repeated_interior_knots_at_and_beside_knots():
    """Interior knots of multiplicity 2 and 3 (empty spans between the copies) and non-uniform
    spacing: at every knot and one ulp either side, every field to order 3 matches the oracle,
    whose span search is the reference's (the kernel's search starts from a uniform-knot guess)."""
    p = q = 3
    ku = np.r_[[0.0] * 4, 0.1, 0.3, 0.3, 0.55, 0.55, 0.55, 0.8, [1.0] * 4]
    kv = np.r_[[0.0] * 4, 0.25, 0.25, 0.6, 0.9, [1.0] * 4]
    nu, nv = len(ku) - p - 1, len(kv) - q - 1
    rng = np.random.default_rng(31)
    X, Y = np.meshgrid(np.linspace(0, 1, nu), np.linspace(0, 1, nv), indexing="ij")
    cp = np.stack([X, Y, 0.2 * rng.standard_normal((nu, nv))], axis=-1)
    w = rng.uniform(0.5, 2.0, (nu, nv))
    sd = synth.SurfaceDef(p, q, ku, kv, cp, w)
    gs, os_ = _pair(sd)
    pts = []
    for k in np.unique(np.r_[ku, kv]):
        pts += [np.nextafter(k, -np.inf), k, np.nextafter(k, np.inf)]
    pts = np.array(pts)
    U, V = np.meshgrid(pts, pts, indexing="ij")
    u, v = U.ravel(), V.ravel()
    got = blocks(_np(gs.Evaluate(u, v, 3)))
    ref = blocks(os_.eval(u, v, 3))
    floor = 1e-2 * np.max(np.abs(ref["S"]))
    for key in FIELDS:
        ok, err = field_close(got[key], ref[key], 1e-12, floor)
        assert ok, (key, err)
    gs.close()


@pytest.mark.parametrize("order", [2, 3])
@pytest.mark.parametrize("name,mk", SURFS)
def test_eval_fp32(name, mk, order):
    sd = mk()
    gs, os_ = _pair(sd)
    u, v = _edge_uv(sd, 20000, 3)
    u32, v32 = u.astype(np.float32), v.astype(np.float32)
    got = blocks(_np(gs.Evaluate(u32, v32, order, precision=G.GN_FP32)).astype(np.float64))
    ref = blocks(os_.eval(u32.astype(np.float64), v32.astype(np.float64), order))
    floor = 1e-2 * np.max(np.abs(ref["S"]))
    for key in FIELDS[: (1, 3, 6, 10)[order]]:
        ok, err = field_close(got[key], ref[key], 1e-5, floor)
        assert ok, (name, key, err)


@pytest.mark.parametrize("name,mk,nu_pts,nv_pts", [
    ("bicubic8", SURFS[0][1], 256, 200),
    ("bicubic64", SURFS[1][1], 333, 517),
    ("p5q5_jitter", SURFS[3][1], 130, 97),
    ("p6q2", SURFS[5][1], 64, 50),
    ("sparse", SURFS[1][1], 40, 30),  # fewer than 4 samples per span: per-point fallback
])
def test_eval_grid_mfma(name, mk, nu_pts, nv_pts):
    sd = mk()
    gs, os_ = _pair(sd)
    us = np.linspace(0, 1, nu_pts).astype(np.float32)
    vs = np.sort(np.random.default_rng(5).random(nv_pts)).astype(np.float32)
    for order in (0, 2):
        got = blocks(_np(gs.EvaluateGrid(us, vs, order)).astype(np.float64))
        U, V = np.meshgrid(us.astype(np.float64), vs.astype(np.float64), indexing="ij")
        ref = blocks(os_.eval(U.ravel(), V.ravel(), order))
        floor = 1e-2 * np.max(np.abs(ref["S"]))
        for key in FIELDS[: (1, 3, 6)[order]]:
            ok, err = field_close(got[key], ref[key], 1e-5, floor)
            assert ok, (name, order, key, err)


@pytest.mark.parametrize("flags", [0, G.TREE_REF_PARENT])
@pytest.mark.parametrize("name,mk", [SURFS[0], SURFS[2], SURFS[5]])
def test_tree_parity(name, mk, flags):
    sd = mk()
    gs, os_ = _pair(sd)
    gs.BuildTree(3, flags)
    gn, gv, gd = gs.GetTree()
    on, ov, od = os_.build_tree(3, flags)
    assert np.array_equal(gn[:, :4], on[:, :4])  # uv rectangles exactly
    scale = np.max(np.abs(on[:, 4:]))
    assert np.max(np.abs(gn[:, 4:] - on[:, 4:])) <= 1e-12 * scale
    assert np.array_equal(gv, ov) and np.array_equal(gd, od)


def test_tree_reference_model(golden_dir):
    p, q, ku, kv, cp, w = read_nu(os.path.join(golden_dir, "center_bk.nu"))
    gs = G.Surface.readNu(os.path.join(golden_dir, "center_bk.nu"))
    os_ = O.OSurf(p, q, ku, kv, cp, w)
    gs.BuildTree(2)
    gn, gv, gd = gs.GetTree()
    on, ov, od = os_.build_tree(2)
    assert np.array_equal(gv, ov) and np.array_equal(gd, od)
    assert np.max(np.abs(gn - on)) <= 1e-12 * np.max(np.abs(on))


def test_seed_nearest_leaf_matches_brute_force():
    sd = synth.surface_G(32, 32, 3, 3, seed=1002)
    gs, os_ = _pair(sd)
    gs.BuildTree(3)
    nodes, _, _ = gs.GetTree()
    rng = np.random.default_rng(4)
    n = 3000
    P = np.concatenate([rng.uniform(-0.3, 1.3, (3, n // 2)), os_.eval(*synth.uniform_uv(n - n // 2, 8), 0)], axis=1)
    leaf, su, sv = (_np(t) for t in gs.Seed(P[0], P[1], P[2], exact=True))
    ref_leaf, ref_su, ref_sv = O.seed_nearest_leaf(nodes, 3, P[0], P[1], P[2])
    C = nodes[:, 10:13]
    d_got = np.sum((C[leaf] - P.T) ** 2, axis=1)
    d_ref = np.sum((C[ref_leaf] - P.T) ** 2, axis=1)
    same = leaf == ref_leaf
    # a different leaf is only allowed on a tie up to rounding
    assert np.all(same | (np.abs(d_got - d_ref) <= 1e-12 * np.maximum(d_ref, 1e-30)))
    assert same.mean() > 0.999
    assert np.array_equal(su[same], ref_su[same]) and np.array_equal(sv[same], ref_sv[same])


def test_seed_greedy_leaf_matches_oracle():
    sd = synth.surface_G(32, 32, 3, 3, seed=1002)
    gs, os_ = _pair(sd)
    gs.BuildTree(3)
    nodes, _, _ = gs.GetTree()
    u0, v0 = synth.uniform_uv(20000, 31)
    b = blocks(os_.eval(u0, v0, 1))
    P, _ = synth.offset_queries(b["S"], b["Su"], b["Sv"], 1e-2, 32)
    leaf, su, sv = (_np(t) for t in gs.Seed(P[0], P[1], P[2], exact=False))
    ref_leaf, ref_su, ref_sv = O.seed_greedy_leaf(nodes, 3, P[0], P[1], P[2])
    same = leaf == ref_leaf
    # identical except where FP32 box distances of the descent tie within rounding
    assert same.mean() > 0.999
    assert np.array_equal(su[same], ref_su[same]) and np.array_equal(sv[same], ref_sv[same])


def _projection_case(n, sigma, seed=1002):
    sd = synth.surface_G(32, 32, 3, 3, seed=seed)
    gs, os_ = _pair(sd)
    u0, v0 = synth.uniform_uv(n, seed + 1)
    b = blocks(os_.eval(u0, v0, 1))
    P, delta = synth.offset_queries(b["S"], b["Su"], b["Sv"], sigma, seed + 2)
    return sd, gs, os_, u0, v0, P


@pytest.mark.parametrize("newton", [0, G.NEWTON_FP64])
def test_invert_fp64_matches_oracle_given_seeds(newton):
    sd, gs, os_, u0, v0, P = _projection_case(4000, 1e-2)
    rng = np.random.default_rng(3)
    su = np.clip(u0 + rng.uniform(-0.01, 0.01, u0.size), 0, 1)
    sv = np.clip(v0 + rng.uniform(-0.01, 0.01, v0.size), 0, 1)
    opts = G.default_opts(max_iter=20, tol_step=1e-10, newton=newton)
    u, v, st = (_np(t) for t in gs.NewInverse(P[0], P[1], P[2], opts, seeds=(su, sv)))
    ru, rv, rst, rit = os_.gauss_newton(P[0], P[1], P[2], su, sv, tol_step=1e-10, max_iter=20,
                                        newton=1 if newton else 0)
    both = (st == G.PT_CONVERGED) & (rst == O.CONV_STEP)
    assert both.mean() > (0.9 if newton else 0.6)
    assert np.max(np.abs(u[both] - ru[both])) < 1e-9 and np.max(np.abs(v[both] - rv[both])) < 1e-9
    assert (st == rst).mean() > 0.99


def test_invert_mixed_reaches_1e9():
    sd, gs, os_, u0, v0, P = _projection_case(20000, 1e-2)
    gs.BuildTree(3)
    opts = G.default_opts(precision=G.GN_MIXED, newton=G.NEWTON_FP32 | G.NEWTON_FP64, fp32_iters=20, fp64_tail=3)
    u, v, st = (_np(t) for t in gs.NewInverse(P[0], P[1], P[2], opts))
    err = np.maximum(np.abs(u - u0), np.abs(v - v0))
    assert np.mean(err <= 1e-9) > 0.9
    # against the FP64 oracle (exact Hessian) from the same greedy-leaf seeds, where both converge
    nodes, _, _ = gs.GetTree()
    _, ssu, ssv = O.seed_greedy_leaf(nodes, 3, P[0], P[1], P[2])
    ru, rv, rst, _ = os_.gauss_newton(P[0], P[1], P[2], ssu, ssv, tol_step=1e-12, max_iter=40, newton=1)
    both = (st == G.PT_CONVERGED) & (rst == O.CONV_STEP)
    d = np.maximum(np.abs(u - ru), np.abs(v - rv))[both]
    same = d < 1e-6  # the few others reached another local minimum of the distance from the seed
    assert both.mean() > 0.95 and same.mean() > 0.99
    assert np.max(d[same]) <= 1e-9


@pytest.mark.parametrize("newton", [0, G.NEWTON_FP32, G.NEWTON_FP64, G.NEWTON_FP32 | G.NEWTON_FP64])
def test_invert_mixed_newton_flags_match_oracle(newton):
    """Every combination of exact-Hessian steps in the FP32 bulk and the FP64 closing pass (the
    streamed first closing step and its listed rest in both forms) reaches the oracle's foot points
    from the same greedy-leaf seeds."""
    sd, gs, os_, u0, v0, P = _projection_case(20000, 1e-2)
    gs.BuildTree(3)
    opts = G.default_opts(precision=G.GN_MIXED, newton=newton, fp32_iters=20, fp64_tail=20,
                          tol_step32=1e-4, tol_step=1e-12,
                          tol_quad=1e-10 if newton & G.NEWTON_FP64 else 0.0)
    u, v, st = (_np(t) for t in gs.NewInverse(P[0], P[1], P[2], opts))
    nodes, _, _ = gs.GetTree()
    _, ssu, ssv = O.seed_greedy_leaf(nodes, 3, P[0], P[1], P[2])
    ru, rv, rst, _ = os_.gauss_newton(P[0], P[1], P[2], ssu, ssv, tol_step=1e-12, max_iter=40, newton=1)
    both = (st == G.PT_CONVERGED) & (rst == O.CONV_STEP)
    assert both.mean() > (0.95 if newton & G.NEWTON_FP64 else 0.6)
    d = np.maximum(np.abs(u - ru), np.abs(v - rv))[both]
    same = d < 1e-6  # the few others reached another local minimum of the distance from the seed
    assert same.mean() > 0.99
    assert np.max(d[same]) <= 1e-9


def test_invert_plain_nearest_leaf_on_surface_round_trip():
    sd = synth.surface_G(32, 32, 3, 3, seed=1002)
    gs, os_ = _pair(sd)
    gs.BuildTree(3)
    u0, v0 = synth.uniform_uv(5000, 11)
    S = os_.eval(u0, v0, 0)
    u, v, st = (_np(t) for t in gs.NewInverse(S[0], S[1], S[2], G.default_opts(max_iter=30)))
    err = np.maximum(np.abs(u - u0), np.abs(v - v0))
    assert np.mean(err <= 1e-9) > 0.97


def test_first_hit_reference_chain(golden_dir):
    # NewInverse on the reference's own model and query point (src/main.cu:134, 283, 310;
    # src/kernel/kernel.cu:454): tree level 2, tol 0.01, 10 iterations, no clamp
    path = os.path.join(golden_dir, "center_bk.nu")
    p, q, ku, kv, cp, w = read_nu(path)
    gs = G.Surface.readNu(path)
    os_ = O.OSurf(p, q, ku, kv, cp, w)
    gs.BuildTree(2)
    nodes, valid, _ = gs.GetTree()
    P = np.array([[0.0], [0.200049], [0.616]])
    P = np.repeat(P, 512, axis=1)
    opts = G.default_opts(seed_mode=G.SEED_REF_FIRST_HIT, tol_dist=0.01, tol_step=0.0, max_iter=10, clamp=0)
    u, v, st = (_np(t) for t in gs.NewInverse(P[0], P[1], P[2], opts))
    ru, rv, rst, _ = os_.invert_first_hit(nodes, valid, P[0], P[1], P[2], tol_dist=0.01, max_iter=10)
    assert np.all(st == rst) and np.all(st == G.PT_CONVERGED_DIST)
    assert np.max(np.abs(u - ru)) < 1e-12 and np.max(np.abs(v - rv)) < 1e-12


def test_readnu_matches_tuple_constructor(golden_dir):
    for fname in ("center_bk.nu", "part3_0.nu", "part3_2.nu", "test_face2.nu", "test.nu"):
        path = os.path.join(golden_dir, fname)
        p, q, ku, kv, cp, w = read_nu(path)
        a = G.Surface.readNu(path)
        b = G.Surface(p, q, ku, kv, cp, w)
        u, v = synth.uniform_uv(1000, 1)
        assert np.array_equal(_np(a.Evaluate(u, v, 2)), _np(b.Evaluate(u, v, 2))), fname


def test_known_answer_points_on_gpu(golden_dir):
    for fname, expect in (("center_bk.nu", (0.0, 0.200049, 0.616)), ("test_face2.nu", (0.7, 0.455556, 0.055556))):
        s = G.Surface.readNu(os.path.join(golden_dir, fname))
        S = _np(s.Evaluate(np.array([0.25]), np.array([0.25]), 0))[:, 0]
        assert np.max(np.abs(S - np.array(expect))) < 5e-7


def test_errors_are_reported_not_fatal():
    sd = synth.surface_G(8, 8, 3, 3, seed=1000)
    with pytest.raises(G.GnError) as e:
        G.Surface(9, 3, sd.ku, sd.kv, sd.cp, sd.w)
    assert e.value.code == G.GN_EINVAL
    gs = G.Surface(sd.p, sd.q, sd.ku, sd.kv, sd.cp, sd.w)
    with pytest.raises(G.GnError) as e:
        gs.NewInverse(np.zeros(4), np.zeros(4), np.zeros(4))
    assert e.value.code == G.GN_ESTATE
    with pytest.raises(G.GnError):
        gs.Evaluate(np.zeros(4), np.zeros(4), 4)
    with pytest.raises(G.GnError):
        G.Surface.readNu("/nonexistent.nu")
    # empty batches are no-ops
    assert gs.Evaluate(np.zeros(0), np.zeros(0), 2).shape == (18, 0)


@pytest.mark.parametrize("precision,newton", [(G.GN_FP64, 0), (G.GN_MIXED, G.NEWTON_FP32 | G.NEWTON_FP64)])
def test_result_independent_of_batch_and_order(precision, newton):
    # a large batch is processed in spatially sorted order with lane refill; a small one in index
    # order: every point's result must be bit-identical either way
    sd, gs, os_, u0, v0, P = _projection_case(30000, 1e-2)
    gs.BuildTree(3)
    opts = G.default_opts(precision=precision, newton=newton)
    big = [_np(t) for t in gs.NewInverse(P[0], P[1], P[2], opts)]
    sel = np.random.default_rng(12).choice(30000, 3000, replace=False)
    small = [_np(t) for t in gs.NewInverse(P[0][sel], P[1][sel], P[2][sel], G.default_opts(precision=precision,
                                                                                           newton=newton))]
    for a, b in zip(big, small):
        assert np.array_equal(a[sel], b)


@pytest.mark.parametrize("precision,newton", [(G.GN_FP64, G.NEWTON_FP64), (G.GN_MIXED, G.NEWTON_FP32 | G.NEWTON_FP64)])
def test_ragged_sizes_match_the_full_run(precision, newton):
    """Calls of 1, 63, 64, 65, 257 and 4,097 points (partial 64-point tiles, a grid of one
    workgroup, one wave holding a few points) give each point the bits it gets in a call of
    30,000; an empty call is a no-op."""
    sd, gs, os_, u0, v0, P = _projection_case(30000, 1e-2)
    gs.BuildTree(3)
    opts = G.default_opts(precision=precision, newton=newton, tol_quad=1e-10)
    big = [_np(t) for t in gs.NewInverse(P[0], P[1], P[2], opts, iters=True)]
    rng = np.random.default_rng(21)
    for n in (1, 63, 64, 65, 257, 4097):
        sel = rng.choice(30000, n, replace=False)
        small = [_np(t) for t in gs.NewInverse(P[0][sel], P[1][sel], P[2][sel], opts, iters=True)]
        for a, b in zip(big, small):
            assert np.array_equal(a[sel], b), n
    e = gs.NewInverse(np.zeros(0), np.zeros(0), np.zeros(0), opts)
    assert all(t.numel() == 0 for t in e)


def test_long_knot_vectors_take_the_global_knot_paths():
    """Knot vectors too long for LDS (603 + 1003 knots): evaluation falls back from the
    LDS-staged kernel and inversion from the exact-degree kernels to the generic kernels that
    read knots through L1; results keep the FP64 and inversion tolerances."""
    nu, nv = 40, 1000  # smooth rational net (no per-point noise, which would be 1000 bumps)
    X, Y = np.meshgrid(np.arange(nu) / (nu - 1), np.arange(nv) / (nv - 1), indexing="ij")
    cp = np.stack([X, Y, 0.2 * np.sin(2 * np.pi * X) * np.cos(2 * np.pi * Y)], axis=-1)
    w = 1.0 + 0.3 * np.sin(3 * X + 5 * Y)
    sd = synth.SurfaceDef(2, 2, synth.clamped_knots(nu, 2), synth.clamped_knots(nv, 2), cp, w)
    g, o = _pair(sd)
    rng = np.random.default_rng(3)
    u, v = rng.random(20000), rng.random(20000)
    got = blocks(_np(g.Evaluate(u, v, 2)))
    ref = blocks(o.eval(u, v, 2))
    floor = 1e-2 * np.max(np.abs(ref["S"]))
    for f in FIELDS[:6]:
        # second derivatives of a 998-span net carry chain factors ~1e6 and a cancelling quotient
        # rule: the two FP64 computations (different basis-matrix rounding) agree to ~2e-12 there
        tol = 1e-12 if f in ("S", "Su", "Sv") else 1e-11
        ok, err = field_close(got[f], ref[f], tol, floor)
        assert ok, (f, err)
    P = o.eval(u[:5000], v[:5000], 0)
    g.BuildTree(3)
    for prec in (G.GN_FP64, G.GN_MIXED):
        opts = G.default_opts(precision=prec, newton=G.NEWTON_FP32 | G.NEWTON_FP64, tol_quad=1e-10,
                              tol_step32=1e-4)
        gu, gv, st = (_np(t) for t in g.NewInverse(P[0], P[1], P[2], opts))
        err = np.maximum(np.abs(gu - u[:5000]), np.abs(gv - v[:5000]))
        assert np.mean(err <= 1e-9) > 0.97, prec


@pytest.mark.parametrize("prec", ["fp64", "mixed_tail0"])
def test_zero_iteration_limit_returns_the_start_point(prec):
    """A lane that may take no step keeps its parameters, also in the kernels whose whole wave
    steps together (LDS-staged patch rows): max_iter = 0 returns the given seeds exactly, and a
    mixed run with no FP64 closing steps returns its FP32 result."""
    sd = synth.surface_G(32, 32, 3, 3, seed=1002)
    g, o = _pair(sd)
    rng = np.random.default_rng(12)
    n = 5000
    u0, v0 = rng.random(n), rng.random(n)
    P = o.eval(u0, v0, 0)
    su, sv = np.clip(u0 + 0.01, 0, 1), np.clip(v0 - 0.01, 0, 1)
    if prec == "fp64":
        opts = G.default_opts(precision=G.GN_FP64, newton=G.NEWTON_FP64, max_iter=0)
        gu, gv, st = (_np(t) for t in g.NewInverse(P[0], P[1], P[2], opts, seeds=(su, sv)))
        assert np.array_equal(gu, su) and np.array_equal(gv, sv)
        assert np.all(st == G.PT_MAXITER)
    else:
        kw = dict(precision=G.GN_MIXED, newton=G.NEWTON_FP32 | G.NEWTON_FP64, fp32_iters=20,
                  tol_step32=1e-4)
        g32 = G.default_opts(fp64_tail=0, **kw)
        gu, gv, st = (_np(t) for t in g.NewInverse(P[0], P[1], P[2], g32, seeds=(su, sv)))
        full = G.default_opts(fp64_tail=3, tol_quad=1e-10, **kw)
        fu, fv, _ = (_np(t) for t in g.NewInverse(P[0], P[1], P[2], full, seeds=(su, sv)))
        err = np.maximum(np.abs(gu - u0), np.abs(gv - v0))
        ferr = np.maximum(np.abs(fu - u0), np.abs(fv - v0))
        # without closing steps the result keeps the FP32 bulk's error (a single FP64 step would
        # take it to ~1e-12); with them it reaches the FP64 tolerance
        assert np.percentile(err, 99) < 1e-3 and np.median(err) > 1e-10
        assert np.median(ferr) < 1e-12
        assert np.mean(st == G.PT_MAXITER) > 0.99  # points whose FP32 bulk broke down restart


def test_seed_greedy_leaf_independent_of_batch_size():
    """A query's greedy-leaf seed does not depend on the other queries of the call: prefixes of
    ragged lengths (where threads carry an odd number of points, or none past n) give the
    prefix of the full call's leaves and seeds, bit for bit."""
    import torch
    sd = synth.surface_G(32, 32, 3, 3, seed=1002)
    gs = G.Surface(sd.p, sd.q, sd.ku, sd.kv, sd.cp, sd.w)
    gs.BuildTree(3)  # depth 6: levels past the ones staged in LDS are read from the tree
    N = 600001
    u0, v0 = synth.uniform_uv(N, seed=33)
    ev = gs.Evaluate(u0, v0, 1)
    nrm = torch.linalg.cross(ev[3:6].T, ev[6:9].T).T
    nrm = nrm / torch.linalg.norm(nrm, dim=0, keepdim=True)
    delta = torch.from_numpy(np.random.default_rng(34).normal(0.0, 1e-2, N)).cuda()
    P = (ev[0:3] + delta * nrm).contiguous()
    full = gs.Seed(P[0].contiguous(), P[1].contiguous(), P[2].contiguous(), exact=False)
    for m in (1, 3, 255, 257, 4097, 262143, 262145, 524289, 524290):
        part = gs.Seed(P[0, :m].contiguous(), P[1, :m].contiguous(), P[2, :m].contiguous(), exact=False)
        assert all(torch.equal(a, b[:m]) for a, b in zip(part, full)), m
    gs.close()


@pytest.mark.parametrize("precision", [G.GN_FP64, G.GN_MIXED])
def test_reference_cpu_case_4096_on_surface_inversions(precision):
    """The reference's CPU-runnable case: bicubic rational 8x8 net (generator G, seed 1000),
    4,096 points sampled exactly on the surface at uniform (u*, v*), inverted from greedy-leaf
    seeds on a depth-6 tree.  Plain FP64 (the reference iteration) equals the oracle's same
    iteration from the oracle's seeds within 1e-9 with the same status; every converged point of
    either mode is within 1e-9 of (u*, v*)."""
    sd = synth.surface_G(8, 8, 3, 3, seed=1000)
    gs, os_ = _pair(sd)
    gs.BuildTree(3)
    u0, v0 = synth.uniform_uv(4096, seed=1001)
    S = os_.eval(u0, v0, 0)
    if precision == G.GN_FP64:
        opts = G.default_opts(precision=G.GN_FP64, newton=0, max_iter=20, tol_step=1e-10)
    else:
        opts = G.default_opts(precision=G.GN_MIXED, newton=G.NEWTON_FP32 | G.NEWTON_FP64, tol_step32=1e-4)
    u, v, st = (_np(t) for t in gs.NewInverse(S[0], S[1], S[2], opts))
    conv = st == G.PT_CONVERGED
    assert conv.mean() > 0.99
    err = np.maximum(np.abs(u - u0), np.abs(v - v0))
    assert np.max(err[conv]) <= 1e-9
    if precision == G.GN_FP64:
        nodes, _, _ = gs.GetTree()
        leaf, ssu, ssv = O.seed_greedy_leaf(nodes, 3, S[0], S[1], S[2])
        gleaf, gsu, gsv = (_np(t) for t in gs.Seed(S[0], S[1], S[2], exact=False))
        assert np.array_equal(leaf, gleaf) and np.array_equal(ssu, gsu) and np.array_equal(ssv, gsv)
        ru, rv, rst, _ = os_.gauss_newton(S[0], S[1], S[2], ssu, ssv, tol_step=1e-10, max_iter=20)
        assert np.array_equal(st, rst)
        assert np.max(np.maximum(np.abs(u - ru), np.abs(v - rv))[conv]) <= 1e-9
