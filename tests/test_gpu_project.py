"""Global projection on the GPU (GN_SEED_GLOBAL: branch and bound over the seed quadtree, pruned by
leaf boxes and normal cones) against the oracle's exhaustive form (a leaf-clamped projected Newton
from every leaf, no pruning), and against properties no algorithm choice can fake: never farther
than any densely sampled surface point, never farther than the local method.  The leaves' normal
cones against the oracle's or_leaf_cones."""
import os

import numpy as np
import pytest

import oracle as O
import gnurbs as G
from gnurbs import synth

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


def _queries(os_, n, scale, seed):
    rng = np.random.default_rng(seed)
    u0, v0 = rng.uniform(0.02, 0.98, n), rng.uniform(0.02, 0.98, n)
    ev = os_.eval(u0, v0, 1)
    nrm = np.cross(ev[3:6].T, ev[6:9].T).T
    nrm /= np.linalg.norm(nrm, axis=0)
    return u0, v0, ev[0:3] + rng.normal(0.0, scale, n) * nrm


def _dist(os_, P, u, v):
    return np.linalg.norm(os_.eval(u, v, 0) - P, axis=0)


LEAVES_SOLVED_MAX = 3.05  # 1M-query workload: boxes alone 3.30, with cones 3.05

GLOBAL = dict(precision=G.GN_FP64, newton=G.NEWTON_FP64, seed_mode=G.SEED_GLOBAL, max_iter=20,
              tol_step=1e-12)


@pytest.fixture(scope="module")
def bumpy():
    sd = synth.surface_G(12, 12, 3, 3, seed=31)
    g = G.Surface(sd.p, sd.q, sd.ku, sd.kv, sd.cp, sd.w)
    g.BuildTree(3)
    os_ = O.OSurf.from_def(sd)
    nodes, _, _ = os_.build_tree(3)
    yield sd, g, os_, nodes
    g.close()


def test_spatial_order_leaves_every_result_unchanged(bumpy):
    """Calls of 16,384 queries or more take the queries in spatial order (k_proj_order); smaller
    calls and the no-allocation entry point take them in input order.  Each query's result is its
    own: the same bits either way, on offset, far and non-finite queries."""
    sd, g, os_, nodes = bumpy
    _, _, P = _queries(os_, 40000, 0.05, 61)
    P[:, :500] = P[:, :500] * 6.0 - 2.0
    P[0, 500] = np.nan
    opts = G.default_opts(**GLOBAL)
    u, v, st, n32, _ = (_np(t) for t in g.NewInverse(P[0], P[1], P[2], opts, iters=True))
    parts = [[_np(t) for t in g.NewInverse(P[0][a:a + 10000], P[1][a:a + 10000], P[2][a:a + 10000], opts,
                                           iters=True)] for a in range(0, 40000, 10000)]
    for k, ref in enumerate((u, v, st, n32)):
        assert np.array_equal(ref, np.concatenate([q[k] for q in parts])), k
    nu, nv, nst = (_np(t) for t in g.NewInverseNoAlloc(P[0], P[1], P[2], opts))
    assert np.array_equal(u, nu) and np.array_equal(v, nv) and np.array_equal(st, nst)


def test_global_matches_exhaustive_oracle_and_dense_samples(bumpy):
    sd, g, os_, nodes = bumpy
    _, _, P = _queries(os_, 200, 0.15, 1)
    gu, gv, gst, ncand, _ = (_np(t) for t in g.NewInverse(P[0], P[1], P[2], G.default_opts(**GLOBAL),
                                                          iters=True))
    ou, ov, ost = os_.project_global(nodes, 3, P[0], P[1], P[2], max_iter=20, leaf_iter=8, tol_step=1e-12)
    d, do = _dist(os_, P, gu, gv), _dist(os_, P, ou, ov)
    assert np.all(np.abs(d - do) <= 1e-12 + 1e-9 * do), np.max(np.abs(d - do))
    gr = np.linspace(0.0, 1.0, 301)
    U, V = np.meshgrid(gr, gr, indexing="ij")
    S = os_.eval(U.ravel(), V.ravel(), 0)
    for i in range(P.shape[1]):
        assert d[i] <= np.min(np.linalg.norm(S - P[:, i:i + 1], axis=0)) + 1e-12, i
    assert np.all(ncand >= 1) and np.mean(ncand) > 1  # pruning leaves some candidates


def test_normal_cones_match_oracle(bumpy):
    cases = [(bumpy[0], bumpy[1], 3)]
    sd5 = synth.surface_G(10, 9, 5, 4, seed=43, jitter=0.2)
    g5 = G.Surface(sd5.p, sd5.q, sd5.ku, sd5.kv, sd5.cp, sd5.w)
    g5.BuildTree(2)
    cases.append((sd5, g5, 2))
    for sd, g, levels in cases:
        got = g.GetNormalCones().astype(np.float64)
        ref = O.OSurf.from_def(sd).leaf_cones(levels)
        bo, bg = ref[:, 3] < 1.5, got[:, 3] < 1.5
        assert np.array_equal(bo, bg) and bo.mean() > 0.5
        # half-angle: the oracle's plus the 1e-6 rad margin, rounded up to FP32
        dth = got[bo, 3] - ref[bo, 3]
        assert np.all(dth >= 1e-6 - 1e-12) and np.all(dth <= 1e-6 + 2e-7), (dth.min(), dth.max())
        # axis: the oracle's, rounded to FP32 per component (~1e-7 rad)
        ga, ra = got[bo, :3], ref[bo, :3]
        ang = np.arctan2(np.linalg.norm(np.cross(ga, ra), axis=1), np.sum(ga * ra, axis=1))
        assert np.all(ang <= 5e-7), ang.max()
        assert np.all(got[~bo, 3] == 4.0)
    g5.close()


def test_cone_pruning_keeps_inner_and_outer_foot_points():
    """An exact quarter cylinder (radius 1): queries inside (radius 0.5, the foot point's normal
    points away from the query) and outside (radius 1.6).  The nearest point is radial, at the
    query's angle; the cones must pass the leaf that holds it whichever side the query is on."""
    r2 = np.sqrt(2) / 2
    cp = np.array([[[1, 0, 0], [1, 0, 2]], [[1, 1, 0], [1, 1, 2]], [[0, 1, 0], [0, 1, 2]]], dtype=float)
    w = np.array([[1, 1], [r2, r2], [1, 1]])
    g = G.Surface(2, 1, np.array([0, 0, 0, 1, 1, 1.0]), np.array([0, 0, 1, 1.0]), cp, w)
    g.BuildTree(3)
    rng = np.random.default_rng(9)
    n = 4000
    phi = rng.uniform(0.05, np.pi / 2 - 0.05, n)
    rad = np.where(rng.random(n) < 0.5, 0.5, 1.6)
    z = rng.uniform(0.1, 1.9, n)
    P = np.stack([rad * np.cos(phi), rad * np.sin(phi), z])
    gu, gv, st, ncand, _ = (_np(t) for t in g.NewInverse(P[0], P[1], P[2], G.default_opts(**GLOBAL), iters=True))
    os_ = O.OSurf(2, 1, np.array([0, 0, 0, 1, 1, 1.0]), np.array([0, 0, 1, 1.0]), cp, w)
    S = os_.eval(gu, gv, 0)
    assert np.max(np.abs(np.arctan2(S[1], S[0]) - phi)) < 1e-9
    assert np.max(np.abs(S[2] - z)) < 1e-9
    assert np.all(st == G.PT_CONVERGED)
    g.close()


def test_cone_pruning_cuts_leaf_solves():
    """The projection workload's surface and query spread (bicubic 32 x 32, offsets N(0, 1e-2) along
    the normal): results equal the exhaustive oracle's on a subset, with few leaves solved."""
    sd = synth.surface_G(32, 32, 3, 3, seed=1002)
    g = G.Surface(sd.p, sd.q, sd.ku, sd.kv, sd.cp, sd.w)
    g.BuildTree(3)
    os_ = O.OSurf.from_def(sd)
    _, _, P = _queries(os_, 20000, 1e-2, 5)
    gu, gv, st, ncand, _ = (_np(t) for t in g.NewInverse(P[0], P[1], P[2], G.default_opts(**GLOBAL), iters=True))
    nodes, _, _ = os_.build_tree(3)
    sel = np.arange(0, P.shape[1], 100)
    Ps = P[:, sel]
    ou, ov, _ = os_.project_global(nodes, 3, Ps[0].copy(), Ps[1].copy(), Ps[2].copy(), max_iter=20, leaf_iter=8,
                                   tol_step=1e-12)
    d, do = _dist(os_, Ps, gu[sel], gv[sel]), _dist(os_, Ps, ou, ov)
    assert np.all(np.abs(d - do) <= 1e-12 + 1e-9 * do), np.max(np.abs(d - do))
    print("leaves solved per query", ncand.mean())
    assert ncand.mean() < LEAVES_SOLVED_MAX
    g.close()


@pytest.mark.parametrize("name,cap", [("test_face2.nu", 2.6), ("center_bk.nu", 2.2)])
def test_cone_pruning_on_reference_models(golden_dir, name, cap):
    """The reference's CAD models, whose leaf cones are a few degrees wide: far fewer leaves solved
    (boxes alone: about 4.2 and 4.7 per query) and the exhaustive oracle's nearest points."""
    from helpers import read_nu
    p, q, ku, kv, cp, w = read_nu(os.path.join(golden_dir, name))
    g = G.Surface(p, q, ku, kv, cp, w)
    g.BuildTree(3)
    os_ = O.OSurf(p, q, ku, kv, cp, w)
    rng = np.random.default_rng(11)
    n = 20000
    u0, v0 = rng.random(n), rng.random(n)
    ev = os_.eval(u0, v0, 1)
    nrm = np.cross(ev[3:6].T, ev[6:9].T).T
    nrm /= np.linalg.norm(nrm, axis=0)
    diag = np.linalg.norm(cp.reshape(-1, 3).max(0) - cp.reshape(-1, 3).min(0))
    P = ev[0:3] + rng.normal(0.0, 1e-2 * diag, n) * nrm
    gu, gv, st, ncand, _ = (_np(t) for t in g.NewInverse(P[0], P[1], P[2], G.default_opts(**GLOBAL), iters=True))
    assert ncand.mean() < cap, ncand.mean()
    nodes, _, _ = os_.build_tree(3)
    sel = np.arange(0, n, 100)
    Ps = P[:, sel]
    ou, ov, _ = os_.project_global(nodes, 3, Ps[0].copy(), Ps[1].copy(), Ps[2].copy(), max_iter=20, leaf_iter=8,
                                   tol_step=1e-12)
    d, do = _dist(os_, Ps, gu[sel], gv[sel]), _dist(os_, Ps, ou, ov)
    assert np.all(np.abs(d - do) <= 1e-12 + 1e-9 * do), np.max(np.abs(d - do))
    g.close()


def test_global_on_a_collapsed_edge_and_non_finite_queries():
    """A control-net edge collapsed to one point (S_v = 0 there: its leaves get no bounded cone and
    are never cone-pruned) against the exhaustive oracle; queries with NaN or infinite coordinates
    end non-converged without disturbing the others."""
    sd = synth.surface_G(10, 10, 3, 3, seed=44)
    cp = sd.cp.copy()
    cp[0, :, :] = cp[0, 0, :]
    g = G.Surface(sd.p, sd.q, sd.ku, sd.kv, cp, sd.w)
    g.BuildTree(3)
    os_ = O.OSurf(sd.p, sd.q, sd.ku, sd.kv, cp, sd.w)
    cones = g.GetNormalCones()
    assert np.all(cones[:64, 3] == 4.0)  # the leaves along the collapsed edge (bx = 0)
    rng = np.random.default_rng(12)
    n = 400
    u0, v0 = rng.random(n), rng.random(n)
    ev = os_.eval(u0, v0, 0)
    P = ev + rng.normal(0.0, 0.05, (3, n))
    P[0, 5], P[1, 17], P[2, 33] = np.nan, np.inf, -np.inf
    bad = np.zeros(n, dtype=bool)
    bad[[5, 17, 33]] = True
    gu, gv, st = (_np(t) for t in g.NewInverse(P[0], P[1], P[2], G.default_opts(**GLOBAL)))
    assert np.all(st[bad] == G.PT_NONFINITE)
    # the local modes report them the same way (FP64 and mixed)
    for opts in (dict(precision=G.GN_FP64, max_iter=20, tol_step=1e-10),
                 dict(precision=G.GN_MIXED, newton=G.NEWTON_FP32 | G.NEWTON_FP64, tol_step32=1e-5,
                      fp32_iters=8, tol_quad=1e-10, seed_mode=G.SEED_GREEDY_STEP)):
        _, _, sl = (_np(t) for t in g.NewInverse(P[0], P[1], P[2], G.default_opts(**opts)))
        assert np.all(sl[bad] == G.PT_NONFINITE), (opts, sl[bad])
    nodes, _, _ = os_.build_tree(3)
    Pg = P[:, ~bad]
    ou, ov, _ = os_.project_global(nodes, 3, Pg[0].copy(), Pg[1].copy(), Pg[2].copy(), max_iter=20, leaf_iter=8,
                                   tol_step=1e-12)
    d, do = _dist(os_, Pg, gu[~bad], gv[~bad]), _dist(os_, Pg, ou, ov)
    assert np.all(np.abs(d - do) <= 1e-12 + 1e-9 * do), np.max(np.abs(d - do))
    g.close()


def test_global_never_farther_than_local_and_sometimes_nearer(bumpy):
    sd, g, os_, nodes = bumpy
    _, _, P = _queries(os_, 2000, 0.2, 2)
    gu, gv, _ = (_np(t) for t in g.NewInverse(P[0], P[1], P[2], G.default_opts(**GLOBAL)))
    local = dict(GLOBAL, seed_mode=G.SEED_NEAREST_LEAF)
    lu, lv, _ = (_np(t) for t in g.NewInverse(P[0], P[1], P[2], G.default_opts(**local)))
    d, dl = _dist(os_, P, gu, gv), _dist(os_, P, lu, lv)
    assert np.all(d <= dl + 1e-12)
    assert np.sum(d < dl - 1e-6) >= 20


def test_near_surface_queries_recover_the_foot_point():
    sd = synth.surface_G(32, 32, 3, 3, seed=1002)
    g = G.Surface(sd.p, sd.q, sd.ku, sd.kv, sd.cp, sd.w)
    g.BuildTree(3)
    os_ = O.OSurf.from_def(sd)
    u0, v0, P = _queries(os_, 5000, 1e-3, 3)
    gu, gv, st = (_np(t) for t in g.NewInverse(P[0], P[1], P[2], G.default_opts(**GLOBAL)))
    assert np.max(np.maximum(np.abs(gu - u0), np.abs(gv - v0))) < 1e-9
    assert np.all(st == G.PT_CONVERGED)
    g.close()


def test_global_argument_errors():
    sd = synth.surface_G(8, 8, 3, 3, seed=1)
    g = G.Surface(sd.p, sd.q, sd.ku, sd.kv, sd.cp, sd.w)
    z = np.zeros(4)
    with pytest.raises(G.GnError) as e:  # no tree
        g.NewInverse(z, z, z, G.default_opts(**GLOBAL))
    assert e.value.code == G.GN_ESTATE
    g.BuildTree(3)
    with pytest.raises(G.GnError) as e:
        g.NewInverse(z, z, z, G.default_opts(**dict(GLOBAL, precision=G.GN_MIXED)))
    assert e.value.code == G.GN_EINVAL
    g.close()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_fallback_resolves_unconverged_points(bumpy, prec):
    """opts.fallback: the points an inversion leaves unconverged are solved again by the global
    projection; converged points are untouched, each of the others is the nearer of its first
    result and the global projection run on it alone with the same options."""
    sd, g, os_, nodes = bumpy
    _, _, P = _queries(os_, 3000, 0.2, 4)
    if prec == "fp64":  # few iterations: many points end unconverged
        base = dict(precision=G.GN_FP64, newton=G.NEWTON_FP64, max_iter=3, tol_step=1e-12)
    else:
        base = dict(precision=G.GN_MIXED, newton=G.NEWTON_FP32 | G.NEWTON_FP64, fp32_iters=3, fp64_tail=1,
                    tol_step32=1e-4, tol_step=1e-12, tol_quad=1e-10, max_iter=3)
    u0, v0, s0 = (_np(t) for t in g.NewInverse(P[0], P[1], P[2], G.default_opts(**base)))
    u1, v1, s1 = (_np(t) for t in g.NewInverse(P[0], P[1], P[2], G.default_opts(fallback=1, **base)))
    bad = (s0 != G.PT_CONVERGED) & (s0 != G.PT_CONVERGED_DIST)
    assert bad.sum() > 50
    assert np.array_equal(u1[~bad], u0[~bad]) and np.array_equal(v1[~bad], v0[~bad])
    assert np.array_equal(s1[~bad], s0[~bad])
    glob = dict(base, precision=G.GN_FP64, seed_mode=G.SEED_GLOBAL)
    Pb = P[:, bad]
    ug, vg, sg = (_np(t) for t in g.NewInverse(Pb[0].copy(), Pb[1].copy(), Pb[2].copy(), G.default_opts(**glob)))
    first = (u1[bad] == u0[bad]) & (v1[bad] == v0[bad]) & (s1[bad] == s0[bad])
    glob_ = (u1[bad] == ug) & (v1[bad] == vg) & (s1[bad] == sg)
    assert np.all(first | glob_)
    d0, d1, dg = _dist(os_, P, u0, v0), _dist(os_, P, u1, v1), _dist(os_, Pb, ug, vg)
    assert np.all(d1 <= d0 + 1e-12)
    assert np.all(d1[bad] <= dg + 1e-12)
    assert np.sum(d1[bad] < d0[bad] - 1e-6) > 10  # the fallback found nearer points


def test_fallback_needs_the_tree():
    sd = synth.surface_G(8, 8, 3, 3, seed=1)
    g = G.Surface(sd.p, sd.q, sd.ku, sd.kv, sd.cp, sd.w)
    z = np.zeros(4)
    with pytest.raises(G.GnError) as e:
        g.NewInverse(z, z, z, G.default_opts(fallback=1), seeds=(z, z))
    assert e.value.code == G.GN_ESTATE
    g.close()
