"""Global projection on the GPU (GN_SEED_GLOBAL: branch and bound over the seed quadtree) against
the oracle's exhaustive form (a leaf-clamped projected Newton from every leaf), and against
properties no algorithm choice can fake: never farther than any densely sampled surface point,
never farther than the local method."""
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
