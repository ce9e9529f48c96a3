"""Global projection in the oracle (or_project_global; the reference's normal-cone projection,
src/kernel/kernel.cu:471-513, 623-704, is unfinished).  Pinned by properties that do not depend on
its algorithm:
  * the closest point is never farther than any point of a dense sampling of the surface;
  * it is never farther than the local method (Newton from the nearest leaf centre), and strictly
    nearer for some far queries on a bumpy surface (several local minima);
  * for queries close to the surface the foot point is the constructed one.
"""
import numpy as np
import pytest

import oracle as O
from gnurbs import synth


def _queries(os_, n, scale, seed):
    rng = np.random.default_rng(seed)
    u0, v0 = rng.uniform(0.02, 0.98, n), rng.uniform(0.02, 0.98, n)
    ev = os_.eval(u0, v0, 1)
    nrm = np.cross(ev[3:6].T, ev[6:9].T).T
    nrm /= np.linalg.norm(nrm, axis=0)
    d = rng.normal(0.0, scale, n)
    return u0, v0, ev[0:3] + d * nrm


def _dist(os_, P, u, v):
    return np.linalg.norm(os_.eval(u, v, 0) - P, axis=0)


@pytest.fixture(scope="module")
def bumpy():
    sd = synth.surface_G(12, 12, 3, 3, seed=31)  # z noise N(0, 0.05): many local minima far away
    os_ = O.OSurf.from_def(sd)
    nodes, _, _ = os_.build_tree(3)  # 64 x 64 leaves: each leaf's piece of surface nearly flat
    return sd, os_, nodes


def test_global_not_farther_than_dense_samples(bumpy):
    sd, os_, nodes = bumpy
    _, _, P = _queries(os_, 100, 0.15, 1)
    u, v, st = os_.project_global(nodes, 3, P[0], P[1], P[2])
    d = _dist(os_, P, u, v)
    g = np.linspace(0.0, 1.0, 301)
    U, V = np.meshgrid(g, g, indexing="ij")
    S = os_.eval(U.ravel(), V.ravel(), 0)
    for i in range(P.shape[1]):
        dmin = np.min(np.linalg.norm(S - P[:, i:i + 1], axis=0))
        assert d[i] <= dmin + 1e-12, (i, d[i], dmin)


def test_global_not_farther_than_local_and_sometimes_nearer(bumpy):
    sd, os_, nodes = bumpy
    _, _, P = _queries(os_, 150, 0.2, 2)
    u, v, st = os_.project_global(nodes, 3, P[0], P[1], P[2])
    _, su, sv = O.seed_nearest_leaf(nodes, 3, P[0], P[1], P[2])
    lu, lv, _, _ = os_.gauss_newton(P[0], P[1], P[2], su, sv, tol_step=1e-12, max_iter=20, newton=1)
    d, dl = _dist(os_, P, u, v), _dist(os_, P, lu, lv)
    assert np.all(d <= dl + 1e-12)
    assert np.sum(d < dl - 1e-6) >= 3  # the local method stopped in another local minimum


def test_near_surface_queries_find_the_constructed_foot_point():
    sd = synth.surface_G(10, 10, 3, 3, seed=7)
    os_ = O.OSurf.from_def(sd)
    nodes, _, _ = os_.build_tree(3)
    u0, v0, P = _queries(os_, 60, 1e-3, 3)
    u, v, st = os_.project_global(nodes, 3, P[0], P[1], P[2])
    assert np.max(np.maximum(np.abs(u - u0), np.abs(v - v0))) < 1e-9
    assert np.all(st == 0)
