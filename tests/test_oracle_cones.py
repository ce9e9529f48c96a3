"""The oracle's leaf normal cones (or_leaf_cones: the normal box of NormalBoundingBox and its cone,
src/kernel/kernel.cu:623-704, 1605-1636, with the normal derivatives of
src/NURBS/Evaluate.cu:1125-1135), pinned without the oracle's own evaluation:

- N = S_u x S_v and N_uu, N_uv, N_vv against scipy's B-spline basis and central differences of it;
- every normal of a dense sample grid of each leaf (scipy) lies inside the leaf's cone;
- on an exact circular cylinder the normal turns with the arc, so a cone holding a leaf's normals
  needs at least half the leaf's arc as half-angle: the cones are that tight (within 15%), which a
  K bound with a wrong power of the cell size or a box around the wrong vectors would break;
- a collapsed control-net edge (N = 0 there) gives the leaves along it no bounded cone."""
import math

import numpy as np

import oracle as O
from gnurbs import synth
from helpers import scipy_eval


def _normals(sd, u, v):
    d = scipy_eval(sd.p, sd.q, sd.ku, sd.kv, sd.cp, sd.w, u, v, order=1)
    return np.cross(d["Su"].T, d["Sv"].T)


def test_normal_derivatives_against_scipy_differences():
    sd = synth.surface_G(12, 10, 3, 4, seed=41, jitter=0.3)
    os_ = O.OSurf.from_def(sd)
    rng = np.random.default_rng(3)
    u = 0.1 + 0.8 * rng.random(200)
    v = 0.1 + 0.8 * rng.random(200)
    nd = os_.normal_ders(u, v)
    N = _normals(sd, u, v)
    assert np.max(np.abs(nd[:, 0] - N)) <= 1e-12 * np.max(np.abs(N))
    h = 1e-4
    f = lambda du, dv: _normals(sd, u + du, v + dv)
    fd_uu = (f(h, 0) - 2 * N + f(-h, 0)) / h ** 2
    fd_vv = (f(0, h) - 2 * N + f(0, -h)) / h ** 2
    fd_uv = (f(h, h) - f(h, -h) - f(-h, h) + f(-h, -h)) / (4 * h * h)
    for got, fd in ((nd[:, 1], fd_uu), (nd[:, 2], fd_uv), (nd[:, 3], fd_vv)):
        scale = np.max(np.abs(fd))
        assert np.max(np.abs(got - fd)) <= 1e-5 * scale


def _angles(axis, n):
    c = np.cross(n, axis[None])
    return np.arctan2(np.linalg.norm(c, axis=1), n @ axis)


def test_dense_normals_lie_in_their_leaf_cones():
    for sd in (synth.surface_G(12, 10, 3, 2, seed=42, jitter=0.3),
               synth.surface_G(10, 9, 5, 4, seed=43, jitter=0.2)):
        os_ = O.OSurf.from_def(sd)
        levels = 2
        g = 8 ** (levels - 1)
        cones = os_.leaf_cones(levels)
        bounded = cones[:, 3] < math.pi / 2
        # near clamped ends with short jittered spans the sampled |N_uu| (and K) is large, and
        # many edge leaves get no bounded cone; the interior ones must
        assert bounded.mean() > 0.5
        u0, u1, v0, v1 = sd.ku[sd.p], sd.ku[-sd.p - 1], sd.kv[sd.q], sd.kv[-sd.q - 1]
        t = np.linspace(0, 1, 17)
        worst = -np.inf
        for bx in range(g):
            for by in range(g):
                a, th = cones[bx * g + by, :3], cones[bx * g + by, 3]
                if th >= math.pi / 2:
                    continue
                U, V = np.meshgrid(u0 + (u1 - u0) * (bx + t) / g, v0 + (v1 - v0) * (by + t) / g, indexing="ij")
                ang = _angles(a, _normals(sd, U.ravel(), V.ravel()))
                worst = max(worst, float(np.max(ang - th)))
        assert worst <= 1e-12, worst


def _quarter_cylinder():
    r2 = math.sqrt(2) / 2
    P = np.array([[[1, 0, 0], [1, 0, 2]], [[1, 1, 0], [1, 1, 2]], [[0, 1, 0], [0, 1, 2]]], dtype=float)
    w = np.array([[1, 1], [r2, r2], [1, 1]])
    return O.OSurf(2, 1, np.array([0, 0, 0, 1, 1, 1.0]), np.array([0, 0, 1, 1.0]), P, w)


def test_cylinder_cones_are_as_tight_as_the_arc_allows():
    os_ = _quarter_cylinder()
    for levels in (2, 3):
        g = 8 ** (levels - 1)
        cones = os_.leaf_cones(levels)
        for bx in range(g):
            e = os_.eval(np.array([bx / g, (bx + 1) / g, (bx + 0.5) / g]), np.full(3, 0.5), 0)
            phi = np.arctan2(e[1], e[0])
            half = (phi[1] - phi[0]) / 2
            c = cones[bx * g:(bx + 1) * g]
            assert np.all(c[:, 3] >= half * (1 - 1e-12)) and np.all(c[:, 3] <= 1.15 * half), (levels, bx)
            # the axis is horizontal and points into the leaf's arc
            assert np.max(np.abs(c[:, 2])) < 1e-12
            ax = np.arctan2(c[:, 1], c[:, 0])
            assert np.all((ax >= phi[0]) & (ax <= phi[1]))


def test_collapsed_edge_has_no_bounded_cone():
    sd = synth.surface_G(8, 8, 3, 3, seed=44)
    cp = sd.cp.copy()
    cp[0, :, :] = cp[0, 0, :]  # the u = u0 edge collapses to one point: S_v = 0 there
    os_ = O.OSurf(sd.p, sd.q, sd.ku, sd.kv, cp, sd.w)
    g = 8
    cones = os_.leaf_cones(2)
    assert np.all(cones[:g, 3] == 4.0)  # leaves bx = 0
    assert np.mean(cones[g:, 3] < math.pi / 2) > 0.75
