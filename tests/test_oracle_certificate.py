"""The Newton-step certificate (gn_invert_opts.tol_quad) is honest: a point it stops is within
tol_quad of the local minimum it converges to.

The certificate is the library's own stopping rule (the reference stops on a distance tolerance
only, src/kernel/kernel.cu:177).  The oracle restates it (or_gn_opts.tol_quad).  Here every point
the oracle certifies is compared with the minimiser reached by 60 further Newton iterations from
it (rounding-level convergence), over every degree class the batched kernels dispatch (p, q in
2..5 with jittered knots, plus degrees 1 and 6 that take the generic kernel), at offsets up to a
few curvature radii (flat minima: small eigenvalue of the Hessian), and on a closed form: an
exact circular cylinder whose queries sit near the axis, where the distance is nearly constant
along the circle and the foot point is known exactly.
"""
import math

import numpy as np
import pytest

import oracle as O
from gnurbs import synth
from helpers import blocks

TOL = 1e-10

CASES = [(p, q) for p in range(2, 6) for q in range(2, 6)] + [(1, 3), (6, 2)]


def _case(p, q, n, sigma, seed):
    nu, nv = 2 * p + 5, 2 * q + 4
    sd = synth.surface_G(nu, nv, p, q, seed=seed, jitter=0.4)
    os_ = O.OSurf.from_def(sd)
    u0, v0 = synth.uniform_uv(n, seed + 1)
    b = blocks(os_.eval(u0, v0, 1))
    P, _ = synth.offset_queries(b["S"], b["Su"], b["Sv"], sigma, seed + 2)
    rng = np.random.default_rng(seed + 3)
    su = np.clip(u0 + rng.uniform(-0.02, 0.02, n), 0, 1)
    sv = np.clip(v0 + rng.uniform(-0.02, 0.02, n), 0, 1)
    return os_, P, su, sv


@pytest.mark.parametrize("p,q", CASES)
def test_certified_points_are_within_tol_quad(p, q):
    os_, P, su, sv = _case(p, q, 3000, 2e-2, 5000 + 10 * p + q)
    u, v, st, it = os_.gauss_newton(P[0], P[1], P[2], su, sv, tol_step=0.0, max_iter=40, newton=1,
                                    tol_quad=TOL)
    cert = st == O.CONV_STEP
    # the certificate does stop points (most within a few steps), so the check below has content
    assert cert.mean() > 0.8 and np.median(it[cert]) <= 6
    tu, tv, _, _ = os_.gauss_newton(P[0], P[1], P[2], u, v, tol_step=0.0, max_iter=60, newton=1)
    err = np.maximum(np.abs(u - tu), np.abs(v - tv))[cert]
    assert np.max(err) <= TOL, (p, q, float(np.max(err)))


def test_certificate_on_flat_minima_of_a_cylinder():
    """Exact circular cylinder of radius 1 (quarter circle with w = sqrt(2)/2, swept in v);
    queries on the inward normal at depth 1 - r for r down to 1e-3: the distance along the circle
    varies by r (1 - cos) only, so the Hessian's smallest eigenvalue is ~ r |S_u|^2.  The foot point
    is the constructed (u*, v*) in closed form.  A step-size rule would stop such points early;
    the certificate may only stop them within TOL of (u*, v*)."""
    r2 = math.sqrt(2) / 2
    Pc = np.array([[[1, 0, 0], [1, 0, 2]], [[1, 1, 0], [1, 1, 2]], [[0, 1, 0], [0, 1, 2]]], dtype=float)
    w = np.array([[1, 1], [r2, r2], [1, 1]])
    os_ = O.OSurf(2, 1, np.array([0, 0, 0, 1, 1, 1.0]), np.array([0, 0, 1, 1.0]), Pc, w)
    rng = np.random.default_rng(77)
    n = 4000
    u0 = rng.uniform(0.05, 0.95, n)
    v0 = rng.uniform(0.05, 0.95, n)
    S = os_.eval(u0, v0, 0)
    r = 10.0 ** rng.uniform(-3, 0, n)          # remaining distance to the axis
    P = S.copy()
    P[0] *= r                                  # inward along the radius (the normal)
    P[1] *= r
    su = np.clip(u0 + rng.uniform(-0.03, 0.03, n), 0, 1)
    sv = np.clip(v0 + rng.uniform(-0.03, 0.03, n), 0, 1)
    u, v, st, it = os_.gauss_newton(P[0], P[1], P[2], su, sv, tol_step=0.0, max_iter=60, newton=1,
                                    tol_quad=TOL)
    cert = st == O.CONV_STEP
    assert cert.mean() > 0.9
    err = np.maximum(np.abs(u - u0), np.abs(v - v0))[cert]
    assert np.max(err) <= TOL, float(np.max(err))
    # the flat cases take more steps before they are certified than the well-conditioned ones
    flat, steep = cert & (r < 3e-3), cert & (r > 0.3)
    assert flat.sum() > 100 and steep.sum() > 100
    assert np.mean(it[flat]) > np.mean(it[steep])


def _parabolic_cylinder_case(n, seed):
    """z = u^2, x = u, y = v as a polynomial patch (degree 2 x 1 Bezier, unit weights).  Queries on
    the normal at (u0, v0), toward the centre of curvature, at (1 - eps) of the curvature radius
    R(u0) = (1 + 4 u0^2)^1.5 / 2 with eps = 10^U(-4, -1): (u0, v0) stays a strict local minimum of
    the distance (the closed-form foot point) while the Hessian's smallest eigenvalue shrinks
    like eps and, away from the vertex, the third derivative does not (the evolute has no cusp
    there), so the Newton error constant grows like 1/eps."""
    Pc = np.array([[[0, 0, 0], [0, 1, 0]], [[0.5, 0, 0], [0.5, 1, 0]], [[1, 0, 1], [1, 1, 1]]], dtype=float)
    os_ = O.OSurf(2, 1, np.array([0, 0, 0, 1, 1, 1.0]), np.array([0, 0, 1, 1.0]), Pc, np.ones((3, 2)))
    rng = np.random.default_rng(seed)
    u0, v0 = rng.uniform(0.2, 0.9, n), rng.uniform(0.05, 0.95, n)
    eps = 10.0 ** rng.uniform(-4, -1, n)
    R = (1 + 4 * u0 ** 2) ** 1.5 / 2
    nrm = np.stack([-2 * u0, np.zeros(n), np.ones(n)]) / np.sqrt(1 + 4 * u0 ** 2)
    P = np.stack([u0, v0, u0 ** 2]) + (1 - eps) * R * nrm
    su = np.clip(u0 + rng.uniform(-1, 1, n) * 0.05 * eps, 0, 1)  # inside the narrow basin
    sv = np.clip(v0 + rng.uniform(-0.02, 0.02, n), 0, 1)
    return os_, P, u0, v0, su, sv, eps


def test_certificate_on_near_evolute_minima():
    os_, P, u0, v0, su, sv, eps = _parabolic_cylinder_case(4000, 79)
    S = os_.eval(u0, v0, 0)
    assert np.max(np.abs(S - np.stack([u0, v0, u0 ** 2]))) < 1e-15  # the patch is z = x^2
    u, v, st, it = os_.gauss_newton(P[0], P[1], P[2], su, sv, tol_step=0.0, max_iter=60, newton=1,
                                    tol_quad=TOL)
    cert = st == O.CONV_STEP
    assert cert.mean() > 0.9
    # the closed-form foot point carries the rounding of the normal, amplified by 1/eps
    err = np.maximum(np.abs(u - u0), np.abs(v - v0))
    assert np.max(err[cert]) <= TOL + 1e-13 / eps[cert].min(), float(np.max(err[cert]))


def test_step_size_rule_alone_is_not_enough_on_flat_minima():
    """Why the certificate scales with the conditioning: on the near-evolute queries a bare
    step-size threshold (the earlier rule, 1e-6) stops points that are still farther than 1e-9
    from the foot point; the certificate stops none of them there."""
    os_, P, u0, v0, su, sv, eps = _parabolic_cylinder_case(4000, 80)
    u, v, st, _ = os_.gauss_newton(P[0], P[1], P[2], su, sv, tol_step=1e-6, max_iter=60, newton=1)
    bad = (st == O.CONV_STEP) & (np.maximum(np.abs(u - u0), np.abs(v - v0)) > 1e-9)
    assert bad.sum() > 10
    u, v, st, _ = os_.gauss_newton(P[0], P[1], P[2], su, sv, tol_step=0.0, max_iter=60, newton=1,
                                   tol_quad=TOL)
    cert = st == O.CONV_STEP
    assert not np.any(cert & (np.maximum(np.abs(u - u0), np.abs(v - v0)) > 1e-9))


def test_no_certificate_for_a_step_across_a_knot():
    """The bound holds on one polynomial piece.  On a degree-2 surface S'' jumps at every interior
    knot, so a Newton step that crosses one can land farther than tol_quad from the minimum while
    the certificate, computed from the span the step started in, passes.  Foot points here sit
    just beside an interior knot and the seeds just on its other side; a step that ends outside
    its evaluation span is never certified, and every certified point is within tol_quad."""
    sd = synth.surface_G(11, 10, 2, 2, seed=97, jitter=0.4)
    os_ = O.OSurf.from_def(sd)
    rng = np.random.default_rng(98)
    n = 4000
    ku = np.asarray(sd.ku)[3:-3]
    kv = np.asarray(sd.kv)[3:-3]
    k0 = ku[rng.integers(0, len(ku), n)]
    off = rng.uniform(1e-7, 3e-5, n) * rng.choice([-1.0, 1.0], n)
    u0 = k0 + off
    v0 = np.where(rng.random(n) < 0.5, kv[rng.integers(0, len(kv), n)] + rng.uniform(-3e-5, 3e-5, n),
                  rng.uniform(0.05, 0.95, n))
    b = blocks(os_.eval(u0, v0, 1))
    P, _ = synth.offset_queries(b["S"], b["Su"], b["Sv"], 2e-2, 99)
    su = k0 - off * rng.uniform(0.2, 3.0, n)  # across the knot from the foot point
    sv = v0 + rng.uniform(-3e-5, 3e-5, n)
    u, v, st, _ = os_.gauss_newton(P[0], P[1], P[2], su, sv, tol_step=0.0, max_iter=40, newton=1,
                                   tol_quad=TOL)
    cert = st == O.CONV_STEP
    assert cert.mean() > 0.8
    tu, tv, _, _ = os_.gauss_newton(P[0], P[1], P[2], u, v, tol_step=0.0, max_iter=60, newton=1)
    err = np.maximum(np.abs(u - tu), np.abs(v - tv))[cert]
    assert np.max(err) <= TOL, float(np.max(err))


def test_knot_crossing_point_of_a_c1_batch_surface():
    """A point of the 224-surface uneven batch (surface 42, degree 2 x 2) whose FP32 bulk result
    lies across a u knot (0.44547494865471027) from the foot point: a certified Newton step from
    there once ended 1.4e-8 from the minimum."""
    defs, _ = synth.uneven_batch(224, seed=1004)
    os_ = O.OSurf.from_def(defs[42])
    P = np.array([[0.45269181064699265], [0.2427569024800431], [0.0012385542991139222]])
    su, sv = np.array([0.4447364762891084]), np.array([0.23347128368914127])
    bu, bv, _, _ = os_.gauss_newton(P[0], P[1], P[2], su, sv, tol_step=1e-4, max_iter=20, newton=1)
    assert bu[0] < 0.44547494865471027 < 0.4454750193089489  # the closing step crosses the knot
    u, v, st, _ = os_.gauss_newton(P[0], P[1], P[2], bu, bv, tol_step=0.0, max_iter=3, newton=1,
                                   tol_quad=TOL)
    assert st[0] == O.CONV_STEP
    assert abs(u[0] - 0.4454750193089489) <= TOL and abs(v[0] - 0.23427055545943834) <= TOL
