"""p-curves (gn_pcurve): sampling a 3-D curve, inverting the samples onto a surface and fitting
the 2-D B-spline (u(t), v(t)) by least squares.

Pinned by
  * a closed form: the v-iso-curve S(u0, t) of a surface, written as a rational B-spline curve
    (homogeneous control points sum_i N_i(u0) (w_ij P_ij, w_ij) on the v knots), has the p-curve
    u = u0, v = t, whose B-spline control points are u0 and the Greville abscissae;
  * the oracle chain: the oracle's curve samples, greedy-leaf seeds and FP64 Gauss-Newton, then an
    independent least-squares fit (scipy's B-spline design matrix and numpy lstsq).
"""
import numpy as np
import pytest
from scipy.interpolate import BSpline

import oracle as O
import gnurbs as G
from gnurbs import synth
from helpers import _basis

pytestmark = pytest.mark.gpu


def smooth_surface(nu=14, nv=12, p=3, q=3):
    X, Y = np.meshgrid(np.arange(nu) / (nu - 1), np.arange(nv) / (nv - 1), indexing="ij")
    cp = np.stack([X, Y, 0.15 * np.sin(2 * np.pi * X) * np.cos(np.pi * Y)], axis=-1)
    w = 1.0 + 0.3 * np.sin(3 * X + 2 * Y)
    return synth.SurfaceDef(p, q, synth.clamped_knots(nu, p), synth.clamped_knots(nv, q), cp, w)


def iso_curve(sd, u0):
    nu = sd.w.shape[0]
    Nu = _basis(sd.ku, sd.p, nu, [u0], 0)[0]
    W = Nu @ sd.w
    Q = np.einsum("i,ijc,ij->jc", Nu, sd.cp, sd.w) / W[:, None]
    return synth.CurveDef(sd.q, sd.kv.copy(), np.ascontiguousarray(Q), np.ascontiguousarray(W))


def clamped_uniform(ncp, d, t0=0.0, t1=1.0):
    ns = ncp - d
    return np.r_[np.full(d + 1, t0), t0 + (t1 - t0) * np.arange(1, ns) / ns, np.full(d + 1, t1)]


def greville(kn, d, ncp):
    return np.array([kn[i + 1:i + d + 1].mean() for i in range(ncp)])


OPTS = {
    "fp64": dict(precision=G.GN_FP64, max_iter=20, tol_step=1e-12),
    "mixed": dict(precision=G.GN_MIXED, newton=G.NEWTON_FP32 | G.NEWTON_FP64, fp32_iters=20, fp64_tail=3,
                  tol_step32=1e-4, tol_step=1e-12, tol_quad=1e-10),
}


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("u0,degree,ncp", [(0.37, 3, 12), (0.81, 2, 7), (0.05, 5, 20)])
def test_iso_curve_pcurve_is_the_parameter_line(prec, u0, degree, ncp):
    sd = smooth_surface()
    s = G.Surface(sd.p, sd.q, sd.ku, sd.kv, sd.cp, sd.w)
    s.BuildTree(3)
    c = G.Curve.from_def(iso_curve(sd, u0))
    cp, kn, dev, nfail = s.PCurve(c, 20000, degree=degree, ncp=ncp, opts=G.default_opts(**OPTS[prec]))
    assert nfail == 0
    assert np.allclose(kn, clamped_uniform(ncp, degree), rtol=0, atol=1e-15)
    assert np.max(np.abs(cp[:, 0] - u0)) < 1e-9
    assert np.max(np.abs(cp[:, 1] - greville(kn, degree, ncp))) < 1e-9
    assert dev < 1e-9
    s.close()
    c.close()


def test_pcurve_matches_the_oracle_chain():
    """An off-surface curve (an iso-curve lifted by a smooth offset): the GPU p-curve equals the
    least-squares fit to the oracle's own inversions of the oracle's own samples."""
    sd = smooth_surface()
    cd = iso_curve(sd, 0.55)
    cd.cp[:, 2] += 0.01 * np.sin(np.arange(len(cd.w)))
    s = G.Surface(sd.p, sd.q, sd.ku, sd.kv, sd.cp, sd.w)
    s.BuildTree(3)
    c = G.Curve.from_def(cd)
    m, degree, ncp = 4000, 3, 10
    cp, kn, dev, nfail = s.PCurve(c, m, degree=degree, ncp=ncp, opts=G.default_opts(**OPTS["fp64"]))
    assert nfail == 0
    # the oracle chain on the same sample parameters
    t = np.arange(m) / (m - 1)
    t[-1] = 1.0
    P = O.OCurve.from_def(cd).eval(t, 0)
    os_ = O.OSurf.from_def(sd)
    nodes, _, _ = os_.build_tree(3)
    _, su, sv = O.seed_greedy_leaf(nodes, 3, P[0], P[1], P[2])
    u, v, st, _ = os_.gauss_newton(P[0], P[1], P[2], su, sv, tol_step=1e-12, max_iter=20)
    assert np.all((st == 0) | (st == 1))
    kref = clamped_uniform(ncp, degree)
    B = BSpline.design_matrix(t, kref, degree).toarray()
    ref, *_ = np.linalg.lstsq(B, np.stack([u, v], axis=1), rcond=None)
    assert np.max(np.abs(cp - ref)) < 1e-8
    # deviation: the fitted p-curve pushed through the surface, against the samples
    uv = B @ ref
    S = os_.eval(uv[:, 0], uv[:, 1], 0)
    ref_dev = np.max(np.linalg.norm(S - P, axis=0))
    assert abs(dev - ref_dev) <= 1e-9 + 1e-6 * ref_dev
    assert 1e-4 < dev < 0.05  # the curve is off the surface: the deviation is the lift
    s.close()
    c.close()


def test_pcurve_argument_errors():
    sd = smooth_surface()
    s = G.Surface(sd.p, sd.q, sd.ku, sd.kv, sd.cp, sd.w)
    c = G.Curve.from_def(iso_curve(sd, 0.5))
    with pytest.raises(G.GnError) as e:  # no tree yet
        s.PCurve(c, 1000)
    assert e.value.code == G.GN_ESTATE
    s.BuildTree(3)
    for kw in (dict(degree=0), dict(degree=6), dict(degree=3, ncp=3), dict(samples=8, ncp=16)):
        args = dict(samples=1000, degree=3, ncp=16)
        args.update(kw)
        with pytest.raises(G.GnError) as e:
            s.PCurve(c, args.pop("samples"), **args)
        assert e.value.code == G.GN_EINVAL
    with pytest.raises(G.GnError) as e:
        s.PCurve(c, 1000, opts=G.default_opts(seed_mode=G.SEED_GIVEN))
    assert e.value.code == G.GN_EINVAL
    s.close()
    c.close()
