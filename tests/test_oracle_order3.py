"""Third-order evaluation of the oracle (MatrixSurfaceRd{UUU,UUV,UVV,VVV} and
MatrixSurfaceDerivative{UUU,UUV,UVV,VVV}, src/NURBS/Evaluate.cu:408-502, 960-1020).

Pinned by checks that do not share the oracle's formulas:
  * the generic quotient rule of The NURBS Book eq. 4.20 on scipy's B-spline basis
    (helpers.scipy_eval_general: a loop over binomial terms, not the four hand-written forms);
  * central finite differences of the order-2 fields at points away from the knots;
  * symbolic derivatives (sympy) of a rational Bezier patch with non-constant weights, where the
    NURBS surface is the Bernstein quotient in closed form.
It also records the documented divergence from the reference: the new-version S_uv, S_uuv and
S_uvv (src/NURBS/Evaluate.cu:958, 966, 982) add the w_uv-type terms, so they agree with the
quotient rule only when the weights are constant.
"""
import math

import numpy as np
import pytest

import oracle as O
from gnurbs import synth
from helpers import DER_NAMES, blocks, field_close, scipy_eval_general

THIRD = ("Suuu", "Suuv", "Suvv", "Svvv")


@pytest.mark.parametrize("nu,nv,p,q", [(8, 8, 3, 3), (12, 7, 2, 4), (9, 11, 5, 2), (6, 6, 5, 5),
                                       (4, 5, 1, 3), (10, 9, 4, 4)])
def test_order3_vs_generic_quotient_rule(nu, nv, p, q):
    sd = synth.surface_G(nu, nv, p, q, seed=21, jitter=0.3)
    os_ = O.OSurf.from_def(sd)
    u, v = synth.uniform_uv(300, seed=5)
    mine = blocks(os_.eval(u, v, order=3))
    ref = scipy_eval_general(p, q, sd.ku, sd.kv, sd.cp, sd.w, u, v, order=3)
    floor = 1e-2 * np.max(np.abs(ref["S"]))
    for key in DER_NAMES:
        ok, err = field_close(mine[key], ref[key], 1e-10, floor)
        assert ok, (key, err)


def test_order3_lower_orders_unchanged():
    sd = synth.surface_G(9, 8, 3, 4, seed=3, jitter=0.2)
    os_ = O.OSurf.from_def(sd)
    u, v = synth.uniform_uv(500, seed=6)
    a = os_.eval(u, v, 2)
    b = os_.eval(u, v, 3)
    assert np.array_equal(a, b[:18])


def _away_from_knots(knots, n, h, rng):
    """Parameters at least 100 h from every distinct knot (third derivatives of a C^2 spline
    jump at knots, so a difference quotient must not straddle one)."""
    k = np.unique(knots)
    out = []
    while len(out) < n:
        x = rng.uniform(k[0], k[-1])
        if np.min(np.abs(k - x)) > 100 * h:
            out.append(x)
    return np.array(out)


def test_order3_finite_differences():
    sd = synth.surface_G(8, 8, 3, 3, seed=1000)
    os_ = O.OSurf.from_def(sd)
    rng = np.random.default_rng(1)
    h = 1e-5
    u = _away_from_knots(sd.ku, 60, h, rng)
    v = _away_from_knots(sd.kv, 60, h, rng)
    b0 = blocks(os_.eval(u, v, 3))
    bu = [blocks(os_.eval(u + s * h, v, 2)) for s in (-1, 1)]
    bv = [blocks(os_.eval(u, v + s * h, 2)) for s in (-1, 1)]
    fd = {
        "Suuu": (bu[1]["Suu"] - bu[0]["Suu"]) / (2 * h),
        "Suuv": (bv[1]["Suu"] - bv[0]["Suu"]) / (2 * h),
        "Suvv": (bu[1]["Svv"] - bu[0]["Svv"]) / (2 * h),
        "Svvv": (bv[1]["Svv"] - bv[0]["Svv"]) / (2 * h),
    }
    # the mixed fields have a second difference route each
    fd2 = {"Suuv": (bu[1]["Suv"] - bu[0]["Suv"]) / (2 * h),
           "Suvv": (bv[1]["Suv"] - bv[0]["Suv"]) / (2 * h)}
    for key in THIRD:
        scale = max(np.max(np.abs(b0[key])), 1.0)
        assert np.max(np.abs(fd[key] - b0[key])) < 1e-5 * scale, key
        if key in fd2:
            assert np.max(np.abs(fd2[key] - b0[key])) < 1e-5 * scale, key


def _bezier_closed_form(p, q, cp, w):
    """Exact S and its derivatives up to order 3 of a rational Bezier patch, from sympy."""
    sp = pytest.importorskip("sympy")
    uu, vv = sp.symbols("u v")
    Bu = [sp.binomial(p, i) * uu ** i * (1 - uu) ** (p - i) for i in range(p + 1)]
    Bv = [sp.binomial(q, j) * vv ** j * (1 - vv) ** (q - j) for j in range(q + 1)]
    W = sum(Bu[i] * Bv[j] * sp.Float(w[i, j], 30) for i in range(p + 1) for j in range(q + 1))
    fns = {}
    for c in range(3):
        A = sum(Bu[i] * Bv[j] * sp.Float(w[i, j] * cp[i, j, c], 30)
                for i in range(p + 1) for j in range(q + 1))
        S = A / W
        for name, (a, b) in (("S", (0, 0)), ("Su", (1, 0)), ("Sv", (0, 1)), ("Suu", (2, 0)),
                             ("Suv", (1, 1)), ("Svv", (0, 2)), ("Suuu", (3, 0)), ("Suuv", (2, 1)),
                             ("Suvv", (1, 2)), ("Svvv", (0, 3))):
            e = S
            if a:
                e = sp.diff(e, uu, a)
            if b:
                e = sp.diff(e, vv, b)
            fns[name, c] = sp.lambdify((uu, vv), e, "numpy")
    return fns


@pytest.mark.parametrize("p,q", [(3, 2), (2, 3), (3, 3)])
def test_order3_rational_bezier_closed_form(p, q):
    rng = np.random.default_rng(40 + p * 10 + q)
    cp = rng.normal(size=(p + 1, q + 1, 3))
    w = rng.uniform(0.5, 2.0, size=(p + 1, q + 1))
    ku = np.r_[np.zeros(p + 1), np.ones(p + 1)]
    kv = np.r_[np.zeros(q + 1), np.ones(q + 1)]
    os_ = O.OSurf(p, q, ku, kv, cp, w)
    u, v = synth.uniform_uv(40, seed=8)
    mine = blocks(os_.eval(u, v, 3))
    fns = _bezier_closed_form(p, q, cp, w)
    for name in DER_NAMES:
        ref = np.stack([np.broadcast_to(np.asarray(fns[name, c](u, v), dtype=np.float64), u.shape)
                        for c in range(3)])
        ok, err = field_close(mine[name], ref, 1e-11)
        assert ok, (name, err)


def _reference_new_version(hom, S):
    """The new-version mixed derivatives as the reference writes them (Evaluate.cu:958, 966, 982),
    in the global parameters; hom = (A, W) homogeneous derivatives keyed by (k, l)."""
    A, W = hom
    uv = (A[1, 1] - W[0, 1] * S["Su"] - W[1, 0] * S["Sv"] + W[1, 1] * S["S"]) / W[0, 0]
    uuv = (A[2, 1] - W[0, 1] * S["Suu"] - 2 * W[1, 0] * S["Suv"] + 2 * W[1, 1] * S["Su"]
           - W[2, 0] * S["Sv"] + W[2, 1] * S["S"]) / W[0, 0]
    uvv = (A[1, 2] - W[1, 0] * S["Svv"] - 2 * W[0, 1] * S["Suv"] + 2 * W[1, 1] * S["Sv"]
           - W[0, 2] * S["Su"] + W[1, 2] * S["S"]) / W[0, 0]
    uuu = (A[3, 0] - 3 * W[1, 0] * S["Suu"] - 3 * W[2, 0] * S["Su"] - W[3, 0] * S["S"]) / W[0, 0]
    return {"Suv": uv, "Suuv": uuv, "Suvv": uvv, "Suuu": uuu}


@pytest.mark.parametrize("constant_weights", [True, False])
def test_reference_new_version_mixed_terms_need_constant_weights(constant_weights):
    sd = synth.surface_G(8, 8, 3, 3, seed=12)
    w = np.full_like(sd.w, 1.3) if constant_weights else sd.w
    sd = synth.SurfaceDef(sd.p, sd.q, sd.ku, sd.kv, sd.cp, w)
    os_ = O.OSurf.from_def(sd)
    u, v = synth.uniform_uv(200, seed=9)
    mine = blocks(os_.eval(u, v, 3))
    S, A, W = scipy_eval_general(sd.p, sd.q, sd.ku, sd.kv, sd.cp, sd.w, u, v, order=3,
                                 return_hom=True)
    ref_new = _reference_new_version((A, W), S)
    floor = 1e-2 * np.max(np.abs(S["S"]))
    # S_uuu has no w_uv-type term: it agrees either way
    assert field_close(mine["Suuu"], ref_new["Suuu"], 1e-10, floor)[0]
    for key in ("Suv", "Suuv", "Suvv"):
        ok, err = field_close(mine[key], ref_new[key], 1e-10, floor)
        assert ok == constant_weights, (key, err)
