"""The Newton certificate on the GPU does not certify a step that crosses a knot.

The certificate's bound (gn_invert_opts.tol_quad) holds on one polynomial piece; on a degree-2
surface S'' jumps at every interior knot.  Foot points here sit just beside an interior u knot and
the seeds just across it, so the closing FP64 steps cross knots often.  Every point the mixed and
the FP64 Newton paths call converged is compared with the minimiser the oracle reaches by 60
further Newton iterations from it (the same check as the oracle's own certificate tests), at the
1e-9 bar of the inversion parity tests.  The last case is the point of the 224-surface uneven batch
(surface 42) whose FP32 bulk result lies across a u knot from its foot point.
"""
import numpy as np
import pytest

import oracle as O
import gnurbs as G
from gnurbs import synth
from helpers import blocks

pytestmark = pytest.mark.gpu

MIXED = dict(precision=G.GN_MIXED, newton=G.NEWTON_FP32 | G.NEWTON_FP64, fp32_iters=20, fp64_tail=3,
             tol_step32=1e-4, tol_step=1e-10, tol_quad=1e-10)
NEWTON64 = dict(precision=G.GN_FP64, newton=G.NEWTON_FP64, max_iter=40, tol_step=0.0, tol_quad=1e-10)


def _np(t):
    return t.detach().cpu().numpy()


def _knot_case(n=4000):
    sd = synth.surface_G(11, 10, 2, 2, seed=97, jitter=0.4)
    os_ = O.OSurf.from_def(sd)
    rng = np.random.default_rng(98)
    ku = np.asarray(sd.ku)[3:-3]
    kv = np.asarray(sd.kv)[3:-3]
    k0 = ku[rng.integers(0, len(ku), n)]
    off = rng.uniform(1e-7, 3e-5, n) * rng.choice([-1.0, 1.0], n)
    u0 = k0 + off
    v0 = np.where(rng.random(n) < 0.5, kv[rng.integers(0, len(kv), n)] + rng.uniform(-3e-5, 3e-5, n),
                  rng.uniform(0.05, 0.95, n))
    b = blocks(os_.eval(u0, v0, 1))
    P, _ = synth.offset_queries(b["S"], b["Su"], b["Sv"], 2e-2, 99)
    su = k0 - off * rng.uniform(0.2, 3.0, n)
    sv = v0 + rng.uniform(-3e-5, 3e-5, n)
    return sd, os_, P, su, sv


@pytest.mark.parametrize("mode", ["mixed", "fp64_newton"])
def test_certified_points_beside_knots_are_within_1e9(mode):
    sd, os_, P, su, sv = _knot_case()
    gs = G.Surface(sd.p, sd.q, sd.ku, sd.kv, sd.cp, sd.w)
    try:
        opts = G.default_opts(**(MIXED if mode == "mixed" else NEWTON64))
        u, v, st = (_np(t) for t in gs.NewInverse(P[0], P[1], P[2], opts, seeds=(su, sv)))
    finally:
        gs.close()
    conv = st == G.PT_CONVERGED
    assert conv.mean() > 0.8, (mode, conv.mean())
    tu, tv, _, _ = os_.gauss_newton(P[0], P[1], P[2], u, v, tol_step=0.0, max_iter=60, newton=1)
    err = np.maximum(np.abs(u - tu), np.abs(v - tv))[conv]
    assert np.max(err) <= 1e-9, (mode, float(np.max(err)))


def test_knot_crossing_point_of_a_c1_batch_surface():
    defs, _ = synth.uneven_batch(224, seed=1004)
    sd = defs[42]
    P = np.array([[0.45269181064699265], [0.2427569024800431], [0.0012385542991139222]])
    su, sv = np.array([0.4447364762891084]), np.array([0.23347128368914127])
    gs = G.Surface(sd.p, sd.q, sd.ku, sd.kv, sd.cp, sd.w)
    try:
        u, v, st = (_np(t) for t in gs.NewInverse(P[0], P[1], P[2], G.default_opts(**MIXED),
                                                 seeds=(su, sv)))
    finally:
        gs.close()
    assert st[0] == G.PT_CONVERGED
    assert abs(u[0] - 0.4454750193089489) <= 1e-10 and abs(v[0] - 0.23427055545943834) <= 1e-10
