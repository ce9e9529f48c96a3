"""The oracle's greedy-leaf seed with one FP32 Gauss-Newton step (the library's GN_SEED_GREEDY_STEP;
no reference counterpart -- the reference seeds Gauss-Newton at a node's uv centre,
src/kernel/kernel.cu:156, 433-469), pinned by an independent numpy restatement on scipy's
B-spline basis, and by what the step is for: seeds nearer the foot points and fewer iterations
after them."""
import numpy as np

import oracle as O
from gnurbs import synth
from helpers import blocks, scipy_eval


def _case(seed, n=20000, levels=3, ku_scale=1.0):
    sd = synth.surface_G(16, 14, 3, 2, seed=seed, jitter=0.3)
    ku = sd.ku * ku_scale
    os_ = O.OSurf(sd.p, sd.q, ku, sd.kv, sd.cp, sd.w)
    nodes, _, _ = os_.build_tree(levels)
    u0, v0 = synth.uniform_uv(n, seed + 1)
    u0 = u0 * ku_scale
    b = blocks(os_.eval(u0, v0, 1))
    P, _ = synth.offset_queries(b["S"], b["Su"], b["Sv"], 1e-2, seed + 2)
    return sd, ku, os_, nodes, u0, v0, P


def test_step_equals_an_independent_gauss_newton_step():
    for scale in (1.0, 5.0):  # also on a knot domain other than [0, 1]
        sd, ku, os_, nodes, u0, v0, P = _case(3, ku_scale=scale)
        leaf, su, sv = os_.seed_greedy_step(nodes, 3, P[0], P[1], P[2])
        gl, gu_, gv_ = O.seed_greedy_leaf(nodes, 3, P[0], P[1], P[2])
        assert np.array_equal(leaf, gl)
        # numpy: J at the leaf's seed from scipy's basis, C from the tree node
        d = scipy_eval(sd.p, sd.q, ku, sd.kv, sd.cp, sd.w, gu_, gv_, order=1)
        C = nodes[leaf, 10:13].T
        r = C - P
        a, b, e = (d["Su"] ** 2).sum(0), (d["Su"] * d["Sv"]).sum(0), (d["Sv"] ** 2).sum(0)
        g1, g2 = (d["Su"] * r).sum(0), (d["Sv"] * r).sum(0)
        det = a * e - b * b
        du, dv = (e * g1 - b * g2) / det, (a * g2 - b * g1) / det
        hu, hv = (ku[-1] - ku[0]) / 64, (sd.kv[-1] - sd.kv[0]) / 64
        sc = np.maximum(1.0, np.maximum(np.abs(du) / hu, np.abs(dv) / hv))
        eu = np.clip(gu_ - du / sc, ku[0], ku[-1])
        ev = np.clip(gv_ - dv / sc, sd.kv[0], sd.kv[-1])
        # the oracle's step is FP32 (as the library's): within FP32 rounding of the FP64 step,
        # whose residual C - P rounds to about 1e-7 of the coordinates (1e-5 of a leaf cell)
        su_step, sv_step = np.abs(du / sc), np.abs(dv / sc)
        assert np.all(np.abs(su - eu) <= 2e-6 * su_step + 1e-5 * hu)
        assert np.all(np.abs(sv - ev) <= 2e-6 * sv_step + 1e-5 * hv)
        # at most one leaf cell from the leaf's seed, inside the domain
        assert np.all(np.abs(su - gu_) <= hu * (1 + 1e-12)) and np.all(np.abs(sv - gv_) <= hv * (1 + 1e-12))
        assert su.min() >= ku[0] and su.max() <= ku[-1]


def test_step_brings_seeds_nearer_and_saves_iterations():
    sd, ku, os_, nodes, u0, v0, P = _case(5, n=40000)
    _, su, sv = os_.seed_greedy_step(nodes, 3, P[0], P[1], P[2])
    _, gu_, gv_ = O.seed_greedy_leaf(nodes, 3, P[0], P[1], P[2])
    e_leaf = np.maximum(np.abs(gu_ - u0), np.abs(gv_ - v0))
    e_step = np.maximum(np.abs(su - u0), np.abs(sv - v0))
    assert np.median(e_step) < 0.3 * np.median(e_leaf)
    its = []
    for a, b in ((gu_, gv_), (su, sv)):
        _, _, _, it = os_.gauss_newton(P[0], P[1], P[2], a, b, tol_step=1e-4, max_iter=20, newton=1)
        its.append(it.mean())
    assert its[1] < its[0] - 0.5  # about one Newton iteration fewer to a 1e-4 step
    # and the same foot points where both runs converge
    fu, fv, fst, _ = os_.gauss_newton(P[0], P[1], P[2], gu_, gv_, tol_step=1e-12, max_iter=40, newton=1)
    tu, tv, tst, _ = os_.gauss_newton(P[0], P[1], P[2], su, sv, tol_step=1e-12, max_iter=40, newton=1)
    both = (fst == O.CONV_STEP) & (tst == O.CONV_STEP)
    same = np.maximum(np.abs(fu - tu), np.abs(fv - tv)) < 1e-9
    assert same[both].mean() > 0.99
