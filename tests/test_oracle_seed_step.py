"""The oracle's greedy-descent seed with one FP32 Gauss-Newton step (the library's
GN_SEED_GREEDY_STEP; no reference counterpart -- the reference seeds Gauss-Newton at a node's uv
centre, src/kernel/kernel.cu:156, 433-469), pinned by independent numpy restatements of the descent
(FP32 boxes built from the tree's nodes) and of the step (scipy's B-spline basis), and by what the
step is for: seeds nearer the foot points and fewer iterations after them."""
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


def _descent(nodes, levels, P):
    """numpy restatement of the greedy descent: quadtree boxes from the leaves' boxes (each widened
    to hold its centre), rounded outward to FP32 after a margin of 2^-20 (1 + the root's largest
    coordinate magnitude); at each level the child of least FP32 squared box distance (the first
    of equals); returns the reference index of the leaf reached."""
    D = 3 * (levels - 1)
    side = 1 << D
    base = sum(64 ** l for l in range(levels - 1))
    nd = nodes[base:base + side * side]
    lo = np.minimum(nd[:, 4:7], nd[:, 10:13]).reshape(side, side, 3)
    hi = np.maximum(nd[:, 7:10], nd[:, 10:13]).reshape(side, side, 3)
    los, his = [lo], [hi]
    for _ in range(D):
        l, h = los[0], his[0]
        s = l.shape[0] // 2
        los.insert(0, l.reshape(s, 2, s, 2, 3).min(axis=(1, 3)))
        his.insert(0, h.reshape(s, 2, s, 2, 3).max(axis=(1, 3)))
    mag = max(np.abs(los[0]).max(), np.abs(his[0]).max())
    m = (1.0 + mag) * 2.0 ** -20

    def down(x):
        f = (x - m).astype(np.float32)
        return np.where(f.astype(np.float64) > x - m, np.nextafter(f, np.float32(-np.inf)), f)

    def up(x):
        f = (x + m).astype(np.float32)
        return np.where(f.astype(np.float64) < x + m, np.nextafter(f, np.float32(np.inf)), f)

    l32 = [down(x) for x in los]
    h32 = [up(x) for x in his]
    p = P.T.astype(np.float32)
    n = p.shape[0]
    ix, iy = np.zeros(n, dtype=np.int64), np.zeros(n, dtype=np.int64)
    for lev in range(1, D + 1):
        best = np.full(n, np.inf, dtype=np.float32)
        bx, by = 2 * ix, 2 * iy
        for a in range(4):
            cx, cy = 2 * ix + (a >> 1), 2 * iy + (a & 1)
            bl, bh = l32[lev][cx, cy], h32[lev][cx, cy]
            e = np.maximum(np.maximum(bl - p, p - bh), np.float32(0))
            s = np.float32(0)
            for k in range(3):  # fmaf(e, e, s) in FP32: exact products of FP32 values, one rounding
                s = (e[:, k].astype(np.float64) ** 2 + np.asarray(s, np.float64)).astype(np.float32)
            take = s < best
            best = np.where(take, s, best)
            bx, by = np.where(take, cx, bx), np.where(take, cy, by)
        ix, iy = bx, by
    return (base + ix * side + iy).astype(np.int32)


def test_step_equals_an_independent_gauss_newton_step():
    """The step restated in FP64 numpy with the derivatives from scipy's basis: one Gauss-Newton
    step from the leaf's seed (at most one leaf cell), then two Gauss-Newton iterations on the
    leaf's quadratic Taylor model (each result at most two cells from the seed)."""
    for scale in (1.0, 5.0):  # also on a knot domain other than [0, 1]
        sd, ku, os_, nodes, u0, v0, P = _case(3, ku_scale=scale)
        leaf, su, sv = os_.seed_greedy_step(nodes, 3, P[0], P[1], P[2])
        gl, gu_, gv_ = O.seed_greedy_descent(nodes, 3, P[0], P[1], P[2])
        assert np.array_equal(leaf, gl)
        assert np.array_equal(gl, _descent(nodes, 3, P))
        # numpy: J and the second derivatives at the leaf's seed from scipy's basis, C from the node
        d = scipy_eval(sd.p, sd.q, ku, sd.kv, sd.cp, sd.w, gu_, gv_, order=2)
        Su, Sv, Suu, Suv, Svv = d["Su"], d["Sv"], d["Suu"], d["Suv"], d["Svv"]
        C = nodes[leaf, 10:13].T
        r = C - P
        a, b, e = (Su ** 2).sum(0), (Su * Sv).sum(0), (Sv ** 2).sum(0)
        g1, g2 = (Su * r).sum(0), (Sv * r).sum(0)
        det = a * e - b * b
        du, dv = (e * g1 - b * g2) / det, (a * g2 - b * g1) / det
        hu, hv = (ku[-1] - ku[0]) / 64, (sd.kv[-1] - sd.kv[0]) / 64
        sc = np.maximum(1.0, np.maximum(np.abs(du) / hu, np.abs(dv) / hv))
        tu, tv = -du / sc, -dv / sc
        for _ in range(2):
            m = r + Su * tu + Sv * tv + 0.5 * (Suu * tu ** 2 + 2 * Suv * tu * tv + Svv * tv ** 2)
            Au, Av = Su + Suu * tu + Suv * tv, Sv + Suv * tu + Svv * tv
            A, B, E = (Au ** 2).sum(0), (Au * Av).sum(0), (Av ** 2).sum(0)
            G1, G2 = (Au * m).sum(0), (Av * m).sum(0)
            D = A * E - B * B
            nu, nv = tu - (E * G1 - B * G2) / D, tv - (A * G2 - B * G1) / D
            s2 = np.maximum(1.0, np.maximum(np.abs(nu) / (2 * hu), np.abs(nv) / (2 * hv)))
            ok = D > 0
            tu, tv = np.where(ok, nu / s2, tu), np.where(ok, nv / s2, tv)
        eu = np.clip(gu_ + tu, ku[0], ku[-1])
        ev = np.clip(gv_ + tv, sd.kv[0], sd.kv[-1])
        # the oracle's step is FP32 (as the library's): within FP32 rounding of the FP64 one, whose
        # residual C - P rounds to about 1e-7 of the coordinates (1e-5 of a leaf cell), amplified
        # where the model's system is poorly conditioned
        err = np.maximum(np.abs(su - eu) / hu, np.abs(sv - ev) / hv)
        assert np.median(err) < 5e-6 and np.percentile(err, 99) < 2e-5 and err.max() < 1e-4, err.max()
        # at most two leaf cells from the leaf's seed, inside the domain
        assert np.all(np.abs(su - gu_) <= 2 * hu * (1 + 1e-6)) and np.all(np.abs(sv - gv_) <= 2 * hv * (1 + 1e-6))
        assert su.min() >= ku[0] and su.max() <= ku[-1]
        # the model iterations pay: seeds nearer the foot points than the linear step's
        lu, lv = np.clip(gu_ - du / sc, ku[0], ku[-1]), np.clip(gv_ - dv / sc, sd.kv[0], sd.kv[-1])
        e_lin = np.maximum(np.abs(lu - u0), np.abs(lv - v0))
        e_mod = np.maximum(np.abs(su - u0), np.abs(sv - v0))
        assert np.median(e_mod) < 0.3 * np.median(e_lin)


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
