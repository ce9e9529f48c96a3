"""Inversion parity on every degree class the inversion kernels dispatch: p, q in 2..5 (the exact
classes of the single-surface and batched kernels) and degrees outside 2..5 (the padded generic
class), each with jittered interior knots and 3,000 offset queries (N(0, 1e-2) along the normal).

Against the oracle from the same seeds (GaussNewton, src/kernel/kernel.cu:151-249, is
degree-generic through findSpan/MatrixSurfacePoint, src/NURBS/Evaluate.cu:159-227, 504-533):
  * seeds: the greedy-leaf seed search equals the oracle's (leaf indices and seed parameters);
  * plain FP64 (the reference iteration, single surface and gn_batch_invert): within 1e-9 of the
    oracle's same iteration wherever both converge, and the same status almost everywhere;
  * mixed (single surface and gn_batch_invert): within 1e-9 of the oracle's FP64 Newton solution
    wherever both converge to the same foot point -- no fraction of points exempt.
"""
import numpy as np
import pytest
import torch

import oracle as O
import gnurbs as G
from gnurbs import synth
from helpers import blocks

pytestmark = pytest.mark.gpu

CLASSES = [(p, q) for p in range(2, 6) for q in range(2, 6)] + [(1, 3), (6, 2), (3, 7)]
MIXED = dict(precision=G.GN_MIXED, newton=G.NEWTON_FP32 | G.NEWTON_FP64, fp32_iters=20, fp64_tail=3,
             tol_step32=1e-4, tol_step=1e-10, tol_quad=1e-10)
N = 3000


def _np(t):
    return t.detach().cpu().numpy()


def _case(p, q, seed):
    sd = synth.surface_G(2 * p + 6, 2 * q + 5, p, q, seed=seed, jitter=0.4)
    os_ = O.OSurf.from_def(sd)
    u0, v0 = synth.uniform_uv(N, seed + 1)
    b = blocks(os_.eval(u0, v0, 1))
    P, _ = synth.offset_queries(b["S"], b["Su"], b["Sv"], 1e-2, seed + 2)
    gs = G.Surface(sd.p, sd.q, sd.ku, sd.kv, sd.cp, sd.w)
    gs.BuildTree(3)
    return sd, os_, gs, P


@pytest.fixture(scope="module")
def cases():
    out = {}
    for i, (p, q) in enumerate(CLASSES):
        sd, os_, gs, P = _case(p, q, 7100 + 10 * i)
        nodes, _, _ = gs.GetTree()
        leaf, su, sv = (_np(t) for t in gs.Seed(P[0], P[1], P[2], exact=False))
        rl, rsu, rsv = O.seed_greedy_leaf(nodes, 3, P[0], P[1], P[2])
        # oracle solutions from the greedy seeds: the reference iteration and the FP64 Newton truth
        gn = os_.gauss_newton(P[0], P[1], P[2], rsu, rsv, tol_step=1e-10, max_iter=20)
        nt = os_.gauss_newton(P[0], P[1], P[2], rsu, rsv, tol_step=1e-12, max_iter=40, newton=1)
        # greedy leaf + one Gauss-Newton step with the leaf's Jacobian (GN_SEED_GREEDY_STEP)
        sleaf, ssu, ssv = (_np(t) for t in gs.Seed(P[0], P[1], P[2], exact="step"))
        tl, tsu, tsv = os_.seed_greedy_step(nodes, 3, P[0], P[1], P[2])
        nts = os_.gauss_newton(P[0], P[1], P[2], tsu, tsv, tol_step=1e-12, max_iter=40, newton=1)
        out[(p, q)] = dict(sd=sd, os=os_, gs=gs, P=P, leaf=leaf, su=su, sv=sv, rleaf=rl, rsu=rsu, rsv=rsv,
                           gn=gn, nt=nt, sleaf=sleaf, ssu=ssu, ssv=ssv, tleaf=tl, tsu=tsu, tsv=tsv, nts=nts)
    yield out
    for c in out.values():
        c["gs"].close()


def _check_plain(u, v, st, gn, tag):
    ru, rv, rst, _ = gn
    both = (st == G.PT_CONVERGED) & (rst == O.CONV_STEP)
    assert both.mean() > 0.5, tag
    d = np.maximum(np.abs(u - ru), np.abs(v - rv))[both]
    assert np.max(d) <= 1e-9, (tag, float(np.max(d)))
    assert (st == rst).mean() > 0.99, tag


def _check_mixed(u, v, st, nt, tag):
    ru, rv, rst, _ = nt
    both = (st == G.PT_CONVERGED) & (rst == O.CONV_STEP)
    d = np.maximum(np.abs(u - ru), np.abs(v - rv))[both]
    same = d < 1e-6  # the others reached another local minimum of the distance from the seed
    assert both.mean() > 0.9 and same.mean() > 0.98, (tag, both.mean(), same.mean())
    assert np.max(d[same]) <= 1e-9, (tag, float(np.max(d[same])))


@pytest.mark.parametrize("pq", CLASSES, ids=lambda pq: f"p{pq[0]}q{pq[1]}")
def test_seeds_match_oracle(cases, pq):
    c = cases[pq]
    same = c["leaf"] == c["rleaf"]
    assert same.mean() > 0.999  # a different leaf only where FP32 box distances tie
    assert np.array_equal(c["su"][same], c["rsu"][same]) and np.array_equal(c["sv"][same], c["rsv"][same])


@pytest.mark.parametrize("pq", CLASSES, ids=lambda pq: f"p{pq[0]}q{pq[1]}")
def test_step_seeds_match_oracle(cases, pq):
    """GN_SEED_GREEDY_STEP: the same leaf as the oracle (the greedy descent's, without the 3 x 3
    refinement of GN_SEED_GREEDY_LEAF), and the same FP32 step: bit-identical seeds except where a
    leaf Jacobian entry, evaluated in FP64 by two computations that round differently, rounds to a
    different FP32 value (then within 1e-9)."""
    c = cases[pq]
    same = c["sleaf"] == c["tleaf"]
    assert same.mean() > 0.999
    # the 3 x 3 refinement of GN_SEED_GREEDY_LEAF moves the leaf of some of the queries
    ref = c["leaf"]
    assert 0.5 < (c["sleaf"] == ref).mean() < 1.0  # about 2/3 on these surfaces
    eq = (c["ssu"] == c["tsu"]) & (c["ssv"] == c["tsv"])
    assert eq[same].mean() > 0.99
    assert np.max(np.abs(c["ssu"] - c["tsu"])[same]) <= 1e-9
    assert np.max(np.abs(c["ssv"] - c["tsv"])[same]) <= 1e-9


@pytest.mark.parametrize("pq", CLASSES, ids=lambda pq: f"p{pq[0]}q{pq[1]}")
def test_single_surface_plain_and_mixed(cases, pq):
    c = cases[pq]
    P, gs = c["P"], c["gs"]
    u, v, st = (_np(t) for t in gs.NewInverse(P[0], P[1], P[2], G.default_opts(max_iter=20, tol_step=1e-10),
                                             seeds=(c["rsu"], c["rsv"])))
    _check_plain(u, v, st, c["gn"], ("plain", pq))
    u, v, st = (_np(t) for t in gs.NewInverse(P[0], P[1], P[2], G.default_opts(**MIXED),
                                             seeds=(c["rsu"], c["rsv"])))
    _check_mixed(u, v, st, c["nt"], ("mixed", pq))


@pytest.mark.parametrize("mode", ["plain", "mixed", "mixed_step"])
def test_batch_all_classes_against_oracle(cases, mode):
    """One gn_batch_invert over all classes (one persistent launch per class group and precision
    phase), seeds found inside the batch: each point against the oracle from the oracle's seeds."""
    keys = list(cases)
    b = G.Batch([cases[k]["gs"] for k in keys])
    P = np.concatenate([cases[k]["P"] for k in keys], axis=1)
    counts = np.full(len(keys), N)
    segs = G.segments(np.arange(len(keys)), counts)
    if mode == "plain":
        opts = G.default_opts(max_iter=20, tol_step=1e-10)
    else:
        opts = G.default_opts(**MIXED)
        if mode == "mixed_step":
            opts.seed_mode = G.SEED_GREEDY_STEP
    u, v, st = (_np(t) for t in b.invert(segs, P[0], P[1], P[2], opts))
    for i, k in enumerate(keys):
        c = cases[k]
        sl = slice(i * N, (i + 1) * N)
        same_seed = c["leaf"] == c["rleaf"]
        ui, vi, sti = u[sl][same_seed], v[sl][same_seed], st[sl][same_seed]
        ref = tuple(a[same_seed] for a in (c["gn"] if mode == "plain" else c["nts"] if mode == "mixed_step"
                                           else c["nt"]))
        if mode == "plain":
            _check_plain(ui, vi, sti, ref, ("batch", k))
        else:
            _check_mixed(ui, vi, sti, ref, ("batch", k))
    b.close()


def test_batch_given_seeds_all_classes():
    """gn_batch_invert with GN_SEED_GIVEN over a mix of every class: mixed results equal the
    oracle's FP64 Newton solution from the same seeds within 1e-9 (same foot point)."""
    sds, gss, Ps, seeds, refs = [], [], [], [], []
    for i, (p, q) in enumerate(CLASSES):
        sd, os_, gs, P = _case(p, q, 8100 + 10 * i)
        rng = np.random.default_rng(i)
        u0, v0 = synth.uniform_uv(N, 8100 + 10 * i + 1)  # the constructed foot points of _case
        su = np.clip(u0 + rng.uniform(-0.01, 0.01, N), 0, 1)
        sv = np.clip(v0 + rng.uniform(-0.01, 0.01, N), 0, 1)
        refs.append(os_.gauss_newton(P[0], P[1], P[2], su, sv, tol_step=1e-12, max_iter=40, newton=1))
        gss.append(gs)
        Ps.append(P)
        seeds.append((su, sv))
    b = G.Batch(gss)
    P = np.concatenate(Ps, axis=1)
    segs = G.segments(np.arange(len(gss)), np.full(len(gss), N))
    u = torch.from_numpy(np.concatenate([s[0] for s in seeds])).cuda()
    v = torch.from_numpy(np.concatenate([s[1] for s in seeds])).cuda()
    st = torch.empty(u.numel(), dtype=torch.uint8, device="cuda")
    opts = G.default_opts(seed_mode=G.SEED_GIVEN, **MIXED)
    b.invert(segs, P[0], P[1], P[2], opts, out=(u, v, st))
    u, v, st = _np(u), _np(v), _np(st)
    for i in range(len(gss)):
        sl = slice(i * N, (i + 1) * N)
        _check_mixed(u[sl], v[sl], st[sl], refs[i], ("given", CLASSES[i]))
    b.close()
    for g in gss:
        g.close()
