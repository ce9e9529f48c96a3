"""Batched inversion of many surfaces of uneven degree in one persistent launch: every point's
result equals the single-surface result (bit-identical: same arithmetic, same seeds), in any mix,
order or tile shape; hipGraph replays equal the eager launch; the autotuner's cache round-trips."""
import os

import numpy as np
import pytest

import gnurbs as G
import oracle as O
from gnurbs import synth
from helpers import blocks

pytestmark = pytest.mark.gpu

MIXED = dict(precision=G.GN_MIXED, newton=G.NEWTON_FP32 | G.NEWTON_FP64, tol_step32=1e-4, tol_quad=1e-10)


def _objects(count, seed, with_generic=True):
    rng = np.random.default_rng(seed)
    defs = []
    for i in range(count):
        p, q = int(rng.integers(2, 6)), int(rng.integers(2, 6))
        if with_generic and i % 7 == 3:
            p = 6 if i % 2 else 1  # degrees outside 2..5 take the padded generic path
        nu, nv = int(rng.integers(max(4, p + 1), 20)), int(rng.integers(max(4, q + 1), 20))
        defs.append(synth.surface_G(nu, nv, p, q, seed=seed * 1000 + i, jitter=0.4))
    surfs = []
    for sd in defs:
        s = G.Surface(sd.p, sd.q, sd.ku, sd.kv, sd.cp, sd.w)
        s.BuildTree(3)
        surfs.append(s)
    return defs, surfs


def _queries(defs, counts, seed):
    xs, u0s, v0s = [], [], []
    for i, (sd, c) in enumerate(zip(defs, counts)):
        os_ = O.OSurf.from_def(sd)
        u0, v0 = synth.uniform_uv(c, seed + i)
        b = blocks(os_.eval(u0, v0, 1))
        P, _ = synth.offset_queries(b["S"], b["Su"], b["Sv"], 1e-2, seed + 7 * i)
        xs.append(P)
        u0s.append(u0)
        v0s.append(v0)
    return np.concatenate(xs, axis=1), np.concatenate(u0s), np.concatenate(v0s)


@pytest.mark.parametrize("mode", ["mixed", "plain"])
def test_batch_equals_single_surface(mode):
    defs, surfs = _objects(14, 5)
    counts = np.random.default_rng(1).integers(100, 3000, len(defs))
    P, u0, v0 = _queries(defs, counts, 50)
    opts = G.default_opts(**MIXED) if mode == "mixed" else G.default_opts()
    b = G.Batch(surfs)
    segs = G.segments(np.arange(len(defs)), counts)
    bu, bv, bst = (t.cpu().numpy() for t in b.invert(segs, P[0], P[1], P[2], opts))
    for i, s in enumerate(surfs):
        a, c = segs["start"][i], segs["start"][i] + counts[i]
        su, sv, sst = (t.cpu().numpy() for t in s.NewInverse(P[0][a:c], P[1][a:c], P[2][a:c],
                                                             G.default_opts(**(MIXED if mode == "mixed" else {}))))
        assert np.array_equal(bu[a:c], su) and np.array_equal(bv[a:c], sv) and np.array_equal(bst[a:c], sst), i
        # against the oracle from the same seeds: the reference iteration (plain) or the FP64
        # Newton solution (mixed), within 1e-9 wherever both converge to the same foot point
        _, ssu, ssv = (t.cpu().numpy() for t in s.Seed(P[0][a:c], P[1][a:c], P[2][a:c], exact=False))
        os_ = O.OSurf.from_def(defs[i])
        if mode == "mixed":
            ru, rv, rst, _ = os_.gauss_newton(P[0][a:c], P[1][a:c], P[2][a:c], ssu, ssv, tol_step=1e-12,
                                              max_iter=40, newton=1)
        else:
            ru, rv, rst, _ = os_.gauss_newton(P[0][a:c], P[1][a:c], P[2][a:c], ssu, ssv, tol_step=1e-10,
                                              max_iter=20)
        both = (bst[a:c] == G.PT_CONVERGED) & (rst == O.CONV_STEP)
        d = np.maximum(np.abs(bu[a:c] - ru), np.abs(bv[a:c] - rv))[both]
        same = d < 1e-6
        assert same.mean() > 0.98, i
        assert np.max(d[same], initial=0.0) <= 1e-9, (i, float(np.max(d[same], initial=0.0)))


def test_batch_order_and_shape_independent():
    defs, surfs = _objects(10, 9)
    counts = np.random.default_rng(2).integers(50, 2000, len(defs))
    P, _, _ = _queries(defs, counts, 70)
    b = G.Batch(surfs)
    opts = G.default_opts(**MIXED)
    segs = G.segments(np.arange(len(defs)), counts)
    ref = [t.cpu().numpy() for t in b.invert(segs, P[0], P[1], P[2], opts)]
    # reversed segment order and split segments, other tile shapes
    rev = segs[::-1].copy()
    for tile, bpc in ((256, 0), (4096, 1), (777, 2)):
        b.set_shape(tile, bpc)
        got = [t.cpu().numpy() for t in b.invert(rev, P[0], P[1], P[2], G.default_opts(**MIXED))]
        for a, c in zip(ref, got):
            assert np.array_equal(a, c), (tile, bpc)


def test_batch_graph_replays_match_eager():
    import torch

    defs, surfs = _objects(8, 11, with_generic=False)
    counts = np.random.default_rng(3).integers(500, 3000, len(defs))
    P, _, _ = _queries(defs, counts, 90)
    b = G.Batch(surfs)
    segs = G.segments(np.arange(len(defs)), counts)
    opts = G.default_opts(**MIXED)
    eager = [t.cpu().numpy() for t in b.invert(segs, P[0], P[1], P[2], opts)]
    x, y, z = (torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in P)
    n = x.numel()
    u = torch.zeros(n, dtype=torch.float64, device="cuda")
    v = torch.zeros_like(u)
    st = torch.zeros(n, dtype=torch.uint8, device="cuda")
    s = torch.cuda.Stream()
    g = b.graph(segs, x, y, z, u, v, st, G.default_opts(**MIXED), stream=s)
    for rep in range(3):
        u.fill_(-1.0)
        v.fill_(-1.0)
        torch.cuda.synchronize()
        g.launch(stream=s)
        s.synchronize()
        got = [u.cpu().numpy(), v.cpu().numpy(), st.cpu().numpy()]
        for a, c in zip(eager, got):
            assert np.array_equal(a, c), rep
        # eager work between replays
        b.invert(segs, x, y, z, G.default_opts(**MIXED))
        torch.cuda.synchronize()
    g.close()


def test_batch_graph_replays_beside_concurrent_eager_inversions():
    """A graph keeps the claim-counter slots it was captured with.  They are its own: replays on
    one stream while more than 256 eager mixed inversions (two counter slots each, so the shared
    ring wraps twice) run on another stream give bit-identical results on both sides."""
    import torch

    defs, surfs = _objects(6, 17, with_generic=False)
    counts = np.random.default_rng(5).integers(500, 3000, len(defs))
    P, _, _ = _queries(defs, counts, 130)
    b = G.Batch(surfs)
    segs = G.segments(np.arange(len(defs)), counts)
    opts = G.default_opts(**MIXED)
    ref = [t.cpu().numpy() for t in b.invert(segs, P[0], P[1], P[2], opts)]
    x, y, z = (torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in P)
    n = x.numel()
    u = torch.zeros(n, dtype=torch.float64, device="cuda")
    v = torch.zeros_like(u)
    st = torch.zeros(n, dtype=torch.uint8, device="cuda")
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    g = b.graph(segs, x, y, z, u, v, st, G.default_opts(**MIXED), stream=s1)
    # eager work: 300 single-surface mixed inversions of one surface's points on s2
    a0, c0 = int(segs["start"][0]), int(counts[0])
    ex, ey, ez = x[a0:a0 + c0].contiguous(), y[a0:a0 + c0].contiguous(), z[a0:a0 + c0].contiguous()
    eref = [t.clone() for t in surfs[0].NewInverse(ex, ey, ez, G.default_opts(**MIXED))]
    outs = [(torch.empty(c0, dtype=torch.float64, device="cuda"), torch.empty(c0, dtype=torch.float64, device="cuda"),
             torch.empty(c0, dtype=torch.uint8, device="cuda")) for _ in range(300)]
    torch.cuda.synchronize()
    for i in range(300):
        if i % 15 == 0:
            g.launch(stream=s1)
        surfs[0].NewInverse(ex, ey, ez, G.default_opts(**MIXED), out=outs[i], stream=s2)
    torch.cuda.synchronize()
    got = [u.cpu().numpy(), v.cpu().numpy(), st.cpu().numpy()]
    for a, c in zip(ref, got):
        assert np.array_equal(a, c)
    for i, o in enumerate(outs):
        assert all(torch.equal(a, c) for a, c in zip(eref, o)), i
    g.close()
    b.close()


def test_batch_tune_and_cache(tmp_path):
    defs, surfs = _objects(6, 13, with_generic=False)
    counts = np.random.default_rng(4).integers(1000, 4000, len(defs))
    P, _, _ = _queries(defs, counts, 110)
    b = G.Batch(surfs)
    segs = G.segments(np.arange(len(defs)), counts)
    path = str(tmp_path / "tiles.txt")
    shape = b.tune(segs, P[0], P[1], P[2], G.default_opts(**MIXED), cache_path=path)
    assert shape[0] in (128, 256, 512, 1024, 4096)  # the candidate tile sizes
    assert os.path.exists(path) and len(open(path).read().splitlines()) == 1
    b2 = G.Batch(surfs)
    assert b2.tune(segs, P[0], P[1], P[2], G.default_opts(**MIXED), cache_path=path) == shape
    assert len(open(path).read().splitlines()) == 1  # reused, not re-tuned


def test_batch_rejects_bad_segments():
    defs, surfs = _objects(2, 15, with_generic=False)
    b = G.Batch(surfs)
    x = np.zeros(10)
    with pytest.raises(G.GnError):
        b.invert(G.segments([5], [3]), x, x, x)
    with pytest.raises(G.GnError):
        b.invert(G.segments([0], [30]), x, x, x)
    with pytest.raises(G.GnError):
        b.invert(G.segments([0], [3]), x, x, x, G.default_opts(seed_mode=G.SEED_REF_FIRST_HIT))
    s = G.Surface(defs[0].p, defs[0].q, defs[0].ku, defs[0].kv, defs[0].cp, defs[0].w)
    with pytest.raises(G.GnError):
        G.Batch([s])  # no tree


def test_batch_tune_with_given_seeds():
    """The autotuner times every candidate shape from the seeds held in u and v on entry (it keeps
    a copy), so it tunes given-seed work too; its final results equal an untuned call's from the
    same seeds."""
    import torch

    defs, surfs = _objects(6, 19, with_generic=False)
    counts = np.random.default_rng(6).integers(500, 3000, len(defs))
    P, u0, v0 = _queries(defs, counts, 150)
    rng = np.random.default_rng(7)
    su = np.clip(u0 + rng.uniform(-0.01, 0.01, u0.size), 0, 1)
    sv = np.clip(v0 + rng.uniform(-0.01, 0.01, v0.size), 0, 1)
    b = G.Batch(surfs)
    segs = G.segments(np.arange(len(defs)), counts)
    opts = G.default_opts(seed_mode=G.SEED_GIVEN, **MIXED)

    def bufs():
        return (torch.from_numpy(su.copy()).cuda(), torch.from_numpy(sv.copy()).cuda(),
                torch.empty(su.size, dtype=torch.uint8, device="cuda"))

    tuned = bufs()
    shape = b.tune(segs, P[0], P[1], P[2], opts, out=tuned)
    assert shape[0] in (128, 256, 512, 1024, 4096)
    ref = bufs()
    b.invert(segs, P[0], P[1], P[2], opts, out=ref)
    for a, c in zip(tuned, ref):
        assert torch.equal(a, c)
    b.close()
