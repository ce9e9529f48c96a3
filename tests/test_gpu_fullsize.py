"""Parity at the workloads' full sizes through size-independent properties (the oracle checks a
random subset; the rest is consistency the whole batch must satisfy):
  * evaluation, 16,777,216 points, bicubic 64x64, order 2: a random subset against the oracle
    (1e-12 of the field scale); the subset evaluated alone is bit-identical to its values in the
    full batch; FP32 within 1e-5 of FP64 at the same parameters over all points;
  * projection, 1,048,576 queries, bicubic 32x32: converged mixed results inside the domain are
    stationary points (residual orthogonal to both tangents, within 1e-9 in parameter units), all
    agree with the same algorithm run in FP64 and with the constructed foot points within 1e-9
    where they reach the same foot point, and a random subset agrees with the oracle's
    Gauss-Newton from the same seeds within 1e-9 where both reach the same foot point (a few per
    mille reach another local minimum of the distance);
  * uneven batch, 1,024 surfaces (~4.95 M points): every sampled surface's results in the full
    batch are bit-identical to a single-surface inversion of its points alone.
"""
import numpy as np
import pytest
import torch

import oracle as O
import gnurbs as G
from gnurbs import synth
from helpers import blocks, field_close

pytestmark = pytest.mark.gpu

FIELDS = ("S", "Su", "Sv", "Suu", "Suv", "Svv")
MIXED = dict(precision=G.GN_MIXED, newton=G.NEWTON_FP32 | G.NEWTON_FP64, fp32_iters=20, fp64_tail=3,
             tol_step32=1e-4, tol_step=1e-10, tol_quad=1e-10)


def test_evaluation_16M_points():
    n = 1 << 24
    sd = synth.surface_G(64, 64, 3, 3, seed=1001)
    gs = G.Surface(sd.p, sd.q, sd.ku, sd.kv, sd.cp, sd.w)
    gen = torch.Generator(device="cuda").manual_seed(1001)
    u = torch.rand(n, dtype=torch.float64, device="cuda", generator=gen)
    v = torch.rand(n, dtype=torch.float64, device="cuda", generator=gen)
    out = gs.Evaluate(u, v, 2)
    rng = np.random.default_rng(5)
    sel = torch.from_numpy(np.sort(rng.choice(n, 50000, replace=False))).cuda()
    us, vs = u[sel], v[sel]
    got = out[:, sel].cpu().numpy()
    ref = blocks(O.OSurf.from_def(sd).eval(us.cpu().numpy(), vs.cpu().numpy(), 2))
    floor = 1e-2 * np.max(np.abs(ref["S"]))
    g = blocks(got)
    for f in FIELDS:
        ok, err = field_close(g[f], ref[f], 1e-12, floor)
        assert ok, (f, err)
    alone = gs.Evaluate(us.contiguous(), vs.contiguous(), 2).cpu().numpy()
    assert np.array_equal(alone, got)
    del out
    u32, v32 = u.float(), v.float()
    o32 = gs.Evaluate(u32, v32, 2, precision=G.GN_FP32)
    o64 = gs.Evaluate(u32.double(), v32.double(), 2)
    for b in range(6):
        r = o64[3 * b:3 * b + 3]
        scale = max(r.abs().max().item(), 1e-2 * o64[0:3].abs().max().item())
        err = (o32[3 * b:3 * b + 3].double() - r).abs().max().item() / scale
        assert err < 1e-5, (FIELDS[b], err)
    gs.close()


def test_projection_1M_queries():
    N = 1 << 20
    sd = synth.surface_G(32, 32, 3, 3, seed=1002)
    gs = G.Surface(sd.p, sd.q, sd.ku, sd.kv, sd.cp, sd.w)
    gs.BuildTree(3)
    u0, v0 = synth.uniform_uv(N, seed=1003)
    ev = gs.Evaluate(u0, v0, 1)
    nrm = torch.linalg.cross(ev[3:6].T, ev[6:9].T).T
    nrm = nrm / torch.linalg.norm(nrm, dim=0, keepdim=True)
    delta = torch.from_numpy(np.random.default_rng(1004).normal(0.0, 1e-2, N)).cuda()
    P = (ev[0:3] + delta * nrm).contiguous()
    x, y, z = P[0].contiguous(), P[1].contiguous(), P[2].contiguous()
    u, v, st = gs.NewInverse(x, y, z, G.default_opts(**MIXED))
    conv = st == G.PT_CONVERGED
    assert conv.float().mean().item() > 0.99
    # stationarity: (S - P) orthogonal to S_u and S_v at converged points inside the domain (on
    # its boundary the clamped minimum leaves a residual across it), measured in parameter units:
    # |t.(S-P)| / |t|^2 is the first-order distance to the foot point along t
    e = gs.Evaluate(u, v, 1)
    r = e[0:3] - P
    inner = conv & (u > 0) & (u < 1) & (v > 0) & (v < 1)
    assert inner.float().mean().item() > 0.99
    for t in (e[3:6], e[6:9]):
        off = (t * r).sum(0).abs() / (t * t).sum(0)
        assert off[inner].max().item() < 1e-9
    # the same algorithm in FP64 reaches the same foot points
    f64 = dict(MIXED, precision=G.GN_FP64, newton=G.NEWTON_FP64, max_iter=20)
    fu, fv, fst = gs.NewInverse(x, y, z, G.default_opts(**f64))
    both = conv & (fst == G.PT_CONVERGED)
    d = torch.maximum((u - fu).abs(), (v - fv).abs())[both]
    same = d < 1e-6
    assert same.float().mean().item() > 0.999
    assert d[same].max().item() <= 1e-9
    # every converged point that reached the constructed foot point is within 1e-9 of it (about
    # 5% of the offset queries have a nearer local minimum elsewhere, which they converge to)
    dt = torch.maximum((u - torch.from_numpy(u0).cuda()).abs(), (v - torch.from_numpy(v0).cuda()).abs())[conv]
    assert (dt < 1e-6).float().mean().item() > 0.9
    assert dt[dt < 1e-6].max().item() <= 1e-9
    # a random subset against the oracle's Gauss-Newton from the same seeds
    _, su, sv = gs.Seed(x, y, z, exact=False)
    rng = np.random.default_rng(9)
    sel = torch.from_numpy(rng.choice(N, 20000, replace=False)).cuda()
    os_ = O.OSurf.from_def(sd)
    ou, ov, ost, _ = os_.gauss_newton(x[sel].cpu().numpy(), y[sel].cpu().numpy(), z[sel].cpu().numpy(),
                                      su[sel].cpu().numpy(), sv[sel].cpu().numpy(), tol_step=1e-12,
                                      max_iter=30, newton=1)
    gu, gv, gst = u[sel].cpu().numpy(), v[sel].cpu().numpy(), st[sel].cpu().numpy()
    ok = (ost == 0) & (gst == G.PT_CONVERGED)
    assert ok.mean() > 0.99
    dd = np.maximum(np.abs(gu - ou), np.abs(gv - ov))[ok]
    # from one seed the FP32 bulk and the FP64 iteration can cross into different basins of the
    # distance (both results stationary, checked above): a few per mille, 1e-2 apart or more
    same = dd < 1e-6
    assert same.mean() > 0.995
    # the same foot point: within 1e-9 everywhere, flat minima included (the certificate scales
    # the step it accepts with the Hessian's conditioning)
    assert dd[same].max() <= 1e-9
    gs.close()


def test_uneven_batch_full_size_matches_single_surfaces():
    defs, counts = synth.uneven_batch(1024, seed=1003)
    gen = torch.Generator(device="cuda").manual_seed(1003)
    surfs, xs = [], []
    for sd, c in zip(defs, counts):
        s = G.Surface(sd.p, sd.q, sd.ku, sd.kv, sd.cp, sd.w)
        s.BuildTree(3)
        surfs.append(s)
        u0 = torch.rand(int(c), dtype=torch.float64, device="cuda", generator=gen)
        v0 = torch.rand(int(c), dtype=torch.float64, device="cuda", generator=gen)
        ev = s.Evaluate(u0, v0, 1)
        nrm = torch.linalg.cross(ev[3:6].T, ev[6:9].T).T
        nrm = nrm / torch.linalg.norm(nrm, dim=0, keepdim=True)
        xs.append(ev[0:3] + torch.randn(int(c), dtype=torch.float64, device="cuda", generator=gen) * 1e-2 * nrm)
    P = torch.cat(xs, dim=1)
    x, y, z = P[0].contiguous(), P[1].contiguous(), P[2].contiguous()
    b = G.Batch(surfs)
    segs = G.segments(np.arange(len(defs)), counts)
    u, v, st = b.invert(segs, x, y, z, G.default_opts(**MIXED))
    starts = np.concatenate([[0], np.cumsum(counts)[:-1]])
    order = np.argsort(counts)
    # the largest and smallest surfaces and a spread of others (all degree classes occur)
    pick = list(order[-4:]) + list(order[:4]) + list(np.random.default_rng(3).choice(len(defs), 8, replace=False))
    for i in pick:
        a, c = int(starts[i]), int(counts[i])
        su_, sv_, sst = surfs[i].NewInverse(x[a:a + c].contiguous(), y[a:a + c].contiguous(),
                                            z[a:a + c].contiguous(), G.default_opts(**MIXED))
        assert torch.equal(su_, u[a:a + c]) and torch.equal(sv_, v[a:a + c]), i
        assert torch.equal(sst, st[a:a + c]), i
    b.close()
    for s in surfs:
        s.close()
