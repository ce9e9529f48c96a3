"""The split FP64 closing pass of the mixed mode: a streamed first step for every point, then a
lane-refill pass over the listed points that continue.  Each of the list's 64 segments holds 2,048
points; past that a lane finishes its point itself.  All three routes must give every point the
same bits, because a point's iteration does not depend on which route or launch carries it:
  * a run where most points continue (one FP32 iteration, up to 10 FP64 steps) over 1,048,576
    points overflows every segment;
  * a 40,000-point subset of it, run alone, fits the list;
  * the subset's results must be bit-identical to its points in the full run, and converged points
    must be stationary points of the distance (residual orthogonal to both tangents);
  * calls in flight at once on different streams keep their lists apart.
"""
import numpy as np
import pytest
import torch

import gnurbs as G
from gnurbs import synth

pytestmark = pytest.mark.gpu

OPTS = dict(precision=G.GN_MIXED, newton=G.NEWTON_FP32 | G.NEWTON_FP64, fp32_iters=1, fp64_tail=10,
            tol_step32=1e-4, tol_step=1e-12, tol_quad=1e-10)


def test_concurrent_mixed_calls_on_two_streams():
    """Two mixed inversions in flight at once on different streams (the same surface and a second
    one) take separate counter slots and continuation lists: each result is bit-identical to the
    same call run alone."""
    n = 200000
    surfs, queries = [], []
    for k, seed in enumerate((1002, 1005)):
        sd = synth.surface_G(32, 32, 3, 3, seed=seed)
        gs = G.Surface(sd.p, sd.q, sd.ku, sd.kv, sd.cp, sd.w)
        gs.BuildTree(3)
        u0, v0 = synth.uniform_uv(n, seed=seed + 1)
        ev = gs.Evaluate(u0, v0, 1)
        nrm = torch.linalg.cross(ev[3:6].T, ev[6:9].T).T
        nrm = nrm / torch.linalg.norm(nrm, dim=0, keepdim=True)
        delta = torch.from_numpy(np.random.default_rng(seed + 2).normal(0.0, 1e-2, n)).cuda()
        P = (ev[0:3] + delta * nrm).contiguous()
        surfs.append(gs)
        queries.append(tuple(P[i].contiguous() for i in range(3)))
    opts = G.default_opts(**dict(OPTS, fp32_iters=20, fp64_tail=3))
    alone = [surfs[k].NewInverse(*queries[k], opts) for k in (0, 1, 0)]
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream() for _ in range(3)]
    res = []
    for k, s in zip((0, 1, 0), streams):
        u = torch.empty(n, dtype=torch.float64, device="cuda")
        v = torch.empty_like(u)
        st = torch.empty(n, dtype=torch.uint8, device="cuda")
        surfs[k].NewInverse(*queries[k], opts, out=(u, v, st), stream=s)
        res.append((u, v, st))
    torch.cuda.synchronize()
    for a, b in zip(alone, res):
        assert all(torch.equal(x, y) for x, y in zip(a, b))
    for s in surfs:
        s.close()


def test_list_overflow_matches_listed_path():
    N = 1 << 20
    sd = synth.surface_G(32, 32, 3, 3, seed=1002)
    gs = G.Surface(sd.p, sd.q, sd.ku, sd.kv, sd.cp, sd.w)
    gs.BuildTree(3)
    u0, v0 = synth.uniform_uv(N, seed=77)
    ev = gs.Evaluate(u0, v0, 1)
    nrm = torch.linalg.cross(ev[3:6].T, ev[6:9].T).T
    nrm = nrm / torch.linalg.norm(nrm, dim=0, keepdim=True)
    delta = torch.from_numpy(np.random.default_rng(78).normal(0.0, 1e-2, N)).cuda()
    P = (ev[0:3] + delta * nrm).contiguous()
    x, y, z = P[0].contiguous(), P[1].contiguous(), P[2].contiguous()
    _, su, sv = gs.Seed(x, y, z, exact=False)
    u, v, st, it32, it64 = gs.NewInverse(x, y, z, G.default_opts(**OPTS), seeds=(su, sv), iters=True)
    # most points took more than one FP64 step, so the 64 segments of 2,048 overflowed
    assert (it64 > 1).float().mean().item() > 0.5
    sel = torch.from_numpy(np.sort(np.random.default_rng(79).choice(N, 40000, replace=False))).cuda()
    xs, ys, zs = x[sel].contiguous(), y[sel].contiguous(), z[sel].contiguous()
    su_s, sv_s = su[sel].contiguous(), sv[sel].contiguous()
    us, vs, sts, i32s, i64s = gs.NewInverse(xs, ys, zs, G.default_opts(**OPTS), seeds=(su_s, sv_s),
                                            iters=True)
    assert torch.equal(us, u[sel]) and torch.equal(vs, v[sel])
    assert torch.equal(sts, st[sel]) and torch.equal(i64s, it64[sel]) and torch.equal(i32s, it32[sel])
    conv = st == G.PT_CONVERGED
    assert conv.float().mean().item() > 0.95
    e = gs.Evaluate(u, v, 1)
    r = e[0:3] - P
    inner = conv & (u > 0) & (u < 1) & (v > 0) & (v < 1)
    # converged through the Newton-step certificate, which bounds the error left by the
    # conditioning of the Hessian: stationary within 1e-9 everywhere, flat minima included (an
    # unconverged point would be 1e-6 or more off)
    for t in (e[3:6], e[6:9]):
        off = (t * r).sum(0).abs() / (t * t).sum(0)
        assert off[inner].max().item() < 1e-9
    gs.close()
