"""The split FP64 closing pass of the mixed mode: a streamed first step for every point, then a
lane-refill pass over the listed points that continue.  Each of the list's 64 segments holds 2,048
points; past that a lane finishes its point itself.  All three routes must give every point the
same bits, because a point's iteration does not depend on which route or launch carries it:
  * a run where most points continue (one FP32 iteration, up to 10 FP64 steps) over 1,048,576
    points overflows every segment;
  * a 40,000-point subset of it, run alone, fits the list;
  * the subset's results must be bit-identical to its points in the full run, and converged points
    must be stationary points of the distance (residual orthogonal to both tangents).
"""
import numpy as np
import pytest
import torch

import gnurbs as G
from gnurbs import synth

pytestmark = pytest.mark.gpu

OPTS = dict(precision=G.GN_MIXED, newton=G.NEWTON_FP32 | G.NEWTON_FP64, fp32_iters=1, fp64_tail=10,
            tol_step32=1e-4, tol_step=1e-12, tol_quad=1e-6)


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
    # converged through the 1e-6 quadratic certificate: within 1e-9 but for a few 1e-9 at flat
    # minima (an unconverged point would be 1e-6 or more off)
    for t in (e[3:6], e[6:9]):
        off = (t * r).sum(0).abs() / (t * t).sum(0)
        assert (off[inner] < 1e-9).float().mean().item() > 0.9999
        assert off[inner].max().item() < 1e-8
    gs.close()
