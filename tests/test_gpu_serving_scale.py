"""The serving workload at its stated scale: 224 uneven surfaces (the uneven-batch generator, seed
1004: degrees 2..5, nets 4..48, jittered knots) and 32 cubic rational helix curves (64 control
points) resident, a 32 GB HBM budget, two step slots, and a Poisson trace of requests of
log-uniform size on [64, 100000] points for uniformly drawn objects at about 80% of the server's
capacity for half a second (continuous batching: requests arrive while others run).

Checks: every request completes; 200 sampled requests are bit-identical to the direct calls on
their points (gn_batch_invert for surfaces, gn_curve_invert for curves), i.e. the same alone and
in the mix; 40 sampled requests (up to 2,000 points each) agree with the oracle within 1e-9 in
(u, v) or t wherever both converge to the same foot point (surfaces: the oracle's FP64 Newton
from its GN_SEED_GREEDY_STEP seeds; curves: its 1-D Newton from its sample seeds)."""
import ctypes

import numpy as np
import pytest
import torch

import gnurbs as G
import oracle as O
from gnurbs import synth

pytestmark = pytest.mark.gpu

MIXED = dict(precision=G.GN_MIXED, newton=G.NEWTON_FP32 | G.NEWTON_FP64, fp32_iters=20, fp64_tail=3,
             tol_step32=1e-4, tol_step=1e-10, tol_quad=1e-10, seed_mode=G.SEED_GREEDY_STEP)
POOL = 100_000
NS, NCURVE = 224, 32


@pytest.fixture(scope="module")
def world():
    defs, _ = synth.uneven_batch(NS, seed=1004)
    cdefs = [synth.helix_curve(64, 3, seed=1004 * 100 + i) for i in range(NCURVE)]
    gen = torch.Generator(device="cuda").manual_seed(1004)
    surfs, curves, pools = [], [], []
    for sd in defs:
        s = G.Surface(sd.p, sd.q, sd.ku, sd.kv, sd.cp, sd.w)
        s.BuildTree(3)
        surfs.append(s)
        u0 = torch.rand(POOL, dtype=torch.float64, device="cuda", generator=gen)
        v0 = torch.rand(POOL, dtype=torch.float64, device="cuda", generator=gen)
        ev = s.Evaluate(u0, v0, 1)
        nrm = torch.linalg.cross(ev[3:6].T, ev[6:9].T).T
        nrm = nrm / torch.linalg.norm(nrm, dim=0, keepdim=True)
        P = ev[0:3] + torch.randn(POOL, dtype=torch.float64, device="cuda", generator=gen) * 1e-2 * nrm
        pools.append(P.contiguous())
    for cd in cdefs:
        c = G.Curve.from_def(cd)
        curves.append(c)
        t0 = torch.rand(POOL, dtype=torch.float64, device="cuda", generator=gen)
        ev = c.Evaluate(t0, 1)
        tan = ev[3:6] / torch.linalg.norm(ev[3:6], dim=0, keepdim=True)
        a = torch.randn((3, POOL), dtype=torch.float64, device="cuda", generator=gen)
        nrm = a - tan * (a * tan).sum(0)
        nrm = nrm / torch.linalg.norm(nrm, dim=0, keepdim=True)
        P = ev[0:3] + torch.randn(POOL, dtype=torch.float64, device="cuda", generator=gen) * 1e-2 * nrm
        pools.append(P.contiguous())
    torch.cuda.synchronize()
    batch = G.Batch(surfs)
    yield dict(defs=defs, cdefs=cdefs, surfs=surfs, curves=curves, pools=pools, batch=batch)
    batch.close()
    for o in surfs + curves:
        o.close()


def _direct(world, obj, P, opts):
    if obj < NS:
        segs = G.segments([obj], [P.shape[1]])
        u, v, st = world["batch"].invert(segs, P[0].contiguous(), P[1].contiguous(), P[2].contiguous(), opts)
        return u, v, st
    t, st = world["curves"][obj - NS].Invert(P[0].contiguous(), P[1].contiguous(), P[2].contiguous(), opts)
    return t, None, st


def test_serving_at_scale(world):
    L = G.lib()
    srv = G.Server(world["batch"], world["curves"], G.default_opts(**MIXED), hbm_budget=32 << 30, streams=2)
    try:
        rng = np.random.default_rng(404)
        lo, hi = 64, 100_000
        # capacity from a saturating burst of 3,000 requests
        def trace(n):
            sizes = np.exp(rng.uniform(np.log(lo), np.log(hi), n)).astype(np.int64)
            objs = rng.integers(0, NS + NCURVE, n)
            offs = (rng.random(n) * (POOL - sizes + 1)).astype(np.int64)
            return sizes, objs, offs

        def submit(obj, off, n):
            P = world["pools"][obj]
            rid = ctypes.c_int64()
            base = P.data_ptr() + 8 * int(off)
            rc = L.gn_server_submit(srv._h, int(obj), base, base + 8 * POOL, base + 16 * POOL, int(n),
                                    ctypes.byref(rid))
            assert rc == 0, L.gn_last_error()
            return rid.value

        sizes, objs, offs = trace(3000)
        t0 = G.now_ns()
        ids = [submit(o, f, n) for o, f, n in zip(objs, offs, sizes)]
        for rid in ids:
            assert srv.wait(rid, 120000) is not None
            srv.release(rid)
        cap = sizes.sum() / ((G.now_ns() - t0) / 1e9)
        # Poisson arrivals at 80% of that capacity for 0.5 s
        mean = (hi - lo) / np.log(hi / lo)
        rate = 0.8 * cap / mean
        n_req = int(rate * 0.5)
        sizes, objs, offs = trace(n_req)
        arr = np.cumsum(rng.exponential(1.0 / rate, n_req))
        sample = set(rng.choice(n_req, 200, replace=False).tolist())
        small = [i for i in sorted(sample) if sizes[i] <= 2000][:40]
        ids = []
        start = G.now_ns() + 1_000_000
        for i in range(n_req):
            due = start + int(arr[i] * 1e9)
            while G.now_ns() < due:
                pass
            ids.append(submit(objs[i], offs[i], sizes[i]))
        kept = {}
        for i, rid in enumerate(ids):
            info = srv.wait(rid, 120000)
            assert info is not None, i
            if i in sample:
                u, v, st = G.Server.results(info)
                kept[i] = (u.clone(), None if v is None else v.clone(), st.clone())
            srv.release(rid)
        stats = srv.stats()
        assert stats.completed == stats.submitted and stats.max_step_requests > 1
    finally:
        srv.close()
    opts = G.default_opts(**MIXED)
    for i, (u, v, st) in kept.items():
        o, f, n = int(objs[i]), int(offs[i]), int(sizes[i])
        P = world["pools"][o][:, f:f + n]
        du, dv, dst = _direct(world, o, P, opts)
        assert torch.equal(u, du) and torch.equal(st, dst), (i, o, n)
        assert (v is None and dv is None) or torch.equal(v, dv), (i, o, n)
    assert len(small) >= 10
    for i in small:
        o, f, n = int(objs[i]), int(offs[i]), int(sizes[i])
        u, v, st = (None if a is None else a.cpu().numpy() for a in kept[i])
        Pn = world["pools"][o][:, f:f + n].cpu().numpy()
        if o < NS:
            sd = world["defs"][o]
            os_ = O.OSurf.from_def(sd)
            nodes, _, _ = world["surfs"][o].GetTree()
            leaf, ssu, ssv = os_.seed_greedy_step(nodes, 3, Pn[0], Pn[1], Pn[2])
            gleaf, _, _ = (t.cpu().numpy() for t in world["surfs"][o].Seed(Pn[0], Pn[1], Pn[2], exact="step"))
            ru, rv, rst, _ = os_.gauss_newton(Pn[0], Pn[1], Pn[2], ssu, ssv, tol_step=1e-12, max_iter=40, newton=1)
            both = (st == G.PT_CONVERGED) & (rst == O.CONV_STEP) & (leaf == gleaf)
            d = np.maximum(np.abs(u - ru), np.abs(v - rv))[both]
        else:
            oc = O.OCurve.from_def(world["cdefs"][o - NS])
            rt, rst, _ = oc.invert(Pn[0], Pn[1], Pn[2], oc.seed(Pn[0], Pn[1], Pn[2]), tol_step=1e-12,
                                   max_iter=40, newton=1)
            both = (st == G.PT_CONVERGED) & (rst == O.CONV_STEP)
            d = np.abs(u - rt)[both]
        same = d < 1e-6
        assert both.mean() > 0.9 and same.mean() > 0.97, (i, o, n, both.mean(), same.mean())
        assert np.max(d[same], initial=0.0) <= 1e-9, (i, o, n, float(np.max(d[same], initial=0.0)))
