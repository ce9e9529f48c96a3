"""Serving loop (gn_server_*): continuous batching of inversion requests on surfaces and curves.

Every request's results are bit-identical alone and in a mix, and equal to the direct calls
(gn_batch_invert for surfaces, gn_curve_invert for curves); admission keeps the result arena
within the HBM budget and makes requests wait for releases; errors are reported per call.
"""
import threading

import numpy as np
import pytest

import gnurbs as G
import oracle as O
from gnurbs import synth
from helpers import blocks

pytestmark = pytest.mark.gpu

MIXED = dict(precision=G.GN_MIXED, newton=G.NEWTON_FP32 | G.NEWTON_FP64, tol_step32=1e-4, tol_quad=1e-10)


@pytest.fixture(scope="module")
def world():
    rng = np.random.default_rng(21)
    sdefs = []
    for i in range(12):
        p, q = int(rng.integers(2, 6)), int(rng.integers(2, 6))
        if i == 5:
            p = 6  # generic (padded) degree class
        nu, nv = int(rng.integers(max(4, p + 1), 24)), int(rng.integers(max(4, q + 1), 24))
        sdefs.append(synth.surface_G(nu, nv, p, q, seed=4000 + i, jitter=0.4))
    surfs = []
    for sd in sdefs:
        s = G.Surface(sd.p, sd.q, sd.ku, sd.kv, sd.cp, sd.w)
        s.BuildTree(3)
        surfs.append(s)
    cdefs = [synth.helix_curve(64, 3, 500 + i) for i in range(3)]
    curves = [G.Curve.from_def(cd) for cd in cdefs]
    batch = G.Batch(surfs)
    return dict(sdefs=sdefs, surfs=surfs, cdefs=cdefs, curves=curves, batch=batch)


def queries(world, obj, n, seed):
    """Offset points near object obj (surfaces first, then curves) as (3, n) float64."""
    ns = len(world["sdefs"])
    rng = np.random.default_rng(seed)
    if obj < ns:
        o = O.OSurf.from_def(world["sdefs"][obj])
        b = blocks(o.eval(rng.random(n), rng.random(n), 1))
        P, _ = synth.offset_queries(b["S"], b["Su"], b["Sv"], 1e-2, seed + 1)
        return P
    c = O.OCurve.from_def(world["cdefs"][obj - ns])
    return c.eval(rng.random(n)) + rng.normal(0, 0.01, (3, n))


def direct(world, obj, P, opts):
    ns = len(world["sdefs"])
    if obj < ns:
        segs = G.segments([obj], [P.shape[1]])
        u, v, st = world["batch"].invert(segs, P[0], P[1], P[2], opts)
        return u.cpu().numpy(), v.cpu().numpy(), st.cpu().numpy()
    t, st = world["curves"][obj - ns].Invert(P[0], P[1], P[2], opts)
    return t.cpu().numpy(), None, st.cpu().numpy()


def fetch(server, rid):
    info = server.wait(rid, 60000)
    assert info is not None, "request did not finish"
    u, v, st = G.Server.results(info)
    out = (u.cpu().numpy(), None if v is None else v.cpu().numpy(), st.cpu().numpy())
    server.release(rid)
    return out, info


def same(a, b):
    return np.array_equal(a[0], b[0]) and np.array_equal(a[2], b[2]) and (
        (a[1] is None and b[1] is None) or np.array_equal(a[1], b[1]))


@pytest.mark.parametrize("mode", ["mixed", "plain"])
def test_results_same_alone_and_in_mix(world, mode):
    opts = G.default_opts(**MIXED) if mode == "mixed" else G.default_opts()
    nobj = len(world["sdefs"]) + len(world["curves"])
    srv = G.Server(world["batch"], world["curves"], opts, hbm_budget=1 << 30, streams=2)
    try:
        probes = [(3, 5000), (5, 777), (nobj - 1, 3000), (0, 64)]
        alone = []
        for k, (obj, n) in enumerate(probes):
            P = queries(world, obj, n, 100 + k)
            res, _ = fetch(srv, srv.submit(obj, *P))
            assert same(res, direct(world, obj, P, opts)), (obj, n)
            alone.append((obj, P, res))
        # the same requests again, each buried in a burst of unrelated ones
        rng = np.random.default_rng(7)
        ids, probe_ids = [], {}
        for k, (obj, P, _) in enumerate(alone):
            for _ in range(int(rng.integers(5, 20))):
                o2 = int(rng.integers(0, nobj))
                ids.append(srv.submit(o2, *queries(world, o2, int(rng.integers(64, 20000)), int(rng.integers(1 << 30)))))
            probe_ids[k] = srv.submit(obj, *P)
        for k, (obj, P, res) in enumerate(alone):
            got, _ = fetch(srv, probe_ids[k])
            assert same(got, res), obj
        for rid in ids:
            fetch(srv, rid)
        st = srv.stats()
        assert st.completed == st.submitted and st.steps >= 1
        assert st.max_step_requests > 1  # requests were batched together
    finally:
        srv.close()


def test_admission_within_budget(world):
    opts = G.default_opts(**MIXED)
    probe = G.Server(world["batch"], world["curves"], opts, hbm_budget=1 << 30)
    resident = (1 << 30) - probe.stats().arena_bytes
    probe.close()
    budget = resident + (4 << 20)  # a 4 MiB result arena: ~240k surface points in flight
    srv = G.Server(world["batch"], world["curves"], opts, hbm_budget=budget, streams=2)
    try:
        arena = srv.stats().arena_bytes
        assert 4 << 20 >= arena >= (4 << 20) - 256
        with pytest.raises(G.GnError) as e:
            srv.submit(0, *np.zeros((3, arena // 17 + 1000)))
        assert e.value.code == G.GN_EBUDGET
        # 30 requests of 60k points need ~30 MB of results: they must wait for releases
        n = 60000
        Ps = [queries(world, k % 12, n, 900 + k) for k in range(30)]
        ids = [srv.submit(k % 12, *P) for k, P in enumerate(Ps)]
        done = 0
        for k, rid in enumerate(ids):
            res, info = fetch(srv, rid)  # releasing lets the next ones in
            assert info.admit_ns >= info.submit_ns and info.done_ns >= info.launch_ns >= info.admit_ns
            # a request completes with its class group, so it may be released (and its arena range
            # reused) while the rest of its step still runs: every result must stay intact
            assert same(res, direct(world, k % 12, Ps[k], opts)), k
            done += 1
        st = srv.stats()
        assert done == 30 and st.completed == 30
        assert st.arena_peak <= arena
    finally:
        srv.close()


def test_concurrent_clients(world):
    opts = G.default_opts(**MIXED)
    srv = G.Server(world["batch"], world["curves"], opts, hbm_budget=1 << 30, streams=3)
    nobj = len(world["sdefs"]) + len(world["curves"])
    errors = []
    data = {}
    for t in range(4):
        rng = np.random.default_rng(t)
        data[t] = [(int(o), queries(world, int(o), int(rng.integers(64, 5000)), t * 1000 + j))
                   for j, o in enumerate(rng.integers(0, nobj, 12))]

    def client(t):
        try:
            for obj, P in data[t]:
                res, _ = fetch(srv, srv.submit(obj, *P))
                if not same(res, direct(world, obj, P, opts)):
                    errors.append((t, obj))
        except Exception as exc:  # reported below
            errors.append(repr(exc))

    try:
        th = [threading.Thread(target=client, args=(t,)) for t in range(4)]
        for x in th:
            x.start()
        for x in th:
            x.join(120)
        assert not errors, errors
    finally:
        srv.close()


@pytest.mark.timeout(120)
def test_close_with_steps_in_flight(world):
    """gn_server_destroy right after submits, while steps run, returns every time: the dispatcher
    releases its lock while it waits for the slots' steps to finish (it once held it, and the slots
    blocked finishing their steps).  A request of a closed server is either done or failed."""
    Ps = [queries(world, k, 2000, 50 + k) for k in range(8)]
    for rep in range(30):
        srv = G.Server(world["batch"], world["curves"], G.default_opts(), hbm_budget=1 << 28)
        for k in range(8):
            srv.submit(k, *Ps[k])
        srv.close()  # returns: no deadlock with the steps in flight


@pytest.mark.timeout(120)
def test_server_errors(world):
    srv = G.Server(world["batch"], world["curves"], G.default_opts(), hbm_budget=1 << 28)
    try:
        with pytest.raises(G.GnError) as e:
            srv.submit(999, *np.zeros((3, 4)))
        assert e.value.code == G.GN_ENOTFOUND
        with pytest.raises(G.GnError) as e:
            srv.poll(123456)
        assert e.value.code == G.GN_ENOTFOUND
        with pytest.raises(G.GnError) as e:
            srv.release(123456)
        assert e.value.code == G.GN_ENOTFOUND
        rid = srv.submit(0, *np.zeros((3, 0)))  # empty request completes
        assert srv.wait(rid, 10000) is not None
        srv.release(rid)
        # requests still queued when the server closes are failed, not lost
        for k in range(5):
            srv.submit(k, *queries(world, k, 2000, k))
    finally:
        srv.close()
    with pytest.raises(G.GnError):
        G.Server(world["batch"], world["curves"], G.default_opts(seed_mode=G.SEED_GIVEN))
    with pytest.raises(G.GnError) as e:
        G.Server(world["batch"], world["curves"], G.default_opts(), hbm_budget=1 << 10)
    assert e.value.code == G.GN_EBUDGET
