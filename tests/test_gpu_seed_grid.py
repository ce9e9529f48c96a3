"""The seed grid (gn_build_seed_grid) and GN_SEED_GRID_STEP against the oracle (or_seed_grid_*),
through the C ABI.  The oracle takes the tree nodes the library built (gn_get_tree), so the grid's
geometry and cells, and the seeds, are compared bit for bit; the inversion from the grid seeds is
compared with the oracle's FP64 exact-Hessian Gauss-Newton from the oracle's own grid seeds,
within 1e-9 wherever both converge to the same foot point."""
import numpy as np
import pytest

import oracle as O
import gnurbs as G
from gnurbs import synth
from helpers import blocks

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


def _surface(kind):
    if kind == "bicubic":
        sd = synth.surface_G(32, 32, 3, 3, seed=1002)
        return sd, sd.ku
    sd = synth.surface_G(16, 14, 3, 2, seed=7, jitter=0.3)
    return sd, sd.ku * 5.0  # a u domain of [0, 5]


def _pair(kind, levels, res):
    sd, ku = _surface(kind)
    gs = G.Surface(sd.p, sd.q, ku, sd.kv, sd.cp, sd.w)
    os_ = O.OSurf(sd.p, sd.q, ku, sd.kv, sd.cp, sd.w)
    gs.BuildTree(levels)
    gs.BuildSeedGrid(res)
    nodes, _, _ = gs.GetTree()
    return sd, ku, gs, os_, nodes


def _queries(sd, ku, os_, n, seed):
    rng = np.random.default_rng(seed)
    u0 = ku[0] + (ku[-1] - ku[0]) * rng.random(n)
    v0 = rng.random(n)
    b = blocks(os_.eval(u0, v0, 1))
    P, _ = synth.offset_queries(b["S"], b["Su"], b["Sv"], 1e-2, seed + 1)
    P = P.copy()
    k = n // 50
    P[:, :k] = P[:, :k] * 10.0 - 3.0             # far outside the grid: clamped to border cells
    P[:, k:2 * k] = b["S"][:, k:2 * k]           # on the surface
    P[0, 2 * k] = np.nan                         # non-finite coordinates: cell 0 on that axis
    P[1, 2 * k + 1] = np.inf
    P[2, 2 * k + 2] = -np.inf
    return u0, v0, P


@pytest.mark.parametrize("kind,levels,res", [("bicubic", 3, 0), ("bicubic", 3, 40), ("domain5", 2, 0),
                                             ("domain5", 3, 24)])
def test_grid_equals_oracle(kind, levels, res):
    sd, ku, gs, os_, nodes = _pair(kind, levels, res)
    dims, org, cell, cells = gs.GetSeedGrid()
    od, oo, oc, ocells = O.seed_grid(nodes, levels, res)
    assert np.array_equal(dims, od) and np.array_equal(org, oo) and cell == oc
    assert np.array_equal(cells, ocells)
    if res == 0 and levels == 3:
        assert int(dims.max()) >= 128  # the default: cells of half a leaf's extent or finer


@pytest.mark.parametrize("kind,levels,res", [("bicubic", 3, 0), ("domain5", 3, 24)])
def test_grid_step_seeds_equal_oracle(kind, levels, res):
    sd, ku, gs, os_, nodes = _pair(kind, levels, res)
    u0, v0, P = _queries(sd, ku, os_, 30000, 11)
    leaf, su, sv = (_np(t) for t in gs.Seed(P[0], P[1], P[2], exact="grid"))
    grid = gs.GetSeedGrid()
    rl, rsu, rsv = os_.seed_grid_step(nodes, levels, grid, P[0], P[1], P[2])
    assert np.array_equal(leaf, rl)
    assert np.array_equal(su, rsu) and np.array_equal(sv, rsv)
    # inside the domain; at the default resolution the grid's leaf is the exact nearest leaf for
    # most offset points (a 24-cell grid is coarser than the 64 x 64 leaves)
    assert su.min() >= ku[0] and su.max() <= ku[-1] and sv.min() >= 0 and sv.max() <= 1
    nl, _, _ = (_np(t) for t in gs.Seed(P[0], P[1], P[2], exact=True))
    assert (nl == leaf)[2000:].mean() > (0.6 if res == 0 else 0.1)


@pytest.mark.parametrize("precision", [G.GN_MIXED, G.GN_FP64])
def test_grid_step_inversion_matches_oracle(precision):
    sd, ku, gs, os_, nodes = _pair("bicubic", 3, 0)
    u0, v0, P = _queries(sd, ku, os_, 20000, 21)
    if precision == G.GN_MIXED:
        opts = G.default_opts(precision=G.GN_MIXED, newton=G.NEWTON_FP32 | G.NEWTON_FP64,
                              seed_mode=G.SEED_GRID_STEP)
    else:
        opts = G.default_opts(max_iter=20, tol_step=1e-10, newton=G.NEWTON_FP64, seed_mode=G.SEED_GRID_STEP)
    u, v, st = (_np(t) for t in gs.NewInverse(P[0], P[1], P[2], opts))
    grid = gs.GetSeedGrid()
    _, rsu, rsv = os_.seed_grid_step(nodes, 3, grid, P[0], P[1], P[2])
    ru, rv, rst, _ = os_.gauss_newton(P[0], P[1], P[2], rsu, rsv, tol_step=1e-12, max_iter=40, newton=1)
    both = (st == G.PT_CONVERGED) & (rst == O.CONV_STEP)
    assert both.mean() > 0.95
    d = np.maximum(np.abs(u - ru), np.abs(v - rv))[both]
    same = d < 1e-6  # the few others reached another local minimum of the distance from the seed
    assert same.mean() > 0.99
    assert np.max(d[same]) <= 1e-9
    # the non-finite queries end GN_PT_NONFINITE
    k = 20000 // 50
    assert np.all(st[2 * k:2 * k + 3] == G.PT_NONFINITE)
    if precision == G.GN_MIXED:
        # as many offset queries reach their constructed foot point within 1e-9 as from the
        # greedy-step seeds (a few percent of them have other local minima of the distance nearby,
        # and different seeds reach different ones)
        g = G.default_opts(precision=G.GN_MIXED, newton=G.NEWTON_FP32 | G.NEWTON_FP64,
                           seed_mode=G.SEED_GREEDY_STEP)
        gu, gv, gst = (_np(t) for t in gs.NewInverse(P[0], P[1], P[2], g))
        off = slice(2 * k + 3, None)
        hit = lambda a, b: np.mean(np.maximum(np.abs(a[off] - u0[off]), np.abs(b[off] - v0[off])) <= 1e-9)
        assert hit(u, v) > 0.9 and hit(u, v) >= hit(gu, gv) - 0.005


def test_grid_errors_and_lifetime():
    sd, ku = _surface("bicubic")
    gs = G.Surface(sd.p, sd.q, ku, sd.kv, sd.cp, sd.w)
    x = np.zeros(8)
    opts = G.default_opts(precision=G.GN_MIXED, seed_mode=G.SEED_GRID_STEP)
    with pytest.raises(G.GnError) as e:
        gs.BuildSeedGrid()
    assert e.value.code == G.GN_ESTATE  # no tree
    gs.BuildTree(3)
    base = gs.device_bytes()
    with pytest.raises(G.GnError) as e:
        gs.NewInverse(x, x, x, opts)
    assert e.value.code == G.GN_ESTATE  # no grid
    with pytest.raises(G.GnError) as e:
        gs.GetSeedGrid()
    assert e.value.code == G.GN_ESTATE
    with pytest.raises(G.GnError) as e:
        gs.BuildSeedGrid(1025)
    assert e.value.code == G.GN_EINVAL
    with pytest.raises(G.GnError) as e:
        gs.BuildSeedGrid(1000)  # above 2^25 cells
    assert e.value.code == G.GN_EINVAL
    gs.BuildSeedGrid()
    dims, _, _, _ = gs.GetSeedGrid()
    assert gs.device_bytes() == base + 2 * int(np.prod(dims))
    gs.BuildSeedGrid(64)  # a rebuild replaces the grid
    dims64, _, _, _ = gs.GetSeedGrid()
    assert gs.device_bytes() == base + 2 * int(np.prod(dims64))
    u, v, st = gs.NewInverse(x + 0.5, x + 0.5, x, opts)
    assert _np(st).max() <= G.PT_MAXITER
    # a tree rebuild discards the grid, and the device byte count returns to the tree's
    gs.BuildTree(3)
    assert gs.device_bytes() == base
    with pytest.raises(G.GnError) as e:
        gs.NewInverse(x, x, x, opts)
    assert e.value.code == G.GN_ESTATE
    # a batch takes grid seeds only when every surface had its grid when the batch was made
    b = G.Batch([gs])
    with pytest.raises(G.GnError) as e:
        b.invert(G.segments(np.array([0]), np.array([8])), x, x, x, opts)
    assert e.value.code == G.GN_ESTATE
    with pytest.raises(G.GnError) as e:
        G.Server(b, opts=opts)
    assert e.value.code == G.GN_ESTATE
    gs.BuildSeedGrid()
    b = G.Batch([gs])
    _, _, st = b.invert(G.segments(np.array([0]), np.array([8])), x + 0.5, x + 0.5, x, opts)
    assert _np(st).max() <= G.PT_MAXITER


def _fp32_d2(c, q):
    """fmaf(ez, ez, fmaf(ey, ey, ex * ex)) of FP32 values: FP32 differences, the products exact in
    FP64, each sum rounded to FP32 (as one fmaf rounds, except in rare double-rounding cases)."""
    e = (c[:, None, :].astype(np.float32) - q[None, :, :].astype(np.float32)).astype(np.float64)
    s = (e[..., 0] ** 2).astype(np.float32).astype(np.float64)
    s = (e[..., 1] ** 2 + s).astype(np.float32).astype(np.float64)
    return (e[..., 2] ** 2 + s).astype(np.float32)


def test_wide_grid_of_a_four_level_tree():
    """262,144 leaves: 4-byte cells, and the default resolution halved until the build stays within
    2^36 cell-leaf pairs.  Geometry against the oracle; sampled cells against a brute-force
    nearest leaf (the oracle's full build would take minutes on the CPU)."""
    sd = synth.surface_G(24, 20, 3, 3, seed=12)
    gs = G.Surface(sd.p, sd.q, sd.ku, sd.kv, sd.cp, sd.w)
    gs.BuildTree(4)
    gs.BuildSeedGrid()
    nodes, _, _ = gs.GetTree()
    dims, org, cell, cells = gs.GetSeedGrid()
    od, oo, oc, _ = O.seed_grid(nodes, 4, 0, build=False)
    assert np.array_equal(dims, od) and np.array_equal(org, oo) and cell == oc
    leaves = 512 * 512
    assert int(np.prod(dims)) * leaves <= 1 << 36
    assert gs.device_bytes() > 4 * int(np.prod(dims))
    base = O.tree_base(3)
    C = nodes[base:base + leaves, 10:13]
    rng = np.random.default_rng(5)
    ids = np.stack([rng.integers(0, d, 400) for d in dims], 1)
    cc = (org[None, :].astype(np.float64) + (ids + 0.5) * np.float64(cell)).astype(np.float32)
    got = cells[ids[:, 0], ids[:, 1], ids[:, 2]]
    agree = 0
    for a in range(0, 400, 50):
        d2 = _fp32_d2(cc[a:a + 50], C)
        best = np.argmin(d2, axis=1)
        agree += np.sum(base + best == got[a:a + 50])
        # where they differ, the grid's leaf is as near to within FP32 rounding
        gd = d2[np.arange(d2.shape[0]), got[a:a + 50] - base]
        assert np.all(gd <= d2.min(axis=1) * (1 + 1e-5) + 1e-30)
    assert agree >= 396


def test_batch_and_server_grid_seeds_equal_single_surface():
    """Surfaces of several degree classes, each with its seed grid, in one batch: grid-seeded
    mixed inversion bit-identical to each surface's own gn_invert; the server's results for the
    same requests bit-identical to the batch's."""
    defs = [synth.surface_G(12, 10, 2, 3, seed=31, jitter=0.2), synth.surface_G(32, 32, 3, 3, seed=1002),
            synth.surface_G(14, 12, 4, 5, seed=33, jitter=0.2), synth.surface_G(9, 9, 6, 2, seed=34)]
    surfs = []
    for i, sd in enumerate(defs):
        s = G.Surface(sd.p, sd.q, sd.ku, sd.kv, sd.cp, sd.w)
        s.BuildTree(3 if i % 2 else 2)
        s.BuildSeedGrid(0 if i else 40)
        surfs.append(s)
    counts = np.array([3000, 20000, 5000, 700])
    P = []
    for i, sd in enumerate(defs):
        os_ = O.OSurf.from_def(sd)
        u0, v0, Pi = _queries(sd, sd.ku, os_, int(counts[i]), 40 + i)
        P.append(Pi)
    P = np.concatenate(P, axis=1)
    opts = G.default_opts(precision=G.GN_MIXED, newton=G.NEWTON_FP32 | G.NEWTON_FP64,
                          seed_mode=G.SEED_GRID_STEP)
    b = G.Batch(surfs)
    segs = G.segments(np.arange(len(defs)), counts)
    bu, bv, bst = (_np(t) for t in b.invert(segs, P[0], P[1], P[2], opts))
    for i, s in enumerate(surfs):
        a, c = segs["start"][i], segs["start"][i] + counts[i]
        su, sv, sst = (_np(t) for t in s.NewInverse(P[0][a:c], P[1][a:c], P[2][a:c], opts))
        assert np.array_equal(bu[a:c], su) and np.array_equal(bv[a:c], sv) and np.array_equal(bst[a:c], sst), i
    srv = G.Server(b, opts=opts)
    try:
        ids = [srv.submit(i, P[0][a:a + n], P[1][a:a + n], P[2][a:a + n])
               for i, (a, n) in enumerate(zip(segs["start"], counts))]
        for i, rid in enumerate(ids):
            a, c = segs["start"][i], segs["start"][i] + counts[i]
            info = srv.wait(rid, 60000)
            assert info is not None, "request did not finish"
            ru, rv, rst = (_np(t) for t in G.Server.results(info))
            srv.release(rid)
            assert np.array_equal(ru, bu[a:c]) and np.array_equal(rv, bv[a:c]) and np.array_equal(rst, bst[a:c]), i
    finally:
        srv.close()
