"""The oracle's seed grid (or_seed_grid_geometry, or_seed_grid_build, or_seed_grid_step: the
library's gn_build_seed_grid and GN_SEED_GRID_STEP; no reference counterpart -- the reference seeds
by scanning its tree, src/kernel/kernel.cu:433-469), pinned without its own code:

- every cell holds the leaf whose centre is nearest to the cell centre, by a brute-force FP64
  search in numpy (equal except where two leaves tie within FP32 rounding);
- the geometry: cubic cells, the leaf centres' box R cells wide along its largest extent, at
  least R/16 cells of margin around it on every side;
- the lookup: each query's cell contains it (FP64 arithmetic), out-of-grid queries go to the
  nearest border cell;
- the step from the grid's leaf is the step of GN_SEED_GREEDY_STEP (pinned by
  test_oracle_seed_step) from that leaf: checked by comparing with that step taken from the same
  leaves;
- what the seeds are for: as few Gauss-Newton iterations as the greedy descent's, the same foot
  points."""
import numpy as np

import oracle as O
from gnurbs import synth
from helpers import blocks


def _case(seed=3, levels=3, res=24, n=20000):
    sd = synth.surface_G(16, 14, 3, 2, seed=seed, jitter=0.3)
    os_ = O.OSurf.from_def(sd)
    nodes, _, _ = os_.build_tree(levels)
    u0, v0 = synth.uniform_uv(n, seed + 1)
    b = blocks(os_.eval(u0, v0, 1))
    P, _ = synth.offset_queries(b["S"], b["Su"], b["Sv"], 1e-2, seed + 2)
    return sd, os_, nodes, u0, v0, P, O.seed_grid(nodes, levels, res)


def _leaf_centres(nodes, levels):
    side = 8 ** (levels - 1)
    base = O.tree_base(levels - 1)
    return base, nodes[base:base + side * side, 10:13].astype(np.float32).astype(np.float64)


def test_cells_hold_their_nearest_leaf():
    for levels, res in ((2, 0), (3, 24)):
        sd, os_, nodes, u0, v0, P, (dims, org, cell, cells) = _case(levels=levels, res=res)
        base, C = _leaf_centres(nodes, levels)
        idx = np.stack(np.meshgrid(*[np.arange(d) for d in dims], indexing="ij"), -1).reshape(-1, 3)
        cc = org.astype(np.float64) + (idx + 0.5) * np.float64(cell)
        got = cells.reshape(-1) - base
        same = 0
        for a in range(0, len(cc), 4096):
            d2 = ((cc[a:a + 4096, None, :] - C[None]) ** 2).sum(-1)
            best = d2.argmin(1)
            same += np.sum(best == got[a:a + 4096])
            gd = d2[np.arange(len(best)), got[a:a + 4096]]
            assert np.all(gd <= d2.min(1) * (1 + 1e-5) + 1e-12)
        assert same >= 0.999 * len(cc)


def test_geometry():
    for levels, res in ((2, 0), (3, 0), (3, 24), (3, 100)):
        sd, os_, nodes, u0, v0, P, g = _case(levels=levels, res=res, n=10)
        dims, org, cell, _ = O.seed_grid(nodes, levels, res, build=False)
        R = res if res else min(256, 2 * 8 ** (levels - 1))
        pad = max(1, R // 16)
        _, C = _leaf_centres(nodes, levels)
        lo, hi = C.min(0), C.max(0)
        assert abs((hi - lo).max() / cell - R) < 1e-4 * R
        rel = (C - org.astype(np.float64)) / np.float64(cell)  # leaf centres in cell units
        assert rel.min() >= pad - 1e-3 and np.all(rel.max(0) <= dims - pad + 1e-3)
        assert np.all(dims <= np.ceil((hi - lo) / cell) + 2 * pad + 1)
    # a resolution whose grid exceeds 2^25 cells is refused
    assert O.seed_grid(nodes, 3, 1000, build=False) is None


def test_lookup_cells_contain_the_queries():
    sd, os_, nodes, u0, v0, P, grid = _case()
    dims, org, cell, cells = grid
    P = P.copy()
    P[:, :100] = P[:, :100] * 8.0 - 2.0  # outside the grid
    leaf, su, sv = os_.seed_grid_step(nodes, 3, grid, P[0], P[1], P[2])
    t = (P.T - org.astype(np.float64)) / np.float64(cell)
    ids = np.clip(np.floor(t), 0, dims - 1).astype(np.int64)
    inside = np.all((t >= 1e-3) & (t <= dims - 1e-3), axis=1) & np.all(np.abs(t - np.round(t)) > 1e-3, axis=1)
    exp = cells[ids[:, 0], ids[:, 1], ids[:, 2]]
    assert np.array_equal(leaf[inside], exp[inside])
    # the clamped ones: nearest border cell along each axis
    out = ~np.all((t >= 0) & (t < dims), axis=1)
    assert out[:100].mean() > 0.9 and np.array_equal(leaf[out], exp[out])


def test_step_is_the_greedy_step_from_the_grid_leaf():
    """Where the grid's leaf and the greedy descent's leaf coincide, the seeds are identical (the
    same step from the same leaf); the step itself is pinned in test_oracle_seed_step."""
    sd, os_, nodes, u0, v0, P, _ = _case(n=40000)
    grid = O.seed_grid(nodes, 3, 0)
    leaf, su, sv = os_.seed_grid_step(nodes, 3, grid, P[0], P[1], P[2])
    gl, gsu, gsv = os_.seed_greedy_step(nodes, 3, P[0], P[1], P[2])
    same = leaf == gl
    assert same.mean() > 0.5
    assert np.array_equal(su[same], gsu[same]) and np.array_equal(sv[same], gsv[same])
    # and the seeds never leave their leaf's uv centre by more than two cells (the model
    # iterations' limit)
    nd = nodes[leaf]
    cu, cv = (nd[:, 0] + nd[:, 2]) / 2, (nd[:, 1] + nd[:, 3]) / 2
    assert np.all(np.abs(su - cu) <= 2 * (nd[:, 2] - nd[:, 0]) * (1 + 1e-6))
    assert np.all(np.abs(sv - cv) <= 2 * (nd[:, 3] - nd[:, 1]) * (1 + 1e-6))


def test_grid_seeds_save_iterations_and_keep_foot_points():
    sd, os_, nodes, u0, v0, P, _ = _case(n=40000)
    grid = O.seed_grid(nodes, 3, 0)
    _, su, sv = os_.seed_grid_step(nodes, 3, grid, P[0], P[1], P[2])
    _, gu, gv = os_.seed_greedy_step(nodes, 3, P[0], P[1], P[2])
    its = []
    for a, b in ((gu, gv), (su, sv)):
        _, _, _, it = os_.gauss_newton(P[0], P[1], P[2], a, b, tol_step=1e-4, max_iter=20, newton=1)
        its.append(it.mean())
    assert its[1] <= its[0] + 0.05
    fu, fv, fst, _ = os_.gauss_newton(P[0], P[1], P[2], gu, gv, tol_step=1e-12, max_iter=40, newton=1)
    tu, tv, tst, _ = os_.gauss_newton(P[0], P[1], P[2], su, sv, tol_step=1e-12, max_iter=40, newton=1)
    both = (fst == O.CONV_STEP) & (tst == O.CONV_STEP)
    same = np.maximum(np.abs(fu - tu), np.abs(fv - tv)) < 1e-9
    assert same[both].mean() > 0.99
