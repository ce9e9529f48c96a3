"""gn_evaluate_sorted_f32: the FP32 span-GEMM form for points in any order (counting sort by span
pair, v_mfma_f32_16x16x4f32 for the V-side contraction of each 16-point group of one span pair).

Checks: perm is a permutation of the input indices, each block of 65,536 consecutive input points
keeps its own output range, grouped by span pair in increasing key order (keys from the oracle's
findSpan, src/NURBS/Evaluate.cu:504-533); every field
at sorted point k equals the FP64 oracle at input point perm[k] within 1e-5 of the field scale
(the FP32 evaluation tolerance), for degrees 1..3 on the matrix cores and higher degrees on the
per-point fallback, for orders 0..2, on jittered and repeated knots, knots outside [0,1], a span
count beyond the LDS histogram (global-atomic path), ragged sizes; and at the 16,777,216-point
workload, against the per-point FP32 kernel over all points and the oracle on a subset."""
import numpy as np
import pytest
import torch

import oracle as O
import gnurbs as G
from gnurbs import synth
from helpers import blocks, field_close

pytestmark = pytest.mark.gpu

FIELDS = ("S", "Su", "Sv", "Suu", "Suv", "Svv")
BLOCK = 65536  # input points per sort block


def _np(t):
    return t.detach().cpu().numpy()


def _keys(sd, u, v):
    m = np.array([O.find_span(sd.p, sd.nu, sd.ku, float(a)) for a in u]) - sd.p
    n = np.array([O.find_span(sd.q, sd.nv, sd.kv, float(a)) for a in v]) - sd.q
    return m * (sd.nv - sd.q) + n


def _check(sd, u, v, order):
    gs = G.Surface(sd.p, sd.q, sd.ku, sd.kv, sd.cp, sd.w)
    u32, v32 = u.astype(np.float32), v.astype(np.float32)
    out, perm = gs.EvaluateSorted(u32, v32, order)
    out, perm = _np(out).astype(np.float64), _np(perm).astype(np.int64)
    n = u.size
    assert np.array_equal(np.sort(perm), np.arange(n))
    us, vs = u32[perm].astype(np.float64), v32[perm].astype(np.float64)
    k = _keys(sd, us, vs)
    for b0 in range(0, n, BLOCK):  # each block of input points sorted within its own range
        assert np.all(perm[b0:b0 + BLOCK] // BLOCK == b0 // BLOCK)
        assert np.all(np.diff(k[b0:b0 + BLOCK]) >= 0)  # grouped by span pair, keys increasing
    ref = blocks(O.OSurf.from_def(sd).eval(us, vs, order))
    got = blocks(out)
    floor = 1e-2 * np.max(np.abs(ref["S"]))
    for f in FIELDS[: (1, 3, 6)[order]]:
        ok, err = field_close(got[f], ref[f], 1e-5, floor)
        assert ok, (f, err)
    gs.close()


def _uv(sd, n, seed):
    rng = np.random.default_rng(seed)
    u = sd.ku[0] + (sd.ku[-1] - sd.ku[0]) * rng.random(n)
    v = sd.kv[0] + (sd.kv[-1] - sd.kv[0]) * rng.random(n)
    ku, kv = np.unique(sd.ku), np.unique(sd.kv)
    m = min(len(ku), len(kv), n // 4)
    u[:m], v[:m] = ku[:m], kv[:m]  # on knots, ends included
    return u, v


@pytest.mark.parametrize("order", [0, 1, 2])
@pytest.mark.parametrize("name,mk", [
    ("bicubic64", lambda: synth.surface_G(64, 64, 3, 3, seed=1001)),
    ("p2q3_jitter", lambda: synth.surface_G(14, 11, 2, 3, seed=17, jitter=0.4)),
    ("p1q2", lambda: synth.surface_G(6, 9, 1, 2, seed=18)),
    ("p5q4_fallback", lambda: synth.surface_G(12, 10, 5, 4, seed=19, jitter=0.3)),
])
def test_sorted_matches_oracle(name, mk, order):
    sd = mk()
    u, v = _uv(sd, 100000, 5)
    _check(sd, u, v, order)


def test_sorted_repeated_knots_and_other_domain():
    p = q = 3
    ku = np.r_[[0.0] * 4, 0.1, 0.3, 0.3, 0.55, 0.55, 0.55, 0.8, [1.0] * 4] * 4.0 - 1.0
    kv = np.r_[[0.0] * 4, 0.25, 0.25, 0.6, 0.9, [1.0] * 4] + 2.0
    nu, nv = len(ku) - p - 1, len(kv) - q - 1
    rng = np.random.default_rng(31)
    X, Y = np.meshgrid(np.linspace(0, 1, nu), np.linspace(0, 1, nv), indexing="ij")
    cp = np.stack([X, Y, 0.2 * rng.standard_normal((nu, nv))], axis=-1)
    w = rng.uniform(0.5, 2.0, (nu, nv))
    sd = synth.SurfaceDef(p, q, ku, kv, cp, w)
    u, v = _uv(sd, 50000, 6)
    _check(sd, u, v, 2)


def test_sorted_many_spans_use_the_global_histogram():
    """223 x 98 = 21,854 span pairs: more keys than the LDS histogram holds."""
    nu, nv = 225, 100
    X, Y = np.meshgrid(np.arange(nu) / (nu - 1), np.arange(nv) / (nv - 1), indexing="ij")
    cp = np.stack([X, Y, 0.1 * np.sin(7 * X) * np.cos(5 * Y)], axis=-1)
    w = 1.0 + 0.3 * np.sin(3 * X + 2 * Y)
    sd = synth.SurfaceDef(2, 2, synth.clamped_knots(nu, 2), synth.clamped_knots(nv, 2), cp, w)
    u, v = _uv(sd, 200000, 7)
    _check(sd, u, v, 1)


@pytest.mark.parametrize("n", [1, 15, 16, 17, 63, 65, 1000])
def test_sorted_ragged_sizes(n):
    sd = synth.surface_G(8, 8, 3, 3, seed=1000)
    u, v = _uv(sd, n, 8 + n)
    _check(sd, u, v, 2)


def test_sorted_16M_points_against_per_point_kernel_and_oracle():
    n = 1 << 24
    sd = synth.surface_G(64, 64, 3, 3, seed=1001)
    gs = G.Surface(sd.p, sd.q, sd.ku, sd.kv, sd.cp, sd.w)
    gen = torch.Generator(device="cuda").manual_seed(1001)
    u = torch.rand(n, dtype=torch.float32, device="cuda", generator=gen)
    v = torch.rand(n, dtype=torch.float32, device="cuda", generator=gen)
    out, perm = gs.EvaluateSorted(u, v, 2)
    p64 = perm.long()
    seen = torch.zeros(n, dtype=torch.int32, device="cuda")
    seen.index_add_(0, p64, torch.ones(n, dtype=torch.int32, device="cuda"))
    assert bool((seen == 1).all())
    del seen
    blk = torch.arange(n, device="cuda") // BLOCK
    assert bool((p64 // BLOCK == blk).all())
    ref = gs.Evaluate(u[p64].contiguous(), v[p64].contiguous(), 2, precision=G.GN_FP32)
    for b in range(6):
        r = ref[3 * b:3 * b + 3]
        scale = max(r.abs().max().item(), 1e-2 * ref[0:3].abs().max().item())
        err = (out[3 * b:3 * b + 3] - r).abs().max().item() / scale
        assert err < 1e-5, (FIELDS[b], err)
    del ref
    sel = torch.from_numpy(np.sort(np.random.default_rng(5).choice(n, 50000, replace=False))).cuda()
    us = u[p64[sel]].double().cpu().numpy()
    vs = v[p64[sel]].double().cpu().numpy()
    ro = blocks(O.OSurf.from_def(sd).eval(us, vs, 2))
    got = blocks(out[:, sel].double().cpu().numpy())
    floor = 1e-2 * np.max(np.abs(ro["S"]))
    for f in FIELDS:
        ok, err = field_close(got[f], ro[f], 1e-5, floor)
        assert ok, (f, err)
    gs.close()
