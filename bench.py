#!/usr/bin/env python3
"""Headline benchmark: point inversions per second at <= 1e-9 (u, v) error, mixed precision
(FP32 bulk + FP64 closing iterations) beside the plain FP64 kernel.

Workload (one step = one pass of the inversion hot path over one batch): a bicubic rational surface
with a 32 x 32 control net from the seeded generator G (x, y on a grid, z = 0.2 sin(2 pi x)
cos(2 pi y) + N(0, 0.05), weights U[0.5, 2], clamped uniform knots), 1,048,576 query points
P = S(u*, v*) + delta n, (u*, v*) ~ U[0,1]^2, delta ~ N(0, 1e-2), seeded from a depth-6 subdivision
tree (64 x 64 leaves), at most 20 iterations.  Inputs are resident in HBM before timing; a step is
seed search + Gauss-Newton.  A point counts when max(|u-u*|, |v-v*|) <= 1e-9.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--points N] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gpu-nurbs_amd"))

PEAK_HBM = 8.0e12
PEAK_F32 = 157.3e12
PEAK_F64 = 78.6e12


# The mixed form's options: the seed grid's leaf (one lookup; gn_build_seed_grid at set-up, like the
# tree), one Gauss-Newton step with the leaf's stored Jacobian and two on its quadratic model as the
# seed; FP32 exact-Hessian Newton iterations to a 3e-5 step, at most 6 (the few points that need
# more are the long dependent chains every wave of the persistent kernel would wait for: interleaved
# runs measured 6.17 against 6.02 G/s for a cap of 8, with 95.02% against 95.12% of points within
# 1e-9; a 1e-4 step 5.83 G/s, a 1e-5 step 5.97 G/s); then at most 3 FP64 exact-Hessian closing
# steps until the Newton certificate bounds the (u,v) error by 1e-10, or a 1e-10 step.  At most 9
# iterations in all (the workload allows 20).
# The batch and serving objects carry 64-cell seed grids (0.4 MB each; 128 cells, the default, would
# take 3.4 MB and 1.5 ms to build per object).
def mixed_options(G):
    return dict(precision=G.GN_MIXED, newton=G.NEWTON_FP32 | G.NEWTON_FP64, fp32_iters=6, fp64_tail=3,
                tol_step32=3e-5, tol_step=1e-10, tol_quad=1e-10, seed_mode=G.SEED_GRID_STEP)


OBJECT_GRID_RES = 64


def contract_flops_per_iter(k1, k2):
    """The per-iteration count the workload definition states: 2 (8 k1 k2 + 12 k1) + 40 flops
    (392 for bicubic): the order-1 contraction and the 2x2 Gauss-Newton solve."""
    return 2 * (8 * k1 * k2 + 12 * k1) + 40


def cpu_info():
    """Cores this process may use, the machine's logical CPU count and the CPU model."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS")
    threads = usable
    if omp and omp.isdigit() and int(omp) > 0:
        threads = min(threads, int(omp))  # the host's CPU share for this GPU, where it is set
    return {"threads": threads, "affinity": usable, "nproc": os.cpu_count(), "omp_num_threads": omp,
            "model": model}


def flops_per_iter(k1, k2, newton):
    """Algorithmic flops of one iteration: contraction FMAs x 2 (point and first derivatives; with
    Newton also the second derivatives) plus the rational quotient and the 2x2 solve."""
    fma1 = 4 * (k1 * k2 + (k1 - 1) * k2 + k2 + k2 + (k2 - 1))
    f = 2 * fma1 + 26 + 45
    if newton:
        fma2 = 4 * ((k1 - 2) * k2 + k2 + (k2 - 1) + (k2 - 2))
        f += 2 * fma2 + 45 + 20
    return f


def bench_cpu_case(G, synth, torch, O, steps, threads):
    """The reference's CPU-runnable case: bicubic rational 8x8 net (generator G, seed 1000).
    Evaluation of point + first derivatives at 65,536 uniform (u, v); inversion of 4,096 points
    sampled exactly on the surface.  GPU (plain FP64 kernels, inputs resident in HBM) beside the
    oracle's FP64 restatement on the host cores (OpenMP)."""
    import numpy as np

    sd = synth.surface_G(8, 8, 3, 3, seed=1000)
    gs = G.Surface(sd.p, sd.q, sd.ku, sd.kv, sd.cp, sd.w)
    gs.BuildTree(3)
    os_ = O.OSurf.from_def(sd)
    ne, ni = 65536, 4096
    ue, ve = synth.uniform_uv(ne, seed=1000)
    ui, vi = synth.uniform_uv(ni, seed=1001)
    ue_t, ve_t = torch.from_numpy(ue).cuda(), torch.from_numpy(ve).cuda()
    ui_t, vi_t = torch.from_numpy(ui).cuda(), torch.from_numpy(vi).cuda()
    Pi = gs.Evaluate(ui_t, vi_t, 0)
    x, y, z = (Pi[k].contiguous() for k in range(3))
    out = torch.empty((9, ne), dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream()
    reps = max(10, steps)

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(reps):
            fn()
        b.record(stream)
        torch.cuda.synchronize()
        return a.elapsed_time(b) / reps / 1e3

    res = {"workload": "bicubic rational 8x8 net (G, seed 1000): evaluate point + first derivatives "
                       "at 65,536 uniform (u,v); invert 4,096 on-surface points",
           "cpu_threads": threads}
    t_gpu = timed(lambda: gs.Evaluate(ue_t, ve_t, 1, out=out))
    t0 = time.perf_counter()
    for _ in range(3):
        ref = os_.eval(ue, ve, 1, nthreads=threads)
    t_cpu = (time.perf_counter() - t0) / 3
    got = out.cpu().numpy()
    scale = np.maximum(np.abs(ref).max(axis=1, keepdims=True), 1e-2 * np.abs(ref[0:3]).max())
    res["eval"] = {"gpu_points_per_s": ne / t_gpu, "cpu_points_per_s": ne / t_cpu,
                   "max_rel_err_gpu_vs_cpu": float((np.abs(got - ref) / scale).max())}
    plain = G.default_opts(precision=G.GN_FP64, newton=0, max_iter=20, tol_step=1e-10)
    u = torch.empty(ni, dtype=torch.float64, device="cuda")
    v = torch.empty(ni, dtype=torch.float64, device="cuda")
    st = torch.empty(ni, dtype=torch.uint8, device="cuda")
    t_gpu = timed(lambda: gs.NewInverse(x, y, z, plain, out=(u, v, st)))
    gerr = torch.maximum((u - ui_t).abs(), (v - vi_t).abs()).cpu().numpy()
    _, su, sv = gs.Seed(x, y, z, exact=False)
    xs, ys, zs, ssu, ssv = (a.cpu().numpy() for a in (x, y, z, su, sv))
    t0 = time.perf_counter()
    for _ in range(3):
        cu, cv, _, _ = os_.gauss_newton(xs, ys, zs, ssu, ssv, tol_step=1e-10, max_iter=20, nthreads=threads)
    t_cpu = (time.perf_counter() - t0) / 3
    cerr = np.maximum(np.abs(cu - ui), np.abs(cv - vi))
    res["invert"] = {"gpu_inversions_per_s": ni / t_gpu, "gpu_ok_frac": float(np.mean(gerr <= 1e-9)),
                     "cpu_inversions_per_s": ni / t_cpu, "cpu_ok_frac": float(np.mean(cerr <= 1e-9)),
                     "note": "GPU: seed search + plain FP64 Gauss-Newton; CPU: the oracle's plain FP64 "
                             "Gauss-Newton from the GPU's greedy-leaf seeds (seed search not timed)"}
    gs.close()
    return res


def bench_eval(G, synth, torch, steps):
    """Evaluation at scale: bicubic 64x64 surface (generator G, seed 1001), 16,777,216 uniform
    (u, v), point + derivatives to order 2.  FP64 per-point (160 B/pt), FP32 per-point (80 B/pt),
    the FP32 MFMA span-GEMM form on the same unordered points (sorted by span pair on the device,
    results in span order with the permutation) and on a 4096 x 4096 parameter grid (72 B/pt
    written).  FP32 error is measured against the FP64 evaluation (itself held to the FP64 oracle
    by the tests)."""
    import numpy as np

    n = 1 << 24
    sd = synth.surface_G(64, 64, 3, 3, seed=1001)
    gs = G.Surface(sd.p, sd.q, sd.ku, sd.kv, sd.cp, sd.w)
    g = torch.Generator(device="cuda").manual_seed(1001)
    u = torch.rand(n, dtype=torch.float64, device="cuda", generator=g)
    v = torch.rand(n, dtype=torch.float64, device="cuda", generator=g)
    u32, v32 = u.float(), v.float()
    out64 = torch.empty((18, n), dtype=torch.float64, device="cuda")
    out32 = torch.empty((18, n), dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream()

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(steps):
            fn()
        b.record(stream)
        torch.cuda.synchronize()
        return a.elapsed_time(b) / steps

    res = {"workload": "bicubic 64x64 net (G, seed 1001), 16,777,216 points, order 2 (18 outputs)"}
    ms = timed(lambda: gs.Evaluate(u, v, 2, out=out64))
    res["fp64"] = {"ms": ms, "points_per_s": n / (ms / 1e3), "GB_per_s": n * 160 / (ms / 1e3) / 1e9,
                   "hbm_frac": n * 160 / (ms / 1e3) / PEAK_HBM}
    ms = timed(lambda: gs.Evaluate(u32, v32, 2, precision=G.GN_FP32, out=out32))
    ref = gs.Evaluate(u32.double(), v32.double(), 2)
    err = []
    for blk in range(6):
        r = ref[3 * blk:3 * blk + 3]
        e = (out32[3 * blk:3 * blk + 3].double() - r).abs()
        scale = max(r.abs().max().item(), 1e-2 * ref[0:3].abs().max().item())
        err.append(e.max().item() / scale)
    res["fp32"] = {"ms": ms, "points_per_s": n / (ms / 1e3), "GB_per_s": n * 80 / (ms / 1e3) / 1e9,
                   "hbm_frac": n * 80 / (ms / 1e3) / PEAK_HBM, "max_rel_err_vs_fp64": max(err)}
    # the MFMA span-GEMM form for the same unordered points: counting sort by span pair, then the
    # V-side contraction of every 16 points of one span pair on v_mfma_f32_16x16x4f32; outputs in
    # span order with the permutation (84 B/pt: 8 in, 72 out, 4 perm; the sort adds 36 B/pt)
    perm = torch.empty(n, dtype=torch.int32, device="cuda")
    ms = timed(lambda: gs.EvaluateSorted(u32, v32, 2, out=out32, perm=perm))
    p64 = perm.long()
    refs = ref[:, p64]
    err = []
    for blk in range(6):
        r = refs[3 * blk:3 * blk + 3]
        e = (out32[3 * blk:3 * blk + 3].double() - r).abs()
        scale = max(r.abs().max().item(), 1e-2 * refs[0:3].abs().max().item())
        err.append(e.max().item() / scale)
    del refs, p64
    res["fp32_span_gemm_mfma_sorted"] = {
        "ms": ms, "points_per_s": n / (ms / 1e3), "GB_per_s": n * 84 / (ms / 1e3) / 1e9,
        "hbm_frac": n * 84 / (ms / 1e3) / PEAK_HBM, "hbm_frac_with_sort_traffic": n * 120 / (ms / 1e3) / PEAK_HBM,
        "max_rel_err_vs_fp64": max(err),
        "note": "outputs in span order + permutation (gn_evaluate_sorted_f32); includes the device counting sort"}
    del ref, perm
    side = 4096
    us = torch.linspace(0, 1, side, dtype=torch.float32, device="cuda")
    ms = timed(lambda: gs.EvaluateGrid(us, us, 2, out=out32))
    res["fp32_grid_mfma"] = {"ms": ms, "points_per_s": side * side / (ms / 1e3),
                             "GB_per_s": side * side * 72 / (ms / 1e3) / 1e9,
                             "hbm_frac": side * side * 72 / (ms / 1e3) / PEAK_HBM}
    del out64, out32, u, v
    torch.cuda.empty_cache()
    return res


def bench_batch(G, synth, torch, steps, count=1024):
    """Uneven batch: `count` rational surfaces, degrees U{2..5} per direction, nets 4x4..48x48,
    jittered clamped knots, weights U[0.5, 2]; points per surface from a power law on
    [256, 65536]; queries offset N(0, 1e-2) along the normal.  One persistent launch per
    precision phase over all surfaces, tile shape from the autotuner, captured in a hipGraph."""
    import numpy as np

    defs, counts = synth.uneven_batch(count, seed=1003)
    gen = torch.Generator(device="cuda").manual_seed(1003)
    surfs, xs, u0s, v0s = [], [], [], []
    for sd, c in zip(defs, counts):
        s = G.Surface(sd.p, sd.q, sd.ku, sd.kv, sd.cp, sd.w)
        s.BuildTree(3)
        s.BuildSeedGrid(OBJECT_GRID_RES)
        surfs.append(s)
        u0 = torch.rand(int(c), dtype=torch.float64, device="cuda", generator=gen)
        v0 = torch.rand(int(c), dtype=torch.float64, device="cuda", generator=gen)
        ev = s.Evaluate(u0, v0, 1)
        nrm = torch.linalg.cross(ev[3:6].T, ev[6:9].T).T
        nrm = nrm / torch.linalg.norm(nrm, dim=0, keepdim=True)
        delta = torch.randn(int(c), dtype=torch.float64, device="cuda", generator=gen) * 1e-2
        xs.append(ev[0:3] + delta * nrm)
        u0s.append(u0)
        v0s.append(v0)
    P = torch.cat(xs, dim=1)
    x, y, z = P[0].contiguous(), P[1].contiguous(), P[2].contiguous()
    u0, v0 = torch.cat(u0s), torch.cat(v0s)
    del xs, u0s, v0s, P
    n = x.numel()
    b = G.Batch(surfs)
    segs = G.segments(np.arange(count), counts)
    u = torch.empty(n, dtype=torch.float64, device="cuda")
    v = torch.empty_like(u)
    st = torch.empty(n, dtype=torch.uint8, device="cuda")
    mixed = mixed_options(G)
    shape = b.tune(segs, x, y, z, G.default_opts(**mixed), out=(u, v, st))
    stream = torch.cuda.Stream()
    res = {"workload": f"{count} surfaces (degrees 2..5, nets 4..48, jittered knots), {n} points "
                       "(power law 256..65536 per surface), offsets N(0,1e-2); depth-6 trees and 64-cell seed "
                       "grids",
           "points": n, "tile_points": shape[0], "blocks_per_cu": shape[1],
           "batch_device_bytes": b.device_bytes()}

    def ok():
        return int(((u - u0).abs().maximum((v - v0).abs()) <= 1e-9).sum().item())

    def timed(fn):
        with torch.cuda.stream(stream):
            fn()
            stream.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(steps):
                fn()
            e1.record(stream)
            stream.synchronize()
        return e0.elapsed_time(e1) / steps

    for name, opts in (("mixed", mixed), ("plain_fp64", dict(precision=G.GN_FP64, max_iter=20, tol_step=1e-10))):
        ms = timed(lambda: b.invert(segs, x, y, z, G.default_opts(**opts), out=(u, v, st), stream=stream))
        k = ok()
        res[name] = {"ms": ms, "ok_frac": k / n, "points_per_s": k / (ms / 1e3)}
    g = b.graph(segs, x, y, z, u, v, st, G.default_opts(**mixed), stream=stream)
    ms = timed(lambda: g.launch(stream=stream))
    k = ok()
    res["mixed_graph"] = {"ms": ms, "ok_frac": k / n, "points_per_s": k / (ms / 1e3)}
    res["speedup_vs_plain_fp64"] = res["mixed_graph"]["points_per_s"] / res["plain_fp64"]["points_per_s"]
    g.close()
    b.close()
    for s in surfs:
        s.close()
    torch.cuda.empty_cache()
    return res


def bench_serve(G, synth, torch, duration=1.0, loads=(0.5, 0.8), budget=32 << 30, streams=2,
                modes=("mixed", "plain_fp64")):
    """Serving: 224 uneven surfaces (the uneven-batch generator, seed 1004) and 32 cubic rational
    helix curves (64 control points) resident; requests of log-uniform size on [64, 100000] points
    for a uniformly drawn object, Poisson arrivals at 50% and 80% of the capacity measured by a
    saturating burst; continuous batching (gn_server, 2 step slots, 32 GB HBM budget).  Latency is
    from the scheduled arrival to the server's completion time, over one measured second at each
    load after 0.3 s of warm-up at the same rate."""
    import ctypes
    import sys

    import numpy as np

    L = G.lib()
    pool = 100_000
    defs, _ = synth.uneven_batch(224, seed=1004)
    gen = torch.Generator(device="cuda").manual_seed(1004)
    surfs, curves, pools = [], [], []
    for sd in defs:
        s = G.Surface(sd.p, sd.q, sd.ku, sd.kv, sd.cp, sd.w)
        s.BuildTree(3)
        s.BuildSeedGrid(OBJECT_GRID_RES)
        surfs.append(s)
        u0 = torch.rand(pool, dtype=torch.float64, device="cuda", generator=gen)
        v0 = torch.rand(pool, dtype=torch.float64, device="cuda", generator=gen)
        ev = s.Evaluate(u0, v0, 1)
        nrm = torch.linalg.cross(ev[3:6].T, ev[6:9].T).T
        nrm = nrm / torch.linalg.norm(nrm, dim=0, keepdim=True)
        P = ev[0:3] + torch.randn(pool, dtype=torch.float64, device="cuda", generator=gen) * 1e-2 * nrm
        pools.append([P[k].contiguous() for k in range(3)])
    for i in range(32):
        c = G.Curve.from_def(synth.helix_curve(64, 3, seed=1004 * 100 + i))
        curves.append(c)
        t0 = torch.rand(pool, dtype=torch.float64, device="cuda", generator=gen)
        ev = c.Evaluate(t0, 1)
        tan = ev[3:6] / torch.linalg.norm(ev[3:6], dim=0, keepdim=True)
        a = torch.randn((3, pool), dtype=torch.float64, device="cuda", generator=gen)
        nrm = a - tan * (a * tan).sum(0)
        nrm = nrm / torch.linalg.norm(nrm, dim=0, keepdim=True)
        P = ev[0:3] + torch.randn(pool, dtype=torch.float64, device="cuda", generator=gen) * 1e-2 * nrm
        pools.append([P[k].contiguous() for k in range(3)])
    torch.cuda.synchronize()
    batch = G.Batch(surfs)
    batch.set_shape(512, 0)
    ptrs = np.array([[t.data_ptr() for t in p] for p in pools], dtype=np.uint64)
    lo, hi = 64, 100_000
    mean_size = (hi - lo) / np.log(hi / lo)
    res = {"workload": "224 surfaces (degrees 2..5, nets 4..48; depth-6 trees, 64-cell seed grids) + 32 cubic "
                       "rational curves (64 cps); requests log-uniform 64..100000 points, uniform object, "
                       "Poisson arrivals",
           "hbm_budget_bytes": budget, "step_slots": streams, "objects_device_bytes": batch.device_bytes() + sum(c.device_bytes() for c in curves)}
    old_switch = sys.getswitchinterval()
    sys.setswitchinterval(1e-4)

    def trace(n_req, rate, seed):
        rng = np.random.default_rng(seed)
        sizes = np.exp(rng.uniform(np.log(lo), np.log(hi), n_req)).astype(np.int64)
        objs = rng.integers(0, 256, n_req)
        offs = (rng.random(n_req) * (pool - sizes + 1)).astype(np.int64)
        arr = np.cumsum(rng.exponential(1.0 / rate, n_req)) if rate > 0 else np.zeros(n_req)
        return sizes, objs, offs, (arr * 1e9).astype(np.int64)

    def run(srv, sizes, objs, offs, arr_ns):
        h = srv._h
        n_req = sizes.size
        ids = np.zeros(n_req, dtype=np.int64)
        rid = ctypes.c_int64()
        sub = L.gn_server_submit
        now = L.gn_now_ns
        t0 = now() + 2_000_000
        for i in range(n_req):
            due = t0 + int(arr_ns[i])
            while now() < due:
                pass
            o, off = int(objs[i]), int(offs[i]) * 8
            rc = sub(h, o, int(ptrs[o, 0]) + off, int(ptrs[o, 1]) + off, int(ptrs[o, 2]) + off, int(sizes[i]),
                     ctypes.byref(rid))
            if rc:
                raise G.GnError(rc, L.gn_last_error().decode())
            ids[i] = rid.value
        info = G.RequestInfo()
        done = np.zeros(n_req, dtype=np.int64)
        subm = np.zeros(n_req, dtype=np.int64)
        launch = np.zeros(n_req, dtype=np.int64)
        for i in range(n_req):
            rc = L.gn_server_wait(h, int(ids[i]), 120000, ctypes.byref(info))
            if rc:
                raise G.GnError(rc, L.gn_last_error().decode())
            done[i], subm[i], launch[i] = info.done_ns, info.submit_ns, info.launch_ns
            L.gn_server_release(h, int(ids[i]))
        lat = (done - (t0 + arr_ns)) / 1e6
        span = (done.max() - t0 - arr_ns.min()) / 1e9
        # steps: requests launched together; duration launch -> completion vs points in the step
        _, step_of = np.unique(launch, return_inverse=True)
        pts = np.bincount(step_of, weights=sizes.astype(np.float64))
        dur = np.zeros(pts.size)
        dur[step_of] = (done - launch) / 1e9
        return lat, (subm - (t0 + arr_ns)) / 1e6, span, (pts, dur)

    def step_model(pts, dur):
        """Least-squares duration = fixed + points / rate over the steps of one run."""
        A = np.stack([np.ones_like(pts), pts], axis=1)
        (a, b), *_ = np.linalg.lstsq(A, dur, rcond=None)
        return {"steps": int(pts.size), "mean_step_points": float(pts.mean()),
                "mean_step_ms": float(dur.mean() * 1e3), "fixed_ms": float(a * 1e3),
                "in_step_points_per_s": float(1.0 / b) if b > 0 else None}

    mixed = mixed_options(G)
    plain = dict(precision=G.GN_FP64, max_iter=20, tol_step=1e-10)
    try:
        for name, opts in (("mixed", mixed), ("plain_fp64", plain)):
            if name not in modes:
                continue
            srv = G.Server(batch, curves, G.default_opts(**opts), hbm_budget=budget, streams=streams)
            sizes, objs, offs, arr = trace(12000, 0, 77)
            run(srv, sizes[:200], objs[:200], offs[:200], arr[:200])  # warm-up
            # saturating burst (about 160 M points, long against the submission loop): capacity
            lat, _, span, _ = run(srv, sizes, objs, offs, arr)
            cap = float(sizes.sum()) / span
            out = {"capacity_points_per_s": cap}
            arena = srv.stats().arena_bytes
            for load in loads:
                rate = load * cap / mean_size
                dur = min(duration, 0.6 * arena / (17.0 * load * cap))
                n_req = max(200, int(rate * dur))
                # steady state: 0.3 s at the same rate first (discarded); the first loaded window
                # after the idle gap has shown one stall of all GPU work of several ms
                n_w = max(100, int(rate * 0.3 * dur / duration))
                run(srv, *trace(n_w, rate, int(2004 + 100 * load)))
                sizes, objs, offs, arr = trace(n_req, rate, int(1004 + 100 * load))
                st0 = srv.stats()
                lat, lag, span, steps = run(srv, sizes, objs, offs, arr)
                st1 = srv.stats()
                mid = (sizes >= 5000) & (sizes <= 20000)
                out[f"load_{int(load * 100)}"] = {
                    "requests": int(n_req), "points": int(sizes.sum()),
                    "offered_points_per_s": load * cap, "served_points_per_s": float(sizes.sum()) / span,
                    "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                    "p50_ms_10k": float(np.percentile(lat[mid], 50)) if mid.any() else None,
                    "p99_ms_10k": float(np.percentile(lat[mid], 99)) if mid.any() else None,
                    "max_ms": float(lat.max()), "submit_lag_p99_ms": float(np.percentile(lag, 99)),
                    "steps": int(st1.steps - st0.steps), "max_step_requests": int(st1.max_step_requests),
                    "step_model": step_model(*steps),
                    "host_launch_ms_mean": (st1.launch_ns_total - st0.launch_ns_total) / 1e6 /
                                           max(1, st1.steps - st0.steps),
                    "host_launch_ms_max_so_far": st1.launch_ns_max / 1e6,
                }
            out["arena_peak_bytes"] = int(srv.stats().arena_peak)
            srv.close()
            res[name] = out
    finally:
        sys.setswitchinterval(old_switch)
        batch.close()
        for o in surfs + curves:
            o.close()
        del pools
        torch.cuda.empty_cache()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--points", type=int, default=1 << 20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-eval", action="store_true", help="skip the evaluation-at-scale measurement")
    ap.add_argument("--no-batch", action="store_true", help="skip the uneven-batch measurement")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving (latency) measurement")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="CPU baseline: repeat the whole query set until this much wall time")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one 512 MiB read + 512 MiB write scale kernel (reference for HBM "
                         "counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (inversion is "
              "independent per point, so points would be sharded across GPUs with no communication)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import gnurbs as G
    from gnurbs import synth

    torch.cuda.set_device(0)
    dev = "cuda"
    if args.copy_calibration:
        # known-size reference for the HBM counters: read 512 MiB, write 512 MiB (after a cache
        # flush-sized fill of another buffer, so the read comes from HBM)
        src = torch.ones(512 << 18, dtype=torch.float32, device=dev)
        flush = torch.zeros(1024 << 18, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        torch.mul(src, 2.0, out=dst)
        torch.cuda.synchronize()
        del src, dst, flush
    N = args.points
    sd = synth.surface_G(32, 32, 3, 3, seed=1002)
    gs = G.Surface(sd.p, sd.q, sd.ku, sd.kv, sd.cp, sd.w)
    torch.cuda.synchronize()
    t_setup = time.perf_counter()
    gs.BuildTree(3)
    gs.BuildSeedGrid(0)  # set-up of the surface, not part of a step
    torch.cuda.synchronize()
    setup_ms = (time.perf_counter() - t_setup) * 1e3
    u0, v0 = synth.uniform_uv(N, seed=1003)
    u0_t = torch.from_numpy(u0).to(dev)
    v0_t = torch.from_numpy(v0).to(dev)
    ev = gs.Evaluate(u0_t, v0_t, 1)
    S, Su, Sv = ev[0:3], ev[3:6], ev[6:9]
    nrm = torch.linalg.cross(Su.T, Sv.T).T
    nrm = nrm / torch.linalg.norm(nrm, dim=0, keepdim=True)
    delta = torch.from_numpy(np.random.default_rng(1004).normal(0.0, 1e-2, N)).to(dev)
    P = (S + delta * nrm).contiguous()
    x, y, z = P[0].contiguous(), P[1].contiguous(), P[2].contiguous()
    u = torch.empty(N, dtype=torch.float64, device=dev)
    v = torch.empty(N, dtype=torch.float64, device=dev)
    st = torch.empty(N, dtype=torch.uint8, device=dev)
    torch.cuda.synchronize()

    # mixed: mixed_options (stepped seeds, <= 6 FP32 Newton iterations, <= 3 FP64 closing steps).
    # It is measured last: the comparison variants before it bring the GPU out of its idle clocks
    # (measured first, its 10 steps ran 5-8% slower than in steady state).
    mixed_opts = mixed_options(G)
    variants = {
        # the reference iteration (J^T J Gauss-Newton, src/kernel/kernel.cu:151-249) in FP64 from the
        # reference's node-centre seeds (greedy leaf)
        "plain_fp64": G.default_opts(precision=G.GN_FP64, newton=0, max_iter=20, tol_step=1e-10),
        # the same reference iteration from the mixed mode's stepped seeds, to separate the seed's
        # share of the gain
        "plain_fp64_step_seeds": G.default_opts(precision=G.GN_FP64, newton=0, max_iter=20, tol_step=1e-10,
                                                seed_mode=G.SEED_GRID_STEP),
        # the mixed mode's algorithm and seeds run entirely in FP64, to separate algorithm from precision
        "fp64_newton": G.default_opts(precision=G.GN_FP64, newton=G.NEWTON_FP64, max_iter=20, tol_step=1e-10,
                                      tol_quad=1e-10, seed_mode=G.SEED_GRID_STEP),
        # the mixed mode seeded by the greedy descent (no seed grid)
        "mixed_greedy_step_seeds": G.default_opts(**dict(mixed_opts, seed_mode=G.SEED_GREEDY_STEP)),
        "mixed": G.default_opts(**mixed_opts),
    }

    def run(opts):
        gs.NewInverse(x, y, z, opts, out=(u, v, st))

    def ok_count():
        err = torch.maximum((u - u0_t).abs(), (v - v0_t).abs())
        return int((err <= 1e-9).sum().item())

    stream = torch.cuda.current_stream()
    res = {}
    for name, opts in variants.items():
        for _ in range(args.warmup):
            run(opts)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(args.steps):
            run(opts)
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / args.steps
        okc = ok_count()
        st_np = st.cpu().numpy()
        res[name] = {"ms_per_step": ms, "ok_frac": okc / N, "points_per_s": okc / (ms / 1e3),
                     "status_hist": np.bincount(st_np, minlength=6).tolist()}

    # per-kernel timing of the dominant kernel (mixed Gauss-Newton from given seeds) and the seed
    # search, both with HIP events on the launch stream
    leaf, su_t, sv_t = gs.Seed(x, y, z, exact="grid")
    torch.cuda.synchronize()
    mixed_given = G.default_opts(**mixed_opts)
    reps = max(3, args.steps)

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(reps):
            fn()
        b.record(stream)
        torch.cuda.synchronize()
        return a.elapsed_time(b) / reps

    def gn_only():
        u.copy_(su_t)
        v.copy_(sv_t)
        mixed_given.seed_mode = G.SEED_GIVEN
        G._check(G.lib().gn_invert(gs.handle, x.data_ptr(), y.data_ptr(), z.data_ptr(), N,
                                   G.ctypes.byref(mixed_given), u.data_ptr(), v.data_ptr(), st.data_ptr(),
                                   G._c_void(stream.cuda_stream)))

    copy_ms = timed(lambda: (u.copy_(su_t), v.copy_(sv_t)))
    gn_ms = timed(gn_only) - copy_ms
    seed_ms = timed(lambda: gs.Seed(x, y, z, exact="grid"))
    _, _, _, it32, it64 = gs.NewInverse(x, y, z, G.default_opts(**{k: getattr(mixed_given, k) for k, _ in
                                                                   G.InvertOpts._fields_}),
                                        seeds=(su_t, sv_t), iters=True)
    it32 = it32.cpu().numpy().astype(np.int64)
    it64 = it64.cpu().numpy().astype(np.int64)
    # contract count: 392 flops per iteration (bicubic); the Newton iteration as built does more
    # (second derivatives and the Hessian term: flops_per_iter), reported beside it
    f32 = float(it32.sum()) * contract_flops_per_iter(4, 4)
    f64 = float(it64.sum()) * contract_flops_per_iter(4, 4)
    f32n = float(it32.sum()) * flops_per_iter(4, 4, True)
    f64n = float(it64.sum()) * flops_per_iter(4, 4, True)
    t = gn_ms / 1e3
    ideal = f32 / PEAK_F32 + f64 / PEAK_F64
    ideal_n = f32n / PEAK_F32 + f64n / PEAK_F64
    peak_eff = (f32 + f64) / ideal
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "inversion_projection_1M_traffic.json")
    if N == 1 << 20 and os.path.exists(tpath):
        # HBM bytes of the two Gauss-Newton launches of one step, from the committed rocprofv3
        # FETCH_SIZE / WRITE_SIZE passes (scaled by the copy-kernel calibration)
        with open(tpath) as f:
            traffic = json.load(f)["gn_step_traffic_bytes"]
    roof = {"bound": "valu", "achieved": (f32 + f64) / t / 1e12, "peak": peak_eff / 1e12, "unit": "TFLOP/s",
            "frac": ideal / t, "traffic": traffic,
            "flops_per_iter": contract_flops_per_iter(4, 4),
            "frac_newton_flops": ideal_n / t, "newton_flops_per_iter": flops_per_iter(4, 4, True),
            "algorithmic_bytes": N * (40 + 17) + N * (41 + 17),
            "kernel": "FP32 bulk (k_gn_refill) + FP64 closing pass (k_close_first, then k_gn_refill over the "
                      "listed rest): Gauss-Newton from given seeds",
            "roof_note": "VALU arithmetic (no MFMA): compute roof from FP32 flops at 157.3 TFLOP/s plus FP64 "
                         "flops at 78.6 TFLOP/s (published MI355X vector peaks); frac = time at those peaks / "
                         "measured time, counting 392 flops per iteration (2(8k1k2+12k1)+40, bicubic); "
                         "frac_newton_flops counts the Newton iteration as built",
            "kernel_ms": gn_ms, "flops_fp32": f32, "flops_fp64": f64,
            "iters_fp32_mean": float(it32.mean()), "iters_fp64_mean": float(it64.mean())}

    # global projection (branch and bound over the seed tree) on the same queries: its cost
    # against the local mixed mode, and the leaves it solves per query
    gopts = G.default_opts(precision=G.GN_FP64, newton=G.NEWTON_FP64, seed_mode=G.SEED_GLOBAL, max_iter=20,
                           tol_step=1e-12)
    gs.NewInverse(x, y, z, gopts, out=(u, v, st))
    torch.cuda.synchronize()
    g0 = time.perf_counter()
    _, _, _, ncand, _ = gs.NewInverse(x, y, z, gopts, out=(u, v, st), iters=True)
    torch.cuda.synchronize()
    g_ms = (time.perf_counter() - g0) * 1e3
    global_res = {"ms": g_ms, "points_per_s": ok_count() / (g_ms / 1e3), "ok_frac": ok_count() / N,
                  "leaves_solved_mean": float(ncand.float().mean().item()),
                  "note": "GN_SEED_GLOBAL: greedy seed leaf + Newton, then leaf-clamped projected Newton in "
                          "every leaf whose box may hold a nearer point and whose normal cone may hold its "
                          "normal (solved in wave-wide rounds), then a polish (FP64); one timed call"}

    cpu = cpu_case = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O

        ci = cpu_info()
        threads = ci["threads"]
        os_ = O.OSurf.from_def(sd)
        nodes, _, _ = os_.build_tree(3)  # set-up, like the GPU tree (not timed on either side)
        xs, ys, zs = (a.cpu().numpy() for a in (x, y, z))
        reps, dt = 0, 0.0
        while dt < args.cpu_seconds and reps < 100:
            t0 = time.perf_counter()
            _, ssu, ssv = O.seed_greedy_leaf(nodes, 3, xs, ys, zs, nthreads=threads)
            cu, cv, cst, cit = os_.gauss_newton(xs, ys, zs, ssu, ssv, tol_step=1e-10, max_iter=20,
                                                nthreads=threads)
            dt += time.perf_counter() - t0
            reps += 1
        cok = int(np.sum(np.maximum(np.abs(cu - u0), np.abs(cv - v0)) <= 1e-9))
        cpu = {"value": cok * reps / dt, "unit": "inversions/s", "cores": threads, "kind": "port",
               "sample": f"all {N} queries x {reps} repetitions ({dt:.1f} s wall): greedy-leaf seed "
                         f"search + plain FP64 Gauss-Newton (oracle restatement of "
                         f"src/kernel/kernel.cu:151-249, <=20 iterations, 1e-10 step)",
               "ok_frac": cok / N, "nproc": ci["nproc"], "affinity_cpus": ci["affinity"],
               "omp_num_threads": ci["omp_num_threads"], "cpu_model": ci["model"],
               "threads_note": "every CPU this process may use (its affinity mask), capped by OMP_NUM_THREADS "
                               "where the host sets the GPU's CPU share"}
        cpu_case = bench_cpu_case(G, synth, torch, O, args.steps, threads)

    eval_res = None if args.no_eval else bench_eval(G, synth, torch, max(3, args.steps))
    batch_res = None if args.no_batch else bench_batch(G, synth, torch, max(3, args.steps))
    serve_res = None if args.no_serve else bench_serve(G, synth, torch)

    mixed = res["mixed"]
    out = {
        "metric": "point inversions/s at <=1e-9 (u,v) err, mixed vs plain FP64; p50/p99 latency",
        "value": mixed["points_per_s"],
        "unit": "inversions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mixed["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": "projection: bicubic rational surface, 32x32 net (generator G, seed 1002), "
                               f"{N} queries offset N(0,1e-2) along the normal, depth-6 tree (64x64 leaves) "
                               "and seed grid, seed-grid leaf + one Gauss-Newton step (leaf Jacobian) + two on the "
                               "leaf's quadratic model, <=6 FP32 + <=3 FP64 iterations",
                   "points": N, "surface": "p=q=3, 32x32", "tree_depth": 6},
        "roofline": roof,
        "cpu_baseline": cpu,
        "mixed": mixed,
        "plain_fp64": res["plain_fp64"],
        "fp64_newton": res["fp64_newton"],
        "plain_fp64_step_seeds": res["plain_fp64_step_seeds"],
        "mixed_greedy_step_seeds": res["mixed_greedy_step_seeds"],
        "surface_setup_ms": setup_ms,
        "speedup_vs_plain_fp64": mixed["points_per_s"] / max(res["plain_fp64"]["points_per_s"], 1e-30),
        "seed_ms": seed_ms,
        "global_projection": global_res,
        "cpu_case_8x8": cpu_case,
        "eval_16M": eval_res,
        "uneven_batch": batch_res,
        "serving": serve_res,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
