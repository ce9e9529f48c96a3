// batch.hip -- inversion of many surfaces of uneven degree and net size (the batched form of
// NewInverse + GaussNewton, src/kernel/kernel.cu:151-249, 433-469).
//
// A batch packs the span tables, knots and seed trees of its surfaces into one allocation and
// describes each surface by a GnBatchObj (pointers into the pack, degree class).  Work is a list
// of tiles (object, contiguous point range, request) sorted by degree class.  The 17 degree
// classes (exact unrolled contraction for degrees 2..5 in each direction, padded generic form
// otherwise) fall into three groups by patch size (K1*K2 <= 16, <= 25, larger or generic); each
// group present gets its own persistent launch of a kernel instantiated for that group alone, so
// a launch has the register footprint of its group -- one kernel holding every class would be
// sized by the largest and run at one wave per SIMD.  Within a launch, waves claim the group's
// tiles from sliced counters and keep their 64 lanes busy with the points of those tiles (lane
// refill); a wave holds points of one class at a time, so the class dispatch is wave-uniform,
// and it drains its lanes before it starts a tile of another class.  Each group runs on its own
// side stream: a seed launch over its tiles (greedy or nearest leaf of each point's own surface,
// one lane per point: found inside the refill loop a seed would be paid by the whole wave at the
// refilling lanes' occupancy), then the FP32 bulk and the FP64 closing pass, launches of the same
// template as in the single-surface mixed mode.  Point arrays are addressed through a request
// table (GnReq): a tile names its request, so one launch serves the points of many independent
// requests (the serving loop) or of one call (a one-entry table).
#include <algorithm>
#include <mutex>

#include "gn_search.hpp"
#include "invert_common.hpp"

namespace gn {

constexpr int kExactMaxK = 6;  // largest k1, k2 of an exact degree class (degree 5)

struct BatchArgs {
    const GnBatchObj *objs;
    const GnTile *tiles;  // the tiles of one degree class
    int ntiles;
    const GnReq *reqs;
    gn_invert_opts o;
};

// Seeds of every point of the tiles into its request's u, v.  Workgroup b takes the points
// (b % per_tile) * kSeedBlock .. + kSeedBlock - 1 of tile b / per_tile, one point per lane, so
// every point has a lane of its own (the seed search is a chain of dependent loads).  All lanes
// of a workgroup search one surface, so the greedy form first stages the FP32 boxes of its tree's
// levels 1..kBatchSeedLdsLevels (340 boxes, 8 KB) in LDS and descends those levels there; the grid
// form (GN_SEED_GRID_STEP) looks the surface's seed grid up instead and stages nothing.
constexpr int kSeedBlock = 256;
constexpr int kBatchSeedLdsLevels = 4;
__global__ __launch_bounds__(kSeedBlock) void k_batch_seed(const GnBatchObj *__restrict__ objs,
                                                           const GnTile *__restrict__ tiles,
                                                           long long nblocks, int per_tile,
                                                           const GnReq *__restrict__ reqs,
                                                           int mode) {
    __shared__ __align__(16) float lbox[(((1 << (2 * (kBatchSeedLdsLevels + 1))) - 1) / 3 - 1) * 6];
    for (long long b = blockIdx.x; b < nblocks; b += gridDim.x) {  // workgroup-uniform
        const GnTile tl = tiles[b / per_tile];
        const GnSearchDev &sd = objs[tl.obj].sd;
        const int ll = mode == GN_SEED_NEAREST_LEAF || mode == GN_SEED_GRID_STEP
                           ? 0
                           : (sd.depth < kBatchSeedLdsLevels ? sd.depth : kBatchSeedLdsLevels);
        __syncthreads();  // the previous tile's descents are done with lbox
        if (ll > 0) {
            // boxes of levels 1..ll start one box (24 bytes) into box32: float2 copies
            const int nf2 = (int)(qoff(ll + 1) - 1) * 3;
            const float2 *src = reinterpret_cast<const float2 *>(sd.box32 + 6);
            float2 *dst = reinterpret_cast<float2 *>(lbox);
            for (int k = threadIdx.x; k < nf2; k += blockDim.x) dst[k] = src[k];
        }
        __syncthreads();
        const int j = (int)(b % per_tile) * kSeedBlock + threadIdx.x;
        if (j >= tl.count) continue;
        const GnReq &R = reqs[tl.req];
        const int64_t i = tl.start + j;
        const double P[3] = {R.x[i], R.y[i], R.z[i]};
        if (mode == GN_SEED_GRID_STEP) {  // the object (and its grid's width) is workgroup-uniform
            const uint32_t c = sd.grid_wide ? grid_leaf<true>(sd, P) : grid_leaf<false>(sd, P);
            double u, v;
            step_seed(sd, c, sd.rec[(size_t)c * kRecVec], P, u, v);
            R.u[i] = u;
            R.v[i] = v;
            continue;
        }
        float4 cq;
        const uint32_t c = mode == GN_SEED_NEAREST_LEAF
                               ? nearest_leaf(sd, P)
                               : greedy_leaf(sd, P, lbox, ll, &cq, mode != GN_SEED_GREEDY_STEP);
        if (mode == GN_SEED_GREEDY_STEP) {
            double u, v;
            step_seed(sd, c, cq, P, u, v);
            R.u[i] = u;
            R.v[i] = v;
        } else {
            R.u[i] = sd.seed_uv[c * 2 + 0];
            R.v[i] = sd.seed_uv[c * 2 + 1];
        }
    }
}

// One inversion step of a point of degree class cls, for the classes of group GRP (other classes
// never reach a kernel of this group).  pr: the lane's register patch (FP32 group-0 classes).
// WAVE (exact classes only): patch rows staged through the wave's LDS buffer with the next row
// prefetched; the whole wave must make the call.
template <int GRP, typename T, bool NEWTON, bool PREG, bool WAVE = false>
__device__ __forceinline__ int step_group(int cls, const GnSurfDev &d, T Px, T Py, T Pz,
                                          double &cu, double &cv, double tol_dist,
                                          double tol_step, double tol_quad, int clamp,
                                          SpanCache &sc, Hom<T> *pr, float4 *wave_lds = nullptr) {
    switch (cls) {
#define GN_CLS(P, Q)                                                                            \
    case (P - 2) * 4 + (Q - 2):                                                                 \
        if constexpr (class_group((P - 2) * 4 + (Q - 2)) == GRP)                                \
            return gn_step<P + 1, Q + 1, true, T, NEWTON,                                       \
                           WAVE ? kPatchWavePipe                                                \
                                : (PREG && (P + 1) * (Q + 1) <= 16) ? kPatchRegs : kPatchGlobal, \
                           false, true>(                                                        \
                d, d.ku, d.kv, Px, Py, Pz, cu, cv, tol_dist, tol_step, tol_quad, clamp,         \
                domain_of(d), sc, pr, wave_lds);                                                \
        break;
        GN_CLS(2, 2) GN_CLS(2, 3) GN_CLS(2, 4) GN_CLS(2, 5)
        GN_CLS(3, 2) GN_CLS(3, 3) GN_CLS(3, 4) GN_CLS(3, 5)
        GN_CLS(4, 2) GN_CLS(4, 3) GN_CLS(4, 4) GN_CLS(4, 5)
        GN_CLS(5, 2) GN_CLS(5, 3) GN_CLS(5, 4) GN_CLS(5, 5)
#undef GN_CLS
        default:
            break;
    }
    if constexpr (GRP == 2 && !WAVE)
        return gn_step<GN_MAXK, GN_MAXK, false, T, NEWTON, kPatchGlobal, false, true>(
            d, d.ku, d.kv, Px, Py, Pz, cu, cv, tol_dist, tol_step, tol_quad, clamp, domain_of(d),
            sc);
    return GN_PT_NONFINITE;  // not reached: the tiles of a group launch are of its classes
}

// Persistent lane-refill kernel over the tiles of one class group.  A wave's lanes always hold
// points of one degree class (wave-uniform dispatch); when its next tile is of another class,
// the wave drains its current points first.
// No minimum occupancy is forced: the group-0 FP32 kernel (six classes behind one switch) needs
// about 220 VGPRs, and holding it to three waves per SIMD spilled its patch registers to scratch.
template <int GRP, typename T, bool NEWTON, int MODE>
__global__ __launch_bounds__(kInvBlock, 1)
void k_batch(BatchArgs a, unsigned long long *counters, int slices, int max_iter,
             int restart_iter, double tol_dist, double tol_step, double tol_quad) {
    const int lane = threadIdx.x & 63;
    const unsigned long long lt = (1ull << lane) - 1ull, all = ~0ull;
    const int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int slice = wave % slices;
    const long long s_lo = (long long)a.ntiles * slice / slices;
    const long long s_hi = (long long)a.ntiles * (slice + 1) / slices;
    unsigned long long *counter = counters + (size_t)slice * kCounterStride;
    // wave-uniform tile state
    long long q_lo = 0, q_hi = 0;
    int cur_obj = 0, cur_cls = -1, cur_req = 0, pcls = -1;
    bool exhausted = false, pend = false;
    GnTile pt{0, 0, 0, 0, 0};
    // lane state
    long long idx = -1;
    int rq = 0, it = 0, lim = max_iter;
    GnSurfDev d{};
    double cu = 0, cv = 0, su = 0, sv = 0;
    T Px = 0, Py = 0, Pz = 0;
    SpanCache sc;
    constexpr bool PREG = GRP == 0 && sizeof(T) == 4;
    PatchRegs<T, 16> pr;
    // FP64 group-1 kernels: exact classes stage patch rows through LDS with the next row in flight,
    // one buffer per wave sized for the largest exact row (q = 5).  Group 0 would drop from two
    // waves per SIMD to one and group 2 would spill; group 1 runs one wave per SIMD either way
    // (uneven batch of 1,024 surfaces: plain FP64 8.09 -> 7.37 ms, mixed 2.52 -> 2.44 ms).
    constexpr bool WAVE = sizeof(T) == 8 && GRP == 1;
    __shared__ float4 wrows[kInvBlock / 64][WAVE ? WaveRowsPipe<kExactMaxK, kExactMaxK, double>::STRIDE * 64 : 1];
    float4 *wave_lds = wrows[threadIdx.x >> 6];
    while (true) {
        while (true) {
            const unsigned long long idle = __ballot(idx < 0);
            if (idle == 0ull) break;
            if (q_lo >= q_hi) {
                if (!pend) {
                    if (exhausted) break;
                    unsigned long long k = 0;
                    if (lane == 0) k = atomicAdd(counter, 1ull);
                    k = bcast64(k, 0);
                    if (s_lo + (long long)k >= s_hi) {
                        exhausted = true;
                        break;
                    }
                    pt = a.tiles[s_lo + (long long)k];
                    pcls = a.objs[pt.obj].cls;
                    pend = true;
                }
                if (pcls != cur_cls && idle != all) break;  // drain the current class first
                cur_obj = pt.obj;
                cur_cls = pcls;
                cur_req = pt.req;
                q_lo = pt.start;
                q_hi = pt.start + pt.count;
                pend = false;
                continue;
            }
            const unsigned long long avail = (unsigned long long)(q_hi - q_lo);
            if (idx < 0) {
                const unsigned long long rank = (unsigned long long)__popcll(idle & lt);
                if (rank < avail) {
                    idx = q_lo + (long long)rank;
                    rq = cur_req;
                    d = a.objs[cur_obj].s;
                    const GnReq &R = a.reqs[cur_req];
                    Px = T(R.x[idx]);
                    Py = T(R.y[idx]);
                    Pz = T(R.z[idx]);
                    cu = R.u[idx];  // the seed, or the FP32 result for the closing pass
                    cv = R.v[idx];
                    it = 0;
                    lim = max_iter;
                    if (MODE == kClose) {
                        const int sb = R.status[idx];
                        if (sb == GN_PT_SINGULAR || sb == GN_PT_NONFINITE) lim = restart_iter;
                    }
                    su = cu;
                    sv = cv;
                    sc = SpanCache();
                }
            }
            const unsigned long long nidle = (unsigned long long)__popcll(idle);
            q_lo += (long long)(nidle < avail ? nidle : avail);
        }
        if (__ballot(idx >= 0) == 0ull) break;
        const bool act = idx >= 0 && it < lim;
        int st = kStepContinue;
        if (WAVE && cur_cls < 16) {
            // the whole wave steps (cur_cls is wave-uniform); a lane with no step due evaluates
            // the wave's current surface at its domain corner and discards the result
            const GnSurfDev dd = act ? d : a.objs[cur_obj].s;
            double tu = act ? cu : dd.ku[dd.p], tv = act ? cv : dd.kv[dd.q];
            const int s2 = step_group<GRP, T, NEWTON, PREG, true>(
                cur_cls, dd, act ? Px : T(0), act ? Py : T(0), act ? Pz : T(0), tu, tv, tol_dist,
                tol_step, tol_quad, a.o.clamp, sc, pr.m, wave_lds);
            if (act) {
                st = s2;
                cu = tu;
                cv = tv;
            }
        } else if (act) {
            st = step_group<GRP, T, NEWTON, PREG>(cur_cls, d, Px, Py, Pz, cu, cv, tol_dist,
                                                  tol_step, tol_quad, a.o.clamp, sc, pr.m);
        }
        if (idx >= 0) {
            if (!act) {
                st = GN_PT_MAXITER;
            } else {
                if (st == kStepContinue || st == GN_PT_CONVERGED) it++;
                if (st == kStepContinue && it >= lim) st = GN_PT_MAXITER;
            }
            if (st != kStepContinue) {
                if (MODE == kBulk && (st == GN_PT_SINGULAR || st == GN_PT_NONFINITE)) {
                    cu = su;
                    cv = sv;
                }
                const GnReq &R = a.reqs[rq];
                R.u[idx] = cu;
                R.v[idx] = cv;
                R.status[idx] = (uint8_t)st;
                if (MODE == kBulk) {
                    if (R.it32) R.it32[idx] = sat8(it);
                } else {
                    if (MODE == kPlain && R.it32) R.it32[idx] = 0;
                    if (R.it64) R.it64[idx] = sat8(it);
                }
                idx = -1;
            }
        }
    }
    if (lane == 0) retire_wave(counters, slices);
}

static void batch_shape(const void *kernel, int ntiles, size_t n, int bpc_cap, unsigned &grid,
                        int &slices) {
    const size_t want = (n + kInvBlock - 1) / kInvBlock;
    int bpc = blocks_per_cu(kernel, kInvBlock);
    if (bpc_cap > 0 && bpc_cap < bpc) bpc = bpc_cap;
    const size_t cap = (size_t)cu_count() * bpc;
    grid = (unsigned)(want < cap ? (want ? want : 1) : cap);
    const int waves = (int)grid * (kInvBlock / 64);
    slices = waves / 16;
    if (slices > ntiles / 4) slices = ntiles / 4;
    if (slices < 1) slices = 1;
    if (slices > kMaxSlices) slices = kMaxSlices;
}

// One launch of a group kernel over the group's tiles.
template <int GRP, typename T, bool NEWTON, int MODE>
static int launch_grp(const BatchArgs &a, size_t n, int bpc_cap, int max_iter, int restart_iter,
                      double tol_dist, double tol_step, double tol_quad, hipStream_t st) {
    unsigned long long *ctr = nullptr;
    int rc = work_counters(st, &ctr);
    if (rc) return rc;
    auto k = k_batch<GRP, T, NEWTON, MODE>;
    unsigned grid;
    int slices;
    batch_shape((const void *)k, a.ntiles, n, bpc_cap, grid, slices);
    k<<<grid, kInvBlock, 0, st>>>(a, ctr, slices, max_iter, restart_iter, tol_dist, tol_step,
                                  tol_quad);
    GN_HIP(hipGetLastError());
    return GN_OK;
}

// The precision phases of one class group.
template <int GRP>
static int group_phases(const BatchArgs &a, size_t n, int bpc_cap, hipStream_t st) {
    const gn_invert_opts &o = a.o;
    const bool n32 = o.newton & GN_NEWTON_FP32, n64 = o.newton & GN_NEWTON_FP64;
    int rc;
    if (o.precision == GN_MIXED) {
        rc = n32 ? launch_grp<GRP, float, true, kBulk>(a, n, bpc_cap, o.fp32_iters, 0, 0.0,
                                                        o.tol_step32, 0.0, st)
                 : launch_grp<GRP, float, false, kBulk>(a, n, bpc_cap, o.fp32_iters, 0, 0.0,
                                                         o.tol_step32, 0.0, st);
        if (rc) return rc;
        return n64 ? launch_grp<GRP, double, true, kClose>(a, n, bpc_cap, o.fp64_tail, o.max_iter,
                                                            o.tol_dist, o.tol_step, o.tol_quad, st)
                   : launch_grp<GRP, double, false, kClose>(a, n, bpc_cap, o.fp64_tail, o.max_iter,
                                                             o.tol_dist, o.tol_step, 0.0, st);
    }
    return n64 ? launch_grp<GRP, double, true, kPlain>(a, n, bpc_cap, o.max_iter, 0, o.tol_dist,
                                                        o.tol_step, o.tol_quad, st)
               : launch_grp<GRP, double, false, kPlain>(a, n, bpc_cap, o.max_iter, 0, o.tol_dist,
                                                         o.tol_step, 0.0, st);
}

int batch_waves(int bpc_cap) {
    auto k = k_batch<0, float, true, kBulk>;
    int bpc = blocks_per_cu((const void *)k, kInvBlock);
    if (bpc_cap > 0 && bpc_cap < bpc) bpc = bpc_cap;
    return cu_count() * bpc * (kInvBlock / 64);
}

int class_tile(size_t points, int waves, int cap) {
    size_t t = points / (2 * (size_t)(waves > 0 ? waves : 1)) / 64 * 64;
    if (t > (size_t)cap) t = (size_t)cap;
    if (t < 64) t = 64;
    return (int)t;
}

std::vector<GroupRange> group_ranges(const GnTile *tiles, size_t ntiles, const std::vector<int> &cls) {
    std::vector<GroupRange> r;
    for (size_t t = 0; t < ntiles; t++) {
        const int g = class_group(cls[tiles[t].obj]);
        if (r.empty() || r.back().grp != g) r.push_back(GroupRange{g, (int)t, (int)t, 0, 0});
        r.back().t1 = (int)t + 1;
        r.back().points += (size_t)tiles[t].count;
        if (tiles[t].count > r.back().max_count) r.back().max_count = tiles[t].count;
    }
    return r;
}

int side_streams_create(SideStreams *ss) {
    for (int k = 0; k < kSideStreams; k++) {
        GN_HIP(hipStreamCreateWithFlags(&ss->s[k], hipStreamNonBlocking));
        GN_HIP(hipEventCreateWithFlags(&ss->join[k], hipEventDisableTiming));
    }
    GN_HIP(hipEventCreateWithFlags(&ss->fork, hipEventDisableTiming));
    return GN_OK;
}

void side_streams_destroy(SideStreams *ss) {
    for (int k = 0; k < kSideStreams; k++) {
        if (ss->s[k]) (void)hipStreamDestroy(ss->s[k]);
        if (ss->join[k]) (void)hipEventDestroy(ss->join[k]);
    }
    if (ss->fork) (void)hipEventDestroy(ss->fork);
    *ss = SideStreams();
}

namespace {
std::mutex g_side_mu;
SideStreams g_side[64];  // per device, created on first use
}  // namespace

// Seeds of one group range (one lane per point; blocks cover tiles in per_tile pieces).
static int seed_range(const GnBatchObj *objs, const GnTile *tiles, const GroupRange &gr,
                      const GnReq *reqs, int seed_mode, hipStream_t st) {
    const int per_tile = (std::max(gr.max_count, 1) + kSeedBlock - 1) / kSeedBlock;
    const long long nb = (long long)(gr.t1 - gr.t0) * per_tile;
    const unsigned grid = (unsigned)(nb < (1 << 20) ? nb : (1 << 20));
    k_batch_seed<<<grid, kSeedBlock, 0, st>>>(objs, tiles + gr.t0, nb, per_tile, reqs, seed_mode);
    GN_HIP(hipGetLastError());
    return GN_OK;
}

static int run_range(const GnBatchObj *objs, const GnTile *tiles, const GroupRange &gr,
                     const GnReq *reqs, const gn_invert_opts &o, int bpc_cap, hipStream_t st) {
    if (o.seed_mode != GN_SEED_GIVEN) {
        const int rc = seed_range(objs, tiles, gr, reqs, o.seed_mode, st);
        if (rc) return rc;
    }
    const BatchArgs a{objs, tiles + gr.t0, gr.t1 - gr.t0, reqs, o};
    return gr.grp == 0 ? group_phases<0>(a, gr.points, bpc_cap, st)
         : gr.grp == 1 ? group_phases<1>(a, gr.points, bpc_cap, st)
                       : group_phases<2>(a, gr.points, bpc_cap, st);
}

static int launch_ranges(const GnBatchObj *objs, const GnTile *tiles,
                         const std::vector<GroupRange> &ranges, const GnReq *reqs,
                         const gn_invert_opts &o, int bpc_cap, hipStream_t st,
                         const SideStreams &ss, const std::function<int(hipStream_t)> &beside,
                         std::vector<std::pair<int, hipEvent_t>> *group_done) {
    if (ranges.size() == 1 && !beside)  // one group, nothing beside it: no fork
        return run_range(objs, tiles, ranges[0], reqs, o, bpc_cap, st);
    if (ranges.size() > (size_t)kSideStreams)
        return fail(GN_EDEVICE, "more group ranges than side streams (tiles not sorted by class)");
    GN_HIP(hipEventRecord(ss.fork, st));
    // st joins every side stream that was forked, also when a later launch fails: the caller
    // may then synchronise st alone (or free buffers in its order) before it reuses what the
    // kernels already enqueued still read or write
    int k = 0, rc = GN_OK;
    for (const GroupRange &gr : ranges) {
        if (gr.points == 0) continue;
        const hipError_t e = hipStreamWaitEvent(ss.s[k], ss.fork, 0);
        if (e != hipSuccess) {
            rc = hip_fail(e, "hipStreamWaitEvent(fork)");
            break;
        }
        rc = run_range(objs, tiles, gr, reqs, o, bpc_cap, ss.s[k]);
        const hipError_t ej = hipEventRecord(ss.join[k], ss.s[k]);
        if (ej != hipSuccess) {
            if (!rc) rc = hip_fail(ej, "hipEventRecord(join)");
            (void)hipStreamSynchronize(ss.s[k]);  // no join event: wait for it here
            break;
        }
        k++;
        if (rc) break;
        if (group_done) group_done->emplace_back(gr.grp, ss.join[k - 1]);
    }
    if (!rc && beside) rc = beside(st);
    for (int j = 0; j < k; j++) {
        const hipError_t e = hipStreamWaitEvent(st, ss.join[j], 0);
        if (e != hipSuccess) {
            (void)hipStreamSynchronize(ss.s[j]);
            if (!rc) rc = hip_fail(e, "hipStreamWaitEvent(join)");
        }
    }
    return rc;
}

static int shared_side_locked(int dev) {
    if (g_side[dev].fork) return GN_OK;
    const int rc = side_streams_create(&g_side[dev]);
    if (rc) side_streams_destroy(&g_side[dev]);
    return rc;
}

int batch_prepare() {
    int dev = 0;
    GN_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return fail(GN_EDEVICE, "device index out of range");
    std::lock_guard<std::mutex> lk(g_side_mu);
    return shared_side_locked(dev);
}

int launch_batch_invert(const GnBatchObj *objs, const GnTile *tiles,
                        const std::vector<GroupRange> &ranges, const GnReq *reqs,
                        const gn_invert_opts &o, int bpc_cap, hipStream_t st,
                        const SideStreams *side, const std::function<int(hipStream_t)> &beside,
                        std::vector<std::pair<int, hipEvent_t>> *group_done) {
    if (ranges.empty()) return beside ? beside(st) : GN_OK;
    if (side)
        return launch_ranges(objs, tiles, ranges, reqs, o, bpc_cap, st, *side, beside, group_done);
    int dev = 0;
    GN_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return fail(GN_EDEVICE, "device index out of range");
    std::lock_guard<std::mutex> lk(g_side_mu);  // one user of the shared set at a time
    const int rc = shared_side_locked(dev);
    if (rc) return rc;
    return launch_ranges(objs, tiles, ranges, reqs, o, bpc_cap, st, g_side[dev], beside, group_done);
}

}  // namespace gn
