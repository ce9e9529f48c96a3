// batch.hip -- inversion of many surfaces of uneven degree and net size in one persistent launch
// (the batched form of NewInverse + GaussNewton, src/kernel/kernel.cu:151-249, 433-469).
//
// A batch packs the span tables, knots and seed trees of its surfaces into one allocation and
// describes each surface by a GnBatchObj (pointers into the pack, degree class).  Work is a list
// of tiles (object, contiguous point range) sorted by degree class.  Waves claim tiles from sliced
// counters and keep their 64 lanes busy with the points of their tiles; all lanes of a wave
// always belong to one degree class, so the per-class template (exact unrolled contraction for
// degrees 2..5 in each direction, padded generic form otherwise) is chosen once per wave step
// without divergence.  When the next tile is of another class, the wave drains its current points
// before switching.  Seeds are found inline when a lane takes a point (greedy leaf of the
// point's own surface).  The FP32 bulk and the FP64 closing pass are two launches of the same
// kernel template, as in the single-surface mixed mode.  Point arrays are addressed through a
// request table (GnReq): a tile names its request, so one launch can serve the points of many
// independent requests (the serving loop) or of one call (a one-entry table).
#include "gn_search.hpp"
#include "invert_common.hpp"

namespace gn {

struct BatchArgs {
    const GnBatchObj *objs;
    const GnTile *tiles;
    int ntiles;
    const GnReq *reqs;
    gn_invert_opts o;
};

template <typename T, bool NEWTON>
__device__ __forceinline__ int step_cls(int cls, const GnSurfDev &d, T Px, T Py, T Pz, double &cu,
                                        double &cv, double tol_dist, double tol_step,
                                        double tol_quad, int clamp, SpanCache &sc) {
    switch (cls) {
#define GN_CLS(P, Q)                                                                            \
    case (P - 2) * 4 + (Q - 2):                                                                 \
        return gn_step<P + 1, Q + 1, true, T, NEWTON>(d, d.ku, d.kv, Px, Py, Pz, cu, cv, tol_dist, \
                                                      tol_step, tol_quad, clamp, domain_of(d), sc);
        GN_CLS(2, 2) GN_CLS(2, 3) GN_CLS(2, 4) GN_CLS(2, 5)
        GN_CLS(3, 2) GN_CLS(3, 3) GN_CLS(3, 4) GN_CLS(3, 5)
        GN_CLS(4, 2) GN_CLS(4, 3) GN_CLS(4, 4) GN_CLS(4, 5)
        GN_CLS(5, 2) GN_CLS(5, 3) GN_CLS(5, 4) GN_CLS(5, 5)
#undef GN_CLS
        default:
            return gn_step<GN_MAXK, GN_MAXK, false, T, NEWTON>(d, d.ku, d.kv, Px, Py, Pz, cu, cv,
                                                               tol_dist, tol_step, tol_quad, clamp,
                                                               domain_of(d), sc);
    }
}

template <typename T, bool NEWTON, int MODE>
__global__ __launch_bounds__(kInvBlock) void k_batch(BatchArgs a, unsigned long long *counters,
                                                     int slices, int max_iter, int restart_iter,
                                                     double tol_dist, double tol_step,
                                                     double tol_quad) {
    const int lane = threadIdx.x & 63;
    const unsigned long long lt = (1ull << lane) - 1ull, all = ~0ull;
    const int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int slice = wave % slices;
    const long long s_lo = (long long)a.ntiles * slice / slices;
    const long long s_hi = (long long)a.ntiles * (slice + 1) / slices;
    unsigned long long *counter = counters + (size_t)slice * kCounterStride;
    // wave-uniform tile state
    long long q_lo = 0, q_hi = 0;
    int cur_obj = -1, cur_cls = -1, cur_req = 0;
    bool exhausted = false, pend = false;
    GnTile pt{0, 0, 0};
    int pcls = -1;
    // lane state
    long long idx = -1;
    int obj = 0, rq = 0, it = 0, lim = max_iter;
    GnSurfDev d{};
    double cu = 0, cv = 0, su = 0, sv = 0;
    T Px = 0, Py = 0, Pz = 0;
    SpanCache sc;
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
                    obj = cur_obj;
                    rq = cur_req;
                    d = a.objs[obj].s;
                    const GnReq &R = a.reqs[cur_req];
                    const double P[3] = {R.x[idx], R.y[idx], R.z[idx]};
                    Px = T(P[0]);
                    Py = T(P[1]);
                    Pz = T(P[2]);
                    it = 0;
                    lim = max_iter;
                    if (MODE == kClose) {
                        cu = R.u[idx];
                        cv = R.v[idx];
                        const int sb = R.status[idx];
                        if (sb == GN_PT_SINGULAR || sb == GN_PT_NONFINITE) lim = restart_iter;
                    } else if (a.o.seed_mode == GN_SEED_GIVEN) {
                        cu = R.u[idx];
                        cv = R.v[idx];
                    } else {
                        const GnSearchDev &sd = a.objs[obj].sd;
                        const uint32_t c = a.o.seed_mode == GN_SEED_NEAREST_LEAF
                                               ? nearest_leaf(sd, P)
                                               : greedy_leaf(sd, P);
                        cu = sd.seed_uv[c * 2 + 0];
                        cv = sd.seed_uv[c * 2 + 1];
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
        if (idx >= 0) {
            int st;
            if (it >= lim) {
                st = GN_PT_MAXITER;
            } else {
                st = step_cls<T, NEWTON>(cur_cls, d, Px, Py, Pz, cu, cv, tol_dist, tol_step,
                                         tol_quad, a.o.clamp, sc);
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

template <typename T, bool NEWTON, int MODE>
static int launch_one(const BatchArgs &a, size_t n, int bpc_cap, int max_iter, int restart_iter,
                      double tol_dist, double tol_step, double tol_quad, hipStream_t st) {
    unsigned long long *ctr = nullptr;
    int rc = work_counters(st, &ctr);
    if (rc) return rc;
    auto k = k_batch<T, NEWTON, MODE>;
    unsigned grid;
    int slices;
    batch_shape((const void *)k, a.ntiles, n, bpc_cap, grid, slices);
    k<<<grid, kInvBlock, 0, st>>>(a, ctr, slices, max_iter, restart_iter, tol_dist, tol_step,
                                  tol_quad);
    GN_HIP(hipGetLastError());
    return GN_OK;
}

int batch_waves(int bpc_cap) {
    auto k = k_batch<float, true, kBulk>;
    int bpc = blocks_per_cu((const void *)k, kInvBlock);
    if (bpc_cap > 0 && bpc_cap < bpc) bpc = bpc_cap;
    return cu_count() * bpc * (kInvBlock / 64);
}

int launch_batch_invert(const GnBatchObj *objs, const GnTile *tiles, int ntiles, const GnReq *reqs,
                        size_t n, const gn_invert_opts &o, int bpc_cap, hipStream_t st) {
    if (n == 0 || ntiles == 0) return GN_OK;
    BatchArgs a{objs, tiles, ntiles, reqs, o};
    const bool n32 = o.newton & GN_NEWTON_FP32, n64 = o.newton & GN_NEWTON_FP64;
    int rc;
    if (o.precision == GN_MIXED) {
        rc = n32 ? launch_one<float, true, kBulk>(a, n, bpc_cap, o.fp32_iters, 0, 0.0, o.tol_step32, 0.0, st)
                 : launch_one<float, false, kBulk>(a, n, bpc_cap, o.fp32_iters, 0, 0.0, o.tol_step32, 0.0, st);
        if (rc) return rc;
        return n64 ? launch_one<double, true, kClose>(a, n, bpc_cap, o.fp64_tail, o.max_iter, o.tol_dist,
                                                      o.tol_step, o.tol_quad, st)
                   : launch_one<double, false, kClose>(a, n, bpc_cap, o.fp64_tail, o.max_iter, o.tol_dist,
                                                       o.tol_step, 0.0, st);
    }
    return n64 ? launch_one<double, true, kPlain>(a, n, bpc_cap, o.max_iter, 0, o.tol_dist, o.tol_step,
                                                  o.tol_quad, st)
               : launch_one<double, false, kPlain>(a, n, bpc_cap, o.max_iter, 0, o.tol_dist, o.tol_step,
                                                   0.0, st);
}

}  // namespace gn
