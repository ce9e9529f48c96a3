// project.hip -- global projection (GN_SEED_GLOBAL): the nearest point of the surface to each
// query, not just the local minimum that Newton reaches from one seed.  The reference's version
// (normal-cone culling, src/kernel/kernel.cu:471-513, 623-704) is unfinished: its test tries the
// box corners only, for one sign of the normal.  This one is a branch and bound over the seed
// quadtree whose leaf boxes contain their pieces of surface:
//   1. bound: the greedy seed leaf (descent + 3 x 3), then projected Newton over the whole domain
//      (any start gives a valid bound; the exact nearest leaf centre cost a second tree walk);
//   2. every leaf whose FP32 box (a conservative lower bound) is not farther than the best foot
//      point so far, and whose normal cone allows a foot point inside it (cone_may_hold; leaves
//      on the domain boundary, where the nearest point need not be a foot point, always pass),
//      gets a projected Newton solve clamped to its uv cell, from the cell centre; a nearer
//      result tightens the bound, which prunes the rest of the depth-first walk;
//   3. the nearest foot point is polished by projected Newton over the whole domain (kept unless it
//      is farther by more than rounding, 1e-9 relative in squared distance).
// The minimum is found whenever the solve inside the leaf that holds it reaches it: leaves of the
// depth-3 tree (64 x 64 cells) are nearly flat, so the distance restricted to one is close to
// convex.  Projected Newton (gn_step with ACTIVE) holds a coordinate on its bound while the
// descent points outward, so minima on cell edges and on the surface boundary are reached too.
// FP64, one lane per query, exact-Hessian steps.  Also the fallback of gn_invert (opts.fallback):
// the queries an inversion left unconverged are listed and solved again this way.
#include "gn_search.hpp"
#include "invert_common.hpp"

namespace gn {

constexpr int kLeafIter = 8;  // projected Newton steps inside one leaf cell
constexpr int kCand = 8;      // leaves collected per lane before the wave solves them

// Whether a leaf can hold an interior foot point of P: a point X of its box b with P - X parallel
// to a normal n of its cone (unit axis a, half-angle theta < 90 degrees), i.e. whether the box meets
// the double cone {X : ((X - P) . a)^2 >= cos^2(theta) |X - P|^2} with apex P.  The angle between
// X - P and the line of a has, over a box that the line misses, its minimum on an edge (inside the
// box or a face it could only have one where the line meets it), so the test is: P in the box, or
// the line P + t a meets the box, or some edge has a point with q(s) >= 0, q being that quadratic
// along the edge.  Exact up to rounding; FP64 on the FP32 box's bounds, with the half-angle widened
// by another 2e-6 rad.
__device__ __forceinline__ bool cone_may_hold(const float4 cn, const float *__restrict__ b,
                                              const double P[3]) {
    if (!(cn.w < 1.5f)) return true;  // no bounded cone
    const double a[3] = {cn.x, cn.y, cn.z};
    const double lo[3] = {b[0], b[1], b[2]}, hi[3] = {b[3], b[4], b[5]};
    if (!(P[0] == P[0] && P[1] == P[1] && P[2] == P[2])) return true;
    // the line P + t a against the slabs
    double t0 = -INFINITY, t1 = INFINITY;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        if (a[k] != 0.0) {
            const double r = 1.0 / a[k];
            const double ta = (lo[k] - P[k]) * r, tb = (hi[k] - P[k]) * r;
            t0 = fmax(t0, fmin(ta, tb));
            t1 = fmin(t1, fmax(ta, tb));
        } else if (P[k] < lo[k] || P[k] > hi[k]) {
            t0 = INFINITY;  // parallel to this slab and outside it
        }
    }
    if (t0 <= t1) return true;  // also P inside the box (t = 0)
    const double c = cos((double)cn.w + 2e-6), c2 = c * c;
    // the 12 edges: along axis k, the other two coordinates at their bounds
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const int k1 = k == 0 ? 1 : 0, k2 = k == 2 ? 1 : 2;
        const double d[3] = {k == 0 ? hi[0] - lo[0] : 0.0, k == 1 ? hi[1] - lo[1] : 0.0,
                             k == 2 ? hi[2] - lo[2] : 0.0};
        const double be = d[k] * a[k], ee = d[k] * d[k];
#pragma unroll
        for (int e = 0; e < 4; e++) {
            double w[3];
            w[k] = lo[k] - P[k];
            w[k1] = ((e & 1) ? hi[k1] : lo[k1]) - P[k1];
            w[k2] = ((e & 2) ? hi[k2] : lo[k2]) - P[k2];
            const double al = w[0] * a[0] + w[1] * a[1] + w[2] * a[2];
            const double ga = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
            const double de = w[k] * d[k];
            // q(s) = A s^2 + 2 B s + C on s in [0, 1]
            const double A = be * be - c2 * ee, B = al * be - c2 * de, C = al * al - c2 * ga;
            if (C >= 0.0 || A + 2.0 * B + C >= 0.0) return true;
            if (A < 0.0 && B > 0.0 && B < -A && C - B * B / A >= 0.0) return true;  // vertex -B/A
        }
    }
    return false;
}

// Projected Newton from (cu, cv) inside dom, with backtracking: a step after which the squared
// distance grew by more than 1e-6 relative is halved, up to kHalvings times in a row, before a new
// step is taken.  (Near a minimum the squared distance carries rounding of the surface point
// relative to the residual, ~1e-12 of it at a 1e-3 offset; smaller growth is not a failed step.)
// Returns the status and leaves in (cu, cv), d2 the nearest point evaluated and its squared
// distance.  Inlined at its three call sites: as a called function its live values spilled
// across the call (256 VGPRs + 224 bytes of scratch per lane, one wave per SIMD); inlined the
// bicubic kernel takes 233 VGPRs, no scratch, two waves per SIMD (1M queries: 25.8 -> 7.7 ms).
constexpr int kHalvings = 8;
template <int K1, int K2, bool EXACT>
__device__ __forceinline__ int solve_in(const GnSurfDev &d, const double *P, double &cu,
                                        double &cv, const Domain &dom, int iters,
                                        double tol_step, double &d2) {
    SpanCache sc;
    int st = GN_PT_MAXITER, back = 0;
    double pu = cu, pv = cv, prr = INFINITY;  // the last accepted point
    for (int it = 0; it < iters; it++) {
        const double su = cu, sv = cv;
        double rr = INFINITY;
        const int s = gn_step<K1, K2, EXACT, double, true, kPatchGlobal, true>(
            d, d.ku, d.kv, P[0], P[1], P[2], cu, cv, 0.0, tol_step, 0.0, 1, dom, sc, nullptr,
            nullptr, &rr);
        if (!(rr <= prr * (1.0 + 1e-6)) && back < kHalvings) {  // grew, beyond rounding
            cu = pu + 0.5 * (su - pu);
            cv = pv + 0.5 * (sv - pv);
            back++;
            continue;
        }
        back = 0;
        pu = su;
        pv = sv;
        prr = rr;
        if (s != kStepContinue) {
            st = s;
            break;
        }
    }
    Cart<double> o;
    eval_point<K1, K2, 0, EXACT, double>(d, d.ku, d.kv, cu, cv, o);
    const double dx = o.S[0] - P[0], dy = o.S[1] - P[1], dz = o.S[2] - P[2];
    d2 = fma(dz, dz, fma(dy, dy, dx * dx));
    if (!(d2 <= prr)) {  // the final step did not help: keep the last accepted point
        cu = pu;
        cv = pv;
        d2 = prr;
    }
    return st;
}

// The branch and bound for query i.  Global mode (counts): also writes the number of leaves
// solved to it32.  Fallback (!counts): the (u, v) already in place is kept when it is nearer.
// cand: the workgroup's LDS candidate slots, kCand per thread (blocks of 256 threads).
template <int K1, int K2, bool EXACT>
__device__ __forceinline__ void project_point(const GnSurfDev &d, const GnSearchDev &sd,
                                              const InvArgs &a, size_t i, bool counts,
                                              uint32_t *cand) {
    const Domain full = domain_of(d);
    const int D = sd.depth;
    // leaf cells: the unit square's 2^D x 2^D cells mapped onto the knot domain, exactly as the
    // tree lays out its nodes' uv rectangles
    const double cell = 1.0 / (double)(1u << D);
    const double wu = full.u1 - full.u0, wv = full.v1 - full.v0;
    const int max_iter = a.o.max_iter;
    const double tol_step = a.o.tol_step;
    const double P[3] = {a.x[i], a.y[i], a.z[i]};
    // the bound solve starts from the greedy leaf's stepped seed (step_seed: one Gauss-Newton
    // step and two on the leaf's quadratic model), a few Newton iterations nearer than its centre
    float4 cq;
    const uint32_t c0 = greedy_leaf(sd, P, nullptr, 0, &cq);
    double bu, bv, best;
    step_seed(sd, c0, cq, P, bu, bv);
    int bst = solve_in<K1, K2, EXACT>(d, P, bu, bv, full, max_iter, tol_step, best);
    int ncand = 1;
    // The bound solve stands for leaf c0's own solve only if it ended inside c0's cell; a solve
    // that wandered off may have passed a nearer minimum in the cell, so c0 then stays a candidate.
    bool c0_done;
    {
        const uint32_t ix = compact2(c0 >> 1), iy = compact2(c0);
        c0_done = bu >= full.u0 + wu * (ix * cell) && bu <= full.u0 + wu * ((ix + 1) * cell) &&
                  bv >= full.v0 + wv * (iy * cell) && bv <= full.v0 + wv * ((iy + 1) * cell);
    }
    if (D > 0) {
        const float p[3] = {(float)P[0], (float)P[1], (float)P[2]};
        // stackless depth-first walk (Morton-code successor), pruned by the current best.  The
        // walk collects up to kCand leaves at a time into this lane's LDS slots, then the wave
        // solves them in rounds, each lane its r-th leaf: lanes walk different paths, but they
        // solve together.  A collected leaf is solved only if its box is still not farther than
        // the best so far; a leaf box is never nearer than its ancestors' boxes, so exactly the
        // leaves of a walk that solved each leaf as it reached it are solved, in the same order.
        const uint32_t side = 1u << D;
        uint32_t *cq = cand + threadIdx.x * kCand;
        int l = 1;
        uint32_t c = 0;
        bool walking = true;
        while (__any(walking)) {  // wave-uniform
            int nq = 0;
            while (walking && nq < kCand) {
                const double lb = (double)box_d2(sd.box32 + (qoff(l) + c) * 6, p);
                bool descend = false;
                if (lb <= best) {
                    if (l < D) {
                        descend = true;
                    } else if ((c != c0 || !c0_done) &&
                               (!sd.cone32 || compact2(c >> 1) - 1u >= side - 2u ||
                                compact2(c) - 1u >= side - 2u ||
                                cone_may_hold(sd.cone32[c], sd.box32 + (qoff(l) + c) * 6, P))) {
                        // (the two index tests pass the leaves on the domain boundary, ix or iy
                        // 0 or side - 1, by unsigned wrap-around)
                        cq[nq++] = c;
                    }
                }
                if (descend) {
                    c <<= 2;
                    l++;
                    continue;
                }
                while (l > 1 && (c & 3u) == 3u) {
                    c >>= 2;
                    l--;
                }
                if (l == 1 && (c & 3u) == 3u) {
                    walking = false;
                    break;
                }
                c++;
            }
            for (int r = 0; __any(r < nq); r++) {  // wave-uniform rounds
                if (r >= nq) continue;
                const uint32_t lc = cq[r];
                if (!((double)box_d2(sd.box32 + (qoff(D) + lc) * 6, p) <= best)) continue;
                const uint32_t ix = compact2(lc >> 1), iy = compact2(lc);
                const Domain cr{full.u0 + wu * (ix * cell), full.u0 + wu * ((ix + 1) * cell),
                                full.v0 + wv * (iy * cell), full.v0 + wv * ((iy + 1) * cell)};
                // from the leaf's stepped seed for P (step_seed from its record), clamped into
                // its cell: fewer Newton steps than from the cell centre
                double cu, cv, d2;
                step_seed(sd, lc, sd.rec[(size_t)lc * kRecVec], P, cu, cv);
                cu = fmin(fmax(cu, cr.u0), cr.u1);
                cv = fmin(fmax(cv, cr.v0), cr.v1);
                const int s2 = solve_in<K1, K2, EXACT>(d, P, cu, cv, cr, kLeafIter, tol_step, d2);
                ncand++;
                if (d2 < best) {
                    best = d2;
                    bu = cu;
                    bv = cv;
                    bst = s2;
                }
            }
        }
    }
    double pu = bu, pv = bv, pd2;
    const int pst = solve_in<K1, K2, EXACT>(d, P, pu, pv, full, max_iter, tol_step, pd2);
    if (pd2 <= best * (1.0 + 1e-9)) {  // not farther beyond rounding: the polished point
        bu = pu;
        bv = pv;
        bst = pst;
        best = pd2;
    }
    if (!counts) {  // fallback: the earlier result stays if it is nearer
        Cart<double> o;
        eval_point<K1, K2, 0, EXACT, double>(d, d.ku, d.kv, a.u[i], a.v[i], o);
        const double dx = o.S[0] - P[0], dy = o.S[1] - P[1], dz = o.S[2] - P[2];
        if (fma(dz, dz, fma(dy, dy, dx * dx)) < best) return;
    }
    a.u[i] = bu;
    a.v[i] = bv;
    a.status[i] = (uint8_t)bst;
    if (counts) {
        if (a.it32) a.it32[i] = sat8(ncand);
        if (a.it64) a.it64[i] = 0;
    }
}

// Two waves per SIMD are forced where the spills that costs stay small: patches of up to 20
// entries (bicubic: 12 bytes per lane; one wave without the bound, 13.3 against 8.2 ms on 1M
// projection queries; degree (3,4): 160 bytes, 14.3 against 9.1 ms) and the generic kernel (80
// bytes, 12.9 against 7.2 ms).  Larger exact patches run one wave with AGPRs as spill room:
// forced to two, degree (5,5) spilled 744 bytes per lane and took 24.5 against 18.3 ms.
template <int K1, int K2, bool EXACT>
constexpr int project_waves() { return (!EXACT || K1 * K2 <= 20) ? 2 : 1; }

// order (may be null): the queries in the order to take them, so that a wave's lanes hold
// queries near one another in space and walk much the same part of the tree (k_proj_order).
template <int K1, int K2, bool EXACT>
__global__ __launch_bounds__(256, (project_waves<K1, K2, EXACT>())) void k_project_global(
    GnSurfDev d, GnSearchDev sd, InvArgs a, const uint32_t *__restrict__ order) {
    __shared__ uint32_t cand[256 * kCand];
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.n;
         i += (size_t)gridDim.x * blockDim.x)
        project_point<K1, K2, EXACT>(d, sd, a, order ? (size_t)order[i] : i, true, cand);
}

// Spatial order of the queries: the key is the 16 x 16 x 16 cell of P within the tree's root box
// (Morton order, 12 bits; points outside the box take its border cells, NaN coordinates cell 0).
// One workgroup per block of kOrdBlock consecutive queries counts its keys in LDS, scans them and
// writes each query's index to its key's next slot within the block's own range of `order`.  The
// order within one key is unspecified; each query's result does not depend on it (project_point
// is per query).
constexpr int kOrdThreads = 1024;
constexpr size_t kOrdBlock = 65536;
constexpr int kOrdKeys = 4096;

__device__ __forceinline__ uint32_t proj_key(const float *__restrict__ root, const double P[3]) {
    uint32_t q[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const float lo = root[k], w = root[3 + k] - root[k];
        const float t = w > 0.f ? ((float)P[k] - lo) * (16.f / w) : 0.f;
        q[k] = (uint32_t)fminf(fmaxf(t, 0.f), 15.f);
    }
    uint32_t m = 0;
#pragma unroll
    for (int b = 0; b < 4; b++)
        m |= ((q[0] >> b & 1u) << (3 * b + 2)) | ((q[1] >> b & 1u) << (3 * b + 1)) |
             ((q[2] >> b & 1u) << (3 * b));
    return m;
}

__global__ __launch_bounds__(kOrdThreads) void k_proj_order(const float *__restrict__ root,
                                                            const double *__restrict__ x,
                                                            const double *__restrict__ y,
                                                            const double *__restrict__ z, size_t n,
                                                            uint32_t *__restrict__ order) {
    __shared__ uint32_t cnt[kOrdKeys];
    __shared__ uint32_t part[kOrdThreads];
    for (int k = threadIdx.x; k < kOrdKeys; k += kOrdThreads) cnt[k] = 0u;
    __syncthreads();
    const size_t i0 = (size_t)blockIdx.x * kOrdBlock;
    const size_t i1 = i0 + kOrdBlock < n ? i0 + kOrdBlock : n;
    for (size_t i = i0 + threadIdx.x; i < i1; i += kOrdThreads) {
        const double P[3] = {x[i], y[i], z[i]};
        atomicAdd(&cnt[proj_key(root, P)], 1u);
    }
    __syncthreads();
    // exclusive scan: each thread a run of kOrdKeys / kOrdThreads keys, then the runs' sums
    constexpr int per = kOrdKeys / kOrdThreads;
    const int k0 = threadIdx.x * per;
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < per; k++) sum += cnt[k0 + k];
    part[threadIdx.x] = sum;
    __syncthreads();
    for (int o = 1; o < kOrdThreads; o <<= 1) {  // inclusive Hillis-Steele scan of the sums
        const uint32_t v = threadIdx.x >= (unsigned)o ? part[threadIdx.x - o] : 0u;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = (uint32_t)i0 + (threadIdx.x ? part[threadIdx.x - 1] : 0u);
#pragma unroll
    for (int k = 0; k < per; k++) {
        const uint32_t c = cnt[k0 + k];
        cnt[k0 + k] = run;
        run += c;
    }
    __syncthreads();
    for (size_t i = i0 + threadIdx.x; i < i1; i += kOrdThreads) {
        const double P[3] = {x[i], y[i], z[i]};
        order[atomicAdd(&cnt[proj_key(root, P)], 1u)] = (uint32_t)i;
    }
}

// The fallback: the queries listed in list[0 .. *count) (iteration counts left as they were).
template <int K1, int K2, bool EXACT>
__global__ __launch_bounds__(256) void k_project_list(GnSurfDev d, GnSearchDev sd, InvArgs a,
                                                      const uint32_t *__restrict__ list,
                                                      const uint32_t *__restrict__ count) {
    __shared__ uint32_t cand[256 * kCand];
    const size_t cnt = *count;
    for (size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x; j < cnt;
         j += (size_t)gridDim.x * blockDim.x)
        project_point<K1, K2, EXACT>(d, sd, a, list[j], false, cand);
}

__global__ void k_fb_zero(uint32_t *count) { *count = 0u; }

// indices of the queries whose status is not converged
__global__ void k_fb_compact(const uint8_t *__restrict__ status, size_t n, uint32_t *__restrict__ list,
                             uint32_t *__restrict__ count) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        if (status[i] != GN_PT_CONVERGED && status[i] != GN_PT_CONVERGED_DIST)
            list[atomicAdd(count, 1u)] = (uint32_t)i;
}

struct ProjectLaunch {
    const GnSurfDev &d;
    const GnSearchDev &sd;
    const InvArgs &a;
    const uint32_t *order;
    hipStream_t st;
    template <int K1, int K2, bool EXACT>
    int operator()() {
        k_project_global<K1, K2, EXACT><<<grid_for(a.n, 256, 256 * 16), 256, 0, st>>>(d, sd, a,
                                                                                    order);
        GN_HIP(hipGetLastError());
        return GN_OK;
    }
};

// Calls of at least kOrdMin queries (fewer than 2^32) take them in spatial order: a stream-ordered
// scratch index array of n words, filled by k_proj_order.  Calls that must not allocate (the
// no-allocation entry point) take them in input order.
constexpr size_t kOrdMin = 16384;

int launch_project_global(const GnSurfDev &d, const GnSearchDev &sd, const InvArgs &a,
                          hipStream_t st) {
    if (a.n == 0) return GN_OK;
    uint32_t *order = nullptr;
    if (a.n >= kOrdMin && a.n < 0xffffffffull && sd.depth > 0 && !slot_pool_active()) {
        GN_HIP(hipMallocAsync((void **)&order, sizeof(uint32_t) * a.n, st));
        k_proj_order<<<(unsigned)((a.n + kOrdBlock - 1) / kOrdBlock), kOrdThreads, 0, st>>>(
            sd.box32, a.x, a.y, a.z, a.n, order);
        if (hipGetLastError() != hipSuccess) {
            (void)hipFreeAsync(order, st);
            return fail(GN_EDEVICE, "query ordering launch");
        }
    }
    const int rc = dispatch_degrees(d.p, d.q, ProjectLaunch{d, sd, a, order, st});
    if (order) GN_HIP(hipFreeAsync(order, st));
    return rc;
}

struct ListLaunch {
    const GnSurfDev &d;
    const GnSearchDev &sd;
    const InvArgs &a;
    const uint32_t *list, *count;
    hipStream_t st;
    template <int K1, int K2, bool EXACT>
    int operator()() {
        k_project_list<K1, K2, EXACT><<<grid_for(a.n, 256, 256 * 16), 256, 0, st>>>(d, sd, a, list,
                                                                                 count);
        GN_HIP(hipGetLastError());
        return GN_OK;
    }
};

// After an inversion on st: the unconverged queries are listed on the device (stream-ordered
// scratch of n + 1 words) and solved again by the global projection.
int launch_project_fallback(const GnSurfDev &d, const GnSearchDev &sd, const InvArgs &a,
                            hipStream_t st) {
    if (a.n == 0) return GN_OK;
    if (a.n > 0xffffffffull) return fail(GN_EINVAL, "the fallback lists at most 2^32 - 1 queries");
    uint32_t *buf = nullptr;
    GN_HIP(hipMallocAsync((void **)&buf, sizeof(uint32_t) * (a.n + 1), st));
    uint32_t *count = buf, *list = buf + 1;
    k_fb_zero<<<1, 1, 0, st>>>(count);
    k_fb_compact<<<grid_for(a.n, 256, 256 * 16), 256, 0, st>>>(a.status, a.n, list, count);
    int rc = hipGetLastError() == hipSuccess ? GN_OK : fail(GN_EDEVICE, "fallback compaction launch");
    if (!rc) rc = dispatch_degrees(d.p, d.q, ListLaunch{d, sd, a, list, count, st});
    GN_HIP(hipFreeAsync(buf, st));
    return rc;
}

}  // namespace gn
