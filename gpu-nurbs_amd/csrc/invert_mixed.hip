// invert_mixed.hip -- mixed-precision inversion: FP32 bulk iterations against the FP32 span
// table, then FP64 closing iterations whose residual, rational weight division and 2x2 solve are
// FP64.  The bulk is a persistent lane-refill kernel; the closing pass streams every point's
// first step and hands the few points that continue to a lane-refill pass over a list.
#include "invert_common.hpp"

// Points per wave of the list pass's grid (the list's length is known only on the device).
#ifndef GN_LIST_WAVE_POINTS
#define GN_LIST_WAVE_POINTS 1024
#endif

namespace gn {

// The first FP64 closing step of every point, streamed: almost every point stops after it (the
// bulk result needs one exact-Hessian step to certify), so each wave takes 64 consecutive points
// at a time, with the next tile's inputs loaded while this one steps, and the whole wave stages
// the patch rows through LDS with the next row in flight.  A point that stops is written out as
// the refill closing pass writes it; a point that continues gets its (u, v) written back and its
// index appended to the wave's segment of `list` (wave w uses segment w % kMaxSlices, one atomic
// per wave and tile on that segment's count), for k_gn_refill<kCloseList>.  Past a segment's
// kListSegCap entries a lane finishes its point itself, with per-lane patch loads (same
// arithmetic).  Indices are 32-bit: the launcher takes this path only for n < 2^32.
// Two waves per SIMD are asked for patches of up to 16 entries: the bicubic Newton kernel sits
// at the 248-register edge of that occupancy and otherwise falls to one wave (74.6 against
// 54.9 us on 1M projection queries).
template <int K1, int K2>
constexpr int close_waves() { return K1 * K2 <= 16 ? 2 : 1; }

template <int K1, int K2, bool NEWTON>
__global__ __launch_bounds__(kInvBlock, (close_waves<K1, K2>())) void k_close_first(GnSurfDev d_in, InvArgs a,
                                                           uint32_t *list,
                                                           unsigned long long *counters,
                                                           int max_iter, int restart_iter,
                                                           double tol_dist, double tol_step,
                                                           double tol_quad) {
    using Rows = WaveRowsPipe<K1, K2, double>;
    __shared__ double sk[2 * kLdsKnotsInv];
    __shared__ float4 wrows[kInvBlock / 64][Rows::STRIDE * 64];
    const GnSurfDev d = stage_knots_lds(d_in, sk);
    const double *ku = sk, *kv = sk + d.nu + d.p + 1;
    const Domain dom = domain_of(d);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const unsigned long long lt = (1ull << lane) - 1ull;
    float4 *wave_lds = wrows[wv];
    const int seg = (int)(((size_t)blockIdx.x * (kInvBlock / 64) + wv) % kMaxSlices);
    unsigned long long *seg_n = counters + (size_t)seg * kCounterStride + kListWord;
    uint32_t *seg_list = list + (size_t)seg * kListSegCap;
    const size_t n = a.n, stride = (size_t)gridDim.x * kInvBlock;
    size_t base = ((size_t)blockIdx.x * (kInvBlock / 64) + wv) * 64;
    double px = 0, py = 0, pz = 0, pu = 0, pv = 0;
    int ps = 0;
    if (base + lane < n) {
        const size_t i = base + lane;
        px = a.x[i];
        py = a.y[i];
        pz = a.z[i];
        pu = a.u[i];
        pv = a.v[i];
        ps = a.status[i];
    }
    for (; base < n; base += stride) {
        const size_t i = base + lane;
        const bool live = i < n;
        const double Px = px, Py = py, Pz = pz, cu = pu, cv = pv;
        const int sb = ps;
        if (i + stride < n) {
            px = a.x[i + stride];
            py = a.y[i + stride];
            pz = a.z[i + stride];
            pu = a.u[i + stride];
            pv = a.v[i + stride];
            ps = a.status[i + stride];
        }
        const int lim = (sb == GN_PT_SINGULAR || sb == GN_PT_NONFINITE) ? restart_iter : max_iter;
        const bool act = live && lim > 0;
        double tu = act ? cu : dom.u0, tv = act ? cv : dom.v0;
        SpanCache sc;
        const int s2 = gn_step<K1, K2, true, double, NEWTON, kPatchWavePipe>(
            d, ku, kv, Px, Py, Pz, tu, tv, tol_dist, tol_step, tol_quad, a.o.clamp, dom, sc,
            nullptr, wave_lds);
        int st = GN_PT_MAXITER, it = 0;
        if (act) {
            st = s2;
            if (st == kStepContinue || st == GN_PT_CONVERGED) it = 1;
            if (st == kStepContinue && it >= lim) st = GN_PT_MAXITER;
        } else {
            tu = cu;
            tv = cv;
        }
        const bool cont = live && st == kStepContinue;
        const unsigned long long m = __ballot(cont);
        unsigned long long pos = 0;
        if (m) {
            unsigned long long b = 0;
            if (lane == 0) b = atomicAdd(seg_n, (unsigned long long)__popcll(m));
            pos = bcast64(b, 0) + (unsigned long long)__popcll(m & lt);
        }
        if (cont && pos < kListSegCap) {
            seg_list[pos] = (uint32_t)i;
        } else if (cont) {  // segment full: finish here
            SpanCache sc2;
            do {
                st = gn_step<K1, K2, true, double, NEWTON, kPatchGlobal>(
                    d, ku, kv, Px, Py, Pz, tu, tv, tol_dist, tol_step, tol_quad, a.o.clamp, dom,
                    sc2);
                if (st == kStepContinue || st == GN_PT_CONVERGED) it++;
                if (st == kStepContinue && it >= lim) st = GN_PT_MAXITER;
            } while (st == kStepContinue);
        }
        if (live) {
            a.u[i] = tu;
            a.v[i] = tv;
            if (st != kStepContinue) {
                a.status[i] = (uint8_t)st;
                if (a.it64) a.it64[i] = sat8(it);
            }
        }
    }
}

struct MixedLaunch {
    const GnSurfDev &d;
    const InvArgs &a;
    hipStream_t st;
    template <int K1, int K2, bool EXACT>
    int operator()() {
        unsigned long long *ctr = nullptr;
        int rc = work_counters(st, &ctr);
        if (rc) return rc;
        unsigned grid;
        int slices;
        if (a.o.newton & GN_NEWTON_FP32) {
            auto k = k_gn_refill<K1, K2, EXACT, float, true, kBulk>;
            refill_shape((const void *)k, a.n, grid, slices);
            k<<<grid, kInvBlock, 0, st>>>(d, a, ctr, slices, a.o.fp32_iters, 0, 0.0,
                                          a.o.tol_step32, 0.0);
        } else {
            auto k = k_gn_refill<K1, K2, EXACT, float, false, kBulk>;
            refill_shape((const void *)k, a.n, grid, slices);
            k<<<grid, kInvBlock, 0, st>>>(d, a, ctr, slices, a.o.fp32_iters, 0, 0.0,
                                          a.o.tol_step32, 0.0);
        }
        if ((rc = work_counters(st, &ctr))) return rc;
        if constexpr (EXACT) {  // streamed first closing step, then the listed rest
            if (a.n < ((size_t)1 << 32)) {
                uint32_t *list = nullptr;
                if ((rc = work_list(ctr, &list))) return rc;
                const bool nt = a.o.newton & GN_NEWTON_FP64;
                const double tq = nt ? a.o.tol_quad : 0.0;
                auto kf = nt ? k_close_first<K1, K2, true> : k_close_first<K1, K2, false>;
                const size_t cap = (size_t)cu_count() * blocks_per_cu((const void *)kf, kInvBlock);
                const size_t want = (a.n + kInvBlock - 1) / kInvBlock;
                const unsigned g1 = (unsigned)(want < cap ? (want ? want : 1) : cap);
                kf<<<g1, kInvBlock, 0, st>>>(d, a, list, ctr, a.o.fp64_tail, a.o.max_iter,
                                             a.o.tol_dist, a.o.tol_step, tq);
                InvArgs al = a;
                al.list = list;
                auto k = nt ? k_gn_refill<K1, K2, EXACT, double, true, kCloseList>
                            : k_gn_refill<K1, K2, EXACT, double, false, kCloseList>;
                // one slice per list segment; about 2% of the points continue: one wave per
                // 1,024 points (at least one per segment; list pass at 1M points: 34 us with
                // the full resident grid and one exit counter, 21 us at one wave per 4,096
                // points, 16 us at one per 1,024 or 512)
                refill_shape((const void *)k, a.n, grid, slices);
                slices = kMaxSlices;
                const size_t gl = a.n / ((size_t)GN_LIST_WAVE_POINTS * (kInvBlock / 64));
                const unsigned gmin = kMaxSlices / (kInvBlock / 64);
                if (gl < grid) grid = (unsigned)gl;
                if (grid < gmin) grid = gmin;
                k<<<grid, kInvBlock, 0, st>>>(d, al, ctr, slices, a.o.fp64_tail, a.o.max_iter,
                                              a.o.tol_dist, a.o.tol_step, tq);
                GN_HIP(hipGetLastError());
                return GN_OK;
            }
        }
        if (a.o.newton & GN_NEWTON_FP64) {
            auto k = k_gn_refill<K1, K2, EXACT, double, true, kClose>;
            refill_shape((const void *)k, a.n, grid, slices);
            k<<<grid, kInvBlock, 0, st>>>(d, a, ctr, slices, a.o.fp64_tail, a.o.max_iter,
                                          a.o.tol_dist, a.o.tol_step, a.o.tol_quad);
        } else {
            auto k = k_gn_refill<K1, K2, EXACT, double, false, kClose>;
            refill_shape((const void *)k, a.n, grid, slices);
            k<<<grid, kInvBlock, 0, st>>>(d, a, ctr, slices, a.o.fp64_tail, a.o.max_iter,
                                          a.o.tol_dist, a.o.tol_step, 0.0);
        }
        GN_HIP(hipGetLastError());
        return GN_OK;
    }
};

int launch_gn_mixed(const GnSurfDev &d, const InvArgs &a, hipStream_t st) {
    return dispatch_inv(d, MixedLaunch{d, a, st});
}

}  // namespace gn
