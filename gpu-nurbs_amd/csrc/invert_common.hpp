// invert_common.hpp -- Gauss-Newton step and the persistent lane-refill driver shared by the
// inversion kernels (GaussNewton, src/kernel/kernel.cu:151-249).
#pragma once
#include "gn_device.hpp"

namespace gn {

constexpr int kInvBlock = 256;
constexpr int kLdsKnotsInv = 1024;
constexpr int kChunk = 64;       // points claimed per atomic by one wave
constexpr int kStepContinue = -1;
constexpr int kMaxSlices = 64;   // claim counters per launch
constexpr int kCounterStride = 16;  // 16 x 8 bytes = one 128-byte line per counter

// kMaxSlices zeroed counters (kCounterStride apart) for one launch on stream st.  The kernel
// that uses them must end every wave with retire_wave, which leaves them zeroed again.
int work_counters(hipStream_t st, unsigned long long **out);
constexpr int kExitWord = 8;  // exit counter: word 8 of slice 0's line (slices use word 0)
// The continuation list paired with a counter slot, for a streamed pass that hands the points it
// has not finished to a lane-refill pass (kCloseList): kMaxSlices segments of kListSegCap point
// indices, segment s filled through the count in word kListWord of slice s's line (appends to
// one address would serialise), which retire_wave clears with the claim counters.
constexpr int kListWord = 9;
constexpr size_t kListSegCap = 2048;
constexpr size_t kListCap = kListSegCap * kMaxSlices;
int work_list(const unsigned long long *counters, uint32_t **list);
int cu_count();

struct InvArgs {
    const double *x, *y, *z;
    size_t n;
    double *u, *v;
    uint8_t *status;
    uint8_t *it32, *it64;  // optional per-point iteration counts (FP32 bulk, FP64), or null
    gn_invert_opts o;
    const uint32_t *list = nullptr;  // kCloseList: the points to continue (kListCap segments)
};

// Lane 0 of every wave calls this once, after the wave's last claim.  A wave has consumed the
// value returned by each of its claim atomics before it gets here, so its claims are performed
// before it counts itself out; the counts themselves need no release ordering (an agent-scope
// release would write back the XCD's L2 once per wave).  Exits are counted in two levels: wave w
// counts into the exit word of slice w % slices (its own 128-byte line), and the last wave of a
// slice counts the slice into kExitWord -- one address taking an add from every wave serialises
// at ~12 ns each, which a grid of thousands of waves retiring together would wait out.  The last
// slice to finish takes an acquire fence and zeroes the claim counters, the list counts and the
// exit counts, which readies the slot for its next launch without a separate clearing kernel (and
// keeps captured graphs replayable).  Needs slices <= waves (every slice has a wave).
constexpr int kSliceExitWord = 10;
__device__ __forceinline__ void retire_wave(unsigned long long *counters, int slices) {
    const unsigned waves = gridDim.x * (blockDim.x / 64);  // 32-bit: a grid is far below 2^32 waves
    const unsigned w = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const unsigned ns = (unsigned)slices, s = w % ns;
    const unsigned long long in_slice = waves / ns + (s < waves % ns ? 1u : 0u);
    const unsigned long long prev = __hip_atomic_fetch_add(
        counters + (size_t)s * kCounterStride + kSliceExitWord, 1ull, __ATOMIC_RELAXED,
        __HIP_MEMORY_SCOPE_AGENT);
    if (prev + 1 != in_slice) return;
    const unsigned long long done = __hip_atomic_fetch_add(counters + kExitWord, 1ull, __ATOMIC_RELAXED,
                                                           __HIP_MEMORY_SCOPE_AGENT);
    if (done + 1 != (unsigned long long)slices) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    for (int k = 0; k < slices; k++) {
        __hip_atomic_store(counters + (size_t)k * kCounterStride, 0ull, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(counters + (size_t)k * kCounterStride + kSliceExitWord, 0ull,
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    for (int k = 0; k < kMaxSlices; k++)
        __hip_atomic_store(counters + (size_t)k * kCounterStride + kListWord, 0ull, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(counters + kExitWord, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint8_t sat8(int x) { return (uint8_t)(x > 255 ? 255 : x); }

// Knots of the exact-degree inversion kernels live in LDS (the launchers route surfaces whose
// knots do not fit to the generic kernel, which reads them from global memory), so the span
// search and the span cache run on LDS loads.
inline bool knots_fit_lds(const GnSurfDev &d) {
    return d.nu + d.p + 1 + d.nv + d.q + 1 <= kLdsKnotsInv;
}

// Knots, then the reciprocal span lengths, into LDS (2 * kLdsKnotsInv doubles hold both when
// knots_fit_lds: a surface has fewer spans than knots in each direction).  Returns the surface
// view whose span lengths are read from LDS.
__device__ __forceinline__ GnSurfDev stage_knots_lds(const GnSurfDev &d, double *sk) {
    const int nku = d.nu + d.p + 1, nkv = d.nv + d.q + 1, nsu = d.nu - d.p, nsv = d.nv - d.q;
    for (int i = threadIdx.x; i < nku; i += blockDim.x) sk[i] = d.ku[i];
    for (int i = threadIdx.x; i < nkv; i += blockDim.x) sk[nku + i] = d.kv[i];
    for (int i = threadIdx.x; i < nsu; i += blockDim.x) sk[nku + nkv + i] = d.ihu[i];
    for (int i = threadIdx.x; i < nsv; i += blockDim.x) sk[nku + nkv + nsu + i] = d.ihv[i];
    __syncthreads();
    GnSurfDev dl = d;
    dl.ihu = sk + nku + nkv;
    dl.ihv = sk + nku + nkv + nsu;
    return dl;
}

// The degree dispatch of the inversion launchers: exact instantiations when the surface's knots
// fit in LDS, the generic (global-knot) kernel otherwise.
template <typename F>
__host__ int dispatch_inv(const GnSurfDev &d, F &&f) {
    if (!knots_fit_lds(d)) return f.template operator()<GN_MAXK, GN_MAXK, false>();
    return dispatch_degrees(d.p, d.q, f);
}

// One Gauss-Newton update in arithmetic T on the parameter state (cu, cv) (always FP64).
// Returns kStepContinue, or the point's final status: GN_PT_CONVERGED_DIST (|S-P| < tol_dist,
// checked before the update as the reference does), GN_PT_CONVERGED (step below tol_step after
// the update), GN_PT_SINGULAR (det(J^T J) == 0) or GN_PT_NONFINITE (a NaN or infinite residual,
// e.g. a query with a non-finite coordinate, or a non-finite update).
// With NEWTON the residual-curvature term sum_k r_k d2S_k is added to J^T J when the sum stays
// positive definite (exact Hessian of |S - P|^2 / 2), otherwise the plain J^T J is used.
// Parameter domain [u0, u1] x [v0, v1] of a surface: k[p], k[nu] and k[q], k[nv].
struct Domain {
    double u0, u1, v0, v1;
};

__device__ __forceinline__ Domain domain_of(const GnSurfDev &d) {
    return Domain{d.ku[d.p], d.ku[d.nu], d.kv[d.q], d.kv[d.nv]};
}

// ACTIVE: projected Newton on the clamp rectangle -- a coordinate on its bound whose descent
// direction points outward is held and the other takes its one-dimensional step, with the exact
// one-dimensional second derivative where that is positive (a plain clamp of the coupled step can
// walk uphill along an edge).  rr_out: the squared residual at the point evaluated.
template <int K1, int K2, bool EXACT, typename T, bool NEWTON, int PSRC = kPatchGlobal,
          bool ACTIVE = false, bool NEAR = false>
__device__ __forceinline__ int gn_step(const GnSurfDev &d, const double *__restrict__ ku,
                                       const double *__restrict__ kv, T Px, T Py, T Pz,
                                       double &cu, double &cv, double tol_dist, double tol_step,
                                       double tol_quad, int clamp, const Domain &dom,
                                       SpanCache &sc, Hom<T> *pr = nullptr,
                                       float4 *wave_lds = nullptr, T *rr_out = nullptr) {
    Cart<T> o;
    eval_point_cached<K1, K2, NEWTON ? 2 : 1, EXACT, T, PSRC, NEAR>(d, ku, kv, cu, cv, sc, o, pr,
                                                              wave_lds);
    const T r0 = o.S[0] - Px, r1 = o.S[1] - Py, r2 = o.S[2] - Pz;
    const T rr = r0 * r0 + r1 * r1 + r2 * r2;
    if (rr_out) *rr_out = rr;
    if (!(rr < T(INFINITY))) return GN_PT_NONFINITE;  // NaN or infinite query (or residual)
    if (tol_dist > 0 && sqrt(rr) < T(tol_dist)) return GN_PT_CONVERGED_DIST;
    T a = o.Su[0] * o.Su[0] + o.Su[1] * o.Su[1] + o.Su[2] * o.Su[2];
    T b = o.Su[0] * o.Sv[0] + o.Su[1] * o.Sv[1] + o.Su[2] * o.Sv[2];
    T c = o.Sv[0] * o.Sv[0] + o.Sv[1] * o.Sv[1] + o.Sv[2] * o.Sv[2];
    const T j2 = a + c;  // |J|_F^2
    const T gu = o.Su[0] * r0 + o.Su[1] * r1 + o.Su[2] * r2;
    const T gv = o.Sv[0] * r0 + o.Sv[1] * r1 + o.Sv[2] * r2;
    T a1 = a, c1 = c;    // one-dimensional curvatures for a held coordinate (ACTIVE)
    bool hexact = false;  // the step uses the exact Hessian (quadratic convergence)
    if (NEWTON) {
        const T a2 = a + r0 * o.Suu[0] + r1 * o.Suu[1] + r2 * o.Suu[2];
        const T b2 = b + r0 * o.Suv[0] + r1 * o.Suv[1] + r2 * o.Suv[2];
        const T c2 = c + r0 * o.Svv[0] + r1 * o.Svv[1] + r2 * o.Svv[2];
        if (ACTIVE) {
            if (a2 > T(0)) a1 = a2;
            if (c2 > T(0)) c1 = c2;
        }
        if (a2 > T(0) && a2 * c2 - b2 * b2 > T(0)) {
            a = a2;
            b = b2;
            c = c2;
            hexact = true;
        }
    }
    const T det = a * c - b * b;
    if (!(det == det)) return GN_PT_NONFINITE;
    if (det == T(0)) return GN_PT_SINGULAR;
    const T idet = recip(det);
    T du = (c * gu - b * gv) * idet;
    T dv = (a * gv - b * gu) * idet;
    if (ACTIVE) {
        const bool hu = (cu <= dom.u0 && gu > T(0)) || (cu >= dom.u1 && gu < T(0));
        const bool hv = (cv <= dom.v0 && gv > T(0)) || (cv >= dom.v1 && gv < T(0));
        if (hu) {
            du = T(0);
            dv = hv ? T(0) : gv / c1;
        } else if (hv) {
            dv = T(0);
            du = gu / a1;
        }
    }
    cu -= double(du);
    cv -= double(dv);
    if (clamp) {
        cu = fmin(fmax(cu, dom.u0), dom.u1);
        cv = fmin(fmax(cv, dom.v0), dom.v1);
    }
    if (!(cu == cu) || !(cv == cv)) return GN_PT_NONFINITE;
    if (tol_step > 0 && fabs(du) < T(tol_step) && fabs(dv) < T(tol_step)) return GN_PT_CONVERGED;
    // Certificate of an exact-Hessian (Newton) step: the step d leaves an error of at most
    // |H^-1| L |d|^2 / 2, where L bounds the third derivative of f = |S-P|^2/2,
    // |f'''| <= 3 |J| |S''| + |r| |S'''|, and |S'''| is taken as |S''| (k1/h_u + k2/h_v) (one more
    // derivative of the span's polynomial pieces).  |H^-1| = 1/lambda_min >= (a+c)/det.  The point
    // is certified when that bound is below tol_quad; squared so that no square root is needed:
    //   (|S''|^2 (3|J| + |r| m)^2 <= 2 |S''|^2 (9 |J|^2 + |r|^2 m^2))
    //   0.5 |S''|^2 (9 |J|^2 + |r|^2 m^2) |d|^4 (a+c)^2 <= tol_quad^2 det^2.
    // A flat minimum (small lambda_min) therefore needs a smaller step than a well-conditioned one.
    // Where the Hessian is not positive definite (a saddle of the distance: the query is farther
    // than the curvature radius) the Gauss-Newton step converges linearly and only tol_step stops it.
    // The bound holds on one polynomial piece only: a step that ends outside the span pair it was
    // evaluated in crosses a knot, where S'' may jump (C1 surfaces of degree 2), and is not
    // certified; the next iteration evaluates in the new span and certifies there.
    if (NEWTON && hexact && tol_quad > 0) {
        const T h2 = o.Suu[0] * o.Suu[0] + o.Suu[1] * o.Suu[1] + o.Suu[2] * o.Suu[2] +
                     T(2) * (o.Suv[0] * o.Suv[0] + o.Suv[1] * o.Suv[1] + o.Suv[2] * o.Suv[2]) +
                     o.Svv[0] * o.Svv[0] + o.Svv[1] * o.Svv[1] + o.Svv[2] * o.Svv[2];
        const T m = T((d.p + 1) * sc.ihu + (d.q + 1) * sc.ihv);
        const T d2 = du * du + dv * dv, tr = a + c;
        const T lhs = T(0.5) * h2 * (T(9) * j2 + rr * m * m) * (d2 * d2) * (tr * tr);
        const T tq = T(tol_quad);
        if (lhs <= tq * tq * det * det && cu >= sc.ulo && cu <= sc.uhi && cv >= sc.vlo &&
            cv <= sc.vhi)
            return GN_PT_CONVERGED;
    }
    return kStepContinue;
}

// Up to max_iter steps from (cu, cv); *iters = updates performed.
template <int K1, int K2, bool EXACT, typename T, bool NEWTON>
__device__ __forceinline__ int gn_loop(const GnSurfDev &d, const double *__restrict__ ku,
                                       const double *__restrict__ kv, const double P[3],
                                       double &cu, double &cv, int max_iter, double tol_dist,
                                       double tol_step, int clamp, int &iters,
                                       double tol_quad = 0.0) {
    int it = 0, st = GN_PT_MAXITER;
    SpanCache sc;
    const Domain dom = domain_of(d);
    while (it < max_iter) {
        const int s = gn_step<K1, K2, EXACT, T, NEWTON>(d, ku, kv, T(P[0]), T(P[1]), T(P[2]), cu,
                                                        cv, tol_dist, tol_step, tol_quad, clamp,
                                                        dom, sc);
        if (s == kStepContinue || s == GN_PT_CONVERGED) it++;
        if (s != kStepContinue) { st = s; break; }
    }
    iters = it;
    return st;
}

__device__ __forceinline__ unsigned long long bcast64(unsigned long long x, int src) {
    const unsigned lo = __shfl((unsigned)(x & 0xffffffffull), src, 64);
    const unsigned hi = __shfl((unsigned)(x >> 32), src, 64);
    return ((unsigned long long)hi << 32) | lo;
}

// Persistent lane-refill Gauss-Newton.  Every wave keeps its 64 lanes busy: a lane whose point
// has stopped takes the next point of the wave's current chunk of kChunk points.  The batch is
// cut into `slices` contiguous slices with one claim counter each (separate 128-byte lines);
// wave w claims chunks from slice w % slices until that slice is used up, then exits.  Atomics on
// one address serialise across the chip (about 12 ns each), so spreading the claims over up to 64
// counters keeps them off the critical path.  Each pass of the loop either advances every active
// point by one step (at most max_iter each) or hands out new points (at most n), so it ends.
//   MODE kPlain : one-pass FP64 (or FP32) iteration; writes u, v, status, it64.
//   MODE kBulk  : FP32 bulk of the mixed mode; writes the bulk state (u, v, status, it32); a
//                 breakdown (singular / non-finite) writes the seed back so the closing pass
//                 restarts from it.
//   MODE kClose : FP64 closing pass of the mixed mode; a point gets max_iter steps from the bulk
//                 result, or restart_iter steps when its bulk status is a breakdown.
//   MODE kCloseList : the rest of a closing pass whose first step was streamed: the points of
//                 a.list (slice s takes list segment s; launched with kMaxSlices slices), one
//                 step done, resume from their (u, v); their status is still the bulk's.
enum { kPlain = 0, kBulk = 1, kClose = 2, kCloseList = 3 };

template <int K1, int K2, bool EXACT, typename T, bool NEWTON, int MODE>
__global__ __launch_bounds__(kInvBlock, (EXACT && sizeof(T) == 4 && K1 * K2 <= 16) ? 3 : 1) void k_gn_refill(GnSurfDev d_in, InvArgs a,
                                                         unsigned long long *counters, int slices,
                                                         int max_iter, int restart_iter,
                                                         double tol_dist, double tol_step,
                                                         double tol_quad) {
    __shared__ double sk[EXACT ? 2 * kLdsKnotsInv : 1];
    const double *ku = d_in.ku, *kv = d_in.kv;
    GnSurfDev d = d_in;
    if constexpr (EXACT) {
        d = stage_knots_lds(d_in, sk);
        ku = sk;
        kv = sk + d.nu + d.p + 1;
    }
    const Domain dom = domain_of(d);
    const int lane = threadIdx.x & 63;
    const unsigned long long lt = (1ull << lane) - 1ull;
    const int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int slice = wave % slices;
    unsigned long long s_lo = (unsigned long long)(a.n * (size_t)slice / slices);
    unsigned long long s_hi = (unsigned long long)(a.n * (size_t)(slice + 1) / slices);
    if (MODE == kCloseList) {
        const unsigned long long listed = __hip_atomic_load(
            counters + (size_t)slice * kCounterStride + kListWord, __ATOMIC_RELAXED,
            __HIP_MEMORY_SCOPE_AGENT);
        s_lo = (unsigned long long)slice * kListSegCap;
        s_hi = s_lo + (listed < kListSegCap ? listed : kListSegCap);
    }
    unsigned long long *counter = counters + (size_t)slice * kCounterStride;
    unsigned long long q_lo = 0, q_hi = 0;  // wave-uniform chunk
    bool exhausted = false;                 // wave-uniform
    long long idx = -1;
    double cu = 0, cv = 0, su = 0, sv = 0;
    T Px = 0, Py = 0, Pz = 0;
    int it = 0, lim = max_iter;
    SpanCache sc;
    // FP32 exact-degree kernels keep the lane's span patch in registers between iterations; FP64
    // exact-degree kernels stage patch rows through LDS with the next row in flight (per-lane
    // gathers: plain FP64 974 us, rows one at a time 717 us, prefetched 531 us at 1M points)
    constexpr int PSRC = !EXACT ? kPatchGlobal
                       : sizeof(T) == 4 ? (K1 * K2 <= 16 ? kPatchRegs : kPatchGlobal)
                                        : kPatchWavePipe;
    PatchRegs<T, K1 * K2> pr;
    constexpr bool kWave = PSRC == kPatchWavePipe;
    constexpr int kWaveVec = kWave ? WaveRowsPipe<K1, K2, T>::STRIDE * 64 : 1;
    __shared__ float4 wrows[kInvBlock / 64][kWaveVec];
    float4 *wave_lds = wrows[threadIdx.x >> 6];
    while (true) {
        while (!exhausted) {
            const unsigned long long idle = __ballot(idx < 0);
            if (idle == 0ull) break;
            if (q_lo >= q_hi) {
                unsigned long long base = 0;
                if (lane == 0) base = atomicAdd(counter, (unsigned long long)kChunk);
                base = s_lo + bcast64(base, 0);
                if (base >= s_hi) { exhausted = true; break; }
                q_lo = base;
                q_hi = base + kChunk < s_hi ? base + kChunk : s_hi;
            }
            const unsigned long long avail = q_hi - q_lo;
            if (idx < 0) {
                const unsigned long long rank = (unsigned long long)__popcll(idle & lt);
                if (rank < avail) {
                    idx = MODE == kCloseList ? (long long)a.list[q_lo + rank] : (long long)(q_lo + rank);
                    Px = T(a.x[idx]);
                    Py = T(a.y[idx]);
                    Pz = T(a.z[idx]);
                    su = cu = a.u[idx];
                    sv = cv = a.v[idx];
                    it = MODE == kCloseList ? 1 : 0;
                    lim = max_iter;
                    if (MODE == kClose || MODE == kCloseList) {
                        const int sb = a.status[idx];
                        if (sb == GN_PT_SINGULAR || sb == GN_PT_NONFINITE) lim = restart_iter;
                    }
                    sc = SpanCache();
                }
            }
            const unsigned long long nidle = (unsigned long long)__popcll(idle);
            q_lo += nidle < avail ? nidle : avail;
        }
        if (__ballot(idx >= 0) == 0ull) break;
        // a step for every lane holding a point; with LDS-staged rows the whole wave calls it
        // (lanes that take no step compute on copies of their parameters, inside the domain, and
        // discard the result)
        int st = kStepContinue;
        const bool act = idx >= 0 && it < lim;
        if (kWave || act) {
            double tu = act ? cu : dom.u0, tv = act ? cv : dom.v0;
            const int s2 = gn_step<K1, K2, EXACT, T, NEWTON, PSRC, false, MODE == kBulk>(
                d, ku, kv, Px, Py, Pz, tu, tv, tol_dist, tol_step, tol_quad, a.o.clamp, dom, sc,
                pr.m, wave_lds);
            if (act) {
                st = s2;
                cu = tu;
                cv = tv;
            }
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
                a.u[idx] = cu;
                a.v[idx] = cv;
                a.status[idx] = (uint8_t)st;
                if (MODE == kBulk) {
                    if (a.it32) a.it32[idx] = sat8(it);
                } else {
                    if (MODE == kPlain && a.it32) a.it32[idx] = 0;
                    if (a.it64) a.it64[idx] = sat8(it);
                }
                idx = -1;
            }
        }
    }
    if (lane == 0) retire_wave(counters, slices);
}

// Resident workgroups per CU of a kernel (runtime occupancy calculator, cached per kernel).
int blocks_per_cu(const void *kernel, int block);

// Persistent grid: as many 256-lane workgroups as can be resident at once, never more than the
// batch needs; and the number of claim slices for that grid (16 or more waves per slice).
inline void refill_shape(const void *kernel, size_t n, unsigned &grid, int &slices) {
    const size_t want = (n + kInvBlock - 1) / kInvBlock;
    const size_t cap = (size_t)cu_count() * blocks_per_cu(kernel, kInvBlock);
    grid = (unsigned)(want < cap ? (want ? want : 1) : cap);
    const int waves = (int)grid * (kInvBlock / 64);
    slices = waves / 16;
    if (slices < 1) slices = 1;
    if (slices > kMaxSlices) slices = kMaxSlices;
}

// launchers, one translation unit each (invert_fp64.hip, invert_mixed.hip, invert_ref.hip,
// project.hip)
int launch_gn_plain(const GnSurfDev &d, const InvArgs &a, hipStream_t st);
int launch_gn_mixed(const GnSurfDev &d, const InvArgs &a, hipStream_t st);
int launch_gn_first_hit(const GnSurfDev &d, const InvArgs &a, const GnTreeNode *tree,
                        const uint8_t *valid, long count, hipStream_t st);
int launch_project_global(const GnSurfDev &d, const GnSearchDev &sd, const InvArgs &a,
                          hipStream_t st);  // project.hip
int launch_project_fallback(const GnSurfDev &d, const GnSearchDev &sd, const InvArgs &a,
                            hipStream_t st);  // project.hip

}  // namespace gn
