// eval.hip -- batch evaluation of a surface and its derivatives up to third order (EvaluateRandom,
// src/NURBS/Evaluate.cu:1023-1038, extended to the derivatives of 288-502 and 535-1020).
//
// One lane per query point, grid-stride over the batch.  Parameters stream in and results stream
// out as SoA arrays, so every global load and store of the batch is a fully coalesced 256-byte
// (FP32) or 512-byte (FP64) wave access.  Span patches come from the L2-resident span table:
// for exact degrees they are staged through LDS one row at a time by cooperative contiguous
// loads, with knots and reciprocal span lengths also in LDS (k_eval_rows); otherwise each lane
// gathers its own patch and reads knots through L1 (k_eval).  The target roof is HBM traffic
// (8 or 16 bytes in, 12..240 bytes out per point).
#include "gn_device.hpp"

namespace gn {

constexpr int kEvalBlock = 256;

// Streams one point's fields into the SoA output: S, then Su Sv, Suu Suv Svv, Suuu Suuv Suvv Svvv
// up to ORDER, each field three blocks (x, y, z) of n values.
template <int ORDER, typename T>
__device__ __forceinline__ void store_fields(const Cart<T> &o, T *__restrict__ out, size_t n,
                                             size_t i) {
#pragma unroll
    for (int c = 0; c < 3; c++) __builtin_nontemporal_store(o.S[c], out + c * n + i);
    if (ORDER >= 1) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            __builtin_nontemporal_store(o.Su[c], out + (3 + c) * n + i);
            __builtin_nontemporal_store(o.Sv[c], out + (6 + c) * n + i);
        }
    }
    if (ORDER >= 2) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            __builtin_nontemporal_store(o.Suu[c], out + (9 + c) * n + i);
            __builtin_nontemporal_store(o.Suv[c], out + (12 + c) * n + i);
            __builtin_nontemporal_store(o.Svv[c], out + (15 + c) * n + i);
        }
    }
    if (ORDER >= 3) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            __builtin_nontemporal_store(o.Suuu[c], out + (18 + c) * n + i);
            __builtin_nontemporal_store(o.Suuv[c], out + (21 + c) * n + i);
            __builtin_nontemporal_store(o.Suvv[c], out + (24 + c) * n + i);
            __builtin_nontemporal_store(o.Svvv[c], out + (27 + c) * n + i);
        }
    }
}

template <int K1, int K2, bool EXACT, int ORDER, typename T, typename TP>
__global__ __launch_bounds__(kEvalBlock) void k_eval(GnSurfDev d, const TP *__restrict__ u,
                                                     const TP *__restrict__ v, size_t n,
                                                     T *__restrict__ out) {
    const double *ku = d.ku, *kv = d.kv;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        Cart<T> o;
        eval_point<K1, K2, ORDER, EXACT, T>(d, ku, kv, (double)u[i], (double)v[i], o);
        store_fields<ORDER>(o, out, n, i);
    }
}

constexpr int kLdsKnotsSmall = 512;

// Row-staged form for exact degrees with knots in LDS: a wave takes 64 points at a time and
// stages their patches one row at a time through a small LDS buffer (WaveRowsPipe: contiguous
// 16-byte loads covering whole rows of several lanes, the next row prefetched while the current
// one is contracted).  The buffer is one row per lane (9 KB per wave for FP64 bicubic), so
// occupancy is set by registers rather than LDS.
constexpr int kRowBlock = 256;
constexpr int kRowWaves = kRowBlock / 64;

template <int K1, int K2, int ORDER, typename T, typename TP>
__global__ __launch_bounds__(kRowBlock) void k_eval_rows(GnSurfDev d, const TP *__restrict__ u,
                                                         const TP *__restrict__ v, size_t n,
                                                         T *__restrict__ out) {
    using Rows = WaveRowsPipe<K1, K2, T>;
    // knots and reciprocal span lengths: nku + nkv + (nu - p) + (nv - q) <= 2 * kLdsKnotsSmall
    __shared__ double sk[2 * kLdsKnotsSmall];
    __shared__ float4 stage[kRowWaves * 64 * Rows::STRIDE];
    const int nku = d.nu + d.p + 1, nkv = d.nv + d.q + 1, nsu = d.nu - d.p, nsv = d.nv - d.q;
    for (int i = threadIdx.x; i < nku; i += blockDim.x) sk[i] = d.ku[i];
    for (int i = threadIdx.x; i < nkv; i += blockDim.x) sk[nku + i] = d.kv[i];
    for (int i = threadIdx.x; i < nsu; i += blockDim.x) sk[nku + nkv + i] = d.ihu[i];
    for (int i = threadIdx.x; i < nsv; i += blockDim.x) sk[nku + nkv + nsu + i] = d.ihv[i];
    __syncthreads();
    const double *ku = sk, *kv = sk + nku, *ihus = sk + nku + nkv, *ihvs = ihus + nsu;
    const double inv_u = (double)(d.nu - d.p) / (ku[d.nu] - ku[d.p]);
    const double inv_v = (double)(d.nv - d.q) / (kv[d.nv] - kv[d.q]);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    float4 *my = stage + (size_t)wv * 64 * Rows::STRIDE;
    const T *tab = sizeof(T) == 8 ? reinterpret_cast<const T *>(d.tab64)
                                  : reinterpret_cast<const T *>(d.tab32);
    const size_t waves = (size_t)gridDim.x * kRowWaves;
    size_t base = ((size_t)blockIdx.x * kRowWaves + wv) * 64;
    // parameters of the next tile are loaded one tile ahead
    TP pu = 0, pv = 0;
    if (base + lane < n) { pu = u[base + lane]; pv = v[base + lane]; }
    for (; base < n; base += waves * 64) {
        const size_t i = base + lane;
        const bool live = i < n;
        const double uu = live ? (double)pu : 0.0, vv = live ? (double)pv : 0.0;
        const size_t inext = base + waves * 64 + lane;
        if (inext < n) { pu = u[inext]; pv = v[inext]; }
        const int m = find_span_guess(ku, d.p, d.nu, uu, inv_u);
        const int nn = find_span_guess(kv, d.q, d.nv, vv, inv_v);
        const int im = m - d.p, in = nn - d.q;
        const T *patch = tab + (size_t)(im * d.sv + in) * (K1 * K2 * 4);
        Rows rows(patch, my);
        const double ihu = ihus[im], ihv = ihvs[in];
        const T t = T((uu - ku[m]) * ihu), s = T((vv - kv[nn]) * ihv);
        HomDer<T> h;
        contract_rows<K1, K2, ORDER, true, T>(rows, K1, K2, t, s, h);
        Cart<T> o;
        quotient<ORDER, T>(h, T(ihu), T(ihv), o);
        if (live) store_fields<ORDER>(o, out, n, i);
    }
}

template <typename T, typename TP>
struct EvalLaunch {
    const GnSurfDev &d;
    const TP *u, *v;
    size_t n;
    int order;
    T *out;
    hipStream_t st;
    template <int K1, int K2, bool EXACT>
    int operator()() {
        const unsigned grid = grid_for(n, kEvalBlock, 256 * 16);
        const bool knots_fit = d.nu + d.p + 1 + d.nv + d.q + 1 <= kLdsKnotsSmall;
        if constexpr (EXACT) if (knots_fit) {
            const unsigned g3 = grid_for(n, kRowBlock, 256 * 16);
            switch (order) {
                case 0: k_eval_rows<K1, K2, 0, T, TP><<<g3, kRowBlock, 0, st>>>(d, u, v, n, out); break;
                case 1: k_eval_rows<K1, K2, 1, T, TP><<<g3, kRowBlock, 0, st>>>(d, u, v, n, out); break;
                case 2: k_eval_rows<K1, K2, 2, T, TP><<<g3, kRowBlock, 0, st>>>(d, u, v, n, out); break;
                default: k_eval_rows<K1, K2, 3, T, TP><<<g3, kRowBlock, 0, st>>>(d, u, v, n, out); break;
            }
            GN_HIP(hipGetLastError());
            return GN_OK;
        }
        switch (order) {
            case 0: k_eval<K1, K2, EXACT, 0, T, TP><<<grid, kEvalBlock, 0, st>>>(d, u, v, n, out); break;
            case 1: k_eval<K1, K2, EXACT, 1, T, TP><<<grid, kEvalBlock, 0, st>>>(d, u, v, n, out); break;
            case 2: k_eval<K1, K2, EXACT, 2, T, TP><<<grid, kEvalBlock, 0, st>>>(d, u, v, n, out); break;
            default: k_eval<K1, K2, EXACT, 3, T, TP><<<grid, kEvalBlock, 0, st>>>(d, u, v, n, out); break;
        }
        GN_HIP(hipGetLastError());
        return GN_OK;
    }
};

int launch_eval64(const GnSurfDev &d, const double *u, const double *v, size_t n, int order,
                  double *out, hipStream_t st) {
    if (n == 0) return GN_OK;
    return dispatch_degrees(d.p, d.q, EvalLaunch<double, double>{d, u, v, n, order, out, st});
}

int launch_eval32(const GnSurfDev &d, const float *u, const float *v, size_t n, int order,
                  float *out, hipStream_t st) {
    if (n == 0) return GN_OK;
    return dispatch_degrees(d.p, d.q, EvalLaunch<float, float>{d, u, v, n, order, out, st});
}

}  // namespace gn
