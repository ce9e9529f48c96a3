// eval.hip -- batch evaluation of a surface and its derivatives (EvaluateRandom,
// src/NURBS/Evaluate.cu:1023-1038, extended to the derivatives of 288-502 and 535-930).
//
// One lane per query point, grid-stride over the batch.  Parameters stream in and results stream
// out as SoA arrays, so every global load and store of the batch is a fully coalesced 256-byte
// (FP32) or 512-byte (FP64) wave access.  Knots are staged in LDS for the span search; span
// patches are gathered from the L2-resident span table.  The kernel is bound by HBM traffic
// (8 or 16 bytes in, 12..144 bytes out per point).
#include "gn_device.hpp"

namespace gn {

constexpr int kEvalBlock = 256;
constexpr int kLdsKnots = 2048;  // knots (both directions) staged in LDS when they fit

// Stages both knot vectors in LDS when they fit and returns the pointers to use.
__device__ __forceinline__ void stage_knots(const GnSurfDev &d, double *sk, const double *&ku,
                                            const double *&kv) {
    const int nku = d.nu + d.p + 1, nkv = d.nv + d.q + 1;
    ku = d.ku;
    kv = d.kv;
    if (nku + nkv <= kLdsKnots) {
        for (int i = threadIdx.x; i < nku; i += blockDim.x) sk[i] = d.ku[i];
        for (int i = threadIdx.x; i < nkv; i += blockDim.x) sk[nku + i] = d.kv[i];
        __syncthreads();
        ku = sk;
        kv = sk + nku;
    }
}

template <int K1, int K2, bool EXACT, int ORDER, typename T, typename TP>
__global__ __launch_bounds__(kEvalBlock) void k_eval(GnSurfDev d, const TP *__restrict__ u,
                                                     const TP *__restrict__ v, size_t n,
                                                     T *__restrict__ out) {
    __shared__ double sk[kLdsKnots];
    const double *ku, *kv;
    stage_knots(d, sk, ku, kv);
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        Cart<T> o;
        eval_point<K1, K2, ORDER, EXACT, T>(d, ku, kv, (double)u[i], (double)v[i], o);
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
        switch (order) {
            case 0: k_eval<K1, K2, EXACT, 0, T, TP><<<grid, kEvalBlock, 0, st>>>(d, u, v, n, out); break;
            case 1: k_eval<K1, K2, EXACT, 1, T, TP><<<grid, kEvalBlock, 0, st>>>(d, u, v, n, out); break;
            default: k_eval<K1, K2, EXACT, 2, T, TP><<<grid, kEvalBlock, 0, st>>>(d, u, v, n, out); break;
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
