// span_gemm.hip -- FP32 evaluation of unordered points on the matrix cores, in span order.
//
// Points that share a span pair share its patch M_c (c = x w, y w, z w, w), so for 16 such points
// the V-side contraction of all four components is one dense product
//     Y[(c, r)][j] = sum_s M_c[r][s] V_j[s]       (16 x 4) x (4 x 16)
// (V_j = [1, s_j, s_j^2, s_j^3] of point j, or its first or second derivative), one
// v_mfma_f32_16x16x4f32 each for V, V' and V''.  What is left per point is the U-side dot
// products R^(a,b)_c = U^(a)_j . Y^(b)_c[., j] and the rational quotient, on the vector ALUs.
// Unordered points rarely share a span within a wave, so they are first grouped by span pair:
//   1. k_sg_sort     one workgroup per block of kSgBlockPoints consecutive input points: an LDS
//                    histogram of the span-pair keys, its exclusive scan, then each point to its
//                    key's next slot within the block's own output range (LDS cursors), writing the
//                    sorted parameters and perm[k] = the input index of output k.  The span search
//                    compares the FP32 parameters, widened to FP64, with the FP64 knots, as every
//                    other path does.  Keeping each block in its own range keeps the scattered
//                    writes inside a window the caches hold (a global counting sort over all the
//                    points wrote runs of a few points per key and cache line, and took 1.2 ms of
//                    write-backs at 16.8M points); a block of 65,536 points still gives runs of ~18
//                    points per key on a 61 x 61-span surface.
//   2. k_sg_eval     one wave per 64 sorted points: per 16-point group and span pair in it, the
//                    three MFMAs, the U-side dots, a transpose through LDS, the quotient; outputs
//                    are written in sorted order (coalesced).
// Outputs follow the sorted order, which the caller maps back through perm: within each block of
// kSgBlockPoints consecutive input points, grouped by span pair in increasing key order; the order
// within one key of a block is unspecified (LDS atomics), each point's values are not.  This form
// takes degrees up to 3 in each direction (one 16 x 4 tile holds the four components' rows);
// higher degrees use the per-point kernel and the sort alone.
#include "gn_device.hpp"

namespace gn {

int cu_count();  // runtime.hip

typedef float f4v __attribute__((ext_vector_type(4)));

constexpr int kSgSortThreads = 1024;
constexpr size_t kSgBlockPoints = 65536;
constexpr int kSgMaxLdsKeys = 12288;  // 48 KB of LDS counters; more keys count in global memory

// Span pair key and local parameters of an FP32 point.
__device__ __forceinline__ int sg_key(const GnSurfDev &d, double inv_u, double inv_v, float uf,
                                      float vf, float *t = nullptr, float *s = nullptr) {
    const double u = (double)uf, v = (double)vf;
    const int m = find_span_guess(d.ku, d.p, d.nu, u, inv_u);
    const int n = find_span_guess(d.kv, d.q, d.nv, v, inv_v);
    const int im = m - d.p, in = n - d.q;
    if (t) *t = (float)((u - d.ku[m]) * d.ihu[im]);
    if (s) *s = (float)((v - d.kv[n]) * d.ihv[in]);
    return im * d.sv + in;
}

__device__ __forceinline__ double sg_inv(const double *k, int p, int cp) {
    return (double)(cp - p) / (k[cp] - k[p]);  // the uniform-knot guess of find_span_guess
}

// One workgroup per block of kSgBlockPoints input points.  cnt: the block's key counters (LDS,
// or, for more than kSgMaxLdsKeys keys, its zeroed slice of gcnt).  Loads are issued four points
// per thread at a time, so the block's latency overlaps.
__global__ __launch_bounds__(kSgSortThreads) void k_sg_sort(GnSurfDev d, const float *__restrict__ u,
                                                            const float *__restrict__ v, size_t n,
                                                            int keys, uint32_t *__restrict__ gcnt,
                                                            float *__restrict__ su,
                                                            float *__restrict__ sv,
                                                            uint32_t *__restrict__ perm) {
    extern __shared__ uint32_t lcnt[];
    __shared__ uint32_t part[kSgSortThreads];
    const double inv_u = sg_inv(d.ku, d.p, d.nu), inv_v = sg_inv(d.kv, d.q, d.nv);
    const bool lds = keys <= kSgMaxLdsKeys;
    uint32_t *cnt = lds ? lcnt : gcnt + (size_t)blockIdx.x * keys;
    if (lds)
        for (int k = threadIdx.x; k < keys; k += kSgSortThreads) cnt[k] = 0u;
    __syncthreads();
    const size_t i0 = (size_t)blockIdx.x * kSgBlockPoints;
    const size_t i1 = i0 + kSgBlockPoints < n ? i0 + kSgBlockPoints : n;
    constexpr int U = 4;
    for (size_t b = i0 + threadIdx.x; b < i1; b += (size_t)U * kSgSortThreads) {
        float uu[U], vv[U];
#pragma unroll
        for (int k = 0; k < U; k++) {
            const size_t i = b + (size_t)k * kSgSortThreads;
            uu[k] = i < i1 ? u[i] : 0.f;
            vv[k] = i < i1 ? v[i] : 0.f;
        }
#pragma unroll
        for (int k = 0; k < U; k++)
            if (b + (size_t)k * kSgSortThreads < i1) atomicAdd(&cnt[sg_key(d, inv_u, inv_v, uu[k], vv[k])], 1u);
    }
    __syncthreads();
    // exclusive scan of the counters: per thread a run of consecutive keys, then the runs' sums
    const int per = (keys + kSgSortThreads - 1) / kSgSortThreads;
    const int k0 = threadIdx.x * per;
    uint32_t sum = 0;
    for (int k = k0; k < k0 + per && k < keys; k++) sum += cnt[k];
    part[threadIdx.x] = sum;
    __syncthreads();
    for (int o = 1; o < kSgSortThreads; o <<= 1) {  // inclusive Hillis-Steele scan of the sums
        const uint32_t x = threadIdx.x >= (unsigned)o ? part[threadIdx.x - o] : 0u;
        __syncthreads();
        part[threadIdx.x] += x;
        __syncthreads();
    }
    uint32_t run = (uint32_t)i0 + (threadIdx.x ? part[threadIdx.x - 1] : 0u);
    for (int k = k0; k < k0 + per && k < keys; k++) {
        const uint32_t c = cnt[k];
        cnt[k] = run;
        run += c;
    }
    __syncthreads();
    for (size_t b = i0 + threadIdx.x; b < i1; b += (size_t)U * kSgSortThreads) {
        float uu[U], vv[U];
#pragma unroll
        for (int k = 0; k < U; k++) {
            const size_t i = b + (size_t)k * kSgSortThreads;
            uu[k] = i < i1 ? u[i] : 0.f;
            vv[k] = i < i1 ? v[i] : 0.f;
        }
#pragma unroll
        for (int k = 0; k < U; k++) {
            const size_t i = b + (size_t)k * kSgSortThreads;
            if (i >= i1) continue;
            const uint32_t pos = atomicAdd(&cnt[sg_key(d, inv_u, inv_v, uu[k], vv[k])], 1u);
            su[pos] = uu[k];
            sv[pos] = vv[k];
            perm[pos] = (uint32_t)i;
        }
    }
}

// A, B operands and results of v_mfma_f32_16x16x4f32: lane l supplies A[l%16][l/16] and
// B[l/16][l%16] and receives D[4*(l/16) + q][l%16] in result register q.
template <int ORDER>
__global__ __launch_bounds__(256) void k_sg_eval(GnSurfDev d, const float *__restrict__ su,
                                                 const float *__restrict__ sv, size_t n,
                                                 float *__restrict__ out) {
    constexpr int NC = ORDER == 0 ? 1 : (ORDER == 1 ? 3 : 6);  // (du,dv): 00 10 01 20 11 02
    constexpr int NV = ORDER + 1;                               // V, V', V''
    __shared__ float xbuf[4][64 * (4 * NC + 1)];                // per wave: [point][c][combo]
    const double inv_u = sg_inv(d.ku, d.p, d.nu), inv_v = sg_inv(d.kv, d.q, d.nv);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    float *xb = xbuf[wv];
    const int k1 = d.p + 1, k2 = d.q + 1;
    const int i16 = lane & 15, k4 = lane >> 4;
    // this lane's A entry: row (c, r) = (i16 / 4, i16 % 4), column s = k4
    const int ac = i16 >> 2, ar = i16 & 3, as = k4;
    const bool arow = ar < k1 && as < k2;
    const size_t nwaves = (size_t)gridDim.x * 4;
    for (size_t base = ((size_t)blockIdx.x * 4 + wv) * 64; base < n; base += nwaves * 64) {
        const size_t i = base + lane;
        const bool live = i < n;
        float t = 0.f, s = 0.f;
        const int key = sg_key(d, inv_u, inv_v, live ? su[i] : su[n - 1], live ? sv[i] : sv[n - 1],
                               &t, &s);
        const int im = key / d.sv, in = key - im * d.sv;
        const float ihu = (float)d.ihu[im], ihv = (float)d.ihv[in];
#pragma unroll
        for (int g = 0; g < 4; g++) {
            const int src = 16 * g + i16;  // the point this lane serves in group g
            const int gkey = __shfl(key, src, 64);
            const float gt = __shfl(t, src, 64), gs = __shfl(s, src, 64);
            // V-side operand of this lane: V^(b)_src[s = k4]
            float B[NV];
            {
                float p = 1.f;
                for (int e = 0; e < k4; e++) p *= gs;  // gs^k4
                const float pm1 = k4 >= 1 ? (k4 == 1 ? 1.f : (k4 == 2 ? gs : gs * gs)) : 0.f;
                const float pm2 = k4 >= 2 ? (k4 == 2 ? 1.f : gs) : 0.f;
                B[0] = k4 < k2 ? p : 0.f;
                if constexpr (NV > 1) B[1] = k4 < k2 ? float(k4) * pm1 : 0.f;
                if constexpr (NV > 2) B[2] = k4 < k2 ? float(k4 * (k4 - 1)) * pm2 : 0.f;
            }
            f4v D[NV];
#pragma unroll
            for (int b = 0; b < NV; b++) D[b] = f4v{0.f, 0.f, 0.f, 0.f};
            // the span pairs of the group's points, one at a time (sorted input: usually one);
            // point p's key is held by lane p, so the masks below are over points
            const unsigned long long gmask = 0xffffull << (16 * g);
            unsigned long long todo = gmask;  // wave-uniform
            while (todo) {
                const int first = __ffsll((long long)todo) - 1;
                const int kk = __shfl(key, first, 64);
                todo &= ~__ballot(key == kk);
                const int km = kk / d.sv, kn = kk - km * d.sv;
                const float *patch = d.tab32 + (size_t)(km * d.sv + kn) * (size_t)(k1 * k2 * 4);
                const float a = arow ? patch[(ar * k2 + as) * 4 + ac] : 0.f;
                const bool bcol = gkey == kk;
#pragma unroll
                for (int b = 0; b < NV; b++)
                    D[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, bcol ? B[b] : 0.f, D[b], 0, 0, 0);
            }
            // U-side dots for component c = k4 of point src: R^(a,b) = U^(a) . Y^(b)
            float U[3][4];
            {
                float p = 1.f;
#pragma unroll
                for (int r = 0; r < 4; r++) { U[0][r] = r < k1 ? p : 0.f; p *= gt; }
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    U[1][r] = (r >= 1 && r < k1) ? float(r) * U[0][r - 1] : 0.f;
                    U[2][r] = (r >= 2 && r < k1) ? float(r * (r - 1)) * U[0][r - 2] : 0.f;
                }
            }
            float R[NC];
            auto dot = [&](int a, int b) {
                return fmaf(U[a][3], D[b][3], fmaf(U[a][2], D[b][2], fmaf(U[a][1], D[b][1], U[a][0] * D[b][0])));
            };
            R[0] = dot(0, 0);
            if constexpr (ORDER >= 1) { R[1] = dot(1, 0); R[2] = dot(0, 1); }
            if constexpr (ORDER >= 2) { R[3] = dot(2, 0); R[4] = dot(1, 1); R[5] = dot(0, 2); }
            float *dst = xb + src * (4 * NC + 1) + k4 * NC;
#pragma unroll
            for (int c = 0; c < NC; c++) dst[c] = R[c];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // this lane's own point: all components and combos
        HomDer<float> h;
        const float *mine = xb + lane * (4 * NC + 1);
        auto hom = [&](int c) { return Hom<float>{mine[c], mine[NC + c], mine[2 * NC + c], mine[3 * NC + c]}; };
        h.R = hom(0);
        if constexpr (ORDER >= 1) { h.Ru = hom(1); h.Rv = hom(2); }
        if constexpr (ORDER >= 2) { h.Ruu = hom(3); h.Ruv = hom(4); h.Rvv = hom(5); }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        Cart<float> o;
        quotient<ORDER, float>(h, ihu, ihv, o);
        if (live) {
#pragma unroll
            for (int q = 0; q < 3; q++) out[q * n + i] = o.S[q];
            if (ORDER >= 1) {
#pragma unroll
                for (int q = 0; q < 3; q++) {
                    out[(3 + q) * n + i] = o.Su[q];
                    out[(6 + q) * n + i] = o.Sv[q];
                }
            }
            if (ORDER >= 2) {
#pragma unroll
                for (int q = 0; q < 3; q++) {
                    out[(9 + q) * n + i] = o.Suu[q];
                    out[(12 + q) * n + i] = o.Suv[q];
                    out[(15 + q) * n + i] = o.Svv[q];
                }
            }
        }
    }
}

int launch_eval_sorted32(const GnSurfDev &d, const float *u, const float *v, size_t n, int order,
                         float *out, uint32_t *perm, hipStream_t st) {
    if (n == 0) return GN_OK;
    if (n > (size_t)UINT32_MAX) return fail(GN_EINVAL, "span-order evaluation takes at most 2^32-1 points");
    const int keys = d.su * d.sv;
    const size_t blocks = (n + kSgBlockPoints - 1) / kSgBlockPoints;
    const bool lds = keys <= kSgMaxLdsKeys;
    // scratch: sorted parameters, and the blocks' counters when they do not fit LDS (stream-ordered)
    const size_t gcnt_words = lds ? 0 : blocks * (size_t)keys;
    const size_t bytes = sizeof(float) * 2 * n + sizeof(uint32_t) * gcnt_words + 256;
    char *buf = nullptr;
    GN_HIP(hipMallocAsync((void **)&buf, bytes, st));
    float *su = (float *)buf, *sv = su + n;
    uint32_t *gcnt = (uint32_t *)(sv + n);
    hipError_t e = hipSuccess;
    if (!lds) e = hipMemsetAsync(gcnt, 0, sizeof(uint32_t) * gcnt_words, st);
    if (e == hipSuccess) {
        k_sg_sort<<<(unsigned)blocks, kSgSortThreads, lds ? sizeof(uint32_t) * keys : 0, st>>>(
            d, u, v, n, keys, gcnt, su, sv, perm);
        e = hipGetLastError();
    }
    int rc = e == hipSuccess ? GN_OK : hip_fail(e, "span sort launch");
    if (!rc) {
        if (d.p <= 3 && d.q <= 3) {
            const unsigned grid = grid_for(n, 256, (unsigned)cu_count() * 8);
            switch (order) {
                case 0: k_sg_eval<0><<<grid, 256, 0, st>>>(d, su, sv, n, out); break;
                case 1: k_sg_eval<1><<<grid, 256, 0, st>>>(d, su, sv, n, out); break;
                default: k_sg_eval<2><<<grid, 256, 0, st>>>(d, su, sv, n, out); break;
            }
            e = hipGetLastError();
            if (e != hipSuccess) rc = hip_fail(e, "k_sg_eval launch");
        } else {
            rc = launch_eval32(d, su, sv, n, order, out, st);  // per-point kernel, sorted order
        }
    }
    GN_HIP(hipFreeAsync(buf, st));
    return rc;
}

}  // namespace gn
