// gn_device.hpp -- device building blocks shared by the HIP kernels of libgnurbs.
#pragma once
#include "gn_internal.hpp"

namespace gn {

// Knot-span lookup.  Returns the unique m with k[m] <= u < k[m+1], clamped to [p, cp-1] exactly
// as findSpan (src/NURBS/Evaluate.cu:504-533): u >= k[cp] -> cp-1, u <= k[p] -> p.  A NaN makes
// the reference's while-loop exit at its first midpoint, (p + cp) / 2; that is reproduced too.
// Any search that finds the unique bracketing span returns the reference's index, so the loop
// here is a plain branch-light bisection on an invariant k[lo] <= u < k[hi].
__device__ __forceinline__ int find_span(const double *__restrict__ k, int p, int cp, double u) {
    if (u >= k[cp]) return cp - 1;
    if (u <= k[p]) return p;
    if (u != u) return (p + cp) >> 1;
    int lo = p, hi = cp;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (u < k[mid]) hi = mid; else lo = mid;
    }
    return lo;
}

// find_span with a first guess: the span that u would fall in if the knots were uniform over the
// domain (inv = (cp - p) / (k[cp] - k[p])).  The guess is taken only when it brackets u, and the
// bracketing span is unique, so the result is find_span's in every case; on uniform knots the
// bisection is skipped.
__device__ __forceinline__ int find_span_guess(const double *__restrict__ k, int p, int cp,
                                               double u, double inv) {
    if (u >= k[cp]) return cp - 1;
    if (u <= k[p]) return p;
    if (u != u) return (p + cp) >> 1;
    int m = p + (int)((u - k[p]) * inv);
    m = m < p ? p : (m > cp - 1 ? cp - 1 : m);
    if (k[m] <= u && u < k[m + 1]) return m;
    int lo = p, hi = cp;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (u < k[mid]) hi = mid; else lo = mid;
    }
    return lo;
}

// find_span for the iteration kernels, where a span search is a chain of dependent loads from
// global memory (batched surfaces) or LDS: the uniform-knot guess (an FP32 estimate, clamped
// to [p, cp-1]) is tested first, then the one neighbour on u's side of it, and only a u outside
// those spans is bisected.  Any bracketing span is the unique one, so the result is find_span's
// in every case.
__device__ __forceinline__ int find_span_near(const double *__restrict__ k, int p, int cp, double u) {
    const double kp = k[p], kc = k[cp];
    if (u >= kc) return cp - 1;
    if (u <= kp) return p;
    if (u != u) return (p + cp) >> 1;
    float f = (float)(u - kp) * __builtin_amdgcn_rcpf((float)(kc - kp)) * (float)(cp - p);
    f = fminf(fmaxf(f, 0.f), (float)(cp - p - 1));  // also maps a NaN estimate to 0
    const int m = p + (int)f;
    const double a1 = k[m], a2 = k[m + 1];
    if (a1 <= u && u < a2) return m;
    if (u < a1) {
        if (m > p && k[m - 1] <= u) return m - 1;
    } else if (m + 2 <= cp && u < k[m + 2]) {
        return m + 1;
    }
    int lo = p, hi = cp;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (u < k[mid]) hi = mid; else lo = mid;
    }
    return lo;
}

template <typename T>
struct Hom {  // homogeneous 4-vector
    T x, y, z, w;
};

template <typename T>
struct HomDer {  // homogeneous point and partial derivatives in the local span parameters
    Hom<T> R, Ru, Rv, Ruu, Ruv, Rvv, Ruuu, Ruuv, Ruvv, Rvvv;
};

template <typename T>
__device__ __forceinline__ void hzero(Hom<T> &a) { a.x = a.y = a.z = a.w = T(0); }
template <typename T>
__device__ __forceinline__ void hfma(Hom<T> &acc, T s, const Hom<T> &m) {
    acc.x = fma(s, m.x, acc.x);
    acc.y = fma(s, m.y, acc.y);
    acc.z = fma(s, m.z, acc.z);
    acc.w = fma(s, m.w, acc.w);
}

template <typename T>
__device__ __forceinline__ Hom<T> load_hom(const T *__restrict__ p) {
    Hom<T> h;
    h.x = p[0]; h.y = p[1]; h.z = p[2]; h.w = p[3];
    return h;
}
template <>
__device__ __forceinline__ Hom<float> load_hom<float>(const float *__restrict__ p) {
    float4 v = *reinterpret_cast<const float4 *>(p);
    Hom<float> h;
    h.x = v.x; h.y = v.y; h.z = v.z; h.w = v.w;
    return h;
}
template <>
__device__ __forceinline__ Hom<double> load_hom<double>(const double *__restrict__ p) {
    double2 a = reinterpret_cast<const double2 *>(p)[0];
    double2 b = reinterpret_cast<const double2 *>(p)[1];
    Hom<double> h;
    h.x = a.x; h.y = a.y; h.z = b.x; h.w = b.y;
    return h;
}

// Contraction of one span patch M[r][s][c] (layout [K1][K2][4]) with the power bases of the
// local parameters (t, s):  R = sum_rs t^r M[r][s] s^s  and its partial derivatives up to
// ORDER (MatrixSurfacePoint / MatrixSurfaceRd*, src/NURBS/Evaluate.cu:159-502).  The loop runs
// over the patch rows r outermost: row(r, m) fills m[0..k2-1] with the 4-vectors of row r (one
// 128-byte line of an FP64 bicubic patch), so only one row is live at a time.  Every path
// (global gathers, register-cached patch, LDS-staged rows) shares this summation order, so they
// round identically.  K1/K2 are compile-time bounds; k1/k2 the runtime sizes (equal to K1/K2 when
// EXACT).
template <int K1, int K2, int ORDER, bool EXACT, typename T, typename ROW>
__device__ __forceinline__ void contract_rows(ROW &&row, int k1, int k2, T t, T s, HomDer<T> &o) {
    if (EXACT) { k1 = K1; k2 = K2; }
    T U[K1], dU[K1], ddU[K1], dddU[K1], V[K2], dV[K2], ddV[K2], dddV[K2];
    {
        T now = T(1);
#pragma unroll
        for (int r = 0; r < K1; r++) { U[r] = now; now *= t; }
        dU[0] = T(0);
        ddU[0] = T(0);
#pragma unroll
        for (int r = 1; r < K1; r++) dU[r] = T(r) * U[r - 1];
#pragma unroll
        for (int r = 0; r < K1; r++) ddU[r] = r >= 2 ? T(r * (r - 1)) * U[r - 2] : T(0);
#pragma unroll
        for (int r = 0; r < K1; r++) dddU[r] = r >= 3 ? T(r * (r - 1) * (r - 2)) * U[r - 3] : T(0);
        now = T(1);
#pragma unroll
        for (int c = 0; c < K2; c++) { V[c] = now; now *= s; }
        dV[0] = T(0);
#pragma unroll
        for (int c = 1; c < K2; c++) dV[c] = T(c) * V[c - 1];
#pragma unroll
        for (int c = 0; c < K2; c++) ddV[c] = c >= 2 ? T(c * (c - 1)) * V[c - 2] : T(0);
#pragma unroll
        for (int c = 0; c < K2; c++) dddV[c] = c >= 3 ? T(c * (c - 1) * (c - 2)) * V[c - 3] : T(0);
    }
    hzero(o.R);
    if (ORDER >= 1) { hzero(o.Ru); hzero(o.Rv); }
    if (ORDER >= 2) { hzero(o.Ruu); hzero(o.Ruv); hzero(o.Rvv); }
    if (ORDER >= 3) { hzero(o.Ruuu); hzero(o.Ruuv); hzero(o.Ruvv); hzero(o.Rvvv); }
#pragma unroll
    for (int r = 0; r < K1; r++) {
        if (!EXACT && r >= k1) break;
        Hom<T> m[K2];
        row(r, m);
        Hom<T> b, b1, b2, b3;
        hzero(b);
        hzero(b1);
        hzero(b2);
        hzero(b3);
#pragma unroll
        for (int c = 0; c < K2; c++) {
            if (!EXACT && c >= k2) break;
            hfma(b, V[c], m[c]);
            if (ORDER >= 1 && c >= 1) hfma(b1, dV[c], m[c]);
            if (ORDER >= 2 && c >= 2) hfma(b2, ddV[c], m[c]);
            if (ORDER >= 3 && c >= 3) hfma(b3, dddV[c], m[c]);
        }
        hfma(o.R, U[r], b);
        if (ORDER >= 1) {
            hfma(o.Rv, U[r], b1);
            if (r >= 1) hfma(o.Ru, dU[r], b);
        }
        if (ORDER >= 2) {
            hfma(o.Rvv, U[r], b2);
            if (r >= 1) hfma(o.Ruv, dU[r], b1);
            if (r >= 2) hfma(o.Ruu, ddU[r], b);
        }
        if (ORDER >= 3) {  // MatrixSurfaceRd{UUU,UUV,UVV,VVV}, src/NURBS/Evaluate.cu:408-502
            hfma(o.Rvvv, U[r], b3);
            if (r >= 1) hfma(o.Ruvv, dU[r], b2);
            if (r >= 2) hfma(o.Ruuv, ddU[r], b1);
            if (r >= 3) hfma(o.Ruuu, dddU[r], b);
        }
    }
}

template <int K1, int K2, int ORDER, bool EXACT, typename T>
__device__ __forceinline__ void contract(const T *__restrict__ patch, int k1, int k2, T t, T s,
                                         HomDer<T> &o) {
    const int kk2 = EXACT ? K2 : k2;
    contract_rows<K1, K2, ORDER, EXACT, T>(
        [&](int r, Hom<T>(&m)[K2]) {
#pragma unroll
            for (int c = 0; c < K2; c++)
                if (EXACT || c < kk2) m[c] = load_hom<T>(patch + (r * kk2 + c) * 4);
        },
        k1, k2, t, s, o);
}

// Wave-cooperative staging of patch rows through LDS: row r of the 64 lanes' patches is copied
// with contiguous 16-byte loads (one load instruction covers whole rows of several lanes, i.e.
// a few cache lines, instead of one line per lane), then each lane reads its own row from LDS.
// The global loads of rows r+1..r+DEPTH are in flight while row r goes through LDS and is
// contracted, so up to DEPTH L2 round trips overlap.  Rows must be requested in order 0, 1, ...,
// K1-1 (contract_rows does).  Every lane of the wave must take part (wave-uniform control flow);
// an idle lane passes any valid patch.  lds: the wave's buffer of 64 * STRIDE 16-byte vectors.
template <int K1, int K2, typename T, int DEPTH = 1>
struct WaveRowsPipe {
    static constexpr int VEC = K2 * 4 * (int)sizeof(T) / 16;  // 16-byte vectors per row
    static constexpr int STRIDE = VEC + 1;                     // padded row stride
    static_assert(K2 * 4 * sizeof(T) % 16 == 0, "row size");
    const char *src[VEC];
    float4 *lds;
    float4 nxt[DEPTH][VEC];
    int lane;
    __device__ __forceinline__ void fetch(int r) {
#pragma unroll
        for (int j = 0; j < VEC; j++)
            nxt[r % DEPTH][j] = *reinterpret_cast<const float4 *>(src[j] + (size_t)r * VEC * 16);
    }
    __device__ __forceinline__ WaveRowsPipe(const T *patch, float4 *buf) : lds(buf) {
        lane = threadIdx.x & 63;
        const unsigned long long b = (unsigned long long)patch;
        const unsigned lo = (unsigned)(b & 0xffffffffull), hi = (unsigned)(b >> 32);
#pragma unroll
        for (int j = 0; j < VEC; j++) {
            const int vec = j * 64 + lane, pp = vec / VEC, o = vec - pp * VEC;
            const unsigned long long pb = (unsigned long long)(unsigned)__shfl((int)lo, pp, 64) |
                                          ((unsigned long long)(unsigned)__shfl((int)hi, pp, 64) << 32);
            src[j] = reinterpret_cast<const char *>(pb) + o * 16;
        }
#pragma unroll
        for (int r = 0; r < DEPTH && r < K1; r++) fetch(r);
    }
    __device__ __forceinline__ void operator()(int r, Hom<T> (&m)[K2]) {
        float4 cur[VEC];
#pragma unroll
        for (int j = 0; j < VEC; j++) cur[j] = nxt[r % DEPTH][j];
        if (r + DEPTH < K1) fetch(r + DEPTH);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");  // earlier reads of the buffer
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int j = 0; j < VEC; j++) {
            const int vec = j * 64 + lane, pp = vec / VEC, o = vec - pp * VEC;
            lds[pp * STRIDE + o] = cur[j];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const T *mine = reinterpret_cast<const T *>(lds + lane * STRIDE);
#pragma unroll
        for (int c = 0; c < K2; c++) m[c] = load_hom<T>(mine + c * 4);
    }
};

// A lane's current span patch held in registers (exact sizes only): reloaded from the table only
// when the lane's span pair changes, so iterations that stay in their span issue no gathers.
// Kernels declare one array of the largest size they use and pass its first element.
template <typename T, int N>
struct PatchRegs {
    Hom<T> m[N];
};

template <typename T>
struct Cart {  // Cartesian point and derivatives in the global parameters
    T S[3], Su[3], Sv[3], Suu[3], Suv[3], Svv[3], Suuu[3], Suuv[3], Suvv[3], Svvv[3];
};

// Reciprocal: correctly rounded in FP64; the hardware reciprocal (1 ulp) in FP32.
__device__ __forceinline__ double recip(double x) { return 1.0 / x; }
__device__ __forceinline__ float recip(float x) { return __builtin_amdgcn_rcpf(x); }

// Rational quotient (the old-version rational derivatives of src/NURBS/Evaluate.cu:535-930,
// written in the equivalent Leibniz form S_u = (A_u - w_u S)/w, ...) and the chain factors
// 1/(k[m+1]-k[m]) of the span (the reference multiplies by (k[last]-k[0])/(k[m+1]-k[m]),
// which is the same for knots on [0, 1]).  Third order follows the Leibniz form of
// src/NURBS/Evaluate.cu:960-1020 with every lower-order term subtracted: the reference adds the
// w_uv, w_uuv and w_uvv terms of S_uuv and S_uvv, which is exact only for constant weights.
template <int ORDER, typename T>
__device__ __forceinline__ void quotient(const HomDer<T> &h, T ihu, T ihv, Cart<T> &o) {
    T iw = recip(h.R.w);
    o.S[0] = h.R.x * iw;
    o.S[1] = h.R.y * iw;
    o.S[2] = h.R.z * iw;
    if (ORDER < 1) return;
    const T Ru[3] = {h.Ru.x, h.Ru.y, h.Ru.z}, Rv[3] = {h.Rv.x, h.Rv.y, h.Rv.z};
#pragma unroll
    for (int c = 0; c < 3; c++) {
        o.Su[c] = (Ru[c] - h.Ru.w * o.S[c]) * iw;
        o.Sv[c] = (Rv[c] - h.Rv.w * o.S[c]) * iw;
    }
    T luu[3], luv[3], lvv[3];  // order-2 derivatives in the local parameters
    if (ORDER >= 2) {
        const T Ruu[3] = {h.Ruu.x, h.Ruu.y, h.Ruu.z}, Ruv[3] = {h.Ruv.x, h.Ruv.y, h.Ruv.z},
                Rvv[3] = {h.Rvv.x, h.Rvv.y, h.Rvv.z};
#pragma unroll
        for (int c = 0; c < 3; c++) {
            T uu = (Ruu[c] - T(2) * h.Ru.w * o.Su[c] - h.Ruu.w * o.S[c]) * iw;
            T uv = (Ruv[c] - h.Ru.w * o.Sv[c] - h.Rv.w * o.Su[c] - h.Ruv.w * o.S[c]) * iw;
            T vv = (Rvv[c] - T(2) * h.Rv.w * o.Sv[c] - h.Rvv.w * o.S[c]) * iw;
            luu[c] = uu;
            luv[c] = uv;
            lvv[c] = vv;
            o.Suu[c] = uu * (ihu * ihu);
            o.Suv[c] = uv * (ihu * ihv);
            o.Svv[c] = vv * (ihv * ihv);
        }
    }
    if (ORDER >= 3) {
        const T A3[4][3] = {{h.Ruuu.x, h.Ruuu.y, h.Ruuu.z}, {h.Ruuv.x, h.Ruuv.y, h.Ruuv.z},
                            {h.Ruvv.x, h.Ruvv.y, h.Ruvv.z}, {h.Rvvv.x, h.Rvvv.y, h.Rvvv.z}};
        const T wu = h.Ru.w, wv = h.Rv.w, wuu = h.Ruu.w, wuv = h.Ruv.w, wvv = h.Rvv.w;
        const T cuuu = ihu * ihu * ihu, cuuv = ihu * ihu * ihv, cuvv = ihu * ihv * ihv,
                cvvv = ihv * ihv * ihv;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const T su = o.Su[c], sv = o.Sv[c], s0 = o.S[c];
            T uuu = (A3[0][c] - T(3) * wu * luu[c] - T(3) * wuu * su - h.Ruuu.w * s0) * iw;
            T uuv = (A3[1][c] - wv * luu[c] - T(2) * wu * luv[c] - T(2) * wuv * su - wuu * sv -
                     h.Ruuv.w * s0) * iw;
            T uvv = (A3[2][c] - wu * lvv[c] - T(2) * wv * luv[c] - T(2) * wuv * sv - wvv * su -
                     h.Ruvv.w * s0) * iw;
            T vvv = (A3[3][c] - T(3) * wv * lvv[c] - T(3) * wvv * sv - h.Rvvv.w * s0) * iw;
            o.Suuu[c] = uuu * cuuu;
            o.Suuv[c] = uuv * cuuv;
            o.Suvv[c] = uvv * cuvv;
            o.Svvv[c] = vvv * cvvv;
        }
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
        o.Su[c] *= ihu;
        o.Sv[c] *= ihv;
    }
}

// Full evaluation at global (u, v): span lookup against FP64 knots, local parameters in FP64,
// contraction and quotient in T (double or float).
template <int K1, int K2, int ORDER, bool EXACT, typename T>
__device__ __forceinline__ void eval_point(const GnSurfDev &d, const double *__restrict__ ku,
                                           const double *__restrict__ kv, double u, double v,
                                           Cart<T> &o, int *pm = nullptr, int *pn = nullptr) {
    const int m = find_span(ku, d.p, d.nu, u);
    const int n = find_span(kv, d.q, d.nv, v);
    const int im = m - d.p, in = n - d.q;
    const double ihu = d.ihu[im], ihv = d.ihv[in];
    const T t = T((u - ku[m]) * ihu);
    const T s = T((v - kv[n]) * ihv);
    const int k1 = d.p + 1, k2 = d.q + 1;
    const T *tab = nullptr;
    if constexpr (sizeof(T) == 8) tab = reinterpret_cast<const T *>(d.tab64);
    else tab = reinterpret_cast<const T *>(d.tab32);
    const T *patch = tab + (size_t)(im * d.sv + in) * (size_t)(k1 * k2 * 4);
    HomDer<T> h;
    contract<K1, K2, ORDER, EXACT, T>(patch, k1, k2, t, s, h);
    quotient<ORDER, T>(h, T(ihu), T(ihv), o);
    if (pm) *pm = m;
    if (pn) *pn = n;
}

// Per-lane cache of the current span pair: a parameter that stays inside [k[m], k[m+1]) has
// exactly the span m (the bracketing span is unique), so the search is skipped.
struct SpanCache {
    double ulo = INFINITY, uhi = -INFINITY, vlo = INFINITY, vhi = -INFINITY;
    double ihu = 0, ihv = 0;
    int im = 0, in = 0;  // span slots m - p, n - q
};

// Patch sources of eval_point_cached: per-lane gathers, the register-cached patch, or LDS rows
// staged by the whole wave with the next row prefetched (wave-uniform call required).
enum { kPatchGlobal = 0, kPatchRegs = 1, kPatchWavePipe = 2 };

// NEAR: the span search is find_span_near, for the batched kernels (knots in global memory, where
// bisection's dependent loads are long) and the single-surface FP32 bulk (+2% on 1M projection
// queries).  Elsewhere the bisection over LDS knots costs fewer registers: with find_span_near
// k_close_first no longer fits two waves per SIMD (55.6 -> 65.6 us).
template <int K1, int K2, int ORDER, bool EXACT, typename T, int PSRC = kPatchGlobal,
          bool NEAR = false>
__device__ __forceinline__ void eval_point_cached(const GnSurfDev &d, const double *__restrict__ ku,
                                                  const double *__restrict__ kv, double u, double v,
                                                  SpanCache &sc, Cart<T> &o,
                                                  Hom<T> *pr = nullptr, float4 *wave_lds = nullptr) {
    constexpr bool PREG = PSRC == kPatchRegs;
    bool moved = false;
    if (!(u >= sc.ulo && u < sc.uhi)) {
        const int m = NEAR ? find_span_near(ku, d.p, d.nu, u) : find_span(ku, d.p, d.nu, u);
        sc.ulo = ku[m];
        sc.uhi = ku[m + 1];
        sc.im = m - d.p;
        sc.ihu = d.ihu[sc.im];
        moved = true;
    }
    if (!(v >= sc.vlo && v < sc.vhi)) {
        const int n = NEAR ? find_span_near(kv, d.q, d.nv, v) : find_span(kv, d.q, d.nv, v);
        sc.vlo = kv[n];
        sc.vhi = kv[n + 1];
        sc.in = n - d.q;
        sc.ihv = d.ihv[sc.in];
        moved = true;
    }
    const T t = T((u - sc.ulo) * sc.ihu);
    const T s = T((v - sc.vlo) * sc.ihv);
    const int k1 = d.p + 1, k2 = d.q + 1;
    const T *tab = nullptr;
    if constexpr (sizeof(T) == 8) tab = reinterpret_cast<const T *>(d.tab64);
    else tab = reinterpret_cast<const T *>(d.tab32);
    const T *patch = tab + (size_t)(sc.im * d.sv + sc.in) * (size_t)(k1 * k2 * 4);
    HomDer<T> h;
    if constexpr (PREG) {
        static_assert(EXACT, "register patch cache needs exact sizes");
        if (moved) {
#pragma unroll
            for (int e = 0; e < K1 * K2; e++) pr[e] = load_hom<T>(patch + e * 4);
        }
        contract_rows<K1, K2, ORDER, true, T>(
            [&](int r, Hom<T>(&m)[K2]) {
#pragma unroll
                for (int c = 0; c < K2; c++) m[c] = pr[r * K2 + c];
            },
            k1, k2, t, s, h);
    } else if constexpr (PSRC == kPatchWavePipe) {
        static_assert(EXACT, "LDS row staging needs exact sizes");
        (void)moved;
        WaveRowsPipe<K1, K2, T> rows(patch, wave_lds);
        contract_rows<K1, K2, ORDER, true, T>(rows, k1, k2, t, s, h);
    } else {
        (void)moved;
        contract<K1, K2, ORDER, EXACT, T>(patch, k1, k2, t, s, h);
    }
    quotient<ORDER, T>(h, T(sc.ihu), T(sc.ihv), o);
}

// M[r*k + a] = coefficient of t^r in N_{span-deg+a, deg}(k[span] + t*(k[span+1]-k[span])).
// Same matrix as the Nrc recursion of src/NURBS/Evaluate.cu:90-103; 0/0 is taken as 0.
__device__ inline void basis_matrix(int deg, const double *__restrict__ kn, int span, double *M) {
    double N[GN_MAXK][GN_MAXK], T[GN_MAXK][GN_MAXK];
    for (int a = 0; a < GN_MAXK; a++)
        for (int e = 0; e < GN_MAXK; e++) N[a][e] = 0.0;
    N[0][0] = 1.0;
    const double h = kn[span + 1] - kn[span];
    for (int d = 1; d <= deg; d++) {
        for (int a = 0; a <= d; a++)
            for (int e = 0; e <= d; e++) T[a][e] = 0.0;
        for (int a = 0; a <= d; a++) {
            const int i = span - d + a;
            if (a >= 1) {
                double den = kn[i + d] - kn[i];
                if (den != 0.0) {
                    double c0 = (kn[span] - kn[i]) / den, c1 = h / den;
                    for (int e = 0; e < d; e++) {
                        T[a][e] += c0 * N[a - 1][e];
                        T[a][e + 1] += c1 * N[a - 1][e];
                    }
                }
            }
            if (a <= d - 1) {
                double den = kn[i + d + 1] - kn[i + 1];
                if (den != 0.0) {
                    double c0 = (kn[i + d + 1] - kn[span]) / den, c1 = -h / den;
                    for (int e = 0; e < d; e++) {
                        T[a][e] += c0 * N[a][e];
                        T[a][e + 1] += c1 * N[a][e];
                    }
                }
            }
        }
        for (int a = 0; a <= d; a++)
            for (int e = 0; e <= d; e++) N[a][e] = T[a][e];
    }
    const int k = deg + 1;
    for (int r = 0; r < k; r++)
        for (int a = 0; a < k; a++) M[r * k + a] = N[a][r];
}

// Dispatch helper: calls F.template operator()<K1, K2, EXACT>() for the degree pair of a
// surface.  Degrees 2..5 in each direction get exact instantiations; everything else uses the
// padded <8, 8> form with runtime sizes.
template <typename F>
__host__ int dispatch_degrees(int p, int q, F &&f) {
#define GN_CASE(P, Q) \
    if (p == P && q == Q) return f.template operator()<P + 1, Q + 1, true>();
    GN_CASE(2, 2) GN_CASE(2, 3) GN_CASE(2, 4) GN_CASE(2, 5)
    GN_CASE(3, 2) GN_CASE(3, 3) GN_CASE(3, 4) GN_CASE(3, 5)
    GN_CASE(4, 2) GN_CASE(4, 3) GN_CASE(4, 4) GN_CASE(4, 5)
    GN_CASE(5, 2) GN_CASE(5, 3) GN_CASE(5, 4) GN_CASE(5, 5)
#undef GN_CASE
    return f.template operator()<GN_MAXK, GN_MAXK, false>();
}

__host__ __device__ inline unsigned grid_for(size_t n, int block, unsigned cap) {
    size_t g = (n + block - 1) / block;
    if (g > cap) g = cap;
    if (g == 0) g = 1;
    return (unsigned)g;
}

}  // namespace gn
