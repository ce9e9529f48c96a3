// grid_mfma.hip -- FP32 span-GEMM evaluation of a tensor-product parameter grid on the matrix
// cores (the MI355X form of the FP16 cuBLAS span-GEMM experiment of src/main.cu:533-663 and
// src/kernel/kernel.cu:1558-1604).
//
// For grid rows a (u-samples) and columns b (v-samples) that fall in one span pair (m, n):
//     R_c[a][b] = U(t_a)^T * M_c(m, n) * V(s_b) = (A * Y_c)[a][b],
//     A[a][r] = t_a^r (or its u-derivatives),  Y_c[r][b] = sum_s M_c[r][s] s_b^s (or v-derivs).
// One wave owns a 16 x 16 block of the grid.  Lane l supplies A[l%16][l/16] and Y_c[l/16][l%16]
// to v_mfma_f32_16x16x4f32 and receives D[4*(l/16) + j][l%16], j = 0..3, so every lane ends with
// all homogeneous components and derivatives of its own four grid points and does their rational
// quotient without any cross-lane exchange.  A block that straddles span boundaries is processed
// once per span pair it touches, with rows/columns of other spans zeroed.  K = k1 <= 4 takes one
// MFMA per product, k1 <= 8 two.  Sparse grids (fewer than 4 samples per span on average) use the
// per-point VALU kernel instead, since their blocks would straddle many spans.
#include "gn_device.hpp"

namespace gn {

typedef float f4v __attribute__((ext_vector_type(4)));

template <int K>
__device__ __forceinline__ float pick(const float (&a)[K], int r) {
    float x = 0.f;
#pragma unroll
    for (int i = 0; i < K; i++) x = (i == r) ? a[i] : x;
    return x;
}

template <int K1, int K2, bool EXACT, int ORDER>
__global__ __launch_bounds__(64) void k_grid_mfma(GnSurfDev d, const float *__restrict__ us,
                                                  int nua, const float *__restrict__ vs, int nvb,
                                                  float *__restrict__ out) {
    constexpr int KS = (K1 + 3) / 4;                       // K steps of 4
    constexpr int NC = ORDER == 0 ? 1 : (ORDER == 1 ? 3 : 6);  // derivative combos
    // combo c -> (du, dv): 0 (0,0) 1 (1,0) 2 (0,1) 3 (2,0) 4 (1,1) 5 (0,2)
    const int k1 = EXACT ? K1 : d.p + 1, k2 = EXACT ? K2 : d.q + 1;
    const int lane = threadIdx.x;
    const int i16 = lane & 15, k4 = lane >> 4;
    const int a0 = blockIdx.x * 16, b0 = blockIdx.y * 16;
    const size_t npts = (size_t)nua * nvb;

    // A role: grid row a0 + i16
    const int a = a0 + i16;
    const bool arow = a < nua;
    const double uu = (double)us[arow ? a : nua - 1];
    const int mu = find_span(d.ku, d.p, d.nu, uu);
    const float t = (float)((uu - d.ku[mu]) * d.ihu[mu - d.p]);
    // B role: grid column b0 + i16 (also the output column of this lane)
    const int b = b0 + i16;
    const bool bcol = b < nvb;
    const double vv = (double)vs[bcol ? b : nvb - 1];
    const int mv = find_span(d.kv, d.q, d.nv, vv);
    const float s = (float)((vv - d.kv[mv]) * d.ihv[mv - d.q]);

    float U[3][K1];  // t^r and its first two derivatives
    {
        float now = 1.f;
#pragma unroll
        for (int r = 0; r < K1; r++) { U[0][r] = now; now *= t; }
#pragma unroll
        for (int r = 0; r < K1; r++) {
            U[1][r] = r >= 1 ? float(r) * U[0][r - 1] : 0.f;
            U[2][r] = r >= 2 ? float(r * (r - 1)) * U[0][r - 2] : 0.f;
        }
    }
    float V[3][K2];
    {
        float now = 1.f;
#pragma unroll
        for (int c = 0; c < K2; c++) { V[0][c] = now; now *= s; }
#pragma unroll
        for (int c = 0; c < K2; c++) {
            V[1][c] = c >= 1 ? float(c) * V[0][c - 1] : 0.f;
            V[2][c] = c >= 2 ? float(c * (c - 1)) * V[0][c - 2] : 0.f;
        }
    }

    // spans of this lane's four output rows 4*k4 + j
    int rspan[4];
#pragma unroll
    for (int j = 0; j < 4; j++) rspan[j] = __shfl(mu, 4 * k4 + j, 64);

    float res[NC][4][4];  // [combo][component][row j]
#pragma unroll
    for (int c = 0; c < NC; c++)
#pragma unroll
        for (int q = 0; q < 4; q++)
#pragma unroll
            for (int j = 0; j < 4; j++) res[c][q][j] = 0.f;

    // span ranges touched by the block (wave-uniform; no ordering of us/vs is assumed)
    int mu_first = arow ? mu : 0x7fffffff, mu_last = arow ? mu : -1;
    int mv_first = bcol ? mv : 0x7fffffff, mv_last = bcol ? mv : -1;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        mu_first = min(mu_first, __shfl_xor(mu_first, o, 64));
        mu_last = max(mu_last, __shfl_xor(mu_last, o, 64));
        mv_first = min(mv_first, __shfl_xor(mv_first, o, 64));
        mv_last = max(mv_last, __shfl_xor(mv_last, o, 64));
    }
    const int kk = k1 * k2 * 4;

    for (int m = mu_first; m <= mu_last; m++) {
        if (!__any(arow && mu == m)) continue;
        const bool ain = arow && mu == m;
        for (int n = mv_first; n <= mv_last; n++) {
            if (!__any(bcol && mv == n)) continue;
            const bool bin = bcol && mv == n;
            const float *patch = d.tab32 + (size_t)((m - d.p) * d.sv + (n - d.q)) * kk;
            f4v D[NC][4];
#pragma unroll
            for (int c = 0; c < NC; c++)
#pragma unroll
                for (int q = 0; q < 4; q++) D[c][q] = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ks = 0; ks < KS; ks++) {
                const int r = k4 + 4 * ks;
                const bool rin = r < k1;
                // A operands for du = 0, 1, 2
                float A[3];
#pragma unroll
                for (int du = 0; du < 3; du++) A[du] = (ain && rin) ? pick<K1>(U[du], r) : 0.f;
                // B operands Y_c^(dv)[r][b] for dv = 0, 1, 2
                float Y[3][4];
#pragma unroll
                for (int dv = 0; dv < 3; dv++)
#pragma unroll
                    for (int q = 0; q < 4; q++) Y[dv][q] = 0.f;
                if (bin && rin) {
#pragma unroll
                    for (int c = 0; c < K2; c++) {
                        if (!EXACT && c >= k2) break;
                        const float4 mm = *reinterpret_cast<const float4 *>(patch + (r * k2 + c) * 4);
                        const float mq[4] = {mm.x, mm.y, mm.z, mm.w};
#pragma unroll
                        for (int q = 0; q < 4; q++) {
                            Y[0][q] = fmaf(mq[q], V[0][c], Y[0][q]);
                            if (ORDER >= 1) Y[1][q] = fmaf(mq[q], V[1][c], Y[1][q]);
                            if (ORDER >= 2) Y[2][q] = fmaf(mq[q], V[2][c], Y[2][q]);
                        }
                    }
                }
                constexpr int cdu[6] = {0, 1, 0, 2, 1, 0};
                constexpr int cdv[6] = {0, 0, 1, 0, 1, 2};
#pragma unroll
                for (int c = 0; c < NC; c++)
#pragma unroll
                    for (int q = 0; q < 4; q++)
                        D[c][q] = __builtin_amdgcn_mfma_f32_16x16x4f32(A[cdu[c]], Y[cdv[c]][q],
                                                                       D[c][q], 0, 0, 0);
            }
            const bool colin = bcol && mv == n;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const bool take = colin && rspan[j] == m;
#pragma unroll
                for (int c = 0; c < NC; c++)
#pragma unroll
                    for (int q = 0; q < 4; q++) res[c][q][j] = take ? D[c][q][j] : res[c][q][j];
            }
        }
    }

    if (!bcol) return;
    const float ihv = (float)d.ihv[mv - d.q];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int ar = a0 + 4 * k4 + j;
        if (ar >= nua) continue;
        const float ihu = (float)d.ihu[rspan[j] - d.p];
        HomDer<float> h;
        h.R = Hom<float>{res[0][0][j], res[0][1][j], res[0][2][j], res[0][3][j]};
        if (ORDER >= 1) {
            h.Ru = Hom<float>{res[1][0][j], res[1][1][j], res[1][2][j], res[1][3][j]};
            h.Rv = Hom<float>{res[2][0][j], res[2][1][j], res[2][2][j], res[2][3][j]};
        }
        if (ORDER >= 2) {
            h.Ruu = Hom<float>{res[3][0][j], res[3][1][j], res[3][2][j], res[3][3][j]};
            h.Ruv = Hom<float>{res[4][0][j], res[4][1][j], res[4][2][j], res[4][3][j]};
            h.Rvv = Hom<float>{res[5][0][j], res[5][1][j], res[5][2][j], res[5][3][j]};
        }
        Cart<float> o;
        quotient<ORDER, float>(h, ihu, ihv, o);
        const size_t i = (size_t)ar * nvb + b;
#pragma unroll
        for (int c = 0; c < 3; c++) __builtin_nontemporal_store(o.S[c], out + c * npts + i);
        if (ORDER >= 1) {
#pragma unroll
            for (int c = 0; c < 3; c++) {
                __builtin_nontemporal_store(o.Su[c], out + (3 + c) * npts + i);
                __builtin_nontemporal_store(o.Sv[c], out + (6 + c) * npts + i);
            }
        }
        if (ORDER >= 2) {
#pragma unroll
            for (int c = 0; c < 3; c++) {
                __builtin_nontemporal_store(o.Suu[c], out + (9 + c) * npts + i);
                __builtin_nontemporal_store(o.Suv[c], out + (12 + c) * npts + i);
                __builtin_nontemporal_store(o.Svv[c], out + (15 + c) * npts + i);
            }
        }
    }
}

// Per-point VALU fallback for sparse grids.
template <int K1, int K2, bool EXACT, int ORDER>
__global__ __launch_bounds__(256) void k_grid_valu(GnSurfDev d, const float *__restrict__ us,
                                                   int nua, const float *__restrict__ vs, int nvb,
                                                   float *__restrict__ out) {
    const size_t npts = (size_t)nua * nvb;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < npts;
         i += (size_t)gridDim.x * blockDim.x) {
        const int a = (int)(i / nvb), b = (int)(i % nvb);
        Cart<float> o;
        eval_point<K1, K2, ORDER, EXACT, float>(d, d.ku, d.kv, (double)us[a], (double)vs[b], o);
#pragma unroll
        for (int c = 0; c < 3; c++) out[c * npts + i] = o.S[c];
        if (ORDER >= 1)
#pragma unroll
            for (int c = 0; c < 3; c++) {
                out[(3 + c) * npts + i] = o.Su[c];
                out[(6 + c) * npts + i] = o.Sv[c];
            }
        if (ORDER >= 2)
#pragma unroll
            for (int c = 0; c < 3; c++) {
                out[(9 + c) * npts + i] = o.Suu[c];
                out[(12 + c) * npts + i] = o.Suv[c];
                out[(15 + c) * npts + i] = o.Svv[c];
            }
    }
}

struct GridLaunch {
    const GnSurfDev &d;
    const float *us;
    int nua;
    const float *vs;
    int nvb, order;
    float *out;
    hipStream_t st;
    bool mfma;
    template <int K1, int K2, bool EXACT>
    int operator()() {
        if (mfma) {
            dim3 grid((nua + 15) / 16, (nvb + 15) / 16);
            if (order == 0) k_grid_mfma<K1, K2, EXACT, 0><<<grid, 64, 0, st>>>(d, us, nua, vs, nvb, out);
            else if (order == 1) k_grid_mfma<K1, K2, EXACT, 1><<<grid, 64, 0, st>>>(d, us, nua, vs, nvb, out);
            else k_grid_mfma<K1, K2, EXACT, 2><<<grid, 64, 0, st>>>(d, us, nua, vs, nvb, out);
        } else {
            const unsigned grid = grid_for((size_t)nua * nvb, 256, 256 * 16);
            if (order == 0) k_grid_valu<K1, K2, EXACT, 0><<<grid, 256, 0, st>>>(d, us, nua, vs, nvb, out);
            else if (order == 1) k_grid_valu<K1, K2, EXACT, 1><<<grid, 256, 0, st>>>(d, us, nua, vs, nvb, out);
            else k_grid_valu<K1, K2, EXACT, 2><<<grid, 256, 0, st>>>(d, us, nua, vs, nvb, out);
        }
        GN_HIP(hipGetLastError());
        return GN_OK;
    }
};

int launch_eval_grid32(const GnSurfDev &d, const float *us, int nu_pts, const float *vs,
                       int nv_pts, int order, float *out, hipStream_t st) {
    if ((size_t)nu_pts * nv_pts == 0) return GN_OK;
    const bool mfma = nu_pts >= 4 * d.su && nv_pts >= 4 * d.sv;
    return dispatch_degrees(d.p, d.q, GridLaunch{d, us, nu_pts, vs, nv_pts, order, out, st, mfma});
}

}  // namespace gn
