// pcurve.hip -- the p-curve of a 3-D curve on a surface: the use the reference's README names
// (README.md:10) and its code does not contain.  The curve is sampled at uniformly spaced
// parameters, the samples are inverted onto the surface (gn_invert, either precision, seeds from
// the surface's tree),
// and a clamped uniform 2-D B-spline (u(t), v(t)) over the curve's parameter domain is fitted to
// the converged (u, v) by least squares.  The fit's normal equations are banded: one workgroup per
// (span, chunk of the span's samples) sums its samples' basis products into one record, with no
// atomics; the host adds the records and solves the banded system by Cholesky.  The fitted curve
// is then pushed through the surface and compared with the curve at every sample.
#include <algorithm>
#include <cmath>
#include <vector>

#include "gn_device.hpp"

namespace gn {

constexpr int kPcMaxDeg = 5;
constexpr int kPcBlock = 256;
constexpr int kPcVals = (kPcMaxDeg + 1) * (kPcMaxDeg + 2) / 2 + 2 * (kPcMaxDeg + 1);

// t[i] uniform on [t0, t1], the last sample exactly t1
__global__ void k_pc_params(double t0, double t1, long long m, double *__restrict__ t) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < m;
         i += (long long)gridDim.x * blockDim.x)
        t[i] = i == m - 1 ? t1 : t0 + (t1 - t0) * ((double)i / (double)(m - 1));
}

// Record of workgroup (span j = d + blockIdx.x, chunk blockIdx.y): over the converged samples with
// parameter in span j and index in the chunk, the packed upper triangle of sum B_a B_b, then
// sum B_a u and sum B_a v (B: the d+1 basis functions that are nonzero on span j).
__global__ __launch_bounds__(kPcBlock) void k_pc_normal(int d, const double *__restrict__ kn,
                                                        int ncp, const double *__restrict__ t,
                                                        const double *__restrict__ u,
                                                        const double *__restrict__ v,
                                                        const uint8_t *__restrict__ st,
                                                        long long m, double *__restrict__ rec) {
    const int j = d + (int)blockIdx.x, k = d + 1;
    const int chunks = (int)gridDim.y, chunk = (int)blockIdx.y;
    __shared__ double M[(kPcMaxDeg + 1) * (kPcMaxDeg + 1)];
    __shared__ double red[kPcBlock / 64][kPcVals];
    const double a = kn[j], b = kn[j + 1], h = b - a;
    if (threadIdx.x == 0 && h > 0.0) basis_matrix(d, kn, j, M);
    __syncthreads();
    double acc[kPcVals];
#pragma unroll
    for (int q = 0; q < kPcVals; q++) acc[q] = 0.0;
    const int ntri = k * (k + 1) / 2;
    if (h > 0.0) {
        // samples of span j: index range from the uniform spacing, widened by 2 against rounding;
        // membership is decided by the same span search the evaluation uses
        const double T = t[m - 1] - t[0];
        long long lo = (long long)std::floor((a - t[0]) / T * (double)(m - 1)) - 2;
        long long hi = (long long)std::ceil((b - t[0]) / T * (double)(m - 1)) + 2;
        lo = lo < 0 ? 0 : lo;
        hi = hi > m - 1 ? m - 1 : hi;
        const long long len = hi - lo + 1;
        const long long c0 = lo + len * chunk / chunks, c1 = lo + len * (chunk + 1) / chunks;
        for (long long i = c0 + threadIdx.x; i < c1; i += kPcBlock) {
            if (find_span(kn, d, ncp, t[i]) != j) continue;
            if (st[i] != GN_PT_CONVERGED && st[i] != GN_PT_CONVERGED_DIST) continue;
            const double tau = (t[i] - a) / h;
            double B[kPcMaxDeg + 1];
            for (int aa = 0; aa < k; aa++) {
                double s = 0.0, pw = 1.0;
                for (int r = 0; r < k; r++) {
                    s += pw * M[r * k + aa];
                    pw *= tau;
                }
                B[aa] = s;
            }
            int q = 0;
            for (int aa = 0; aa < k; aa++)
                for (int bb = aa; bb < k; bb++) acc[q++] += B[aa] * B[bb];
            for (int aa = 0; aa < k; aa++) {
                acc[ntri + 2 * aa] += B[aa] * u[i];
                acc[ntri + 2 * aa + 1] += B[aa] * v[i];
            }
        }
    }
    // block sum: within each wave by shuffles, then across the waves in LDS
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < kPcVals; q++) {
        double x = acc[q];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
        if (lane == 0) red[wv][q] = x;
    }
    __syncthreads();
    if (threadIdx.x < kPcVals) {
        double s = 0.0;
        for (int w = 0; w < kPcBlock / 64; w++) s += red[w][threadIdx.x];
        rec[((size_t)blockIdx.x * chunks + chunk) * kPcVals + threadIdx.x] = s;
    }
}

// Power-basis coefficients of the fitted curve on each span: coef[span][0..d] for u, [d+1..] for v.
__global__ void k_pc_coef(int d, const double *__restrict__ kn, int ncp,
                          const double *__restrict__ cp, double *__restrict__ coef) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x, nspan = ncp - d, k = d + 1;
    if (s >= nspan) return;
    const int j = s + d;
    double *o = coef + (size_t)s * 2 * k;
    for (int r = 0; r < 2 * k; r++) o[r] = 0.0;
    if (kn[j + 1] - kn[j] <= 0.0) return;
    double M[GN_MAXK * GN_MAXK];
    basis_matrix(d, kn, j, M);
    for (int r = 0; r < k; r++)
        for (int aa = 0; aa < k; aa++) {
            o[r] += M[r * k + aa] * cp[2 * (j - d + aa)];
            o[k + r] += M[r * k + aa] * cp[2 * (j - d + aa) + 1];
        }
}

// (u, v) of the fitted curve at every sample
__global__ void k_pc_eval(int d, const double *__restrict__ kn, int ncp,
                          const double *__restrict__ coef, const double *__restrict__ t,
                          long long m, double *__restrict__ uu, double *__restrict__ vv) {
    const int k = d + 1;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < m;
         i += (long long)gridDim.x * blockDim.x) {
        const int j = find_span(kn, d, ncp, t[i]);
        const double h = kn[j + 1] - kn[j];
        const double tau = h > 0.0 ? (t[i] - kn[j]) / h : 0.0;
        const double *c = coef + (size_t)(j - d) * 2 * k;
        double a = c[d], b = c[k + d];
        for (int r = d - 1; r >= 0; r--) {
            a = fma(a, tau, c[r]);
            b = fma(b, tau, c[k + r]);
        }
        uu[i] = a;
        vv[i] = b;
    }
}

// per-block maximum of |S - C| over the samples (S, C: SoA blocks of m)
__global__ __launch_bounds__(kPcBlock) void k_pc_dev(const double *__restrict__ S,
                                                     const double *__restrict__ C, long long m,
                                                     double *__restrict__ bmax) {
    double best = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < m;
         i += (long long)gridDim.x * blockDim.x) {
        const double dx = S[i] - C[i], dy = S[m + i] - C[m + i], dz = S[2 * m + i] - C[2 * m + i];
        const double e = sqrt(dx * dx + dy * dy + dz * dz);
        best = e > best || e != e ? e : best;  // a NaN deviation is reported as NaN
    }
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_down(best, off, 64);
        best = o > best || o != o ? o : best;
    }
    __shared__ double red[kPcBlock / 64];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = red[0];
        for (int w = 1; w < kPcBlock / 64; w++) b = red[w] > b || red[w] != red[w] ? red[w] : b;
        bmax[blockIdx.x] = b;
    }
}

// Banded Cholesky solve of A x = rhs (2 right-hand sides), A symmetric with half-bandwidth d,
// stored band[i][e] = A[i][i + e], e = 0..d.  False if A is not numerically positive definite.
static bool band_solve(int n, int d, std::vector<double> &band, std::vector<double> &rhs) {
    auto at = [&](int i, int e) -> double & { return band[(size_t)i * (d + 1) + e]; };
    double dmax = 0.0;
    for (int i = 0; i < n; i++) dmax = std::max(dmax, at(i, 0));
    if (!(dmax > 0.0)) return false;
    for (int i = 0; i < n; i++) {  // L stored in place of the upper band (row i holds L[i+e][i])
        double s = at(i, 0);
        for (int kk = std::max(0, i - d); kk < i; kk++) s -= at(kk, i - kk) * at(kk, i - kk);
        if (!(s > 1e-13 * dmax)) return false;
        const double lii = std::sqrt(s);
        at(i, 0) = lii;
        for (int e = 1; e <= d && i + e < n; e++) {
            double x = at(i, e);
            for (int kk = std::max(0, i + e - d); kk < i; kk++)
                x -= at(kk, i - kk) * at(kk, i + e - kk);
            at(i, e) = x / lii;
        }
    }
    for (int c = 0; c < 2; c++) {
        for (int i = 0; i < n; i++) {  // L y = b
            double s = rhs[2 * i + c];
            for (int kk = std::max(0, i - d); kk < i; kk++) s -= at(kk, i - kk) * rhs[2 * kk + c];
            rhs[2 * i + c] = s / at(i, 0);
        }
        for (int i = n - 1; i >= 0; i--) {  // L^T x = y
            double s = rhs[2 * i + c];
            for (int e = 1; e <= d && i + e < n; e++) s -= at(i, e) * rhs[2 * (i + e) + c];
            rhs[2 * i + c] = s / at(i, 0);
        }
    }
    return true;
}

}  // namespace gn

using gn::fail;

extern "C" int gn_pcurve(const gn_surface *s, const gn_curve *c, size_t samples, int degree,
                         int ncp, const gn_invert_opts *opts, double *cp_uv, double *knots,
                         double *max_dev, size_t *failed, gn_stream stream) {
    using namespace gn;
    if (!s || !c || !cp_uv || !knots) return fail(GN_EINVAL, "null argument");
    if (degree < 1 || degree > kPcMaxDeg) return fail(GN_EINVAL, "p-curve degree must be 1..5");
    if (ncp < degree + 1) return fail(GN_EINVAL, "p-curve needs at least degree + 1 control points");
    if (samples < 2 || samples < (size_t)ncp)
        return fail(GN_EINVAL, "need at least 2 samples and at least as many samples as control points");
    if (samples > ((size_t)1 << 31)) return fail(GN_EINVAL, "at most 2^31 samples");
    if (opts && opts->seed_mode == GN_SEED_GIVEN)
        return fail(GN_EINVAL, "p-curve samples have no given seeds: use a tree seed mode");
    const int d = degree, nspan = ncp - d, k = d + 1;
    const double t0 = c->k[c->p], t1 = c->k[c->n];
    if (!(t1 > t0)) return fail(GN_EINVAL, "curve has an empty parameter domain");
    // clamped uniform knots of the p-curve over the curve's domain
    std::vector<double> kn((size_t)ncp + d + 1);
    for (int i = 0; i <= d; i++) {
        kn[i] = t0;
        kn[ncp + i] = t1;
    }
    for (int i = 1; i < nspan; i++) kn[d + i] = t0 + (t1 - t0) * ((double)i / (double)nspan);
    const long long m = (long long)samples;
    const int chunks = (int)std::min<long long>(64, std::max<long long>(1, m / nspan / 8192));
    const int dev_blocks = (int)grid_for((size_t)m, kPcBlock, 1024);
    // scratch: t, C (3m), u, v, uu, vv, S (3m), status, knots, records, coefficients, block maxima
    const size_t nd = (size_t)m * 11 + kn.size() + (size_t)nspan * chunks * kPcVals +
                      (size_t)nspan * 2 * k + 2 * (size_t)ncp + dev_blocks;
    char *buf = nullptr;
    GN_HIP(hipMalloc((void **)&buf, nd * sizeof(double) + (size_t)m + 256));
    double *pt = (double *)buf, *pC = pt + m, *pu = pC + 3 * m, *pv = pu + m, *puu = pv + m,
           *pvv = puu + m, *pS = pvv + m, *pkn = pS + 3 * m, *prec = pkn + kn.size(),
           *pcoef = prec + (size_t)nspan * chunks * kPcVals, *pcp = pcoef + (size_t)nspan * 2 * k,
           *pmax = pcp + 2 * (size_t)ncp;
    uint8_t *pst = (uint8_t *)(pmax + dev_blocks);
    hipStream_t st = (hipStream_t)stream;
    auto done = [&](int rc) {
        (void)hipStreamSynchronize(st);
        (void)hipFree(buf);
        return rc;
    };
    int rc = GN_OK;
#define PC_HIP(call)                                                   \
    do {                                                               \
        hipError_t _e = (call);                                        \
        if (_e != hipSuccess) return done(gn::hip_fail(_e, #call));    \
    } while (0)
    PC_HIP(hipMemcpyAsync(pkn, kn.data(), kn.size() * sizeof(double), hipMemcpyHostToDevice, st));
    k_pc_params<<<grid_for((size_t)m, 256, 4096), 256, 0, st>>>(t0, t1, m, pt);
    PC_HIP(hipGetLastError());
    if ((rc = gn_curve_evaluate(c, pt, (size_t)m, 0, pC, stream))) return done(rc);
    if ((rc = gn_invert(s, pC, pC + m, pC + 2 * m, (size_t)m, opts, pu, pv, pst, stream)))
        return done(rc);
    k_pc_normal<<<dim3(nspan, chunks), kPcBlock, 0, st>>>(d, pkn, ncp, pt, pu, pv, pst, m, prec);
    PC_HIP(hipGetLastError());
    std::vector<double> rec((size_t)nspan * chunks * kPcVals);
    std::vector<uint8_t> hst((size_t)m);
    PC_HIP(hipMemcpyAsync(rec.data(), prec, rec.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    PC_HIP(hipMemcpyAsync(hst.data(), pst, (size_t)m, hipMemcpyDeviceToHost, st));
    PC_HIP(hipStreamSynchronize(st));
    size_t nfail = 0;
    for (uint8_t x : hst) nfail += x != GN_PT_CONVERGED && x != GN_PT_CONVERGED_DIST;
    // assemble the banded normal equations from the span records and solve
    std::vector<double> band((size_t)ncp * k, 0.0), rhs(2 * (size_t)ncp, 0.0);
    const int ntri = k * (k + 1) / 2;
    for (int sp = 0; sp < nspan; sp++)
        for (int ch = 0; ch < chunks; ch++) {
            const double *r = rec.data() + ((size_t)sp * chunks + ch) * kPcVals;
            int q = 0;
            for (int aa = 0; aa < k; aa++)
                for (int bb = aa; bb < k; bb++) band[(size_t)(sp + aa) * k + (bb - aa)] += r[q++];
            for (int aa = 0; aa < k; aa++) {
                rhs[2 * (sp + aa)] += r[ntri + 2 * aa];
                rhs[2 * (sp + aa) + 1] += r[ntri + 2 * aa + 1];
            }
        }
    if (!band_solve(ncp, d, band, rhs))
        return done(fail(GN_EINVAL, "too few converged samples to determine every p-curve "
                                    "control point (use more samples or fewer control points)"));
    for (size_t i = 0; i < 2 * (size_t)ncp; i++) cp_uv[i] = rhs[i];
    for (size_t i = 0; i < kn.size(); i++) knots[i] = kn[i];
    if (failed) *failed = nfail;
    if (max_dev) {
        PC_HIP(hipMemcpyAsync(pcp, rhs.data(), rhs.size() * sizeof(double), hipMemcpyHostToDevice, st));
        k_pc_coef<<<(nspan + 63) / 64, 64, 0, st>>>(d, pkn, ncp, pcp, pcoef);
        PC_HIP(hipGetLastError());
        k_pc_eval<<<grid_for((size_t)m, 256, 4096), 256, 0, st>>>(d, pkn, ncp, pcoef, pt, m, puu, pvv);
        PC_HIP(hipGetLastError());
        if ((rc = gn_evaluate(s, puu, pvv, (size_t)m, 0, pS, stream))) return done(rc);
        k_pc_dev<<<dev_blocks, kPcBlock, 0, st>>>(pS, pC, m, pmax);
        PC_HIP(hipGetLastError());
        std::vector<double> bm(dev_blocks);
        PC_HIP(hipMemcpyAsync(bm.data(), pmax, bm.size() * sizeof(double), hipMemcpyDeviceToHost, st));
        PC_HIP(hipStreamSynchronize(st));
        double best = 0.0;
        for (double x : bm) best = x > best || x != x ? x : best;
        *max_dev = best;
    }
#undef PC_HIP
    return done(GN_OK);
}
