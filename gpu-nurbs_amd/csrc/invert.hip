// invert.hip -- point inversion / projection by Gauss-Newton (GaussNewton and NewInverse,
// src/kernel/kernel.cu:151-249, 433-469).
//
// Three kernels, one lane per query point:
//   k_invert_fp64       the plain FP64 drop-in: the reference iteration (J = [S_u S_v],
//                       du = (J^T J)^-1 J^T (S - P)) from a nearest-leaf or caller seed;
//   k_invert_first_hit  the reference's seeding too: scan valid tree nodes in index order and run
//                       Gauss-Newton from the centre of each box holding P until one succeeds;
//   k_invert_mixed      FP32 bulk iterations against the FP32 span table, then closing FP64
//                       iterations (optionally with the exact Hessian of |S - P|^2 / 2) whose
//                       residual, weight division and 2x2 solve are FP64.
// The parameter state (u, v) is always held in FP64, so the FP32 phase is limited only by the
// FP32 evaluation, not by the resolution of an FP32 parameter.  A wave leaves its loop as soon
// as every lane has met its stopping rule (the loop condition is evaluated per lane, and the
// hardware exec mask retires finished lanes).
#include "gn_device.hpp"

namespace gn {

constexpr int kInvBlock = 256;
constexpr int kLdsKnotsInv = 1024;

__device__ __forceinline__ void stage_knots_inv(const GnSurfDev &d, double *sk, const double *&ku,
                                                const double *&kv) {
    const int nku = d.nu + d.p + 1, nkv = d.nv + d.q + 1;
    ku = d.ku;
    kv = d.kv;
    if (nku + nkv <= kLdsKnotsInv) {
        for (int i = threadIdx.x; i < nku; i += blockDim.x) sk[i] = d.ku[i];
        for (int i = threadIdx.x; i < nkv; i += blockDim.x) sk[nku + i] = d.kv[i];
        __syncthreads();
        ku = sk;
        kv = sk + nku;
    }
}

// Gauss-Newton on one point in arithmetic T.  Returns the point status; (cu, cv) updated.
template <int K1, int K2, bool EXACT, typename T, bool NEWTON>
__device__ int gn_loop(const GnSurfDev &d, const double *__restrict__ ku,
                       const double *__restrict__ kv, const double P[3], double &cu, double &cv,
                       int max_iter, double tol_dist, double tol_step, int clamp, int &iters) {
    const T Px = T(P[0]), Py = T(P[1]), Pz = T(P[2]);
    const double ulo = ku[d.p], uhi = ku[d.nu], vlo = kv[d.q], vhi = kv[d.nv];
    const T tdist = T(tol_dist), tstep = T(tol_step);
    int st = GN_PT_MAXITER;
    int it = 0;
    for (; it < max_iter; ++it) {
        Cart<T> o;
        eval_point<K1, K2, NEWTON ? 2 : 1, EXACT, T>(d, ku, kv, cu, cv, o);
        const T r0 = o.S[0] - Px, r1 = o.S[1] - Py, r2 = o.S[2] - Pz;
        const T rr = r0 * r0 + r1 * r1 + r2 * r2;
        if (!(rr == rr)) { st = GN_PT_NONFINITE; break; }
        if (tol_dist > 0 && sqrt(rr) < tdist) { st = GN_PT_CONVERGED_DIST; break; }
        T a = o.Su[0] * o.Su[0] + o.Su[1] * o.Su[1] + o.Su[2] * o.Su[2];
        T b = o.Su[0] * o.Sv[0] + o.Su[1] * o.Sv[1] + o.Su[2] * o.Sv[2];
        T c = o.Sv[0] * o.Sv[0] + o.Sv[1] * o.Sv[1] + o.Sv[2] * o.Sv[2];
        const T gu = o.Su[0] * r0 + o.Su[1] * r1 + o.Su[2] * r2;
        const T gv = o.Sv[0] * r0 + o.Sv[1] * r1 + o.Sv[2] * r2;
        if (NEWTON) {
            const T a2 = a + r0 * o.Suu[0] + r1 * o.Suu[1] + r2 * o.Suu[2];
            const T b2 = b + r0 * o.Suv[0] + r1 * o.Suv[1] + r2 * o.Suv[2];
            const T c2 = c + r0 * o.Svv[0] + r1 * o.Svv[1] + r2 * o.Svv[2];
            if (a2 > T(0) && a2 * c2 - b2 * b2 > T(0)) { a = a2; b = b2; c = c2; }
        }
        const T det = a * c - b * b;
        if (!(det == det)) { st = GN_PT_NONFINITE; break; }
        if (det == T(0)) { st = GN_PT_SINGULAR; break; }
        const T idet = T(1) / det;
        const T du = (c * gu - b * gv) * idet;
        const T dv = (a * gv - b * gu) * idet;
        cu -= double(du);
        cv -= double(dv);
        if (clamp) {
            cu = fmin(fmax(cu, ulo), uhi);
            cv = fmin(fmax(cv, vlo), vhi);
        }
        if (!(cu == cu) || !(cv == cv)) { st = GN_PT_NONFINITE; break; }
        if (tol_step > 0 && fabs(du) < tstep && fabs(dv) < tstep) {
            st = GN_PT_CONVERGED;
            ++it;
            break;
        }
    }
    iters = it;
    return st;
}

struct InvArgs {
    const double *x, *y, *z;
    size_t n;
    double *u, *v;
    uint8_t *status;
    uint8_t *it32, *it64;  // optional per-point iteration counts (FP32 bulk, FP64), or null
    gn_invert_opts o;
};

__device__ __forceinline__ uint8_t sat8(int x) { return (uint8_t)(x > 255 ? 255 : x); }

template <int K1, int K2, bool EXACT>
__global__ __launch_bounds__(kInvBlock) void k_invert_fp64(GnSurfDev d, InvArgs a) {
    __shared__ double sk[kLdsKnotsInv];
    const double *ku, *kv;
    stage_knots_inv(d, sk, ku, kv);
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const double P[3] = {a.x[i], a.y[i], a.z[i]};
    double cu = a.u[i], cv = a.v[i];
    int it, st;
    if (a.o.newton & GN_NEWTON_FP64)
        st = gn_loop<K1, K2, EXACT, double, true>(d, ku, kv, P, cu, cv, a.o.max_iter, a.o.tol_dist,
                                                  a.o.tol_step, a.o.clamp, it);
    else
        st = gn_loop<K1, K2, EXACT, double, false>(d, ku, kv, P, cu, cv, a.o.max_iter, a.o.tol_dist,
                                                   a.o.tol_step, a.o.clamp, it);
    a.u[i] = cu;
    a.v[i] = cv;
    a.status[i] = (uint8_t)st;
    if (a.it32) a.it32[i] = 0;
    if (a.it64) a.it64[i] = sat8(it);
}

template <int K1, int K2, bool EXACT>
__global__ __launch_bounds__(kInvBlock) void k_invert_mixed(GnSurfDev d, InvArgs a) {
    __shared__ double sk[kLdsKnotsInv];
    const double *ku, *kv;
    stage_knots_inv(d, sk, ku, kv);
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const double P[3] = {a.x[i], a.y[i], a.z[i]};
    const double su = a.u[i], sv = a.v[i];
    double cu = su, cv = sv;
    int it, st;
    if (a.o.newton & GN_NEWTON_FP32)
        st = gn_loop<K1, K2, EXACT, float, true>(d, ku, kv, P, cu, cv, a.o.fp32_iters, 0.0,
                                                 a.o.tol_step32, a.o.clamp, it);
    else
        st = gn_loop<K1, K2, EXACT, float, false>(d, ku, kv, P, cu, cv, a.o.fp32_iters, 0.0,
                                                  a.o.tol_step32, a.o.clamp, it);
    const int it_bulk = it;
    if (st == GN_PT_SINGULAR || st == GN_PT_NONFINITE) {
        // FP32 breakdown: restart from the seed with the plain FP64 iteration
        cu = su;
        cv = sv;
        st = gn_loop<K1, K2, EXACT, double, false>(d, ku, kv, P, cu, cv, a.o.max_iter,
                                                   a.o.tol_dist, a.o.tol_step, a.o.clamp, it);
    } else if (a.o.newton & GN_NEWTON_FP64) {
        st = gn_loop<K1, K2, EXACT, double, true>(d, ku, kv, P, cu, cv, a.o.fp64_tail,
                                                  a.o.tol_dist, a.o.tol_step, a.o.clamp, it);
    } else {
        st = gn_loop<K1, K2, EXACT, double, false>(d, ku, kv, P, cu, cv, a.o.fp64_tail,
                                                   a.o.tol_dist, a.o.tol_step, a.o.clamp, it);
    }
    a.u[i] = cu;
    a.v[i] = cv;
    a.status[i] = (uint8_t)st;
    if (a.it32) a.it32[i] = sat8(it_bulk);
    if (a.it64) a.it64[i] = sat8(it);
}

// Reference seeding (NewInverse): index-order scan of the valid tree nodes; Gauss-Newton from
// the centre of every closed box that holds P until one run succeeds.
template <int K1, int K2, bool EXACT>
__global__ __launch_bounds__(kInvBlock) void k_invert_first_hit(GnSurfDev d, InvArgs a,
                                                                const GnTreeNode *__restrict__ tree,
                                                                const uint8_t *__restrict__ valid,
                                                                long count) {
    __shared__ double sk[kLdsKnotsInv];
    const double *ku, *kv;
    stage_knots_inv(d, sk, ku, kv);
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const double P[3] = {a.x[i], a.y[i], a.z[i]};
    int st = GN_PT_NOSEED, total = 0;
    double ru = NAN, rv = NAN;
    for (long t = 0; t < count; t++) {
        if (!valid[t]) continue;
        const GnTreeNode &nd = tree[t];
        if (!((P[0] >= nd.lb[0]) && (P[1] >= nd.lb[1]) && (P[2] >= nd.lb[2]) &&
              (P[0] <= nd.rt[0]) && (P[1] <= nd.rt[1]) && (P[2] <= nd.rt[2])))
            continue;
        double cu = (nd.uv_lb[0] + nd.uv_rt[0]) / 2, cv = (nd.uv_lb[1] + nd.uv_rt[1]) / 2;
        int it;
        st = gn_loop<K1, K2, EXACT, double, false>(d, ku, kv, P, cu, cv, a.o.max_iter,
                                                   a.o.tol_dist, a.o.tol_step, a.o.clamp, it);
        ru = cu;
        rv = cv;
        total += it;
        if (st == GN_PT_CONVERGED || st == GN_PT_CONVERGED_DIST) break;
    }
    a.u[i] = ru;
    a.v[i] = rv;
    a.status[i] = (uint8_t)st;
    if (a.it32) a.it32[i] = 0;
    if (a.it64) a.it64[i] = sat8(total);
}

struct InvLaunch {
    const GnSurfDev &d;
    InvArgs a;
    const GnTreeNode *tree;
    const uint8_t *valid;
    long count;
    hipStream_t st;
    template <int K1, int K2, bool EXACT>
    int operator()() {
        const unsigned grid = (unsigned)((a.n + kInvBlock - 1) / kInvBlock);
        if (a.o.seed_mode == GN_SEED_REF_FIRST_HIT)
            k_invert_first_hit<K1, K2, EXACT><<<grid, kInvBlock, 0, st>>>(d, a, tree, valid, count);
        else if (a.o.precision == GN_MIXED)
            k_invert_mixed<K1, K2, EXACT><<<grid, kInvBlock, 0, st>>>(d, a);
        else
            k_invert_fp64<K1, K2, EXACT><<<grid, kInvBlock, 0, st>>>(d, a);
        GN_HIP(hipGetLastError());
        return GN_OK;
    }
};

int launch_invert(const GnSurfDev &d, const GnSearchDev &sd, const GnTreeNode *tree,
                  const uint8_t *valid, int tree_levels, const double *x, const double *y,
                  const double *z, size_t n, const gn_invert_opts &o, double *u, double *v,
                  uint8_t *status, uint8_t *it32, uint8_t *it64, hipStream_t st) {
    if (n == 0) return GN_OK;
    if (o.seed_mode == GN_SEED_NEAREST_LEAF) {
        int rc = launch_seed(sd, x, y, z, n, nullptr, u, v, st);
        if (rc) return rc;
    }
    long count = 0, g = 1;
    for (int l = 0; l < tree_levels; l++) { count += g * g; g *= 8; }
    InvArgs a{x, y, z, n, u, v, status, it32, it64, o};
    return dispatch_degrees(d.p, d.q, InvLaunch{d, a, tree, valid, count, st});
}

}  // namespace gn
