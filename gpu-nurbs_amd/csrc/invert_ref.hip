// invert_ref.hip -- the reference's inversion chain (NewInverse, src/kernel/kernel.cu:433-469):
// scan the valid tree nodes in index order; Gauss-Newton (plain FP64) from the uv centre of every
// closed box that holds P, until one run succeeds.
#include "invert_common.hpp"

namespace gn {

template <int K1, int K2, bool EXACT>
__global__ __launch_bounds__(kInvBlock) void k_invert_first_hit(GnSurfDev d_in, InvArgs a,
                                                                const GnTreeNode *__restrict__ tree,
                                                                const uint8_t *__restrict__ valid,
                                                                long count) {
    __shared__ double sk[EXACT ? 2 * kLdsKnotsInv : 1];
    const double *ku = d_in.ku, *kv = d_in.kv;
    GnSurfDev d = d_in;
    if constexpr (EXACT) {
        d = stage_knots_lds(d_in, sk);
        ku = sk;
        kv = sk + d.nu + d.p + 1;
    }
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

struct FirstHitLaunch {
    const GnSurfDev &d;
    const InvArgs &a;
    const GnTreeNode *tree;
    const uint8_t *valid;
    long count;
    hipStream_t st;
    template <int K1, int K2, bool EXACT>
    int operator()() {
        const unsigned grid = (unsigned)((a.n + kInvBlock - 1) / kInvBlock);
        k_invert_first_hit<K1, K2, EXACT><<<grid, kInvBlock, 0, st>>>(d, a, tree, valid, count);
        GN_HIP(hipGetLastError());
        return GN_OK;
    }
};

int launch_gn_first_hit(const GnSurfDev &d, const InvArgs &a, const GnTreeNode *tree,
                        const uint8_t *valid, long count, hipStream_t st) {
    return dispatch_inv(d, FirstHitLaunch{d, a, tree, valid, count, st});
}

}  // namespace gn
