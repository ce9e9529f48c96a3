// invert_mixed.hip -- mixed-precision inversion: FP32 bulk iterations against the FP32 span
// table, then FP64 closing iterations whose residual, rational weight division and 2x2 solve are
// FP64; both passes are persistent lane-refill kernels.
#include "invert_common.hpp"

namespace gn {

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
