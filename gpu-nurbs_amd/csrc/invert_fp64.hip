// invert_fp64.hip -- the plain FP64 inversion (the drop-in form of GaussNewton,
// src/kernel/kernel.cu:151-249): persistent lane-refill Gauss-Newton in FP64.
#include "invert_common.hpp"

namespace gn {

struct PlainLaunch {
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
        if (a.o.newton & GN_NEWTON_FP64) {
            auto k = k_gn_refill<K1, K2, EXACT, double, true, kPlain>;
            refill_shape((const void *)k, a.n, grid, slices);
            k<<<grid, kInvBlock, 0, st>>>(d, a, ctr, slices, a.o.max_iter, 0, a.o.tol_dist,
                                          a.o.tol_step, a.o.tol_quad);
        } else {
            auto k = k_gn_refill<K1, K2, EXACT, double, false, kPlain>;
            refill_shape((const void *)k, a.n, grid, slices);
            k<<<grid, kInvBlock, 0, st>>>(d, a, ctr, slices, a.o.max_iter, 0, a.o.tol_dist,
                                          a.o.tol_step, 0.0);
        }
        GN_HIP(hipGetLastError());
        return GN_OK;
    }
};

int launch_gn_plain(const GnSurfDev &d, const InvArgs &a, hipStream_t st) {
    return dispatch_inv(d, PlainLaunch{d, a, st});
}

}  // namespace gn
