// invert_mixed.hip -- mixed-precision inversion: FP32 bulk iterations against the FP32 span
// table (persistent lane-refill kernel), then FP64 closing iterations whose residual, rational
// weight division and 2x2 solve are FP64 (one lane per point).
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
            auto k = k_gn_refill<K1, K2, EXACT, float, true, true>;
            refill_shape((const void *)k, a.n, grid, slices);
            k<<<grid, kInvBlock, 0, st>>>(d, a, ctr, slices, a.o.fp32_iters, 0.0, a.o.tol_step32);
        } else {
            auto k = k_gn_refill<K1, K2, EXACT, float, false, true>;
            refill_shape((const void *)k, a.n, grid, slices);
            k<<<grid, kInvBlock, 0, st>>>(d, a, ctr, slices, a.o.fp32_iters, 0.0, a.o.tol_step32);
        }
        const unsigned tgrid = grid_for(a.n, kInvBlock, (unsigned)cu_count() * 16);
        if (a.o.newton & GN_NEWTON_FP64)
            k_gn_tail<K1, K2, EXACT, true><<<tgrid, kInvBlock, 0, st>>>(d, a);
        else
            k_gn_tail<K1, K2, EXACT, false><<<tgrid, kInvBlock, 0, st>>>(d, a);
        GN_HIP(hipGetLastError());
        return GN_OK;
    }
};

int launch_gn_mixed(const GnSurfDev &d, const InvArgs &a, hipStream_t st) {
    return dispatch_degrees(d.p, d.q, MixedLaunch{d, a, st});
}

}  // namespace gn
