// invert.hip -- point inversion / projection entry (NewInverse + GaussNewton,
// src/kernel/kernel.cu:151-249, 433-469): seeding, then one of
//   plain FP64    k_gn_refill<double>          (invert_fp64.hip)   the drop-in iteration
//   mixed         k_gn_refill<float> + k_gn_tail (invert_mixed.hip) FP32 bulk, FP64 closing steps
//   reference     k_invert_first_hit           (invert_ref.hip)    index-order first-hit seeding
#include "invert_common.hpp"

namespace gn {

int launch_invert(const GnSurfDev &d, const GnSearchDev &sd, const GnTreeNode *tree,
                  const uint8_t *valid, int tree_levels, const double *x, const double *y,
                  const double *z, size_t n, const gn_invert_opts &o, double *u, double *v,
                  uint8_t *status, uint8_t *it32, uint8_t *it64, hipStream_t st) {
    if (n == 0) return GN_OK;
    InvArgs a{x, y, z, n, u, v, status, it32, it64, o};
    if (o.seed_mode == GN_SEED_GLOBAL) return launch_project_global(d, sd, a, st);
    int rc;
    if (o.seed_mode == GN_SEED_REF_FIRST_HIT) {
        long count = 0, g = 1;
        for (int l = 0; l < tree_levels; l++) { count += g * g; g *= 8; }
        rc = launch_gn_first_hit(d, a, tree, valid, count, st);
    } else {
        if (o.seed_mode != GN_SEED_GIVEN) {
            const int mode = o.seed_mode == GN_SEED_NEAREST_LEAF ? 0
                           : o.seed_mode == GN_SEED_GREEDY_STEP  ? 2
                           : o.seed_mode == GN_SEED_GRID_STEP    ? 3 : 1;
            rc = launch_seed(sd, mode, x, y, z, n, nullptr, u, v, st);
            if (rc) return rc;
        }
        rc = o.precision == GN_MIXED ? launch_gn_mixed(d, a, st) : launch_gn_plain(d, a, st);
    }
    if (rc || !o.fallback) return rc;
    return launch_project_fallback(d, sd, a, st);
}

}  // namespace gn
