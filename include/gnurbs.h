/*
 * gnurbs.h -- C ABI of libgnurbs: matrix-form NURBS evaluation and inverse evaluation on one
 * AMD Instinct MI355X (gfx950).
 *
 * The reference (Qiang-Zou/GPU-NURBS) has no library ABI: its operator API is a set of CUDA
 * kernels launched by src/main.cu.  Each entry point below names the reference interface it
 * replaces.  Conventions kept from the reference:
 *   - the caller owns query and result buffers, which are device memory (src/main.cu:176-233);
 *   - control points are u-major, cp[(iu*nv + iv)*3 + c] (src/main.cu:97-110), weights separate;
 *   - span indices are the knot indices returned by findSpan (src/NURBS/Evaluate.cu:504-533);
 *   - tree node indices are base_L + bx*8^L + by with base_L = (64^L - 1)/63
 *     (src/kernel/kernel.cu:406-407, 546-548; src/kernel/kernel.cuh:13-28).
 * Differences: errors are returned as gn_status codes, never exit(); per-point inversion outcome
 * is a status byte instead of an uninitialised flag; all work runs on the caller's stream.
 *
 * All functions are thread-safe for distinct objects; a surface may be used concurrently by
 * several threads for read-only calls (evaluate, invert) once its tables are built.
 */
#ifndef GNURBS_H
#define GNURBS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gn_surface gn_surface; /* opaque: knots, span tables and tree, resident in HBM */
typedef struct gn_curve gn_curve;     /* opaque: rational B-spline curve, resident in HBM    */
typedef struct gn_batch gn_batch;     /* opaque: packed span tables of many surfaces          */
typedef struct gn_server gn_server;   /* opaque: continuous-batching inversion server         */
typedef void *gn_stream;              /* hipStream_t; NULL is the null stream                 */

typedef enum {
    GN_OK = 0,
    GN_EINVAL = 1,     /* bad argument (sizes, degrees, knots, null pointers)           */
    GN_ENOMEM = 2,     /* device or host allocation failed                              */
    GN_EDEVICE = 3,    /* HIP runtime error                                             */
    GN_EIO = 4,        /* file could not be read or parsed                              */
    GN_ESTATE = 5,     /* object not ready (e.g. tree not built)                        */
    GN_EBUDGET = 6,    /* request can never fit the server's memory budget              */
    GN_ENOTFOUND = 7,  /* unknown request or object id                                  */
    GN_EAGAIN = 8      /* request not finished yet (gn_server_poll)                     */
} gn_status;

typedef enum { GN_FP64 = 0, GN_FP32 = 1, GN_MIXED = 2 } gn_precision;

/* per-point inversion outcome (status byte) */
enum {
    GN_PT_CONVERGED = 0,      /* step below tol_step                                      */
    GN_PT_CONVERGED_DIST = 1, /* |S - P| below tol_dist (the reference's rule)            */
    GN_PT_MAXITER = 2,        /* iteration cap reached                                    */
    GN_PT_SINGULAR = 3,       /* det(J^T J) == 0                                          */
    GN_PT_NONFINITE = 4,      /* NaN/Inf met                                              */
    GN_PT_NOSEED = 5          /* first-hit seeding found no box containing P              */
};

typedef enum {
    GN_SEED_NEAREST_LEAF = 0, /* leaf of the tree whose centre point is nearest to P       */
    GN_SEED_REF_FIRST_HIT = 1,/* reference NewInverse: index-order scan, first box hit     */
    GN_SEED_GIVEN = 2,        /* caller supplies seeds in u_out / v_out                    */
    GN_SEED_GREEDY_LEAF = 3,  /* greedy box-distance descent, then the nearest leaf centre */
    GN_SEED_GLOBAL = 4,       /* global projection (FP64): Newton from every leaf whose box
                                 could hold a nearer point and whose normal cone could hold the
                                 foot point's normal (branch and bound over the seed tree), each
                                 clamped to its leaf cell, then polished; the nearest foot point
                                 wins.  iters_fp32 of gn_invert_iters receives the number of
                                 leaves solved.  The reference's normal-cone projection
                                 (src/kernel/kernel.cu:471-513, 623-704) is unfinished. */
    GN_SEED_GREEDY_STEP = 5,  /* the greedy descent's leaf (GN_SEED_GREEDY_LEAF without its 3 x 3
                                 neighbour refinement), then one Gauss-Newton step from the
                                 leaf's seed with the Jacobian stored for that leaf (S_u, S_v at
                                 its seed), limited to one leaf cell, and two Gauss-Newton
                                 iterations on the leaf's stored quadratic Taylor model (S_uu,
                                 S_uv, S_vv too), limited to two cells, clamped to the domain:
                                 no surface evaluation, and about 1.5 iterations fewer after it */
    GN_SEED_GRID_STEP = 6     /* the leaf that gn_build_seed_grid's uniform grid holds for the
                                 cell of P (one lookup instead of the descent), then the step of
                                 GN_SEED_GREEDY_STEP from that leaf; needs the grid */
} gn_seed_mode;

/* tree build flags */
enum { GN_TREE_REF_PARENT = 1 /* parent = (index-1)/64 as src/kernel/kernel.cu:380,599 */ };

typedef struct {
    double tol_dist;    /* stop when |S - P| < tol_dist, checked before the update; <= 0 off  */
    double tol_step;    /* stop when max(|du|,|dv|) < tol_step after an update; <= 0 off      */
    int max_iter;       /* FP64 (plain) iteration cap; the reference uses 10                  */
    int clamp;          /* clamp (u,v) to the knot domain after every update                  */
    int precision;      /* GN_FP64 (plain) or GN_MIXED                                        */
    int fp32_iters;     /* GN_MIXED: cap of FP32 bulk iterations                              */
    int fp64_tail;      /* GN_MIXED: FP64 closing iterations                                  */
    double tol_step32;  /* GN_MIXED: FP32 bulk stops below this step                          */
    int seed_mode;      /* gn_seed_mode                                                       */
    int newton;         /* bitmask of GN_NEWTON_*: iterations that add the residual-curvature  */
                        /* term sum_k r_k d2S_k to J^T J (exact Hessian of |S-P|^2/2, falling  */
                        /* back to J^T J where that is not positive definite)                  */
    double tol_quad;    /* FP64 exact-Hessian (Newton) steps also stop (converged) once a      */
                        /* bound on the (u,v) error the step leaves is below tol_quad: the     */
                        /* bound is |H^-1| L |step|^2 / 2 with L bounding the third derivative */
                        /* of |S-P|^2/2 from S', S'' and the span length, so a flat minimum    */
                        /* (small eigenvalue of H) needs a smaller step; points that never     */
                        /* meet it or tol_step end GN_PT_MAXITER.  Not applied to a step that   */
                        /* fell back to J^T J, nor to a step that leaves the knot span pair it */
                        /* was evaluated in (S'' may jump at a knot).  Default 1e-10; <= 0 off */
    int fallback;       /* gn_invert: points that end unconverged are solved again by the      */
                        /* global projection (GN_SEED_GLOBAL, FP64, same max_iter / tol_step), */
                        /* keeping the nearer of the two results -- the role of the legacy     */
                        /* subdivision inversion's robust seeding (src/kernel/kernel.cu:4-149, */
                        /* commented out at src/main.cu:270); needs the tree.  0: off          */
} gn_invert_opts;

enum { GN_NEWTON_FP32 = 1, /* FP32 bulk iterations of GN_MIXED                                */
       GN_NEWTON_FP64 = 2  /* FP64 iterations (GN_FP64, and the closing iterations of GN_MIXED) */
};

/* Defaults: the reference iteration in plain FP64 (tol_dist 0, tol_step 1e-10, max_iter 20,
 * clamp 1, greedy-leaf seeds, newton 0, tol_quad 1e-10 for Newton steps); mixed: <= 20 FP32
 * iterations to a 1e-6 step, then <= 3 FP64 closing iterations. */
void gn_invert_opts_default(gn_invert_opts *opts);

const char *gn_last_error(void);       /* message of the last failure on this thread */
const char *gn_version(void);

/* ---- surfaces ---------------------------------------------------------------------------- */

/* Tuple constructor: (degree_u, degree_v, cp_u, cp_v, knots_u, knots_v, control_points,
 * weights) as passed to GetMatrixN (src/NURBS/Evaluate.cu:3-4).  Host pointers; copied.
 * Degrees 1..7, knots non-decreasing.  Builds the FP64 and FP32 span tables on `stream`. */
int gn_surface_create(int degree_u, int degree_v, int cp_u, int cp_v, const double *knots_u,
                      const double *knots_v, const double *xyz, const double *weights,
                      gn_stream stream, gn_surface **out);
/* readNu (src/main.cu:45-113): NU/CP/KNu/KNv/DP text format. */
int gn_surface_from_nu(const char *path, gn_stream stream, gn_surface **out);
/* The writer the reference lacks (its models are read-only inputs, src/main.cu:45-113): the same
 * tagged format, 17 significant digits, so gn_surface_from_nu reads back the identical surface. */
int gn_surface_to_nu(const gn_surface *s, const char *path);
int gn_surface_destroy(gn_surface *s);
int gn_surface_info(const gn_surface *s, int *degree_u, int *degree_v, int *cp_u, int *cp_v);
/* HBM bytes held by the object (span tables, knots, tree). */
size_t gn_surface_device_bytes(const gn_surface *s);

/* findSpan (src/NURBS/Evaluate.cu:504-533) for n parameters: dir 0 = u, 1 = v.
 * u, span: device pointers. */
int gn_find_span(const gn_surface *s, int dir, const double *u, size_t n, int32_t *span,
                 gn_stream stream);

/* GetMatrixN (src/NURBS/Evaluate.cu:3-88): rebuilds the span tables (done by create). */
int gn_build_span_table(gn_surface *s, gn_stream stream);
/* One span matrix in the reference's MatrixN order [c][r][s] (c = xw, yw, zw, w), FP64,
 * host buffer of 4*(p+1)*(q+1) doubles; (i, j) are knot-span indices.  Synchronous. */
int gn_get_span_matrix(const gn_surface *s, int i, int j, double *out);

/* EvaluateRandom / MatrixSurfacePoint / MatrixSurfaceRd* / MatrixSurfaceDerivative{U,V,UU,UV,VV,
 * UUU,UUV,UVV,VVV} (src/NURBS/Evaluate.cu:159-502, 535-1020, 1023-1038).  u, v, out: device
 * pointers.  out is SoA: blocks of n values S.x S.y S.z | Su.xyz | Sv.xyz | Suu.xyz | Suv.xyz |
 * Svv.xyz | Suuu.xyz | Suuv.xyz | Suvv.xyz | Svvv.xyz; max_order 0, 1, 2, 3 writes 3, 9, 18, 30
 * blocks.  Mixed derivatives use the exact quotient rule (the reference's new-version S_uv,
 * S_uuv and S_uvv, src/NURBS/Evaluate.cu:958-990, differ from it unless the weights are constant). */
int gn_evaluate(const gn_surface *s, const double *u, const double *v, size_t n, int max_order,
                double *out, gn_stream stream);
/* FP32 form: parameters are floats (span lookup still compares against FP64 knots). */
int gn_evaluate_f32(const gn_surface *s, const float *u, const float *v, size_t n,
                    int max_order, float *out, gn_stream stream);
/* Tensor-product grid (us[nu_pts] x vs[nv_pts]), FP32 MFMA span-GEMM form; out SoA blocks of
 * nu_pts*nv_pts values, point (a, b) at a*nv_pts + b.  us, vs ascending; max_order 0..2. */
int gn_evaluate_grid_f32(const gn_surface *s, const float *us, int nu_pts, const float *vs,
                         int nv_pts, int max_order, float *out, gn_stream stream);

/* FP32 span-GEMM form for points in any order (the MFMA counterpart of the FP16 cuBLAS span-GEMM
 * experiment, src/main.cu:533-663, src/kernel/kernel.cu:1558-1604, for unordered points): the
 * points are bucketed by span pair on the device (counting sort), each group of 16 points of one
 * span pair gets its V-side contraction M_c * [V V' V''] from v_mfma_f32_16x16x4f32, and the
 * results are written in bucket order: column k of out (SoA blocks of n values, as
 * gn_evaluate_f32) holds the point whose input index is perm[k]; perm is a permutation of
 * 0..n-1, and the order within one span pair is unspecified.  Degrees <= 3 use the matrix cores;
 * higher degrees are evaluated per point in the same order.  max_order 0..2, n <= 2^32-1;
 * allocates its scratch (16 bytes per point) stream-ordered. */
int gn_evaluate_sorted_f32(const gn_surface *s, const float *u, const float *v, size_t n,
                           int max_order, float *out, uint32_t *perm, gn_stream stream);

/* ---- subdivision tree (GetPoints / GetSecondDerivatives / Max_Sequential_Addressing_Shared /
 * TightBoundingBox, src/kernel/kernel.cu:375-431, 544-621, 707-733) ------------------------- */

/* levels: 8x8-fan-out levels (reference level = 2 builds levels 0 and 1; 3 gives 64x64 leaves,
 * i.e. a quadtree of depth 6).  All nodes are computed; valid/divided follow the reference. */
int gn_build_tree(gn_surface *s, int levels, int flags, gn_stream stream);
/* Host copy of the tree: nodes[13*count] = uv_lb(2) uv_rt(2) box_lb(3) box_rt(3) centre(3);
 * valid[count], divided[count]; count = (64^levels - 1)/63.  Synchronous. */
int gn_get_tree(const gn_surface *s, double *nodes, uint8_t *valid, uint8_t *divided,
                size_t count);
/* Host copy of the leaves' normal cones (NormalBoundingBox + NormalBoxToNormalTreeNode,
 * src/kernel/kernel.cu:623-704, 1605-1636): per leaf of the finest level, in leaf order bx*g+by
 * (g = 8^(levels-1)), 4 floats: unit axis and half-angle in radians; a half-angle of 4 marks a
 * leaf with no bounded cone.  The cone holds the normal S_u x S_v box of the leaf's 9 x 9
 * samples enlarged by the tree's K bound applied to the normal's second derivatives.
 * count >= g*g.  Synchronous. */
int gn_get_normal_cones(const gn_surface *s, float *cones, size_t count);
/* Seed grid for GN_SEED_GRID_STEP (no reference counterpart: the reference seeds by scanning its
 * tree, src/kernel/kernel.cu:433-469).  A uniform grid of cubic cells over the bounding box of
 * the finest leaves' centres (FP32), cell size = its largest extent / resolution, with
 * max(1, resolution/16) cells of margin on every side; each cell holds the leaf whose FP32 centre
 * is nearest to the cell's centre (ties to the lower reference index).  The build compares every
 * cell with every leaf, so a grid may have at most 2^25 cells and 2^36 cell-leaf pairs (about a
 * second); a larger explicit resolution (at most 1024) is GN_EINVAL.  resolution 0 starts at
 * 2 * 2^depth capped at 256 (128 for 3 levels) and halves while the grid is too large.  Needs the
 * tree; gn_build_tree discards the grid.  Cells take 2 bytes (4 beyond 65,536 leaves). */
int gn_build_seed_grid(gn_surface *s, int resolution, gn_stream stream);
/* Host copy of the seed grid: dims, origin and cell size (any may be NULL), and, when leaf is not
 * NULL, the reference index of each cell's leaf at (i * dims[1] + j) * dims[2] + k; count >= the
 * cell count.  Synchronous. */
int gn_get_seed_grid(const gn_surface *s, int32_t *leaf, size_t count, int dims[3],
                     float origin[3], float *cell);
/* Seed search alone: leaf index (reference numbering) and seed (u,v) per point; exact == 0
 * selects GN_SEED_GREEDY_LEAF, exact == 2 GN_SEED_GREEDY_STEP, exact == 3 GN_SEED_GRID_STEP,
 * any other value GN_SEED_NEAREST_LEAF. */
int gn_seed(const gn_surface *s, int exact, const double *x, const double *y, const double *z,
            size_t n, int32_t *leaf, double *seed_u, double *seed_v, gn_stream stream);

/* ---- inversion (NewInverse + GaussNewton, src/kernel/kernel.cu:151-249, 433-469) ----------- */

/* x, y, z: query points (device, SoA).  u_out, v_out, status: device.  With GN_SEED_GIVEN the
 * seeds are read from u_out/v_out.  Requires gn_build_tree unless seeds are given. */
int gn_invert(const gn_surface *s, const double *x, const double *y, const double *z, size_t n,
              const gn_invert_opts *opts, double *u_out, double *v_out, uint8_t *status,
              gn_stream stream);
/* gn_invert plus per-point iteration counts (saturated at 255) of the FP32 bulk and of the FP64
 * iterations; either pointer may be NULL.  Used for algorithmic flop accounting. */
int gn_invert_iters(const gn_surface *s, const double *x, const double *y, const double *z,
                    size_t n, const gn_invert_opts *opts, double *u_out, double *v_out,
                    uint8_t *status, uint8_t *iters_fp32, uint8_t *iters_fp64, gn_stream stream);

/* gn_invert without allocation.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The inversion kernels' scratch -- claim counters for two persistent launches and one
 * continuation list of 2^17 point indices, 528 KB -- lives in a fixed device array of the library
 * (one per GPU, held until the process ends) instead of memory the library allocates; the caller
 * passes inputs and outputs as for gn_invert and gets the same results bit for bit.  Every call on
 * a GPU shares that GPU's array: calls that overlap on one GPU (other streams, other threads) race
 * and must not be issued.  Limits: n <= 2^32-1 points per call (GN_EINVAL beyond), opts->fallback
 * must be 0 (the fallback allocates its list; GN_EINVAL). */
int gn_invert_noalloc(const gn_surface *s, const double *x, const double *y, const double *z,
                      size_t n, const gn_invert_opts *opts, double *u_out, double *v_out,
                      uint8_t *status, gn_stream stream);

/* ---- batched inversion: many surfaces in one persistent launch ---------------------------- */

/* A run of points that belong to one object of a batch: points start .. start+count-1 of the
 * query arrays are inverted on object `obj` (index into the batch's surface list). */
typedef struct {
    int32_t obj;
    int32_t count;
    int64_t start;
} gn_segment;

typedef struct gn_graph gn_graph; /* an instantiated hipGraph of one batched inversion */

/* Packs copies of the surfaces' span tables, knots and seed trees (and seed grids, where built)
 * into one device allocation.  Every surface needs its tree (gn_build_tree); GN_SEED_GRID_STEP
 * needs every surface's seed grid (gn_build_seed_grid) built before this call.  The surfaces may
 * be destroyed afterwards. */
int gn_batch_create(gn_surface *const *surfaces, int count, gn_stream stream, gn_batch **out);
int gn_batch_destroy(gn_batch *b);
size_t gn_batch_device_bytes(const gn_batch *b);
/* Tile shape: points per tile (work unit claimed by one wave) and a cap on resident workgroups
 * per CU (0 = occupancy limit). */
int gn_batch_set_shape(gn_batch *b, int tile_points, int blocks_per_cu);
int gn_batch_get_shape(const gn_batch *b, int *tile_points, int *blocks_per_cu);
/* Inverts the points of all segments (device arrays x, y, z, u, v, status of length n) with one
 * persistent launch per precision phase; seeds: GN_SEED_GREEDY_LEAF, GN_SEED_NEAREST_LEAF,
 * GN_SEED_GREEDY_STEP, GN_SEED_GRID_STEP (each point on its own surface's tree or grid) or
 * GN_SEED_GIVEN. */
int gn_batch_invert(gn_batch *b, const gn_segment *segs, int nseg, const double *x,
                    const double *y, const double *z, size_t n, const gn_invert_opts *opts,
                    double *u, double *v, uint8_t *status, gn_stream stream);
/* Autotuner: times the candidate tile shapes on this work (outputs are overwritten) and keeps
 * the fastest; with cache_path, a shape stored for the same work signature is reused and a newly
 * tuned one is appended to the file.  With GN_SEED_GIVEN every timed run starts from the seeds
 * u and v held on entry (a device copy is kept for the call). */
int gn_batch_tune(gn_batch *b, const gn_segment *segs, int nseg, const double *x, const double *y,
                  const double *z, size_t n, const gn_invert_opts *opts, double *u, double *v,
                  uint8_t *status, const char *cache_path, gn_stream stream);
/* Captures gn_batch_invert on these buffers into a hipGraph (stream must not be NULL); launching
 * it re-runs the inversion on whatever the buffers hold.  One launch of a graph at a time. */
int gn_batch_graph_create(gn_batch *b, const gn_segment *segs, int nseg, const double *x,
                          const double *y, const double *z, size_t n, const gn_invert_opts *opts,
                          double *u, double *v, uint8_t *status, gn_stream stream,
                          gn_graph **out);
int gn_graph_launch(gn_graph *g, gn_stream stream);
int gn_graph_destroy(gn_graph *g);

/* ---- curves ------------------------------------------------------------------------------- */
/* No counterpart in the reference: its README (README.md:10) names curves but src/ has no curve
 * code.  Same matrix form in one parameter: per span, patch = M * P_w with M the span's basis
 * matrix (as GetMatrixN, src/NURBS/Evaluate.cu:3-88, in one direction). */

/* degree 1..7, cp control points (host arrays: knots cp+degree+1, xyz cp*3, w cp); copied. */
int gn_curve_create(int degree, int cp, const double *knots, const double *xyz, const double *w,
                    gn_stream stream, gn_curve **out);
int gn_curve_destroy(gn_curve *c);
int gn_curve_info(const gn_curve *c, int *degree, int *cp);
size_t gn_curve_device_bytes(const gn_curve *c);
/* out (device, SoA blocks of n doubles): C.x C.y C.z, then C' (max_order >= 1), then C''
 * (max_order 2); derivatives with respect to t. */
int gn_curve_evaluate(const gn_curve *c, const double *t, size_t n, int max_order, double *out,
                      gn_stream stream);
/* Point inversion onto the curve: 1-D Gauss-Newton (opts->newton adds the curvature term),
 * seeds from 8 samples per span (or GN_SEED_GIVEN: seeds in t_out).  GN_FP64 or GN_MIXED. */
int gn_curve_invert(const gn_curve *c, const double *x, const double *y, const double *z, size_t n,
                    const gn_invert_opts *opts, double *t_out, uint8_t *status, gn_stream stream);

/* ---- p-curves ---- */
/* The p-curve of curve c on surface s: the use README.md:10 names (inverting a 3-D curve onto a
 * surface and fitting its 2-D pre-image), with no code in the reference.  Samples c at `samples`
 * uniformly spaced parameters over its knot domain, inverts them onto s with opts (the tree must
 * be built unless opts->seed_mode is GN_SEED_GIVEN, which is rejected here: there are no seeds),
 * and fits by least squares a clamped uniform B-spline (u(t), v(t)) of `degree` (1..5) with `ncp`
 * control points over the same domain to the samples that converged.  Host outputs:
 * cp_uv[2*ncp] (u, v per control point), knots[ncp+degree+1]; optional max_dev (largest distance
 * between S(u(t_i), v(t_i)) and c(t_i) over all samples) and failed (samples that did not
 * converge).  Synchronous; allocates its scratch (about 90 bytes per sample). */
int gn_pcurve(const gn_surface *s, const gn_curve *c, size_t samples, int degree, int ncp,
              const gn_invert_opts *opts, double *cp_uv, double *knots, double *max_dev,
              size_t *failed, gn_stream stream);

/* ---- serving: continuous batching of inversion requests ------------------------------------ */
/* No counterpart in the reference (its driver inverts one fixed batch, src/main.cu:277-327).
 * A server owns a worker thread that gathers submitted requests into steps: each step is one
 * batched launch over every admitted request (surfaces: the gn_batch kernel; curves: one
 * workgroup per tile), launched as soon as one of its `streams` step slots is free, so requests
 * that arrive while a step runs join the next one.  Admission: a request's results (u, v, status
 * for surfaces; t in u and status for curves) live in a server-owned HBM arena of size
 * hbm_budget minus the resident objects; requests are admitted in arrival order while their
 * results fit, and stay admitted until gn_server_release.  Results do not depend on what else
 * shares a step. */

typedef struct {
    int64_t submit_ns, admit_ns, launch_ns, done_ns; /* CLOCK_MONOTONIC (gn_now_ns)           */
    const double *u, *v;     /* device results (curves: t in u, v NULL); valid until release */
    const uint8_t *status;
    size_t n;
    int obj;
} gn_request_info;

typedef struct {
    int64_t submitted, completed, steps, points; /* totals since creation                    */
    int64_t max_step_requests;                   /* most requests in one step                */
    size_t arena_bytes, arena_peak;              /* result arena capacity and peak use       */
    int64_t launch_ns_total, launch_ns_max;      /* host time spent launching steps (building */
                                                 /* the step, its copy and its kernel launches) */
} gn_server_stats;

/* Objects 0 .. nsurf-1 are the batch's surfaces (batch may be NULL), nsurf .. nsurf+ncurves-1
 * the curves.  The batch and the curves must outlive the server.  opts: as gn_batch_invert
 * (GN_FP64 or GN_MIXED; GIVEN seeds are not available).  streams: concurrent steps, 1..8.
 * GN_EBUDGET if the objects alone exceed hbm_budget. */
int gn_server_create(gn_batch *batch, gn_curve *const *curves, int ncurves,
                     const gn_invert_opts *opts, size_t hbm_budget, int streams,
                     gn_server **out);
/* Queues n points (device arrays x, y, z, caller-owned, valid until the request is done) for
 * object obj.  GN_EBUDGET if its results could never fit the arena. */
int gn_server_submit(gn_server *s, int obj, const double *x, const double *y, const double *z,
                     size_t n, int64_t *id);
/* GN_OK and info when done; GN_EAGAIN while queued or running; GN_ENOTFOUND for an unknown id;
 * GN_EDEVICE if its step failed. */
int gn_server_poll(gn_server *s, int64_t id, gn_request_info *info);
/* As poll, blocking up to timeout_ms (< 0: no limit). */
int gn_server_wait(gn_server *s, int64_t id, int timeout_ms, gn_request_info *info);
/* Frees a finished request's results (GN_ESTATE while it is still queued or running). */
int gn_server_release(gn_server *s, int64_t id);
int gn_server_stats_get(gn_server *s, gn_server_stats *stats);
/* Finishes the steps in flight, fails the requests not yet launched, frees the arena. */
int gn_server_destroy(gn_server *s);
int64_t gn_now_ns(void);

#ifdef __cplusplus
}
#endif
#endif /* GNURBS_H */
