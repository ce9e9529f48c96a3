// gn_internal.hpp -- shared definitions of libgnurbs (host and device).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <functional>
#include <string>
#include <utility>
#include <vector>

#include "gnurbs.h"

#define GN_MAXK 8  // k = degree + 1; degrees 1..7 (MAX_DEGREE of src/NURBS/Evaluate.cuh:15)

// ---------------------------------------------------------------------------------------------
// Device view of one surface.  Span slots are indexed by (m - p, n - q) for knot spans
// m in [p, nu-1], n in [q, nv-1]; each slot holds the patch M_u * P_w * M_v^T as
// [r][s][c] with c = (x*w, y*w, z*w, w) fastest, so one (r, s) entry is one 4-vector.
// ---------------------------------------------------------------------------------------------
struct GnSurfDev {
    int p, q, nu, nv;       // degrees and control-net size
    int su, sv;             // span slots per direction: nu - p, nv - q
    const double *ku, *kv;  // knots (nu + p + 1, nv + q + 1)
    const double *ihu, *ihv;  // 1 / (k[m+1] - k[m]) per slot (0 for empty spans)
    const double *tab64;    // su*sv*k1*k2*4 doubles
    const float *tab32;     // same, FP32
};

// Tree node record, FP64, 13 values (matches gn_get_tree)
struct GnTreeNode {
    double uv_lb[2], uv_rt[2], lb[3], rt[3], c[3];
};

// Quadtree search structure used by the nearest-leaf seed search: levels 0..D of 4^l nodes in
// Morton order; boxes contain every leaf centre below them.
constexpr int kRecVec = 8;  // 16-byte vectors per leaf record (GnSearchDev::rec)

struct GnSearchDev {
    int depth;                // D = 3*(levels-1)
    const double *box;        // per node 6 doubles lo.xyz hi.xyz, level offsets (4^l-1)/3
    const float *box32;       // the same boxes in FP32, rounded outward and widened by a margin
    const double *centre;     // per leaf (Morton order) centre point xyz
    const double *seed_uv;    // per leaf (Morton order) seed (u, v)
    const int32_t *ref_index; // per leaf (Morton order) reference tree index
    const float4 *leaf32;     // per leaf (Morton order) centre in FP32 and the reference index
                              // (bits of w), one 16-byte load for the greedy 3 x 3 search
    const float4 *rec;        // per leaf (Morton order) one 128-byte record (kRecVec vectors) for
                              // the stepped seeds, one cache line: [0] as leaf32, [1] the seed
                              // (u, v) as two doubles, [2..6] S_u, S_v, S_uu, S_uv, S_vv at the
                              // seed in FP32 (w = 0), [7] zero
    const float4 *cone32;     // per leaf (Morton order) normal cone: unit axis, half-angle
                              // (radians, rounded up; 4 when unbounded); null: no cone pruning
    double u0, u1, v0, v1;    // the knot domain the leaves tile
    // Seed grid (GN_SEED_GRID_STEP): per cell of a uniform grid over the leaf centres, the Morton
    // index of the leaf whose FP32 centre is nearest to the cell centre; uint16 cells when
    // grid_wide == 0, uint32 otherwise; null: not built.  Cell (i, j, k) is entry
    // (i * gdim[1] + j) * gdim[2] + k and covers gorg + [i, i + 1) / ginv along x (y, z alike).
    const void *grid;
    int grid_wide;
    int gdim[3];
    float gorg[3], ginv;
};

// One object of a batch: the surface and its seed tree (pointers into the batch's pack) and its
// degree class: (p-2)*4 + (q-2) for p, q in 2..5, 16 otherwise.
struct GnBatchObj {
    GnSurfDev s;
    GnSearchDev sd;
    int cls;
    int pad_;
};

// Device view of one curve: span slot m - p holds patch[r][c] = sum_a M[r][a] P_w[m-p+a][c].
struct GnCurveDev {
    int p, n, ns;          // degree, control points, span slots (n - p)
    const double *k, *ih;  // knots, 1/h per slot
    const double *tab64;   // [slot][r][4]
    const float *tab32;
    const float *samp;     // [ns * 8][3] seed sample points (FP32)
    const double *samp_t;  // their parameters
};

// The point arrays of one request (or of one whole batched call): inputs x, y, z and outputs
// u, v (curves: t in u, v unused), status and optional iteration counts, all indexed from 0.
struct GnReq {
    const double *x, *y, *z;
    double *u, *v;
    uint8_t *status, *it32, *it64;
};

// A tile of work: `count` points of object `obj`, points start .. start+count-1 of request `req`.
struct GnTile {
    int32_t obj;
    int32_t count;
    int64_t start;
    int32_t req;
    int32_t pad_;
};

// ---------------------------------------------------------------------------------------------
// error handling
// ---------------------------------------------------------------------------------------------
namespace gn {
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);
int hip_fail(hipError_t e, const char *what);
}  // namespace gn

#define GN_HIP(call)                                              \
    do {                                                          \
        hipError_t _e = (call);                                   \
        if (_e != hipSuccess) return gn::hip_fail(_e, #call);     \
    } while (0)

// ---------------------------------------------------------------------------------------------
// host-side surface object
// ---------------------------------------------------------------------------------------------
struct gn_surface {
    int p, q, nu, nv;
    std::vector<double> ku, kv, xyz, w;  // host copies of the definition
    // device buffers
    double *d_ku = nullptr, *d_kv = nullptr, *d_ihu = nullptr, *d_ihv = nullptr;
    double *d_xyz = nullptr, *d_w = nullptr;
    double *d_tab64 = nullptr;
    float *d_tab32 = nullptr;
    size_t tab_entries = 0;  // su*sv*k1*k2*4
    // tree
    int tree_levels = 0;
    GnTreeNode *d_tree = nullptr;
    uint8_t *d_valid = nullptr, *d_divided = nullptr;
    double *d_sbox = nullptr, *d_scentre = nullptr, *d_sseed = nullptr;
    float *d_sbox32 = nullptr;
    int32_t *d_sref = nullptr;
    float4 *d_sleaf32 = nullptr;
    float4 *d_srec = nullptr;
    float4 *d_scone32 = nullptr;
    // seed grid (gn_build_seed_grid), freed with the tree
    void *d_grid = nullptr;
    int grid_wide = 0, gdim[3] = {0, 0, 0};
    float gorg[3] = {0.f, 0.f, 0.f}, gcell = 0.f, ginv = 0.f;
    size_t device_bytes = 0, tree_bytes = 0, grid_bytes = 0;

    GnSurfDev dev() const {
        GnSurfDev d;
        d.p = p; d.q = q; d.nu = nu; d.nv = nv;
        d.su = nu - p; d.sv = nv - q;
        d.ku = d_ku; d.kv = d_kv; d.ihu = d_ihu; d.ihv = d_ihv;
        d.tab64 = d_tab64; d.tab32 = d_tab32;
        return d;
    }
    GnSearchDev search() const {
        GnSearchDev s;
        s.depth = 3 * (tree_levels - 1);
        s.box = d_sbox; s.box32 = d_sbox32; s.centre = d_scentre; s.seed_uv = d_sseed;
        s.ref_index = d_sref;
        s.leaf32 = d_sleaf32;
        s.rec = d_srec;
        s.cone32 = d_scone32;
        s.grid = d_grid;
        s.grid_wide = grid_wide;
        for (int k = 0; k < 3; k++) {
            s.gdim[k] = gdim[k];
            s.gorg[k] = gorg[k];
        }
        s.ginv = ginv;
        s.u0 = ku.empty() ? 0.0 : ku[p];
        s.u1 = ku.empty() ? 1.0 : ku[nu];
        s.v0 = kv.empty() ? 0.0 : kv[q];
        s.v1 = kv.empty() ? 1.0 : kv[nv];
        return s;
    }
};

// host-side curve object
struct gn_curve {
    int p, n;
    std::vector<double> k, xyz, w;
    double *d_k = nullptr, *d_ih = nullptr, *d_xyz = nullptr, *d_w = nullptr, *d_tab64 = nullptr,
           *d_samp_t = nullptr;
    float *d_tab32 = nullptr, *d_samp = nullptr;
    size_t bytes = 0;
    GnCurveDev dev() const {
        GnCurveDev c;
        c.p = p; c.n = n; c.ns = n - p;
        c.k = d_k; c.ih = d_ih; c.tab64 = d_tab64; c.tab32 = d_tab32;
        c.samp = d_samp; c.samp_t = d_samp_t;
        return c;
    }
};

// ---------------------------------------------------------------------------------------------
// kernel launchers (defined in the .hip files)
// ---------------------------------------------------------------------------------------------
namespace gn {
constexpr int kCurveSamples = 8;  // curve seed samples per span
int launch_span_table(const gn_surface *s, const double *d_xyz, const double *d_w,
                      hipStream_t st);
int launch_find_span(const GnSurfDev &d, int dir, const double *u, size_t n, int32_t *span,
                     hipStream_t st);
int launch_eval64(const GnSurfDev &d, const double *u, const double *v, size_t n, int order,
                  double *out, hipStream_t st);
int launch_eval32(const GnSurfDev &d, const float *u, const float *v, size_t n, int order,
                  float *out, hipStream_t st);
int launch_eval_grid32(const GnSurfDev &d, const float *us, int nu_pts, const float *vs,
                       int nv_pts, int order, float *out, hipStream_t st);
// span-order FP32 evaluation (span_gemm.hip): out in sorted order, perm[k] = input index
int launch_eval_sorted32(const GnSurfDev &d, const float *u, const float *v, size_t n, int order,
                         float *out, uint32_t *perm, hipStream_t st);
int launch_tree_level(const GnSurfDev &d, int level, int flags, GnTreeNode *tree,
                      uint8_t *valid, uint8_t *divided, hipStream_t st);
// Seed grid over the leaves' FP32 centres: cell (i, j, k) of dims gets the Morton index of the
// leaf nearest to its centre org + (i + 1/2, j + 1/2, k + 1/2) * cell (ties to the lower
// reference index), as uint16 (wide == 0) or uint32.
int launch_seed_grid_build(const float4 *leaf32, int depth, const int dims[3], const float org[3],
                           float cell, int wide, void *grid, hipStream_t st);
int launch_search_build(const GnSurfDev &d, int levels, const GnTreeNode *tree, double *sbox,
                        float *sbox32, double *scentre, double *sseed, int32_t *sref,
                        float4 *sleaf32, float4 *srec, float4 *scone32, hipStream_t st);
// mode 0: exact nearest leaf centre; mode 1: greedy descent + 3x3 leaf neighbourhood;
// mode 2: the greedy descent's leaf, then one Gauss-Newton step from the leaf's seed with its
// stored Jacobian; mode 3: the seed grid's leaf, then the same step
int launch_seed(const GnSearchDev &sd, int mode, const double *x, const double *y,
                const double *z, size_t n, int32_t *leaf, double *su, double *sv,
                hipStream_t st);
// Degree classes: (p-2)*4 + (q-2) for p, q in 2..5 (exact kernels), 16 for any other degrees
// (padded generic kernel).  Classes form three groups by patch size (k1*k2 <= 16, <= 25, larger
// and generic); each group has its own kernel, whose register footprint is that of its largest
// class.  Tile lists are sorted by class_order (group, then class).
constexpr int kClasses = 17;
__host__ __device__ constexpr int class_group(int cls) {
    return cls < 0 || cls >= 16 ? 2
         : (cls / 4 + 3) * (cls % 4 + 3) <= 16 ? 0
         : (cls / 4 + 3) * (cls % 4 + 3) <= 25 ? 1 : 2;
}
inline int class_order(int cls) { return class_group(cls) * 32 + cls; }
// A run of tiles of one class group in a class_order-sorted tile list, and its point count.
struct GroupRange {
    int grp, t0, t1;
    size_t points;
    int max_count;  // largest tile
};
std::vector<GroupRange> group_ranges(const GnTile *tiles, size_t ntiles, const std::vector<int> &cls);
// Points per tile for a launch of `points` points: about two tiles per resident wave, between 64
// and cap (a tile is the unit a wave claims; too large a tile leaves waves idle).
int class_tile(size_t points, int waves, int cap);
// Side streams and fork/join events for running independent launches concurrently.  A set is
// used by one caller at a time (the batched entry points share one per device under a lock; a
// server owns one per step slot).
constexpr int kSideStreams = 3;  // one per class group
struct SideStreams {
    hipStream_t s[kSideStreams] = {};
    hipEvent_t fork = nullptr, join[kSideStreams] = {};
};
int side_streams_create(SideStreams *ss);
// Creates the device's shared side-stream set (call before a graph capture).
int batch_prepare();
void side_streams_destroy(SideStreams *ss);
// Per group range: seeds (unless GN_SEED_GIVEN), then one persistent launch per precision
// phase.  The groups run concurrently on the side streams (forked from and joined back into st);
// `beside`, if given, enqueues further work on st that runs alongside them.  With side ==
// nullptr the device's shared set is used.  group_done, if given, receives (group, event) for
// every group that ran on a side stream: the event completes when that group's work has (it is
// the side set's join event, valid until the set's next use); a group run on st itself is not
// listed.
int launch_batch_invert(const GnBatchObj *objs, const GnTile *tiles,
                        const std::vector<GroupRange> &ranges, const GnReq *reqs,
                        const gn_invert_opts &o, int bpc_cap, hipStream_t st,
                        const SideStreams *side = nullptr,
                        const std::function<int(hipStream_t)> &beside = {},
                        std::vector<std::pair<int, hipEvent_t>> *group_done = nullptr);
// resident waves of one persistent batch launch (sizes tiles for short serving steps)
int batch_waves(int bpc_cap);
// tiles of curve objects (obj indexes `curves`); max_samples: largest seed sample count
int launch_curve_batch(const GnCurveDev *curves, const GnTile *tiles, int ntiles, const GnReq *reqs,
                       int max_samples, const gn_invert_opts &o, hipStream_t st);
int ensure_runtime();  // allocates the per-device claim counters (before any graph capture)
// Claim-counter slots (and continuation lists) owned by one client instead of the device's shared
// ring: a captured graph keeps the slots it was captured with for every replay, so they must never
// be handed to another launch; the no-allocation inversion uses a fixed device array.  While a
// pool is set for the calling thread, every launch that thread enqueues takes its slots from the
// pool in order (an error once it is used up).  The counters must start zeroed; every launch
// leaves its slot zeroed again.
struct SlotPool {
    unsigned long long *ctr = nullptr;  // nslots slots of slot_words() counters
    uint32_t *list = nullptr;           // nlists continuation lists of list_entries() indices
    int nslots = 0, nlists = 0;
    int used = 0, lists_used = 0;       // taken so far
};
size_t slot_words();    // 64-bit words per counter slot
size_t list_entries();  // 32-bit indices per continuation list
// Whether the calling thread's launches go to a pool (a graph capture or the no-allocation entry
// point): such launches must not allocate.
bool slot_pool_active();
// Routes the calling thread's launches to `pool` for the lifetime of the object.
class PoolScope {
  public:
    explicit PoolScope(SlotPool *pool);
    ~PoolScope();
    PoolScope(const PoolScope &) = delete;
    PoolScope &operator=(const PoolScope &) = delete;

  private:
    SlotPool *prev_;
};
// Copies `bytes` from pinned host memory (src: its device-side address) to device memory with an
// ordinary kernel on st; both buffers 16-byte aligned and readable/writable to bytes rounded up
// to 16.
int launch_stage_copy(const void *src, void *dst, size_t bytes, hipStream_t st);
// Argument checks of gn_invert (a failing check sets the thread's error message).
int check_invert(const gn_surface *s, const double *x, const double *y, const double *z, size_t n,
                 const gn_invert_opts &o, const double *u, const double *v, const uint8_t *status);
int launch_invert(const GnSurfDev &d, const GnSearchDev &sd, const GnTreeNode *tree,
                  const uint8_t *valid, int tree_levels, const double *x, const double *y,
                  const double *z, size_t n, const gn_invert_opts &o, double *u, double *v,
                  uint8_t *status, uint8_t *it32, uint8_t *it64, hipStream_t st);
}  // namespace gn
