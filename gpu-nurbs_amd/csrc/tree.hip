// tree.hip -- subdivision-tree build and nearest-leaf seed search.
//
// Tree build follows GetPoints / GetSecondDerivatives / Max_Sequential_Addressing_Shared /
// TightBoundingBox / NeedDivide (src/kernel/kernel.cu:375-431, 544-621, 707-733, 859-901):
// node (bx, by) of level L (g = 8^L per side) covers [bx/g, (bx+1)/g] x [by/g, (by+1)/g], is
// sampled on a 9 x 9 grid, and gets the AABB of the samples enlarged by
// K = (M1 + 2 M2 + M3) sqrt(3) / (8 (8g)^2), M* = max over samples of |S_uu|inf, |S_uv|inf,
// |S_vv|inf, then is clamped to its parent's box.  The reference lays the tree over [0,1]^2
// (src/kernel/kernel.cu:383-387, 406-407), which is the parameter domain only when the knots
// run from 0 to 1; here the unit square is mapped onto the knot domain [u0,u1] x [v0,v1]
// (u = u0 + (u1-u0) x), and K takes the cell sizes (u1-u0)/(8g), (v1-v0)/(8g):
// K = (M1 Wu^2 + 2 M2 Wu Wv + M3 Wv^2) sqrt(3) / (8 (8g)^2).  On [0,1] knots every mapped value
// and K are bit-identical to the reference's (0 + 1 x = x, M 1 1 = M).  The reference runs five launches per level
// (samples, three derivative passes, two LDS reductions) with a device-wide sync after each; here
// one wave per node evaluates its 81 samples (lanes 0..63, then lanes 0..16 again) and reduces
// them with cross-lane shuffles in a single launch per level.
//
// Seed search (new): a quadtree over the finest level, built bottom-up, whose boxes contain the
// leaf centre points below them.  Two modes:
//   exact   -- each lane descends greedily for an initial bound, then walks the tree depth-first
//              without a stack (Morton-code successor), pruning every box whose squared distance
//              exceeds the best centre distance: the exact nearest leaf centre;
//   greedy  -- the greedy descent alone, refined over the 3 x 3 neighbouring leaves: fixed work
//              and uniform control flow per lane (agrees with the exact leaf for ~98% of the
//              offset queries of the projection workload, with the same Newton iteration count).
#include "gn_search.hpp"

namespace gn {

__device__ __forceinline__ double wave_max(double x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x = fmax(x, __shfl_xor(x, o, 64));
    return x;
}
__device__ __forceinline__ double wave_min(double x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x = fmin(x, __shfl_xor(x, o, 64));
    return x;
}

__device__ __forceinline__ bool need_divide(const double lb[3], const double rt[3]) {
    // NeedDivide, src/kernel/kernel.cu:859-878; Threshold 0.5 (src/kernel/kernel.cuh:28)
    double dx = rt[0] - lb[0], dy = rt[1] - lb[1], dz = rt[2] - lb[2], t;
    if (dx > dy) { t = dx; dx = dy; dy = t; }
    if (dx > dz) { t = dx; dx = dz; dz = t; }
    if ((dx < 0.5 * dy) && (dx < 0.5 * dz)) return false;
    return true;
}

// One 64-lane workgroup per node of level `level`.
template <int K1, int K2, bool EXACT>
__global__ __launch_bounds__(64) void k_tree_level(GnSurfDev d, int level, int flags,
                                                   GnTreeNode *__restrict__ tree,
                                                   uint8_t *__restrict__ valid,
                                                   uint8_t *__restrict__ divided) {
    long g = 1, base = 0;
    for (int l = 0; l < level; l++) { base += g * g; g *= 8; }
    const long idx = blockIdx.x;  // bx * g + by
    const long bx = idx / g, by = idx % g;
    const long ti = base + idx;
    const double base_unit = 1.0 / g, unit = base_unit / 8;
    const double u0 = d.ku[d.p], wu = d.ku[d.nu] - u0, v0 = d.kv[d.q], wv = d.kv[d.nv] - v0;
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    double M1 = 0, M2 = 0, M3 = 0;
    for (int sidx = threadIdx.x; sidx < 81; sidx += 64) {
        const int tx = sidx / 9, ty = sidx % 9;
        const double su = u0 + wu * (base_unit * bx + tx * unit),
                     sv = v0 + wv * (base_unit * by + ty * unit);
        Cart<double> o;
        eval_point<K1, K2, 2, EXACT, double>(d, d.ku, d.kv, su, sv, o);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            lo[c] = fmin(lo[c], o.S[c]);
            hi[c] = fmax(hi[c], o.S[c]);
        }
        M1 = fmax(M1, fmax(fabs(o.Suu[0]), fmax(fabs(o.Suu[1]), fabs(o.Suu[2]))));
        M2 = fmax(M2, fmax(fabs(o.Suv[0]), fmax(fabs(o.Suv[1]), fabs(o.Suv[2]))));
        M3 = fmax(M3, fmax(fabs(o.Svv[0]), fmax(fabs(o.Svv[1]), fabs(o.Svv[2]))));
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
        lo[c] = wave_min(lo[c]);
        hi[c] = wave_max(hi[c]);
    }
    M1 = wave_max(M1);
    M2 = wave_max(M2);
    M3 = wave_max(M3);
    if (threadIdx.x != 0) return;
    const double K = (M1 * wu * wu + 2 * M2 * wu * wv + M3 * wv * wv) * sqrt(3.0) /
                     (8.0 * 8.0 * (double)g * 8.0 * (double)g);
    GnTreeNode nd;
    const double unitg = 1.0 / g;
    nd.uv_lb[0] = u0 + wu * (bx * unitg);
    nd.uv_lb[1] = v0 + wv * (by * unitg);
    nd.uv_rt[0] = u0 + wu * ((bx + 1) * unitg);
    nd.uv_rt[1] = v0 + wv * ((by + 1) * unitg);
    for (int c = 0; c < 3; c++) {
        nd.lb[c] = lo[c] - K;
        nd.rt[c] = hi[c] + K;
    }
    long parent = -1;
    if (level > 0) {
        if (flags & GN_TREE_REF_PARENT) parent = (ti - 1) / 64;
        else {
            long gp = g / 8, bp = base - gp * gp;
            parent = bp + (bx / 8) * gp + by / 8;
        }
    }
    uint8_t vld = 1;
    if (parent >= 0) {
        const GnTreeNode pn = tree[parent];
        for (int c = 0; c < 3; c++) {
            nd.rt[c] = fmin(nd.rt[c], pn.rt[c]);
            nd.lb[c] = fmax(nd.lb[c], pn.lb[c]);
        }
        vld = valid[parent] && divided[parent];
    }
    Cart<double> oc;
    eval_point<K1, K2, 0, EXACT, double>(d, d.ku, d.kv, (nd.uv_lb[0] + nd.uv_rt[0]) / 2,
                                          (nd.uv_lb[1] + nd.uv_rt[1]) / 2, oc);
    for (int c = 0; c < 3; c++) nd.c[c] = oc.S[c];
    tree[ti] = nd;
    valid[ti] = vld;
    divided[ti] = need_divide(nd.lb, nd.rt) ? 1 : 0;
}

template <typename L>
struct TreeLaunch {
    const GnSurfDev &d;
    int level, flags;
    GnTreeNode *tree;
    uint8_t *valid, *divided;
    hipStream_t st;
    template <int K1, int K2, bool EXACT>
    int operator()() {
        long g = 1;
        for (int l = 0; l < level; l++) g *= 8;
        k_tree_level<K1, K2, EXACT><<<(unsigned)(g * g), 64, 0, st>>>(d, level, flags, tree, valid,
                                                                      divided);
        GN_HIP(hipGetLastError());
        return GN_OK;
    }
};

int launch_tree_level(const GnSurfDev &d, int level, int flags, GnTreeNode *tree,
                      uint8_t *valid, uint8_t *divided, hipStream_t st) {
    return dispatch_degrees(d.p, d.q, TreeLaunch<void>{d, level, flags, tree, valid, divided, st});
}

// ---------------------------------------------------------------------------------------------
// quadtree search structure
// ---------------------------------------------------------------------------------------------
// leaves: one thread per finest-level node
__global__ void k_search_leaves(int levels, const GnTreeNode *__restrict__ tree,
                                double *__restrict__ sbox, double *__restrict__ scentre,
                                double *__restrict__ sseed, int32_t *__restrict__ sref,
                                float4 *__restrict__ sleaf32, float4 *__restrict__ srec) {
    long g = 1, base = 0;
    for (int l = 0; l < levels - 1; l++) { base += g * g; g *= 8; }
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= g * g) return;
    const long bx = idx / g, by = idx % g;
    const int D = 3 * (levels - 1);
    const GnTreeNode nd = tree[base + idx];
    const uint32_t c = morton2((uint32_t)bx, (uint32_t)by);
    double *b = sbox + (qoff(D) + c) * 6;
    for (int k = 0; k < 3; k++) {
        b[k] = fmin(nd.lb[k], nd.c[k]);
        b[3 + k] = fmax(nd.rt[k], nd.c[k]);
        scentre[c * 3 + k] = nd.c[k];
    }
    const double su = (nd.uv_lb[0] + nd.uv_rt[0]) / 2, sv = (nd.uv_lb[1] + nd.uv_rt[1]) / 2;
    sseed[c * 2 + 0] = su;
    sseed[c * 2 + 1] = sv;
    sref[c] = (int32_t)(base + idx);
    const float4 l32 = make_float4((float)nd.c[0], (float)nd.c[1], (float)nd.c[2],
                                   __int_as_float((int32_t)(base + idx)));
    sleaf32[c] = l32;
    srec[c * kRecVec] = l32;
    const long long bu = __double_as_longlong(su), bv = __double_as_longlong(sv);
    srec[c * kRecVec + 1] = make_float4(__int_as_float((int)(bu & 0xffffffff)), __int_as_float((int)(bu >> 32)),
                                  __int_as_float((int)(bv & 0xffffffff)), __int_as_float((int)(bv >> 32)));
}

// internal level l from level l+1
__global__ void k_search_level(int l, double *__restrict__ sbox) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (1L << (2 * l))) return;
    const double *ch = sbox + (qoff(l + 1) + 4 * c) * 6;
    double b[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int a = 0; a < 4; a++)
        for (int k = 0; k < 3; k++) {
            b[k] = fmin(b[k], ch[a * 6 + k]);
            b[3 + k] = fmax(b[3 + k], ch[a * 6 + 3 + k]);
        }
    double *o = sbox + (qoff(l) + c) * 6;
    for (int k = 0; k < 6; k++) o[k] = b[k];
}

// FP32 copy of the search boxes for pruning: every bound is rounded outward and widened by a
// margin of 2^-20 of the root box's largest coordinate magnitude, which covers the rounding of the
// query point to FP32 for points inside the root box's scale.
__global__ void k_search_box32(long total, const double *__restrict__ sbox,
                               float *__restrict__ sbox32) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    double mag = 0;
    for (int k = 0; k < 6; k++) mag = fmax(mag, fabs(sbox[k]));
    const double margin = (1.0 + mag) * 0x1p-20;
    for (int k = 0; k < 3; k++) {
        float lo = (float)(sbox[i * 6 + k] - margin);
        if ((double)lo > sbox[i * 6 + k] - margin) lo = nextafterf(lo, -INFINITY);
        float hi = (float)(sbox[i * 6 + 3 + k] + margin);
        if ((double)hi < sbox[i * 6 + 3 + k] + margin) hi = nextafterf(hi, INFINITY);
        sbox32[i * 6 + k] = lo;
        sbox32[i * 6 + 3 + k] = hi;
    }
}

// S_u, S_v, S_uu, S_uv, S_vv at every leaf's seed (u, v), for the stepped seeds: one thread per
// leaf, into vectors 2..6 of the leaf's record (7 zero).
template <int K1, int K2, bool EXACT>
__global__ void k_search_jac(GnSurfDev d, long leaves, const double *__restrict__ sseed,
                             float4 *__restrict__ srec) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= leaves) return;
    Cart<double> o;
    eval_point<K1, K2, 2, EXACT, double>(d, d.ku, d.kv, sseed[c * 2], sseed[c * 2 + 1], o);
    const double *f[5] = {o.Su, o.Sv, o.Suu, o.Suv, o.Svv};
    float4 *r = srec + c * kRecVec;
    for (int k = 0; k < 5; k++) r[2 + k] = make_float4((float)f[k][0], (float)f[k][1], (float)f[k][2], 0.f);
    r[7] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// Normal cone of every leaf (NormalBoundingBox + NormalBoxToNormalTreeNode,
// src/kernel/kernel.cu:623-704, 1605-1636; normal derivatives src/NURBS/Evaluate.cu:1125-1135):
// one wave per leaf samples N = S_u x S_v on the tree's 9 x 9 grid of the leaf cell, takes the
// box of those normals enlarged by K = (M1 Wu^2 + 2 M2 Wu Wv + M3 Wv^2) sqrt(3) / (8 (8g)^2) with
// M* the sample maxima of |N_uu|inf, |N_uv|inf, |N_vv|inf (the position boxes' K applied to N),
// and the cone around the box centre m through the farthest box corner.  A box that holds the
// origin, or a corner 90 degrees or more from m, gives no bounded cone (half-angle 4).  FP64;
// stored in FP32 with the half-angle rounded up past 1e-6 rad, more than the axis' rounding.
__device__ __forceinline__ void cross3(const double *a, const double *b, double *o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

template <int K1, int K2, bool EXACT>
__global__ __launch_bounds__(64) void k_search_cone(GnSurfDev d, int levels,
                                                    float4 *__restrict__ cone32) {
    long g = 1;
    for (int l = 1; l < levels; l++) g *= 8;
    const long idx = blockIdx.x;  // bx * g + by
    const long bx = idx / g, by = idx % g;
    const double base_unit = 1.0 / g, unit = base_unit / 8;
    const double u0 = d.ku[d.p], wu = d.ku[d.nu] - u0, v0 = d.kv[d.q], wv = d.kv[d.nv] - v0;
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    double M1 = 0, M2 = 0, M3 = 0;
    for (int sidx = threadIdx.x; sidx < 81; sidx += 64) {
        const int tx = sidx / 9, ty = sidx % 9;
        const double su = u0 + wu * (base_unit * bx + tx * unit),
                     sv = v0 + wv * (base_unit * by + ty * unit);
        Cart<double> o;
        eval_point<K1, K2, 3, EXACT, double>(d, d.ku, d.kv, su, sv, o);
        double N[3], a[3], b[3], c[3], D[3];
        cross3(o.Su, o.Sv, N);
#pragma unroll
        for (int k = 0; k < 3; k++) {
            lo[k] = fmin(lo[k], N[k]);
            hi[k] = fmax(hi[k], N[k]);
        }
        cross3(o.Suuu, o.Sv, a);  // N_uu = S_uuu x S_v + 2 S_uu x S_uv + S_u x S_uuv
        cross3(o.Suu, o.Suv, b);
        cross3(o.Su, o.Suuv, c);
        for (int k = 0; k < 3; k++) D[k] = a[k] + 2.0 * b[k] + c[k];
        M1 = fmax(M1, fmax(fabs(D[0]), fmax(fabs(D[1]), fabs(D[2]))));
        cross3(o.Suuv, o.Sv, a);  // N_uv = S_uuv x S_v + S_uu x S_vv + S_u x S_uvv
        cross3(o.Suu, o.Svv, b);
        cross3(o.Su, o.Suvv, c);
        for (int k = 0; k < 3; k++) D[k] = a[k] + b[k] + c[k];
        M2 = fmax(M2, fmax(fabs(D[0]), fmax(fabs(D[1]), fabs(D[2]))));
        cross3(o.Suvv, o.Sv, a);  // N_vv = S_uvv x S_v + 2 S_uv x S_vv + S_u x S_vvv
        cross3(o.Suv, o.Svv, b);
        cross3(o.Su, o.Svvv, c);
        for (int k = 0; k < 3; k++) D[k] = a[k] + 2.0 * b[k] + c[k];
        M3 = fmax(M3, fmax(fabs(D[0]), fmax(fabs(D[1]), fabs(D[2]))));
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
        lo[k] = wave_min(lo[k]);
        hi[k] = wave_max(hi[k]);
    }
    M1 = wave_max(M1);
    M2 = wave_max(M2);
    M3 = wave_max(M3);
    if (threadIdx.x != 0) return;
    const double K = (M1 * wu * wu + 2 * M2 * wu * wv + M3 * wv * wv) * sqrt(3.0) /
                     (8.0 * 8.0 * (double)g * 8.0 * (double)g);
    bool zero = true;
    double m[3];
    for (int k = 0; k < 3; k++) {
        lo[k] -= K;
        hi[k] += K;
        if (!(lo[k] <= 0.0 && hi[k] >= 0.0)) zero = false;
        m[k] = 0.5 * (lo[k] + hi[k]);
    }
    double th = 0.0;
    for (int cn = 0; cn < 8 && !zero; cn++) {
        const double x[3] = {(cn & 4) ? hi[0] : lo[0], (cn & 2) ? hi[1] : lo[1],
                             (cn & 1) ? hi[2] : lo[2]};
        double cr[3];
        cross3(m, x, cr);
        th = fmax(th, atan2(sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]),
                            m[0] * x[0] + m[1] * x[1] + m[2] * x[2]));
    }
    const double mm = sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
    float4 out = make_float4(0.f, 0.f, 0.f, 4.f);
    if (!zero && th < 0.5 * M_PI && mm > 0.0) {
        float t32 = (float)(th + 1e-6);
        if ((double)t32 < th + 1e-6) t32 = nextafterf(t32, INFINITY);
        out = make_float4((float)(m[0] / mm), (float)(m[1] / mm), (float)(m[2] / mm), t32);
    }
    cone32[morton2((uint32_t)bx, (uint32_t)by)] = out;
}

struct ConeLaunch {
    const GnSurfDev &d;
    int levels;
    float4 *cone32;
    hipStream_t st;
    template <int K1, int K2, bool EXACT>
    int operator()() {
        long g = 1;
        for (int l = 1; l < levels; l++) g *= 8;
        k_search_cone<K1, K2, EXACT><<<(unsigned)(g * g), 64, 0, st>>>(d, levels, cone32);
        GN_HIP(hipGetLastError());
        return GN_OK;
    }
};

struct JacLaunch {
    const GnSurfDev &d;
    long leaves;
    const double *sseed;
    float4 *srec;
    hipStream_t st;
    template <int K1, int K2, bool EXACT>
    int operator()() {
        k_search_jac<K1, K2, EXACT><<<grid_for(leaves, 256, 1u << 30), 256, 0, st>>>(d, leaves, sseed,
                                                                                   srec);
        GN_HIP(hipGetLastError());
        return GN_OK;
    }
};

int launch_search_build(const GnSurfDev &d, int levels, const GnTreeNode *tree, double *sbox,
                        float *sbox32, double *scentre, double *sseed, int32_t *sref,
                        float4 *sleaf32, float4 *srec, float4 *scone32, hipStream_t st) {
    const int D = 3 * (levels - 1);
    const long leaves = 1L << (2 * D);
    k_search_leaves<<<grid_for(leaves, 256, 1u << 30), 256, 0, st>>>(levels, tree, sbox, scentre,
                                                                       sseed, sref, sleaf32, srec);
    GN_HIP(hipGetLastError());
    int rc = dispatch_degrees(d.p, d.q, JacLaunch{d, leaves, sseed, srec, st});
    if (!rc) rc = dispatch_degrees(d.p, d.q, ConeLaunch{d, levels, scone32, st});
    if (rc) return rc;
    for (int l = D - 1; l >= 0; l--)
        k_search_level<<<grid_for(1L << (2 * l), 256, 1u << 30), 256, 0, st>>>(l, sbox);
    const long total = qoff(D + 1);
    k_search_box32<<<grid_for(total, 256, 1u << 30), 256, 0, st>>>(total, sbox, sbox32);
    GN_HIP(hipGetLastError());
    return GN_OK;
}

// One lane per query; mode 0 exact nearest leaf, mode 1 greedy + 3x3, mode 2 greedy descent and
// the Gauss-Newton step of step_seed.
// Levels of the seed quadtree whose FP32 boxes the greedy seed kernel stages in LDS: levels
// 1..5 are 1364 boxes, 32 KB.
// Staged levels and workgroups per CU of the greedy seed kernel (GN_SEED_GREEDY_STEP on 1M
// projection queries: 5 levels x 4 workgroups 41 us; 5 x 5 45 us; 4 x 8 43 us; 4 x 16 45 us;
// 3 x 16 47 us).
#ifndef GN_SEED_LDS_LEVELS
#define GN_SEED_LDS_LEVELS 5
#endif
#ifndef GN_SEED_WG_PER_CU
#define GN_SEED_WG_PER_CU 4
#endif
constexpr int kSeedLdsLevels = GN_SEED_LDS_LEVELS;

__global__ __launch_bounds__(256) void k_seed(GnSearchDev sd, int mode,
                                              const double *__restrict__ x,
                                              const double *__restrict__ y,
                                              const double *__restrict__ z, size_t n,
                                              int32_t *__restrict__ leaf, double *__restrict__ su,
                                              double *__restrict__ sv) {
    __shared__ __align__(16) float lbox[(((1 << (2 * (kSeedLdsLevels + 1))) - 1) / 3 - 1) * 6];
    const int ll = mode != 0 ? (sd.depth < kSeedLdsLevels ? sd.depth : kSeedLdsLevels) : 0;
    if (ll > 0) {
        // boxes of levels 1..ll start one box (24 bytes) into box32: float2 copies
        const int nf2 = (int)(qoff(ll + 1) - 1) * 3;
        const float2 *src = reinterpret_cast<const float2 *>(sd.box32 + 6);
        float2 *dst = reinterpret_cast<float2 *>(lbox);
#pragma unroll 4
        for (int i = threadIdx.x; i < nf2; i += blockDim.x) dst[i] = src[i];
        __syncthreads();
    }
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const double P[3] = {x[i], y[i], z[i]};
        float4 cq;
        const uint32_t c = mode == 0 ? nearest_leaf(sd, P)
                                     : greedy_leaf(sd, P, lbox, ll, &cq, mode != 2);
        if (leaf) leaf[i] = sd.ref_index[c];
        if (mode == 2) {
            double u, v;
            step_seed(sd, c, cq, P, u, v);
            su[i] = u;
            sv[i] = v;
        } else {
            su[i] = sd.seed_uv[c * 2 + 0];
            sv[i] = sd.seed_uv[c * 2 + 1];
        }
    }
}

// GN_SEED_GRID_STEP: one lane per query; the grid cell's leaf (one 2- or 4-byte gather), its
// 64-byte record (one cache line), the step of step_seed.  The leaf's reference index is the w
// bits of the record's centre.  The cell width is a template parameter, so no branch separates
// the loads.  (Several points per lane, their loads issued together, measured 19.5 us for 1M
// queries against 18.3 us for one.)
template <bool WIDE>
__global__ __launch_bounds__(256) void k_seed_grid(GnSearchDev sd, const double *__restrict__ x,
                                                   const double *__restrict__ y,
                                                   const double *__restrict__ z, size_t n,
                                                   int32_t *__restrict__ leaf,
                                                   double *__restrict__ su,
                                                   double *__restrict__ sv) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const double P[3] = {x[i], y[i], z[i]};
        const uint32_t c = grid_leaf<WIDE>(sd, P);
        const float4 *r = sd.rec + (size_t)c * kRecVec;
        const float4 cq = r[0];
        const double2 uv = *reinterpret_cast<const double2 *>(r + 1);
        double u = uv.x, v = uv.y;
        step_seed_from(sd, cq, r[2], r[3], r[4], r[5], r[6], P, u, v);
        if (leaf) leaf[i] = __float_as_int(cq.w);
        su[i] = u;
        sv[i] = v;
    }
}

// Seed grid build: one thread per cell against every leaf.  The leaves' FP32 centres pass through
// LDS in chunks, in reference order (bx * side + by), so a strict "nearer" keeps the lower
// reference index among equals; the winner is stored as its Morton index.  Distances are the
// greedy refinement's: fmaf(ez, ez, fmaf(ey, ey, ex * ex)) in FP32.
constexpr int kGridChunk = 2048;

__global__ __launch_bounds__(256) void k_seed_grid_build(const float4 *__restrict__ leaf32,
                                                         int depth, int nx, int ny, int nz,
                                                         float ox, float oy, float oz, float cell,
                                                         int wide, void *__restrict__ grid) {
    __shared__ float4 sl[kGridChunk];
    const long cells = (long)nx * ny * nz, leaves = 1L << (2 * depth);
    const int side = 1 << depth;
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long cid = idx < cells ? idx : cells - 1;
    const int k = (int)(cid % nz), j = (int)((cid / nz) % ny), i = (int)(cid / ((long)nz * ny));
    const float cx = fmaf((float)i + 0.5f, cell, ox), cy = fmaf((float)j + 0.5f, cell, oy),
                cz = fmaf((float)k + 0.5f, cell, oz);
    float best = INFINITY;
    long best_t = 0;
    for (long t0 = 0; t0 < leaves; t0 += kGridChunk) {
        const int m = (int)(leaves - t0 < kGridChunk ? leaves - t0 : kGridChunk);
        __syncthreads();
        for (int t = threadIdx.x; t < m; t += blockDim.x) {
            const long r = t0 + t;
            sl[t] = leaf32[morton2((uint32_t)(r >> depth), (uint32_t)(r & (side - 1)))];
        }
        __syncthreads();
        for (int t = 0; t < m; t++) {
            const float4 q = sl[t];
            const float ex = cx - q.x, ey = cy - q.y, ez = cz - q.z;
            const float dd = fmaf(ez, ez, fmaf(ey, ey, ex * ex));
            if (dd < best) {
                best = dd;
                best_t = t0 + t;
            }
        }
    }
    if (idx >= cells) return;
    const uint32_t c = morton2((uint32_t)(best_t >> depth), (uint32_t)(best_t & (side - 1)));
    if (wide) static_cast<uint32_t *>(grid)[idx] = c;
    else static_cast<uint16_t *>(grid)[idx] = (uint16_t)c;
}

int launch_seed_grid_build(const float4 *leaf32, int depth, const int dims[3], const float org[3],
                           float cell, int wide, void *grid, hipStream_t st) {
    const long cells = (long)dims[0] * dims[1] * dims[2];
    k_seed_grid_build<<<(unsigned)((cells + 255) / 256), 256, 0, st>>>(
        leaf32, depth, dims[0], dims[1], dims[2], org[0], org[1], org[2], cell, wide, grid);
    GN_HIP(hipGetLastError());
    return GN_OK;
}

int launch_seed(const GnSearchDev &sd, int mode, const double *x, const double *y,
                const double *z, size_t n, int32_t *leaf, double *su, double *sv,
                hipStream_t st) {
    if (n == 0) return GN_OK;
    if (mode == 3) {
        const unsigned g = grid_for(n, 256, 256 * 32);
        if (sd.grid_wide) k_seed_grid<true><<<g, 256, 0, st>>>(sd, x, y, z, n, leaf, su, sv);
        else k_seed_grid<false><<<g, 256, 0, st>>>(sd, x, y, z, n, leaf, su, sv);
        GN_HIP(hipGetLastError());
        return GN_OK;
    }
    // greedy (with or without the step): a few long-lived workgroups per CU, each staging the
    // upper tree levels in LDS once (a workgroup per 256 points would copy 32 KB per 256 points)
    const unsigned cap = mode != 0 ? 256 * GN_SEED_WG_PER_CU : 256 * 32;
    k_seed<<<grid_for(n, 256, cap), 256, 0, st>>>(sd, mode, x, y, z, n, leaf, su, sv);
    GN_HIP(hipGetLastError());
    return GN_OK;
}

}  // namespace gn
