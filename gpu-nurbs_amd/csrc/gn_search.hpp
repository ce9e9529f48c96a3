// gn_search.hpp -- device functions of the seed search over the quadtree of tree leaves
// (exact nearest leaf centre, and greedy descent + 3 x 3 neighbourhood), shared by the
// single-surface seed kernel and the batched inversion kernels.
#pragma once
#include "gn_device.hpp"

namespace gn {

__host__ __device__ inline long qoff(int l) { return ((1L << (2 * l)) - 1) / 3; }

__device__ __forceinline__ uint32_t spread2(uint32_t x) {  // bit b of x -> bit 2b (16-bit x)
    x &= 0xffffu;
    x = (x | (x << 8)) & 0x00ff00ffu;
    x = (x | (x << 4)) & 0x0f0f0f0fu;
    x = (x | (x << 2)) & 0x33333333u;
    x = (x | (x << 1)) & 0x55555555u;
    return x;
}

__device__ __forceinline__ uint32_t morton2(uint32_t x, uint32_t y) {  // x bits odd, y bits even
    return (spread2(x) << 1) | spread2(y);
}

// Squared distance from p to an FP32 box (0 inside); a lower bound of the FP64 distance from P to
// the FP64 box, because the FP32 box contains it with a margin.
__device__ __forceinline__ float box_d2(const float *__restrict__ b, const float p[3]) {
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const float e = fmaxf(fmaxf(b[k] - p[k], p[k] - b[3 + k]), 0.f);
        s = fmaf(e, e, s);
    }
    return s;
}

// Exact nearest leaf centre (ties to the lower reference index).  Returns the Morton index.
// Pruning uses the FP32 boxes (a conservative bound); leaf distances are FP64.
__device__ inline uint32_t nearest_leaf(const GnSearchDev &sd, const double P[3]) {
    const int D = sd.depth;
    if (D == 0) return 0;
    const float p[3] = {(float)P[0], (float)P[1], (float)P[2]};
    // greedy descent for an initial bound
    uint32_t c = 0;
    for (int l = 0; l < D; l++) {
        float bd = INFINITY;
        uint32_t bc = 4 * c;
        const float *b = sd.box32 + (qoff(l + 1) + 4 * c) * 6;
#pragma unroll
        for (int a = 0; a < 4; a++) {
            const float dd = box_d2(b + a * 6, p);
            if (dd < bd) { bd = dd; bc = 4 * c + a; }
        }
        c = bc;
    }
    auto cdist = [&](uint32_t leaf) {
        const double *q = sd.centre + leaf * 3;
        const double dx = P[0] - q[0], dy = P[1] - q[1], dz = P[2] - q[2];
        return fma(dz, dz, fma(dy, dy, dx * dx));
    };
    uint32_t best_c = c;
    double best = cdist(c);
    // stackless depth-first walk with pruning
    int l = 1;
    c = 0;
    while (true) {
        const double lb = (double)box_d2(sd.box32 + (qoff(l) + c) * 6, p);
        bool descend = false;
        if (lb <= best) {
            if (l == D) {
                const double dd = cdist(c);
                if (dd < best || (dd == best && sd.ref_index[c] < sd.ref_index[best_c])) {
                    best = dd;
                    best_c = c;
                }
            } else {
                descend = true;
            }
        }
        if (descend) {
            c <<= 2;
            l++;
            continue;
        }
        while (l > 1 && (c & 3u) == 3u) {
            c >>= 2;
            l--;
        }
        if (l == 1 && (c & 3u) == 3u) break;
        c++;
    }
    return best_c;
}

__device__ __forceinline__ uint32_t compact2(uint32_t c) {  // even bits of c -> bits 0..15
    c &= 0x55555555u;
    c = (c | (c >> 1)) & 0x33333333u;
    c = (c | (c >> 2)) & 0x0f0f0f0fu;
    c = (c | (c >> 4)) & 0x00ff00ffu;
    c = (c | (c >> 8)) & 0x0000ffffu;
    return c;
}

// One level of the greedy descent: the child of node c (level l) whose box is nearest to p
// (ties to the lower child); `level` points at the boxes of level l + 1.
__device__ __forceinline__ uint32_t greedy_child(const float *__restrict__ level, uint32_t c,
                                                 const float p[3]) {
    float bd = INFINITY;
    uint32_t bc = 4 * c;
    const float *b = level + 4 * c * 6;
#pragma unroll
    for (int a = 0; a < 4; a++) {
        const float dd = box_d2(b + a * 6, p);
        if (dd < bd) { bd = dd; bc = 4 * c + a; }
    }
    return bc;
}

// Fast seed: greedy descent by FP32 box distance (ties to the lower child), then the nearest
// leaf centre (FP32 centres and distances, ties to the lower reference index) among the 3 x 3
// leaves around the descent's leaf in the uv grid.  Fixed work per point: 4*D box tests and 9
// distances.
// lds, when given, holds the FP32 boxes of levels 1 .. lds_levels (the first part of box32 after
// the root), so those levels of the descent read LDS.
// cbest, when given, receives the chosen leaf's FP32 centre.
// refine == false: the descent's leaf itself, without the 3 x 3 neighbourhood (GN_SEED_GREEDY_STEP,
// whose step moves up to a leaf cell anyway: 1M projection queries 4.65 -> 4.90 G inversions/s).
__device__ inline uint32_t greedy_leaf(const GnSearchDev &sd, const double P[3],
                                       const float *lds = nullptr, int lds_levels = 0,
                                       float4 *cbest = nullptr, bool refine = true) {
    const int D = sd.depth;
    if (D == 0) return 0;
    const float p[3] = {(float)P[0], (float)P[1], (float)P[2]};
    uint32_t c = 0;
    int l = 0;
    for (; l < lds_levels && l < D; l++) c = greedy_child(lds + (qoff(l + 1) - 1) * 6, c, p);
    for (; l < D; l++) c = greedy_child(sd.box32 + qoff(l + 1) * 6, c, p);
    if (!refine) {
        if (cbest) *cbest = sd.rec[(size_t)c * kRecVec];  // the leaf32 entry, on the step's line
        return c;
    }
    const int side = 1 << D;
    const int gx = (int)compact2(c >> 1), gy = (int)compact2(c);
    // the 9 neighbour leaves (FP32 centre + reference index, 16 bytes each) are loaded before
    // any is compared (addresses clamped into the grid, out-of-grid neighbours skipped after)
    uint32_t mk[9];
    float4 q[9];
#pragma unroll
    for (int k = 0; k < 9; k++) {
        const int x = min(max(gx + k / 3 - 1, 0), side - 1), y = min(max(gy + k % 3 - 1, 0), side - 1);
        mk[k] = morton2((uint32_t)x, (uint32_t)y);
        q[k] = sd.leaf32[mk[k]];
    }
    float best = INFINITY;
    int best_ref = 0x7fffffff;
    uint32_t best_c = c;
    float4 bq = q[4];
#pragma unroll
    for (int k = 0; k < 9; k++) {
        const int x = gx + k / 3 - 1, y = gy + k % 3 - 1;
        if (x < 0 || y < 0 || x >= side || y >= side) continue;
        const float ex = p[0] - q[k].x, ey = p[1] - q[k].y, ez = p[2] - q[k].z;
        const float dd = fmaf(ez, ez, fmaf(ey, ey, ex * ex));
        const int ref = __float_as_int(q[k].w);
        if (dd < best || (dd == best && ref < best_ref)) {
            best = dd;
            best_ref = ref;
            best_c = mk[k];
            bq = q[k];
        }
    }
    if (cbest) *cbest = bq;
    return best_c;
}

// The seed grid's leaf for P (GN_SEED_GRID_STEP): the cell holding P rounded to FP32, clamped
// into the grid (points outside it take the nearest border cell; NaN coordinates cell 0).
// (p - org) * inv has no multiply-add to contract, so it rounds twice, as the oracle's does.
__device__ __forceinline__ uint32_t grid_cell(const GnSearchDev &sd, const double P[3]) {
    int id[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const float t = ((float)P[k] - sd.gorg[k]) * sd.ginv;
        id[k] = (int)fminf(fmaxf(floorf(t), 0.f), (float)(sd.gdim[k] - 1));
    }
    return ((uint32_t)id[0] * (uint32_t)sd.gdim[1] + (uint32_t)id[1]) * (uint32_t)sd.gdim[2] +
           (uint32_t)id[2];  // below 2^25 cells
}

template <bool WIDE>
__device__ __forceinline__ uint32_t grid_leaf(const GnSearchDev &sd, const double P[3]) {
    const uint32_t cell = grid_cell(sd, P);
    return WIDE ? static_cast<const uint32_t *>(sd.grid)[cell]
                : static_cast<const uint16_t *>(sd.grid)[cell];
}

// Seed refinement of GN_SEED_GREEDY_STEP and GN_SEED_GRID_STEP, from leaf c's seed (u, v) and
// its stored derivatives (S_u, S_v, S_uu, S_uv, S_vv evaluated at the seed, whose surface point is
// the leaf's centre C), with no surface evaluation:
//   1. one Gauss-Newton step t = -(J^T J)^-1 J^T (C - P), J = [S_u S_v], scaled down uniformly to
//      at most one leaf cell in each direction (a singular J^T J or a non-finite scale keeps the
//      leaf's seed);
//   2. two Gauss-Newton iterations on the leaf's quadratic model
//      m(t) = C + S_u t_u + S_v t_v + (S_uu t_u^2 + 2 S_uv t_u t_v + S_vv t_v^2) / 2,
//      whose Jacobian is [S_u + S_uu t_u + S_uv t_v, S_v + S_uv t_u + S_vv t_v], each result scaled
//      down to at most two leaf cells (an iteration with a singular system keeps its start);
//   3. the offset t added to the FP64 seed and clamped to the domain.
// In FP32, from C, P rounded to FP32 and the FP32 derivatives.  On the headline projection queries
// the model iterations take the FP32 Newton iterations after the seed from 2.68 to 2.15 per point
// (CPU restatement).  Every multiply-add is an explicit fmaf and the reciprocals are IEEE
// divisions, so the oracle restates the step operation for operation.
__device__ __forceinline__ void step_seed_from(const GnSearchDev &sd, const float4 &C,
                                               const float4 &ju, const float4 &jv,
                                               const float4 &huu, const float4 &huv,
                                               const float4 &hvv, const double P[3], double &su,
                                               double &sv) {
    const float r[3] = {C.x - (float)P[0], C.y - (float)P[1], C.z - (float)P[2]};
    const float a = fmaf(ju.z, ju.z, fmaf(ju.y, ju.y, ju.x * ju.x));
    const float b = fmaf(ju.z, jv.z, fmaf(ju.y, jv.y, ju.x * jv.x));
    const float e = fmaf(jv.z, jv.z, fmaf(jv.y, jv.y, jv.x * jv.x));
    const float gu = fmaf(ju.z, r[2], fmaf(ju.y, r[1], ju.x * r[0]));
    const float gv = fmaf(jv.z, r[2], fmaf(jv.y, r[1], jv.x * r[0]));
    const float det = fmaf(a, e, -(b * b));
    // everything is computed whatever det is, and kept only when det > 0 and the scale is
    // finite.  Divisions are reciprocals (IEEE) times the numerators: 4 divisions per point
    // instead of 14.
    const float idet = 1.f / det;
    const float du = fmaf(e, gu, -(b * gv)) * idet;
    const float dv = fmaf(a, gv, -(b * gu)) * idet;
    const float cells = (float)(1u << sd.depth);
    const float hu = (float)(sd.u1 - sd.u0) / cells, hv = (float)(sd.v1 - sd.v0) / cells;
    const float ihu = 1.f / hu, ihv = 1.f / hv, ih2u = 0.5f * ihu, ih2v = 0.5f * ihv;
    const float sc = fmaxf(1.f, fmaxf(fabsf(du) * ihu, fabsf(dv) * ihv));
    const bool ok0 = det > 0.f && sc < INFINITY;
    const float isc = 1.f / sc;
    float tu = ok0 ? -(du * isc) : 0.f, tv = ok0 ? -(dv * isc) : 0.f;
    const float Ju[3] = {ju.x, ju.y, ju.z}, Jv[3] = {jv.x, jv.y, jv.z};
    const float Huu[3] = {huu.x, huu.y, huu.z}, Huv[3] = {huv.x, huv.y, huv.z},
                Hvv[3] = {hvv.x, hvv.y, hvv.z};
#pragma unroll
    for (int it = 0; it < 2; it++) {
        const float tuu = tu * tu, tuv = tu * tv, tvv = tv * tv;
        float m[3], au[3], av[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const float t = fmaf(Ju[k], tu, fmaf(Jv[k], tv, r[k]));
            const float w = fmaf(Huu[k], tuu, fmaf(2.f * Huv[k], tuv, Hvv[k] * tvv));
            m[k] = fmaf(0.5f, w, t);
            au[k] = fmaf(Huu[k], tu, fmaf(Huv[k], tv, Ju[k]));
            av[k] = fmaf(Huv[k], tu, fmaf(Hvv[k], tv, Jv[k]));
        }
        const float A = fmaf(au[2], au[2], fmaf(au[1], au[1], au[0] * au[0]));
        const float B = fmaf(au[2], av[2], fmaf(au[1], av[1], au[0] * av[0]));
        const float E = fmaf(av[2], av[2], fmaf(av[1], av[1], av[0] * av[0]));
        const float Gu = fmaf(au[2], m[2], fmaf(au[1], m[1], au[0] * m[0]));
        const float Gv = fmaf(av[2], m[2], fmaf(av[1], m[1], av[0] * m[0]));
        const float D = fmaf(A, E, -(B * B));
        const float iD = 1.f / D;
        // each product in a statement of its own: with the subtraction in the same expression
        // the compiler would contract them into one fused multiply-add
        const float qu = fmaf(E, Gu, -(B * Gv)) * iD;
        const float qv = fmaf(A, Gv, -(B * Gu)) * iD;
        const float nu = tu - qu;
        const float nv = tv - qv;
        const float s2 = fmaxf(1.f, fmaxf(fabsf(nu) * ih2u, fabsf(nv) * ih2v));
        const bool ok = D > 0.f && s2 < INFINITY;
        const float is2 = 1.f / s2;
        tu = ok ? nu * is2 : tu;
        tv = ok ? nv * is2 : tv;
    }
    const double nsu = fmin(fmax(su + (double)tu, sd.u0), sd.u1);
    const double nsv = fmin(fmax(sv + (double)tv, sd.v0), sd.v1);
    su = ok0 ? nsu : su;
    sv = ok0 ? nsv : sv;
}

__device__ __forceinline__ void step_seed(const GnSearchDev &sd, uint32_t c, const float4 &C,
                                          const double P[3], double &su, double &sv) {
    const float4 *r = sd.rec + (size_t)c * kRecVec;
    const double2 uv = *reinterpret_cast<const double2 *>(r + 1);
    su = uv.x;
    sv = uv.y;
    step_seed_from(sd, C, r[2], r[3], r[4], r[5], r[6], P, su, sv);
}

}  // namespace gn
