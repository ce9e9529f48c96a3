// curve.hip -- rational B-spline curves in matrix form: span table, evaluation to order 2 and
// point inversion.  The reference has no curve code (its README claims curves); this is the same
// matrix representation in one parameter: per span, patch[r][c] = sum_a M[r][a] P_w[m-p+a][c]
// with M the basis matrix of the span (span.hip) and P_w = (x w, y w, z w, w), so that
// C(t) = (sum_r s^r patch[r]).xyz / (sum_r s^r patch[r]).w with s the local parameter.
//
// Inversion minimises |C(t) - P|^2 / 2 by Newton on f(t) = C'.(C - P) (f' = C''.(C - P) + C'.C',
// falling back to the Gauss-Newton f' = C'.C' where f' <= 0), seeded by the nearest of 8 samples
// per span (FP32 distances, samples staged in LDS).  Modes as for surfaces: plain FP64, or FP32
// iterations followed by FP64 closing iterations.  One lane per point.
#include "gn_device.hpp"

namespace gn {


// one thread per slot: patch = M * P_w
__global__ void k_curve_table(GnCurveDev c, const double *__restrict__ xyz,
                              const double *__restrict__ w, double *__restrict__ ih,
                              double *__restrict__ tab64, float *__restrict__ tab32) {
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= c.ns) return;
    const int span = slot + c.p, k = c.p + 1;
    const double h = c.k[span + 1] - c.k[span];
    double M[GN_MAXK * GN_MAXK];
    if (h != 0.0) basis_matrix(c.p, c.k, span, M);
    ih[slot] = h != 0.0 ? 1.0 / h : 0.0;
    for (int r = 0; r < k; r++) {
        double acc[4] = {0, 0, 0, 0};
        if (h != 0.0)
            for (int a = 0; a < k; a++) {
                const int g = slot + a;  // control point span - p + a
                const double wt = w[g], m = M[r * k + a];
                acc[0] += m * (xyz[g * 3 + 0] * wt);
                acc[1] += m * (xyz[g * 3 + 1] * wt);
                acc[2] += m * (xyz[g * 3 + 2] * wt);
                acc[3] += m * wt;
            }
        for (int q = 0; q < 4; q++) {
            tab64[((size_t)slot * k + r) * 4 + q] = acc[q];
            tab32[((size_t)slot * k + r) * 4 + q] = (float)acc[q];
        }
    }
}

template <typename T>
struct CurvePt {
    T C[3], C1[3], C2[3];
    T ih;            // 1 / length of the span evaluated
    double lo, hi;   // the span evaluated, [k[m], k[m+1]]
};

template <int ORDER, typename T>
__device__ __forceinline__ void curve_eval(const GnCurveDev &c, double t, CurvePt<T> &o) {
    const int m = find_span(c.k, c.p, c.n, t);
    const int slot = m - c.p, k = c.p + 1;
    const double ih = c.ih[slot];
    const T s = T((t - c.k[m]) * ih);
    const T *tab = nullptr;
    if constexpr (sizeof(T) == 8) tab = reinterpret_cast<const T *>(c.tab64);
    else tab = reinterpret_cast<const T *>(c.tab32);
    const T *patch = tab + (size_t)slot * k * 4;
    Hom<T> R, R1, R2;
    hzero(R);
    hzero(R1);
    hzero(R2);
    // Horner in s for the value and the first two derivatives
    for (int r = k - 1; r >= 0; r--) {
        const Hom<T> a = load_hom<T>(patch + r * 4);
        if (ORDER >= 2) {
            R2.x = fma(R2.x, s, T(2) * R1.x); R2.y = fma(R2.y, s, T(2) * R1.y);
            R2.z = fma(R2.z, s, T(2) * R1.z); R2.w = fma(R2.w, s, T(2) * R1.w);
        }
        if (ORDER >= 1) {
            R1.x = fma(R1.x, s, R.x); R1.y = fma(R1.y, s, R.y);
            R1.z = fma(R1.z, s, R.z); R1.w = fma(R1.w, s, R.w);
        }
        R.x = fma(R.x, s, a.x); R.y = fma(R.y, s, a.y); R.z = fma(R.z, s, a.z); R.w = fma(R.w, s, a.w);
    }
    const T iw = recip(R.w), h1 = T(ih);
    o.ih = h1;
    o.lo = c.k[m];
    o.hi = c.k[m + 1];
    const T A[3] = {R.x, R.y, R.z}, A1[3] = {R1.x, R1.y, R1.z}, A2[3] = {R2.x, R2.y, R2.z};
#pragma unroll
    for (int q = 0; q < 3; q++) o.C[q] = A[q] * iw;
    if (ORDER < 1) return;
#pragma unroll
    for (int q = 0; q < 3; q++) o.C1[q] = (A1[q] - R1.w * o.C[q]) * iw;
    if (ORDER >= 2) {
#pragma unroll
        for (int q = 0; q < 3; q++) o.C2[q] = (A2[q] - T(2) * R1.w * o.C1[q] - R2.w * o.C[q]) * iw * (h1 * h1);
    }
#pragma unroll
    for (int q = 0; q < 3; q++) o.C1[q] *= h1;
}

__global__ void k_curve_samples(GnCurveDev c, float *__restrict__ samp, double *__restrict__ samp_t) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.ns * kCurveSamples) return;
    const int slot = i / kCurveSamples, j = i % kCurveSamples;
    const double lo = c.k[slot + c.p], hi = c.k[slot + c.p + 1];
    const double t = lo + (j + 0.5) * (hi - lo) / kCurveSamples;
    CurvePt<double> o;
    curve_eval<0, double>(c, t, o);
    for (int q = 0; q < 3; q++) samp[i * 3 + q] = (float)o.C[q];
    samp_t[i] = t;
}

// Per span: the FP32 bounding box of its seed samples (lo xyz, hi xyz), stored after them.
__global__ void k_curve_boxes(GnCurveDev c, float *__restrict__ box) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= c.ns) return;
    for (int q = 0; q < 3; q++) {
        float lo = INFINITY, hi = -INFINITY;
        for (int j = 0; j < kCurveSamples; j++) {
            const float v = c.samp[(s * kCurveSamples + j) * 3 + q];
            lo = fminf(lo, v);
            hi = fmaxf(hi, v);
        }
        box[s * 6 + q] = lo;
        box[s * 6 + 3 + q] = hi;
    }
}

template <int ORDER>
__global__ void k_curve_eval(GnCurveDev c, const double *__restrict__ t, size_t n,
                             double *__restrict__ out) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        CurvePt<double> o;
        curve_eval<ORDER, double>(c, t[i], o);
#pragma unroll
        for (int q = 0; q < 3; q++) out[q * n + i] = o.C[q];
        if (ORDER >= 1)
#pragma unroll
            for (int q = 0; q < 3; q++) out[(3 + q) * n + i] = o.C1[q];
        if (ORDER >= 2)
#pragma unroll
            for (int q = 0; q < 3; q++) out[(6 + q) * n + i] = o.C2[q];
    }
}

// One 1-D Newton (or Gauss-Newton) iteration loop in arithmetic T; returns the status.
template <typename T, bool NEWTON>
__device__ int curve_loop(const GnCurveDev &c, const double P[3], double &t, int max_iter,
                          double tol_dist, double tol_step, double tol_quad, int clamp, int &iters) {
    const T Px = T(P[0]), Py = T(P[1]), Pz = T(P[2]);
    const double lo = c.k[c.p], hi = c.k[c.n];
    int it = 0, st = GN_PT_MAXITER;
    while (it < max_iter) {
        CurvePt<T> o;
        curve_eval<NEWTON ? 2 : 1, T>(c, t, o);
        const T r0 = o.C[0] - Px, r1 = o.C[1] - Py, r2 = o.C[2] - Pz;
        const T rr = r0 * r0 + r1 * r1 + r2 * r2;
        if (!(rr < T(INFINITY))) { st = GN_PT_NONFINITE; break; }  // NaN or infinite residual
        if (tol_dist > 0 && sqrt(rr) < T(tol_dist)) { st = GN_PT_CONVERGED_DIST; break; }
        const T f = o.C1[0] * r0 + o.C1[1] * r1 + o.C1[2] * r2;
        const T c1 = o.C1[0] * o.C1[0] + o.C1[1] * o.C1[1] + o.C1[2] * o.C1[2];
        T fp = c1;
        bool exact = false;  // Newton step (exact second derivative of |C-P|^2/2)
        if (NEWTON) {
            const T f2 = fp + o.C2[0] * r0 + o.C2[1] * r1 + o.C2[2] * r2;
            if (f2 > T(0)) {
                fp = f2;
                exact = true;
            }
        }
        if (!(fp == fp)) { st = GN_PT_NONFINITE; break; }
        if (fp == T(0)) { st = GN_PT_SINGULAR; break; }
        const T dt = f * recip(fp);
        t -= double(dt);
        if (clamp) t = fmin(fmax(t, lo), hi);
        it++;
        if (!(t == t)) { st = GN_PT_NONFINITE; break; }
        if (tol_step > 0 && fabs(dt) < T(tol_step)) { st = GN_PT_CONVERGED; break; }
        // certificate of a Newton step dt: the error left is at most L dt^2 / (2 f''), with
        // L = 3 |C'| |C''| + |r| |C'''| bounding f''' and |C'''| taken as |C''| (p+1)/h (one more
        // derivative of the span's polynomial); squared, with L^2 <= 2 |C''|^2 (9 |C'|^2 + |r|^2 m^2).
        // The bound holds on one polynomial piece: a step that leaves the span crosses a knot,
        // where C'' may jump, and is not certified.
        if (NEWTON && exact && tol_quad > 0 && t >= o.lo && t <= o.hi) {
            const T c2 = o.C2[0] * o.C2[0] + o.C2[1] * o.C2[1] + o.C2[2] * o.C2[2];
            const T m = T(c.p + 1) * o.ih, t2 = dt * dt, tq = T(tol_quad);
            if (T(0.5) * c2 * (T(9) * c1 + rr * m * m) * (t2 * t2) <= tq * tq * fp * fp) {
                st = GN_PT_CONVERGED;
                break;
            }
        }
    }
    iters = it;
    return st;
}

__device__ __forceinline__ float d2_f32(const float *q, const float p[3]) {
    const float dx = q[0] - p[0], dy = q[1] - p[1], dz = q[2] - p[2];
    return fmaf(dz, dz, fmaf(dy, dy, dx * dx));
}

// Lower bound of d2_f32 over every sample inside the box (same rounding and operation order, so
// it never exceeds a sample's own distance).
__device__ __forceinline__ float box_d2_f32(const float *b, const float p[3]) {
    const float ex = fmaxf(fmaxf(b[0] - p[0], p[0] - b[3]), 0.f);
    const float ey = fmaxf(fmaxf(b[1] - p[1], p[1] - b[4]), 0.f);
    const float ez = fmaxf(fmaxf(b[2] - p[2], p[2] - b[5]), 0.f);
    return fmaf(ez, ez, fmaf(ey, ey, ex * ex));
}

// Nearest seed sample (smallest FP32 squared distance, lowest index among equals -- the result
// of a full scan): the samples of the span with the nearest box first, then every other span
// whose box is not farther than the best sample so far.
__device__ __forceinline__ int nearest_sample(const float *samp, const float *box, int ns,
                                              const float p[3]) {
    float bb = INFINITY;
    int bs = 0;
    for (int s = 0; s < ns; s++) {
        const float b = box_d2_f32(box + 6 * s, p);
        if (b < bb) { bb = b; bs = s; }
    }
    float best = INFINITY;
    int bj = bs * kCurveSamples;
    for (int j = bs * kCurveSamples; j < (bs + 1) * kCurveSamples; j++) {
        const float d2 = d2_f32(samp + 3 * j, p);
        if (d2 < best) { best = d2; bj = j; }
    }
    for (int s = 0; s < ns; s++) {
        if (s == bs || box_d2_f32(box + 6 * s, p) > best) continue;
        for (int j = s * kCurveSamples; j < (s + 1) * kCurveSamples; j++) {
            const float d2 = d2_f32(samp + 3 * j, p);
            if (d2 < best || (d2 == best && j < bj)) { best = d2; bj = j; }
        }
    }
    return bj;
}

// Seed (nearest sample, or the given t) and the iterations of one point; returns the status.
// sdat: the curve's samples followed by its span boxes (LDS or global).
__device__ __forceinline__ int curve_invert_point(const GnCurveDev &c, const float *sdat,
                                                  const double P[3], const gn_invert_opts &o,
                                                  double &t) {
    if (o.seed_mode != GN_SEED_GIVEN) {
        const float p[3] = {(float)P[0], (float)P[1], (float)P[2]};
        t = c.samp_t[nearest_sample(sdat, sdat + 3 * kCurveSamples * c.ns, c.ns, p)];
    }
    const double ts = t;
    int it, st;
    const bool n32 = o.newton & GN_NEWTON_FP32, n64 = o.newton & GN_NEWTON_FP64;
    int lim = o.max_iter;
    if (o.precision == GN_MIXED) {
        st = n32 ? curve_loop<float, true>(c, P, t, o.fp32_iters, 0.0, o.tol_step32, 0.0, o.clamp, it)
                 : curve_loop<float, false>(c, P, t, o.fp32_iters, 0.0, o.tol_step32, 0.0, o.clamp, it);
        lim = o.fp64_tail;
        if (st == GN_PT_SINGULAR || st == GN_PT_NONFINITE) {  // FP32 breakdown: FP64 restart
            t = ts;
            lim = o.max_iter;
        }
    }
    return n64 ? curve_loop<double, true>(c, P, t, lim, o.tol_dist, o.tol_step, o.tol_quad, o.clamp, it)
               : curve_loop<double, false>(c, P, t, lim, o.tol_dist, o.tol_step, 0.0, o.clamp, it);
}

// Floats of a curve's seed data: samples (3 per sample) then span boxes (6 per span).
__host__ __device__ inline size_t seed_floats(int ns) { return (size_t)ns * (3 * kCurveSamples + 6); }

// Stages the curve's seed data in LDS when use_lds is set; returns the array to read.
__device__ __forceinline__ const float *stage_samples(const GnCurveDev &c, float *lds, int use_lds) {
    if (!use_lds) return c.samp;
    const int nf = (int)seed_floats(c.ns);
    for (int i = threadIdx.x; i < nf; i += blockDim.x) lds[i] = c.samp[i];
    __syncthreads();
    return lds;
}

__global__ __launch_bounds__(256) void k_curve_invert(GnCurveDev c, const double *__restrict__ x,
                                                      const double *__restrict__ y,
                                                      const double *__restrict__ z, size_t n,
                                                      gn_invert_opts o, double *__restrict__ t_out,
                                                      uint8_t *__restrict__ status, int use_lds) {
    extern __shared__ float lds_samp[];  // seed data when use_lds
    const float *ssamp = stage_samples(c, lds_samp, use_lds);
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const double P[3] = {x[i], y[i], z[i]};
        double t = o.seed_mode == GN_SEED_GIVEN ? t_out[i] : 0.0;
        const int st = curve_invert_point(c, ssamp, P, o, t);
        t_out[i] = t;
        status[i] = (uint8_t)st;
    }
}

// Curve tiles of many requests: one workgroup per tile (grid-stride over the tile list), the
// tile's curve samples staged in LDS when they fit.  Results: t in the request's u array.
__global__ __launch_bounds__(256) void k_curve_batch(const GnCurveDev *__restrict__ curves,
                                                     const GnTile *__restrict__ tiles, int ntiles,
                                                     const GnReq *__restrict__ reqs,
                                                     gn_invert_opts o, int use_lds) {
    extern __shared__ float lds_samp[];
    for (int ti = blockIdx.x; ti < ntiles; ti += gridDim.x) {
        const GnTile tl = tiles[ti];
        const GnCurveDev c = curves[tl.obj];
        const GnReq r = reqs[tl.req];
        __syncthreads();  // the previous tile's samples are no longer read
        const float *ssamp = stage_samples(c, lds_samp, use_lds);
        for (int j = threadIdx.x; j < tl.count; j += blockDim.x) {
            const int64_t i = tl.start + j;
            const double P[3] = {r.x[i], r.y[i], r.z[i]};
            double t = o.seed_mode == GN_SEED_GIVEN ? r.u[i] : 0.0;
            const int st = curve_invert_point(c, ssamp, P, o, t);
            r.u[i] = t;
            r.status[i] = (uint8_t)st;
        }
    }
}

int launch_curve_batch(const GnCurveDev *curves, const GnTile *tiles, int ntiles, const GnReq *reqs,
                       int max_samples, const gn_invert_opts &o, hipStream_t st) {
    if (ntiles <= 0) return GN_OK;
    const size_t lds = sizeof(float) * seed_floats(max_samples / kCurveSamples);
    const int use_lds = lds <= 32 * 1024;
    const unsigned grid = (unsigned)(ntiles < 256 * 8 ? ntiles : 256 * 8);
    k_curve_batch<<<grid, 256, use_lds ? lds : 0, st>>>(curves, tiles, ntiles, reqs, o, use_lds);
    GN_HIP(hipGetLastError());
    return GN_OK;
}

}  // namespace gn

// ---- host side ---------------------------------------------------------------------------------
#include <vector>

using gn::fail;

static void free_curve(gn_curve *c) {
    if (!c) return;
    (void)hipFree(c->d_k);
    (void)hipFree(c->d_ih);
    (void)hipFree(c->d_xyz);
    (void)hipFree(c->d_w);
    (void)hipFree(c->d_tab64);
    (void)hipFree(c->d_tab32);
    (void)hipFree(c->d_samp);
    (void)hipFree(c->d_samp_t);
    delete c;
}

extern "C" {

int gn_curve_create(int degree, int cp, const double *knots, const double *xyz, const double *w,
                    gn_stream stream, gn_curve **out) {
    if (!out) return fail(GN_EINVAL, "out is null");
    *out = nullptr;
    if (!knots || !xyz || !w) return fail(GN_EINVAL, "null curve array");
    if (degree < 1 || degree > GN_MAXK - 1) return fail(GN_EINVAL, "degree must be in 1..7");
    if (cp < degree + 1 || cp > (1 << 24)) return fail(GN_EINVAL, "need degree+1 .. 2^24 control points");
    for (int i = 0; i < cp + degree + 1; i++)
        if (!std::isfinite(knots[i]) || (i > 0 && knots[i] < knots[i - 1]))
            return fail(GN_EINVAL, "knots must be finite and non-decreasing");
    if (!(knots[cp] > knots[degree])) return fail(GN_EINVAL, "empty parameter domain");
    for (int i = 0; i < cp; i++) {
        if (!std::isfinite(xyz[3 * i]) || !std::isfinite(xyz[3 * i + 1]) || !std::isfinite(xyz[3 * i + 2]))
            return fail(GN_EINVAL, "non-finite control point");
        if (!(w[i] > 0) || !std::isfinite(w[i])) return fail(GN_EINVAL, "weights must be > 0");
    }
    gn_curve *c = new gn_curve();
    c->p = degree;
    c->n = cp;
    c->k.assign(knots, knots + cp + degree + 1);
    c->xyz.assign(xyz, xyz + 3 * cp);
    c->w.assign(w, w + cp);
    const int ns = cp - degree, k = degree + 1;
    const size_t tabn = (size_t)ns * k * 4, nsamp = (size_t)ns * gn::kCurveSamples;
    hipStream_t st = (hipStream_t)stream;
    hipError_t e = hipSuccess;
    auto alloc = [&](void **p, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(p, bytes ? bytes : 1);
        c->bytes += bytes;
    };
    alloc((void **)&c->d_k, sizeof(double) * c->k.size());
    alloc((void **)&c->d_ih, sizeof(double) * ns);
    alloc((void **)&c->d_xyz, sizeof(double) * 3 * cp);
    alloc((void **)&c->d_w, sizeof(double) * cp);
    alloc((void **)&c->d_tab64, sizeof(double) * tabn);
    alloc((void **)&c->d_tab32, sizeof(float) * tabn);
    alloc((void **)&c->d_samp, sizeof(float) * gn::seed_floats(ns));
    alloc((void **)&c->d_samp_t, sizeof(double) * nsamp);
    if (e == hipSuccess)
        e = hipMemcpyAsync(c->d_k, c->k.data(), sizeof(double) * c->k.size(), hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(c->d_xyz, c->xyz.data(), sizeof(double) * 3 * cp, hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(c->d_w, c->w.data(), sizeof(double) * cp, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
        const GnCurveDev d = c->dev();
        gn::k_curve_table<<<gn::grid_for(ns, 64, 1u << 20), 64, 0, st>>>(d, c->d_xyz, c->d_w, c->d_ih,
                                                                         c->d_tab64, c->d_tab32);
        gn::k_curve_samples<<<gn::grid_for(nsamp, 64, 1u << 20), 64, 0, st>>>(d, c->d_samp, c->d_samp_t);
        gn::k_curve_boxes<<<gn::grid_for(ns, 64, 1u << 20), 64, 0, st>>>(d, c->d_samp + 3 * nsamp);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        free_curve(c);
        return gn::hip_fail(e, "curve set-up");
    }
    *out = c;
    return GN_OK;
}

int gn_curve_destroy(gn_curve *c) {
    if (!c) return GN_OK;
    (void)hipDeviceSynchronize();
    free_curve(c);
    return GN_OK;
}

int gn_curve_info(const gn_curve *c, int *degree, int *cp) {
    if (!c) return fail(GN_EINVAL, "null curve");
    if (degree) *degree = c->p;
    if (cp) *cp = c->n;
    return GN_OK;
}

size_t gn_curve_device_bytes(const gn_curve *c) { return c ? c->bytes : 0; }

int gn_curve_evaluate(const gn_curve *c, const double *t, size_t n, int max_order, double *out,
                      gn_stream stream) {
    if (!c) return fail(GN_EINVAL, "null curve");
    if (max_order < 0 || max_order > 2) return fail(GN_EINVAL, "max_order must be 0..2");
    if (n && (!t || !out)) return fail(GN_EINVAL, "null buffer");
    if (!n) return GN_OK;
    const unsigned grid = gn::grid_for(n, 256, 256 * 16);
    hipStream_t st = (hipStream_t)stream;
    const GnCurveDev d = c->dev();
    if (max_order == 0) gn::k_curve_eval<0><<<grid, 256, 0, st>>>(d, t, n, out);
    else if (max_order == 1) gn::k_curve_eval<1><<<grid, 256, 0, st>>>(d, t, n, out);
    else gn::k_curve_eval<2><<<grid, 256, 0, st>>>(d, t, n, out);
    GN_HIP(hipGetLastError());
    return GN_OK;
}

int gn_curve_invert(const gn_curve *c, const double *x, const double *y, const double *z, size_t n,
                    const gn_invert_opts *opts, double *t_out, uint8_t *status, gn_stream stream) {
    if (!c) return fail(GN_EINVAL, "null curve");
    gn_invert_opts o;
    if (opts) o = *opts;
    else gn_invert_opts_default(&o);
    if (o.precision != GN_FP64 && o.precision != GN_MIXED)
        return fail(GN_EINVAL, "precision must be GN_FP64 or GN_MIXED");
    if (o.seed_mode == GN_SEED_REF_FIRST_HIT) return fail(GN_EINVAL, "no tree seeding for curves");
    if (o.max_iter < 0 || o.fp32_iters < 0 || o.fp64_tail < 0 || o.max_iter > 10000 ||
        o.fp32_iters > 10000 || o.fp64_tail > 10000)
        return fail(GN_EINVAL, "iteration counts must be 0..10000");
    if (n && (!x || !y || !z || !t_out || !status)) return fail(GN_EINVAL, "null buffer");
    if (!n) return GN_OK;
    const size_t lds = sizeof(float) * gn::seed_floats(c->n - c->p);
    const int use_lds = lds <= 32 * 1024;  // longer curves read their samples from L2
    const unsigned grid = gn::grid_for(n, 256, 256 * 8);
    gn::k_curve_invert<<<grid, 256, use_lds ? lds : 0, (hipStream_t)stream>>>(c->dev(), x, y, z, n, o,
                                                                              t_out, status, use_lds);
    GN_HIP(hipGetLastError());
    return GN_OK;
}

}  // extern "C"
