// gnurbs.hpp -- C++ host interface over the C ABI of libgnurbs (include/gnurbs.h).
//
// The reference passes a surface around as a loose tuple (degree_u, degree_v, cp_u, cp_v,
// knots_u, knots_v, control_points, weights) and exposes CUDA kernels launched from
// src/main.cu.  NurbsSurface wraps that tuple in a class, owns its device-resident span tables
// and tree, and keeps the reference's entry-point names as methods:
//   readNu        src/main.cu:45-113
//   findSpan      src/NURBS/Evaluate.cu:504-533
//   GetMatrixN    src/NURBS/Evaluate.cu:3-88       (span table; host copy of one span matrix)
//   EvaluateRandom src/NURBS/Evaluate.cu:1023-1038 (plus the derivatives of 535-930)
//   BuildTree     src/kernel/kernel.cu:375-431, 544-621, 707-733 (launch sequence main.cu:280-309)
//   NewInverse    src/kernel/kernel.cu:433-469 with GaussNewton 151-249
// NurbsCurve (rational curves: the reference's README names them, its code has none) and
// NurbsSurface::PCurve (the p-curve use of README.md:10) have no reference entry points.
// Errors throw gnurbs::Error carrying the gn_status code and message (the reference exits the
// process in HANDLE_ERROR, src/main.cu:34-43).  Query and result buffers are device pointers
// owned by the caller, as in the reference.
#ifndef GNURBS_HPP
#define GNURBS_HPP

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "gnurbs.h"

namespace gnurbs {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc) {
    if (rc != GN_OK) throw Error(rc, gn_last_error());
}

inline gn_invert_opts default_opts() {
    gn_invert_opts o;
    gn_invert_opts_default(&o);
    return o;
}

class NurbsCurve;

class NurbsSurface {
  public:
    NurbsSurface(int degree_u, int degree_v, int cp_u, int cp_v, const std::vector<double> &knots_u,
                 const std::vector<double> &knots_v, const std::vector<double> &control_points_xyz,
                 const std::vector<double> &weights, gn_stream stream = nullptr) {
        if ((long)knots_u.size() != cp_u + degree_u + 1 || (long)knots_v.size() != cp_v + degree_v + 1 ||
            (long)control_points_xyz.size() != 3L * cp_u * cp_v || (long)weights.size() != (long)cp_u * cp_v)
            throw Error(GN_EINVAL, "surface arrays do not match the degrees and net size");
        check(gn_surface_create(degree_u, degree_v, cp_u, cp_v, knots_u.data(), knots_v.data(),
                                control_points_xyz.data(), weights.data(), stream, &s_));
        check(gn_surface_info(s_, &p_, &q_, &nu_, &nv_));
    }
    static NurbsSurface readNu(const std::string &path, gn_stream stream = nullptr) {
        gn_surface *s = nullptr;
        check(gn_surface_from_nu(path.c_str(), stream, &s));
        return NurbsSurface(s);
    }
    // the same text format, read back identically by readNu
    void writeNu(const std::string &path) const { check(gn_surface_to_nu(s_, path.c_str())); }
    NurbsSurface(const NurbsSurface &) = delete;
    NurbsSurface &operator=(const NurbsSurface &) = delete;
    NurbsSurface(NurbsSurface &&o) noexcept { *this = std::move(o); }
    NurbsSurface &operator=(NurbsSurface &&o) noexcept {
        if (this != &o) {
            reset();
            s_ = o.s_;
            p_ = o.p_; q_ = o.q_; nu_ = o.nu_; nv_ = o.nv_; levels_ = o.levels_;
            o.s_ = nullptr;
        }
        return *this;
    }
    ~NurbsSurface() { reset(); }

    int degree_u() const { return p_; }
    int degree_v() const { return q_; }
    int cp_u() const { return nu_; }
    int cp_v() const { return nv_; }
    gn_surface *handle() const { return s_; }

    // findSpan for n device parameters (dir 0 = u, 1 = v) into device int32 spans
    void findSpan(int dir, const double *u, size_t n, int32_t *span, gn_stream st = nullptr) const {
        check(gn_find_span(s_, dir, u, n, span, st));
    }
    // span matrix (i, j) of MatrixN, [c][r][s], 4*(p+1)*(q+1) doubles
    std::vector<double> GetMatrixN(int i, int j) const {
        std::vector<double> m(4 * (p_ + 1) * (q_ + 1));
        check(gn_get_span_matrix(s_, i, j, m.data()));
        return m;
    }
    // points (order 0) and derivatives up to order 3, SoA blocks of n on the device
    void EvaluateRandom(const double *u, const double *v, size_t n, int order, double *out,
                        gn_stream st = nullptr) const {
        check(gn_evaluate(s_, u, v, n, order, out, st));
    }
    void EvaluateRandomF32(const float *u, const float *v, size_t n, int order, float *out,
                           gn_stream st = nullptr) const {
        check(gn_evaluate_f32(s_, u, v, n, order, out, st));
    }
    void EvaluateGrid(const float *us, int nu, const float *vs, int nv, int order, float *out,
                      gn_stream st = nullptr) const {
        check(gn_evaluate_grid_f32(s_, us, nu, vs, nv, order, out, st));
    }
    // FP32 points in any order on the matrix cores: column k of out is input point perm[k]
    void EvaluateSorted(const float *u, const float *v, size_t n, int order, float *out,
                        uint32_t *perm, gn_stream st = nullptr) const {
        check(gn_evaluate_sorted_f32(s_, u, v, n, order, out, perm, st));
    }
    // tree of `levels` 8x8 levels (the reference's `level` loop of src/main.cu:283)
    void BuildTree(int levels, int flags = 0, gn_stream st = nullptr) {
        check(gn_build_tree(s_, levels, flags, st));
        levels_ = levels;
    }
    // host copy of the tree: 13 doubles per node (uv box, xyz box, centre), valid, divided
    void GetTree(std::vector<double> &nodes, std::vector<uint8_t> &valid,
                 std::vector<uint8_t> &divided) const {
        size_t count = 0, g = 1;
        for (int l = 0; l < levels_; l++, g *= 8) count += g * g;
        nodes.resize(13 * count);
        valid.resize(count);
        divided.resize(count);
        check(gn_get_tree(s_, nodes.data(), valid.data(), divided.data(), count));
    }
    // normal cones of the finest-level leaves: axis xyz and half-angle (4: unbounded) per leaf
    std::vector<float> GetNormalCones() const {
        size_t g = 1;
        for (int l = 1; l < levels_; l++) g *= 8;
        std::vector<float> c(4 * g * g);
        check(gn_get_normal_cones(s_, c.data(), g * g));
        return c;
    }
    // seed grid for GN_SEED_GRID_STEP (after BuildTree; 0: the default resolution)
    void BuildSeedGrid(int resolution = 0, gn_stream st = nullptr) {
        check(gn_build_seed_grid(s_, resolution, st));
    }
    // host copy of the seed grid: cells per axis, origin, cell size and each cell's leaf
    // (reference index) at (i * dims[1] + j) * dims[2] + k
    struct SeedGrid {
        int dims[3] = {0, 0, 0};
        float origin[3] = {0.f, 0.f, 0.f}, cell = 0.f;
        std::vector<int32_t> leaf;
    };
    SeedGrid GetSeedGrid() const {
        SeedGrid g;
        check(gn_get_seed_grid(s_, nullptr, 0, g.dims, g.origin, &g.cell));
        g.leaf.resize((size_t)g.dims[0] * g.dims[1] * g.dims[2]);
        check(gn_get_seed_grid(s_, g.leaf.data(), g.leaf.size(), nullptr, nullptr, nullptr));
        return g;
    }
    // seed search alone (exact: 0 greedy leaf, 2 greedy descent + one Gauss-Newton step, 3 seed
    // grid leaf + the step, other values the exact nearest leaf) into device leaf indices and
    // seeds
    void Seed(const double *x, const double *y, const double *z, size_t n, int32_t *leaf,
              double *su, double *sv, int exact = 0, gn_stream st = nullptr) const {
        check(gn_seed(s_, exact, x, y, z, n, leaf, su, sv, st));
    }
    // inversion of device points (x, y, z) into device (u, v, status)
    void NewInverse(const double *x, const double *y, const double *z, size_t n, double *u,
                    double *v, uint8_t *status, const gn_invert_opts &o = default_opts(),
                    gn_stream st = nullptr) const {
        check(gn_invert(s_, x, y, z, n, &o, u, v, status, st));
    }
    // NewInverse with per-point iteration counts (FP32 bulk, FP64); either may be null
    void NewInverseIters(const double *x, const double *y, const double *z, size_t n, double *u,
                         double *v, uint8_t *status, uint8_t *iters_fp32, uint8_t *iters_fp64,
                         const gn_invert_opts &o = default_opts(), gn_stream st = nullptr) const {
        check(gn_invert_iters(s_, x, y, z, n, &o, u, v, status, iters_fp32, iters_fp64, st));
    }
    // NewInverse for callers that cannot allocate (scratch is a fixed device array; calls that
    // overlap on one GPU race)
    void NewInverseNoAlloc(const double *x, const double *y, const double *z, size_t n, double *u,
                           double *v, uint8_t *status, const gn_invert_opts &o = default_opts(),
                           gn_stream st = nullptr) const {
        check(gn_invert_noalloc(s_, x, y, z, n, &o, u, v, status, st));
    }
    // p-curve of `c` on this surface: control points (u, v) and knots of the fitted B-spline
    struct PCurveFit {
        std::vector<double> cp_uv, knots;
        double max_dev = 0;
        size_t failed = 0;
    };
    inline PCurveFit PCurve(const NurbsCurve &c, size_t samples, int degree, int ncp,
                            const gn_invert_opts &o = default_opts(), gn_stream st = nullptr) const;

  private:
    explicit NurbsSurface(gn_surface *s) : s_(s) { check(gn_surface_info(s_, &p_, &q_, &nu_, &nv_)); }
    void reset() {
        if (s_) gn_surface_destroy(s_);
        s_ = nullptr;
    }
    gn_surface *s_ = nullptr;
    int p_ = 0, q_ = 0, nu_ = 0, nv_ = 0, levels_ = 0;
};

// A rational B-spline curve resident in HBM (degree, control points xyz[cp][3], weights,
// clamped knots[cp + degree + 1]).
class NurbsCurve {
  public:
    NurbsCurve(int degree, const std::vector<double> &knots,
               const std::vector<double> &control_points_xyz, const std::vector<double> &weights,
               gn_stream stream = nullptr) {
        const int cp = (int)weights.size();
        if ((long)knots.size() != (long)cp + degree + 1 || (long)control_points_xyz.size() != 3L * cp)
            throw Error(GN_EINVAL, "curve arrays do not match the degree and control point count");
        check(gn_curve_create(degree, cp, knots.data(), control_points_xyz.data(), weights.data(),
                              stream, &c_));
    }
    NurbsCurve(const NurbsCurve &) = delete;
    NurbsCurve &operator=(const NurbsCurve &) = delete;
    NurbsCurve(NurbsCurve &&o) noexcept : c_(o.c_) { o.c_ = nullptr; }
    NurbsCurve &operator=(NurbsCurve &&o) noexcept {
        if (this != &o) {
            if (c_) gn_curve_destroy(c_);
            c_ = o.c_;
            o.c_ = nullptr;
        }
        return *this;
    }
    ~NurbsCurve() {
        if (c_) gn_curve_destroy(c_);
    }
    // C (and C', C'') at device parameters t, SoA blocks of n on the device
    void Evaluate(const double *t, size_t n, int order, double *out, gn_stream st = nullptr) const {
        check(gn_curve_evaluate(c_, t, n, order, out, st));
    }
    // nearest-point parameters t of device points (x, y, z)
    void Invert(const double *x, const double *y, const double *z, size_t n, double *t,
                uint8_t *status, const gn_invert_opts &o = default_opts(),
                gn_stream st = nullptr) const {
        check(gn_curve_invert(c_, x, y, z, n, &o, t, status, st));
    }
    const gn_curve *handle() const { return c_; }

  private:
    gn_curve *c_ = nullptr;
};

inline NurbsSurface::PCurveFit NurbsSurface::PCurve(const NurbsCurve &c, size_t samples, int degree,
                                                    int ncp, const gn_invert_opts &o,
                                                    gn_stream st) const {
    PCurveFit f;
    f.cp_uv.resize(2 * (size_t)ncp);
    f.knots.resize((size_t)ncp + degree + 1);
    check(gn_pcurve(s_, c.handle(), samples, degree, ncp, &o, f.cp_uv.data(), f.knots.data(),
                    &f.max_dev, &f.failed, st));
    return f;
}

}  // namespace gnurbs

#endif  // GNURBS_HPP
