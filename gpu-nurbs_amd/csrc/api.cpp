// api.cpp -- C ABI of libgnurbs (host side): object management, argument checks, the .nu reader
// and dispatch to the HIP kernels.  Errors are returned as gn_status codes with a per-thread
// message (gn_last_error); nothing here exits the process (the reference's HANDLE_ERROR,
// src/main.cu:34-43, calls exit()).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "gn_internal.hpp"

namespace gn {
static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }
int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}
int hip_fail(hipError_t e, const char *what) {
    g_err = std::string("HIP error ") + hipGetErrorString(e) + " in " + what;
    return e == hipErrorOutOfMemory ? GN_ENOMEM : GN_EDEVICE;
}
}  // namespace gn

using gn::fail;

template <typename T>
static int dalloc(T **p, size_t count, size_t &acc) {
    GN_HIP(hipMalloc((void **)p, sizeof(T) * (count ? count : 1)));
    acc += sizeof(T) * count;
    return GN_OK;
}

extern "C" {

const char *gn_last_error(void) { return gn::g_err.c_str(); }
const char *gn_version(void) { return "gnurbs 0.1 (gfx950)"; }

void gn_invert_opts_default(gn_invert_opts *o) {
    if (!o) return;
    o->tol_dist = 0.0;
    o->tol_step = 1e-10;
    o->max_iter = 20;
    o->clamp = 1;
    o->precision = GN_FP64;
    o->fp32_iters = 20;
    o->fp64_tail = 3;
    o->tol_step32 = 1e-6;
    o->seed_mode = GN_SEED_GREEDY_LEAF;
    o->newton = 0;
    o->tol_quad = 1e-10;
    o->fallback = 0;
}

static void free_tree(gn_surface *s) {
    (void)hipFree(s->d_tree);
    (void)hipFree(s->d_valid);
    (void)hipFree(s->d_divided);
    (void)hipFree(s->d_sbox);
    (void)hipFree(s->d_scentre);
    (void)hipFree(s->d_sseed);
    (void)hipFree(s->d_sbox32);
    (void)hipFree(s->d_sref);
    (void)hipFree(s->d_sleaf32);
    (void)hipFree(s->d_srec);
    (void)hipFree(s->d_scone32);
    (void)hipFree(s->d_grid);
    s->d_grid = nullptr;
    s->device_bytes -= s->tree_bytes + s->grid_bytes;
    s->tree_bytes = s->grid_bytes = 0;
    s->d_sleaf32 = nullptr;
    s->d_srec = nullptr;
    s->d_scone32 = nullptr;
    s->d_tree = nullptr;
    s->d_valid = s->d_divided = nullptr;
    s->d_sbox = s->d_scentre = s->d_sseed = nullptr;
    s->d_sbox32 = nullptr;
    s->d_sref = nullptr;
    s->tree_levels = 0;
}

static void free_surface(gn_surface *s) {
    if (!s) return;
    free_tree(s);
    (void)hipFree(s->d_ku);
    (void)hipFree(s->d_kv);
    (void)hipFree(s->d_ihu);
    (void)hipFree(s->d_ihv);
    (void)hipFree(s->d_xyz);
    (void)hipFree(s->d_w);
    (void)hipFree(s->d_tab64);
    (void)hipFree(s->d_tab32);
    delete s;
}

static int check_knots(const double *k, int n, const char *name) {
    for (int i = 0; i < n; i++) {
        if (!std::isfinite(k[i])) return fail(GN_EINVAL, std::string(name) + ": non-finite knot");
        if (i > 0 && k[i] < k[i - 1])
            return fail(GN_EINVAL, std::string(name) + ": knots must be non-decreasing");
    }
    return GN_OK;
}

int gn_surface_create(int p, int q, int nu, int nv, const double *ku, const double *kv,
                      const double *xyz, const double *w, gn_stream stream, gn_surface **out) {
    if (!out) return fail(GN_EINVAL, "out is null");
    *out = nullptr;
    if (!ku || !kv || !xyz || !w) return fail(GN_EINVAL, "null surface array");
    if (p < 1 || q < 1 || p > GN_MAXK - 1 || q > GN_MAXK - 1)
        return fail(GN_EINVAL, "degrees must be in 1..7");
    if (nu < p + 1 || nv < q + 1) return fail(GN_EINVAL, "need at least degree+1 control points");
    if ((long)nu * nv > (1L << 28)) return fail(GN_EINVAL, "control net too large");
    int rc = check_knots(ku, nu + p + 1, "knots_u");
    if (rc) return rc;
    rc = check_knots(kv, nv + q + 1, "knots_v");
    if (rc) return rc;
    if (!(ku[nu] > ku[p]) || !(kv[nv] > kv[q]))
        return fail(GN_EINVAL, "empty parameter domain");
    for (long i = 0; i < (long)nu * nv; i++) {
        if (!std::isfinite(xyz[3 * i]) || !std::isfinite(xyz[3 * i + 1]) ||
            !std::isfinite(xyz[3 * i + 2]))
            return fail(GN_EINVAL, "non-finite control point");
        if (!(w[i] > 0) || !std::isfinite(w[i])) return fail(GN_EINVAL, "weights must be > 0");
    }
    hipStream_t st = (hipStream_t)stream;
    gn_surface *s = new gn_surface();
    s->p = p; s->q = q; s->nu = nu; s->nv = nv;
    s->ku.assign(ku, ku + nu + p + 1);
    s->kv.assign(kv, kv + nv + q + 1);
    s->xyz.assign(xyz, xyz + 3L * nu * nv);
    s->w.assign(w, w + (long)nu * nv);
    const int su = nu - p, sv = nv - q;
    s->tab_entries = (size_t)su * sv * (p + 1) * (q + 1) * 4;
    size_t acc = 0;
    if ((rc = dalloc(&s->d_ku, s->ku.size(), acc)) || (rc = dalloc(&s->d_kv, s->kv.size(), acc)) ||
        (rc = dalloc(&s->d_ihu, (size_t)su, acc)) || (rc = dalloc(&s->d_ihv, (size_t)sv, acc)) ||
        (rc = dalloc(&s->d_xyz, s->xyz.size(), acc)) || (rc = dalloc(&s->d_w, s->w.size(), acc)) ||
        (rc = dalloc(&s->d_tab64, s->tab_entries, acc)) ||
        (rc = dalloc(&s->d_tab32, s->tab_entries, acc))) {
        free_surface(s);
        return rc;
    }
    s->device_bytes = acc;
    hipError_t e;
    if ((e = hipMemcpyAsync(s->d_ku, s->ku.data(), sizeof(double) * s->ku.size(),
                            hipMemcpyHostToDevice, st)) != hipSuccess ||
        (e = hipMemcpyAsync(s->d_kv, s->kv.data(), sizeof(double) * s->kv.size(),
                            hipMemcpyHostToDevice, st)) != hipSuccess ||
        (e = hipMemcpyAsync(s->d_xyz, s->xyz.data(), sizeof(double) * s->xyz.size(),
                            hipMemcpyHostToDevice, st)) != hipSuccess ||
        (e = hipMemcpyAsync(s->d_w, s->w.data(), sizeof(double) * s->w.size(),
                            hipMemcpyHostToDevice, st)) != hipSuccess) {
        free_surface(s);
        return gn::hip_fail(e, "hipMemcpyAsync(surface)");
    }
    rc = gn::launch_span_table(s, s->d_xyz, s->d_w, st);
    if (rc) {
        free_surface(s);
        return rc;
    }
    *out = s;
    return GN_OK;
}

// readNu, src/main.cu:45-113: tagged lines NU p q / CP nu nv / KNu ... / KNv ... / DP x y z w
// (u-major); other lines (comments "**", blanks) are skipped.
int gn_surface_from_nu(const char *path, gn_stream stream, gn_surface **out) {
    if (!path || !out) return fail(GN_EINVAL, "null argument");
    std::ifstream in(path);
    if (!in) return fail(GN_EIO, std::string("cannot open ") + path);
    int p = -1, q = -1, nu = -1, nv = -1;
    std::vector<double> ku, kv, xyz, w;
    bool have_ku = false, have_kv = false;
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::string tag;
        if (!(ls >> tag)) continue;
        if (tag == "NU") {
            ls >> p >> q;
        } else if (tag == "CP") {
            ls >> nu >> nv;
        } else if (tag == "KNu" || tag == "KNv") {
            if (p < 0 || nu < 0) return fail(GN_EIO, "knots before NU/CP");
            const int cnt = tag == "KNu" ? nu + p + 1 : nv + q + 1;
            std::vector<double> &k = tag == "KNu" ? ku : kv;
            k.resize(cnt);
            for (int i = 0; i < cnt; i++)
                if (!(ls >> k[i])) return fail(GN_EIO, "short knot vector " + tag);
            (tag == "KNu" ? have_ku : have_kv) = true;
        } else if (tag == "DP") {
            double x, y, z, ww;
            if (!(ls >> x >> y >> z >> ww)) return fail(GN_EIO, "bad DP line");
            xyz.push_back(x);
            xyz.push_back(y);
            xyz.push_back(z);
            w.push_back(ww);
        }
    }
    if (p < 0 || nu < 0 || !have_ku || !have_kv) return fail(GN_EIO, "missing NU/CP/KNu/KNv");
    if ((long)w.size() < (long)nu * nv) return fail(GN_EIO, "missing DP lines");
    xyz.resize(3L * nu * nv);
    w.resize((long)nu * nv);
    return gn_surface_create(p, q, nu, nv, ku.data(), kv.data(), xyz.data(), w.data(), stream,
                             out);
}

// Writes the definition in the same tagged format (NU / CP / KNu / KNv / DP, u-major control
// points), 17 significant digits, so gn_surface_from_nu reads back the identical surface.
int gn_surface_to_nu(const gn_surface *s, const char *path) {
    if (!s || !path) return fail(GN_EINVAL, "null argument");
    FILE *f = std::fopen(path, "w");
    if (!f) return fail(GN_EIO, std::string("cannot open ") + path + " for writing");
    std::fprintf(f, "** NURBS degree\nNU %d %d\nCP %d %d\n\n** NURBS Knots\nKNu", s->p, s->q, s->nu,
                 s->nv);
    for (double k : s->ku) std::fprintf(f, " %.17g", k);
    std::fprintf(f, "\nKNv");
    for (double k : s->kv) std::fprintf(f, " %.17g", k);
    std::fprintf(f, "\n\n** control points and weights\n");
    for (size_t i = 0; i < s->w.size(); i++)
        std::fprintf(f, "DP %.17g %.17g %.17g %.17g\n", s->xyz[3 * i], s->xyz[3 * i + 1],
                     s->xyz[3 * i + 2], s->w[i]);
    const bool ok = std::ferror(f) == 0;
    if (std::fclose(f) != 0 || !ok) return fail(GN_EIO, std::string("write failed: ") + path);
    return GN_OK;
}

int gn_surface_destroy(gn_surface *s) {
    if (!s) return GN_OK;
    (void)hipDeviceSynchronize();
    free_surface(s);
    return GN_OK;
}

int gn_surface_info(const gn_surface *s, int *p, int *q, int *nu, int *nv) {
    if (!s) return fail(GN_EINVAL, "null surface");
    if (p) *p = s->p;
    if (q) *q = s->q;
    if (nu) *nu = s->nu;
    if (nv) *nv = s->nv;
    return GN_OK;
}

size_t gn_surface_device_bytes(const gn_surface *s) { return s ? s->device_bytes : 0; }

int gn_find_span(const gn_surface *s, int dir, const double *u, size_t n, int32_t *span,
                 gn_stream stream) {
    if (!s) return fail(GN_EINVAL, "null surface");
    if (dir != 0 && dir != 1) return fail(GN_EINVAL, "dir must be 0 (u) or 1 (v)");
    if (n && (!u || !span)) return fail(GN_EINVAL, "null buffer");
    return gn::launch_find_span(s->dev(), dir, u, n, span, (hipStream_t)stream);
}

int gn_build_span_table(gn_surface *s, gn_stream stream) {
    if (!s) return fail(GN_EINVAL, "null surface");
    return gn::launch_span_table(s, s->d_xyz, s->d_w, (hipStream_t)stream);
}

int gn_get_span_matrix(const gn_surface *s, int i, int j, double *out) {
    if (!s || !out) return fail(GN_EINVAL, "null argument");
    if (i < s->p || i >= s->nu || j < s->q || j >= s->nv)
        return fail(GN_EINVAL, "span index outside the domain spans");
    const int k1 = s->p + 1, k2 = s->q + 1, sv = s->nv - s->q;
    const size_t slot = (size_t)(i - s->p) * sv + (j - s->q);
    std::vector<double> buf((size_t)k1 * k2 * 4);
    GN_HIP(hipDeviceSynchronize());
    GN_HIP(hipMemcpy(buf.data(), s->d_tab64 + slot * k1 * k2 * 4, sizeof(double) * buf.size(),
                     hipMemcpyDeviceToHost));
    for (int c = 0; c < 4; c++)
        for (int r = 0; r < k1; r++)
            for (int t = 0; t < k2; t++) out[(c * k1 + r) * k2 + t] = buf[(r * k2 + t) * 4 + c];
    return GN_OK;
}

int gn_evaluate(const gn_surface *s, const double *u, const double *v, size_t n, int max_order,
                double *out, gn_stream stream) {
    if (!s) return fail(GN_EINVAL, "null surface");
    if (max_order < 0 || max_order > 3) return fail(GN_EINVAL, "max_order must be 0..3");
    if (n && (!u || !v || !out)) return fail(GN_EINVAL, "null buffer");
    return gn::launch_eval64(s->dev(), u, v, n, max_order, out, (hipStream_t)stream);
}

int gn_evaluate_f32(const gn_surface *s, const float *u, const float *v, size_t n, int max_order,
                    float *out, gn_stream stream) {
    if (!s) return fail(GN_EINVAL, "null surface");
    if (max_order < 0 || max_order > 3) return fail(GN_EINVAL, "max_order must be 0..3");
    if (n && (!u || !v || !out)) return fail(GN_EINVAL, "null buffer");
    return gn::launch_eval32(s->dev(), u, v, n, max_order, out, (hipStream_t)stream);
}

int gn_evaluate_grid_f32(const gn_surface *s, const float *us, int nu_pts, const float *vs,
                         int nv_pts, int max_order, float *out, gn_stream stream) {
    if (!s) return fail(GN_EINVAL, "null surface");
    if (max_order < 0 || max_order > 2) return fail(GN_EINVAL, "max_order must be 0..2");
    if (nu_pts < 0 || nv_pts < 0) return fail(GN_EINVAL, "negative grid size");
    if (nu_pts > (1 << 20) || nv_pts > (1 << 20))
        return fail(GN_EINVAL, "grid sides are limited to 2^20 samples");
    if ((size_t)nu_pts * nv_pts && (!us || !vs || !out)) return fail(GN_EINVAL, "null buffer");
    return gn::launch_eval_grid32(s->dev(), us, nu_pts, vs, nv_pts, max_order, out,
                                  (hipStream_t)stream);
}

int gn_evaluate_sorted_f32(const gn_surface *s, const float *u, const float *v, size_t n,
                           int max_order, float *out, uint32_t *perm, gn_stream stream) {
    if (!s) return fail(GN_EINVAL, "null surface");
    if (max_order < 0 || max_order > 2) return fail(GN_EINVAL, "max_order must be 0..2");
    if (n > (size_t)UINT32_MAX) return fail(GN_EINVAL, "at most 2^32-1 points per call");
    if (n && (!u || !v || !out || !perm)) return fail(GN_EINVAL, "null buffer");
    return gn::launch_eval_sorted32(s->dev(), u, v, n, max_order, out, perm, (hipStream_t)stream);
}

int gn_build_tree(gn_surface *s, int levels, int flags, gn_stream stream) {
    if (!s) return fail(GN_EINVAL, "null surface");
    if (levels < 1 || levels > 4) return fail(GN_EINVAL, "levels must be 1..4");
    hipStream_t st = (hipStream_t)stream;
    GN_HIP(hipStreamSynchronize(st));
    free_tree(s);
    long count = 0, g = 1;
    for (int l = 0; l < levels; l++) { count += g * g; g *= 8; }
    const int D = 3 * (levels - 1);
    const long qnodes = ((1L << (2 * (D + 1))) - 1) / 3, leaves = 1L << (2 * D);
    size_t acc = 0;
    int rc;
    if ((rc = dalloc(&s->d_tree, (size_t)count, acc)) ||
        (rc = dalloc(&s->d_valid, (size_t)count, acc)) ||
        (rc = dalloc(&s->d_divided, (size_t)count, acc)) ||
        (rc = dalloc(&s->d_sbox, (size_t)qnodes * 6, acc)) ||
        (rc = dalloc(&s->d_sbox32, (size_t)qnodes * 6, acc)) ||
        (rc = dalloc(&s->d_scentre, (size_t)leaves * 3, acc)) ||
        (rc = dalloc(&s->d_sseed, (size_t)leaves * 2, acc)) ||
        (rc = dalloc(&s->d_sref, (size_t)leaves, acc)) ||
        (rc = dalloc(&s->d_sleaf32, (size_t)leaves, acc)) ||
        (rc = dalloc(&s->d_srec, (size_t)leaves * kRecVec, acc)) ||
        (rc = dalloc(&s->d_scone32, (size_t)leaves, acc))) {
        free_tree(s);
        return rc;
    }
    const GnSurfDev d = s->dev();
    for (int l = 0; l < levels; l++) {
        rc = gn::launch_tree_level(d, l, flags, s->d_tree, s->d_valid, s->d_divided, st);
        if (rc) { free_tree(s); return rc; }
    }
    rc = gn::launch_search_build(d, levels, s->d_tree, s->d_sbox, s->d_sbox32, s->d_scentre,
                                 s->d_sseed, s->d_sref, s->d_sleaf32, s->d_srec, s->d_scone32,
                                 st);
    if (rc) { free_tree(s); return rc; }
    s->tree_levels = levels;
    s->tree_bytes = acc;
    s->device_bytes += acc;
    return GN_OK;
}

int gn_build_seed_grid(gn_surface *s, int resolution, gn_stream stream) {
    if (!s) return fail(GN_EINVAL, "null surface");
    if (!s->tree_levels) return fail(GN_ESTATE, "tree not built");
    if (resolution < 0 || resolution > 1024) return fail(GN_EINVAL, "resolution must be 0..1024");
    hipStream_t st = (hipStream_t)stream;
    const int D = 3 * (s->tree_levels - 1);
    const long leaves = 1L << (2 * D);
    std::vector<float4> lf((size_t)leaves);
    GN_HIP(hipMemcpyAsync(lf.data(), s->d_sleaf32, sizeof(float4) * leaves, hipMemcpyDeviceToHost, st));
    GN_HIP(hipStreamSynchronize(st));
    // geometry: cubic cells of the largest extent of the leaf centres / R, R/16 cells (at least
    // one) of margin on every side.  Plain FP32 operations, one per statement, so that the
    // oracle's restatement rounds identically.
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (const float4 &c : lf) {
        const float e[3] = {c.x, c.y, c.z};
        for (int k = 0; k < 3; k++) {
            lo[k] = std::fmin(lo[k], e[k]);
            hi[k] = std::fmax(hi[k], e[k]);
        }
    }
    float ext = 0.f;
    for (int k = 0; k < 3; k++) {
        if (!std::isfinite(lo[k]) || !std::isfinite(hi[k]))
            return fail(GN_EINVAL, "leaf centres are not finite");
        const float w = hi[k] - lo[k];
        ext = std::fmax(ext, w);
    }
    if (!(ext > 0.f) || !std::isfinite(ext)) ext = 1.f;
    int dims[3];
    float org[3], cell = 0.f;
    size_t cells = 0;
    auto shape = [&](int R) {
        const int pad = std::max(1, R / 16);
        cell = ext / (float)R;
        cells = 1;
        for (int k = 0; k < 3; k++) {
            const float w = hi[k] - lo[k];
            const float span = w / cell;
            const float off = (float)pad * cell;
            org[k] = lo[k] - off;
            dims[k] = (int)std::ceil(span) + 2 * pad;
            cells *= (size_t)dims[k];
        }
    };
    // the build compares every cell with every leaf: at most 2^36 pairs (about a second)
    const size_t max_pairs = (size_t)1 << 36;
    if (resolution) {
        shape(resolution);
        if (cells > ((size_t)1 << 25) || cells * (size_t)leaves > max_pairs)
            return fail(GN_EINVAL, "seed grid above 2^25 cells or 2^36 cell-leaf pairs: lower the resolution");
    } else {
        int R = std::min(256, 2 << D);
        shape(R);
        while (R > 8 && (cells > ((size_t)1 << 25) || cells * (size_t)leaves > max_pairs)) shape(R /= 2);
        if (cells > ((size_t)1 << 25) || cells * (size_t)leaves > max_pairs)
            return fail(GN_EINVAL, "seed grid too large for this tree");
    }
    const int wide = leaves > 65536 ? 1 : 0;
    const size_t bytes = cells * (wide ? 4 : 2);
    GN_HIP(hipStreamSynchronize(st));
    (void)hipFree(s->d_grid);
    s->d_grid = nullptr;
    s->device_bytes -= s->grid_bytes;
    s->grid_bytes = 0;
    GN_HIP(hipMalloc(&s->d_grid, bytes));
    const int rc = gn::launch_seed_grid_build(s->d_sleaf32, D, dims, org, cell, wide, s->d_grid, st);
    if (rc) {
        (void)hipFree(s->d_grid);
        s->d_grid = nullptr;
        return rc;
    }
    s->grid_wide = wide;
    s->gcell = cell;
    s->ginv = 1.f / cell;
    for (int k = 0; k < 3; k++) {
        s->gdim[k] = dims[k];
        s->gorg[k] = org[k];
    }
    s->grid_bytes = bytes;
    s->device_bytes += bytes;
    return GN_OK;
}

int gn_get_seed_grid(const gn_surface *s, int32_t *leaf, size_t count, int dims[3],
                     float origin[3], float *cell) {
    if (!s) return fail(GN_EINVAL, "null surface");
    if (!s->d_grid) return fail(GN_ESTATE, "seed grid not built");
    const size_t cells = (size_t)s->gdim[0] * s->gdim[1] * s->gdim[2];
    for (int k = 0; k < 3; k++) {
        if (dims) dims[k] = s->gdim[k];
        if (origin) origin[k] = s->gorg[k];
    }
    if (cell) *cell = s->gcell;
    if (!leaf) return GN_OK;
    if (count < cells) return fail(GN_EINVAL, "count smaller than the cell count");
    const int D = 3 * (s->tree_levels - 1);
    std::vector<uint32_t> m(cells);
    GN_HIP(hipDeviceSynchronize());
    if (s->grid_wide) {
        GN_HIP(hipMemcpy(m.data(), s->d_grid, sizeof(uint32_t) * cells, hipMemcpyDeviceToHost));
    } else {
        std::vector<uint16_t> h(cells);
        GN_HIP(hipMemcpy(h.data(), s->d_grid, sizeof(uint16_t) * cells, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < cells; i++) m[i] = h[i];
    }
    // Morton (x bits odd) -> reference index of the finest level
    long base = 0, g = 1;
    for (int l = 0; l < s->tree_levels - 1; l++) { base += g * g; g *= 8; }
    for (size_t i = 0; i < cells; i++) {
        uint32_t bx = 0, by = 0;
        for (int b = 0; b < D; b++) {
            bx |= ((m[i] >> (2 * b + 1)) & 1u) << b;
            by |= ((m[i] >> (2 * b)) & 1u) << b;
        }
        leaf[i] = (int32_t)(base + (long)bx * g + by);
    }
    return GN_OK;
}

int gn_get_tree(const gn_surface *s, double *nodes, uint8_t *valid, uint8_t *divided,
                size_t count) {
    if (!s) return fail(GN_EINVAL, "null surface");
    if (!s->tree_levels) return fail(GN_ESTATE, "tree not built");
    long c = 0, g = 1;
    for (int l = 0; l < s->tree_levels; l++) { c += g * g; g *= 8; }
    if (count < (size_t)c) return fail(GN_EINVAL, "count smaller than the tree");
    GN_HIP(hipDeviceSynchronize());
    if (nodes) GN_HIP(hipMemcpy(nodes, s->d_tree, sizeof(GnTreeNode) * c, hipMemcpyDeviceToHost));
    if (valid) GN_HIP(hipMemcpy(valid, s->d_valid, c, hipMemcpyDeviceToHost));
    if (divided) GN_HIP(hipMemcpy(divided, s->d_divided, c, hipMemcpyDeviceToHost));
    return GN_OK;
}

int gn_get_normal_cones(const gn_surface *s, float *cones, size_t count) {
    if (!s) return fail(GN_EINVAL, "null surface");
    if (!s->tree_levels) return fail(GN_ESTATE, "tree not built");
    if (!cones) return fail(GN_EINVAL, "null buffer");
    long g = 1;
    for (int l = 1; l < s->tree_levels; l++) g *= 8;
    const size_t leaves = (size_t)(g * g);
    if (count < leaves) return fail(GN_EINVAL, "count smaller than the leaf count");
    std::vector<float4> m(leaves);
    GN_HIP(hipDeviceSynchronize());
    GN_HIP(hipMemcpy(m.data(), s->d_scone32, sizeof(float4) * leaves, hipMemcpyDeviceToHost));
    // Morton order (x bits odd) -> leaf index bx * g + by
    for (long bx = 0; bx < g; bx++)
        for (long by = 0; by < g; by++) {
            uint32_t c = 0;
            for (int b = 0; b < 16; b++)
                c |= (((uint32_t)bx >> b & 1u) << (2 * b + 1)) | (((uint32_t)by >> b & 1u) << (2 * b));
            const float4 &f = m[c];
            float *o = cones + (size_t)(bx * g + by) * 4;
            o[0] = f.x; o[1] = f.y; o[2] = f.z; o[3] = f.w;
        }
    return GN_OK;
}

int gn_seed(const gn_surface *s, int exact, const double *x, const double *y, const double *z,
            size_t n, int32_t *leaf, double *su, double *sv, gn_stream stream) {
    if (!s) return fail(GN_EINVAL, "null surface");
    if (!s->tree_levels) return fail(GN_ESTATE, "tree not built");
    if (n && (!x || !y || !z || !su || !sv)) return fail(GN_EINVAL, "null buffer");
    if (exact == 3 && !s->d_grid) return fail(GN_ESTATE, "seed grid not built (gn_build_seed_grid)");
    // exact: 0 greedy leaf, 2 greedy leaf + Gauss-Newton step, 3 grid leaf + the step, any other
    // value the exact search
    return gn::launch_seed(s->search(), exact == 0 ? 1 : exact == 2 ? 2 : exact == 3 ? 3 : 0, x, y,
                           z, n, leaf, su, sv, (hipStream_t)stream);
}

int gn_invert(const gn_surface *s, const double *x, const double *y, const double *z, size_t n,
              const gn_invert_opts *opts, double *u, double *v, uint8_t *status,
              gn_stream stream) {
    return gn_invert_iters(s, x, y, z, n, opts, u, v, status, nullptr, nullptr, stream);
}

int gn_invert_iters(const gn_surface *s, const double *x, const double *y, const double *z,
                    size_t n, const gn_invert_opts *opts, double *u, double *v, uint8_t *status,
                    uint8_t *iters_fp32, uint8_t *iters_fp64, gn_stream stream) {
    gn_invert_opts o;
    if (opts) o = *opts;
    else gn_invert_opts_default(&o);
    const int rc = gn::check_invert(s, x, y, z, n, o, u, v, status);
    if (rc) return rc;
    return gn::launch_invert(s->dev(), s->search(), s->d_tree, s->d_valid, s->tree_levels, x, y,
                             z, n, o, u, v, status, iters_fp32, iters_fp64, (hipStream_t)stream);
}

}  // extern "C"

namespace gn {

int check_invert(const gn_surface *s, const double *x, const double *y, const double *z, size_t n,
                 const gn_invert_opts &o, const double *u, const double *v, const uint8_t *status) {
    if (!s) return fail(GN_EINVAL, "null surface");
    if (n && (!x || !y || !z || !u || !v || !status)) return fail(GN_EINVAL, "null buffer");
    if (o.precision != GN_FP64 && o.precision != GN_MIXED)
        return fail(GN_EINVAL, "precision must be GN_FP64 or GN_MIXED");
    if (o.seed_mode < GN_SEED_NEAREST_LEAF || o.seed_mode > GN_SEED_GRID_STEP)
        return fail(GN_EINVAL, "bad seed_mode");
    if (o.seed_mode == GN_SEED_GRID_STEP && !s->d_grid)
        return fail(GN_ESTATE, "seed grid not built (gn_build_seed_grid)");
    if (o.seed_mode == GN_SEED_GLOBAL && o.precision != GN_FP64)
        return fail(GN_EINVAL, "global projection runs in FP64 (precision GN_FP64)");
    if (o.fallback && !s->tree_levels)
        return fail(GN_ESTATE, "the global fallback needs the tree (gn_build_tree)");
    if (o.max_iter < 0 || o.fp32_iters < 0 || o.fp64_tail < 0 || o.max_iter > 10000 ||
        o.fp32_iters > 10000 || o.fp64_tail > 10000)
        return fail(GN_EINVAL, "iteration counts must be 0..10000");
    if (o.seed_mode != GN_SEED_GIVEN && !s->tree_levels)
        return fail(GN_ESTATE, "tree not built (gn_build_tree) and no seeds given");
    if (o.seed_mode == GN_SEED_REF_FIRST_HIT && o.precision != GN_FP64)
        return fail(GN_EINVAL, "first-hit seeding is the plain FP64 reference path");
    return GN_OK;
}

}  // namespace gn
