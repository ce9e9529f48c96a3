/*
 * gnurbs_oracle.c -- CPU FP64 oracle for matrix-form NURBS evaluation and inversion.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load it (through ctypes, as oracle/liboracle.so) to compare the HIP
 * product path against.  Nothing in the product library links, calls or falls back to it.
 *
 * It restates, in plain C, the algorithms of Qiang-Zou/GPU-NURBS (paths relative to the reference
 * root).  Every function names the file:line it follows.  Deliberate differences from the reference:
 *   - derivative chain factor is 1/(k[m+1]-k[m]); the reference uses (k[last]-k[0])/(k[m+1]-k[m])
 *     (src/NURBS/Evaluate.cu:120, src/kernel/kernel.cu:167-168).  Identical for knots on [0,1].
 *   - only domain spans m in [p, n-1] are built (the reference also visits, and skips, the
 *     zero-length end slots; src/NURBS/Evaluate.cu:10-11).
 *   - the tree uses the geometric parent at every level unless OR_TREE_REF_PARENT is set, which
 *     reproduces (tree_index-1)/64 of src/kernel/kernel.cu:380,599.
 *   - Gauss-Newton reports a singular 2x2 system instead of dividing by det == 0
 *     (src/kernel/kernel.cu:222-226 has no guard).
 *
 * Pinning (every exported function is checked by tests/test_oracle.py):
 *   or_find_span        brute-force linear scan, boundary cases
 *   or_nrc / basis      independent Cox-de Boor power-basis build (or_basis_matrix_cdb), Bernstein
 *                       closed form, partition of unity; scipy BSpline
 *   or_get_matrix_n     contraction against scipy NURBS evaluation; known-answer points of
 *                       src/kernel/kernel.cu:451-454 on model/center_bk.nu and model/test_face2.nu
 *   or_eval             scipy NURBS value, finite differences of order k -> k+1, Leibniz form,
 *                       closed forms (bilinear patch, exact quarter cylinder); third order against
 *                       the generic quotient rule (The NURBS Book eq. 4.20) on scipy's basis and
 *                       sympy derivatives of rational Bezier patches (tests/test_oracle_order3.py)
 *   or_gauss_newton     round trips on on-surface points, orthogonality of the residual
 *   or_build_tree       containment of dense samples, index formula, NeedDivide truth table
 *   or_seed_*           brute force
 *   or_leaf_cones       every normal of a dense sample grid of each leaf lies in its cone; exact
 *                       half-angles of circular-cylinder leaves (the normal turns with the arc);
 *                       the normal derivatives against finite differences of S_u x S_v
 *                       (tests/test_oracle_cones.py)
 *   or_project_global   never farther than any densely sampled surface point; never farther than
 *                       the local method; GPU branch and bound equals this exhaustive form
 *   or_curve_*          (no reference counterpart) exact circle arcs, scipy BSpline with textbook
 *                       derivative control points, finite differences, closed-form foot points
 *                       on the circle, brute-force nearest sample (tests/test_oracle_curve.py)
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_MAXK 8 /* MAX_DEGREE of src/NURBS/Evaluate.cuh:15 (k = degree + 1) */

/* ------------------------------------------------------------------------------------------ */
/* a1: knot-span lookup -- src/NURBS/Evaluate.cu:504-533                                        */
/* ------------------------------------------------------------------------------------------ */
int or_find_span(int degree, int cp, const double *knots, double u)
{
    int n = cp - 1;
    if (u >= knots[n + 1]) return n;
    if (u <= knots[degree]) return degree;
    int low = degree, high = n + 1;
    int mid = (int)((low + high) / 2.0);
    while (u < knots[mid] || u >= knots[mid + 1]) {
        if (u < knots[mid]) high = mid;
        else low = mid;
        mid = (int)((low + high) / 2.0);
    }
    return mid;
}

/* ------------------------------------------------------------------------------------------ */
/* a2, a3: basis-matrix recursion -- src/NURBS/Evaluate.cu:90-108                                */
/* ------------------------------------------------------------------------------------------ */
double or_nabla(const double *knot, int d, int i, int k)
{
    int index = i + d - 1;
    return knot[index + k] - knot[index];
}

double or_nrc(const double *knot, int m, int r, int c, int j, int i, int d)
{
    if (m == 0) return ((i + c - 1) == (j + d) && r == 1) ? 1.0 : 0.0;
    double den = or_nabla(knot, d, j, d - m + 1);
    double a = or_nabla(knot, d, i, 1) / den;
    double b = or_nabla(knot, d, j, i - j) / den;
    return a * or_nrc(knot, m - 1, r - 1, c, j, i, d) + b * or_nrc(knot, m - 1, r, c, j, i, d) +
           (1 - b) * or_nrc(knot, m - 1, r, c, j - 1, i, d) +
           (-a) * or_nrc(knot, m - 1, r - 1, c, j - 1, i, d);
}

/* M[r*k + s] = coefficient of t^r in N_{span-degree+s, degree}(k[span] + t*(k[span+1]-k[span])),
 * built the way GetMatrixN fills Nu (src/NURBS/Evaluate.cu:15-20). */
void or_basis_matrix_nrc(int degree, const double *knots, int span, double *M)
{
    int k = degree + 1;
    int ij = span - (degree - 1);
    for (int r = 0; r < k; r++)
        for (int s = 0; s < k; s++) M[r * k + s] = or_nrc(knots, degree, r + 1, s + 1, ij, ij, degree);
}

/* Independent path (not in the reference): Cox-de Boor recursion carried out on power-basis
 * polynomials in the local parameter t of the span.  0/0 is taken as 0. */
void or_basis_matrix_cdb(int degree, const double *knots, int span, double *M)
{
    int k = degree + 1;
    double h = knots[span + 1] - knots[span];
    /* N[a][e]: polynomial coefficients (power e) of N_{span-d+a, d} at the current degree d */
    double N[OR_MAXK][OR_MAXK], T[OR_MAXK][OR_MAXK];
    memset(N, 0, sizeof N);
    N[0][0] = 1.0;
    for (int d = 1; d <= degree; d++) {
        memset(T, 0, sizeof T);
        for (int a = 0; a <= d; a++) {
            int i = span - d + a; /* global basis index */
            /* left term: (u - k_i)/(k_{i+d} - k_i) * N_{i,d-1}; N_{i,d-1} is N[a-1] at degree d-1 */
            if (a >= 1) {
                double den = knots[i + d] - knots[i];
                if (den != 0.0) {
                    double c0 = (knots[span] - knots[i]) / den, c1 = h / den;
                    for (int e = 0; e < d; e++) {
                        T[a][e] += c0 * N[a - 1][e];
                        T[a][e + 1] += c1 * N[a - 1][e];
                    }
                }
            }
            /* right term: (k_{i+d+1} - u)/(k_{i+d+1} - k_{i+1}) * N_{i+1,d-1} = N[a] at degree d-1 */
            if (a <= d - 1) {
                double den = knots[i + d + 1] - knots[i + 1];
                if (den != 0.0) {
                    double c0 = (knots[i + d + 1] - knots[span]) / den, c1 = -h / den;
                    for (int e = 0; e < d; e++) {
                        T[a][e] += c0 * N[a][e];
                        T[a][e + 1] += c1 * N[a][e];
                    }
                }
            }
        }
        memcpy(N, T, sizeof N);
    }
    for (int r = 0; r < k; r++)
        for (int s = 0; s < k; s++) M[r * k + s] = N[s][r];
}

/* ------------------------------------------------------------------------------------------ */
/* a4, a5: span table -- MatrixNPosition src/NURBS/Evaluate.cu:110-112, GetMatrixN 3-88          */
/* ------------------------------------------------------------------------------------------ */
int or_matrix_n_position(int i, int j, int p, int q, int nu, int nv, int c)
{
    (void)nu;
    return i * (q + nv) * 4 * (p + 1) * (q + 1) + j * 4 * (p + 1) * (q + 1) + c * (p + 1) * (q + 1);
}

/* MatrixN in the reference's padded layout [(p+nu)][(q+nv)][4][k1][k2]; cp is u-major xyz
 * (cp[(iu*nv+iv)*3+c], src/main.cu:97-110).  valid[i*(q+nv)+j] = 1 for built slots. */
void or_get_matrix_n(int p, int q, int nu, int nv, const double *ku, const double *kv,
                     const double *cp, const double *w, int use_cdb, double *MatrixN,
                     unsigned char *valid)
{
    int k1 = p + 1, k2 = q + 1;
    size_t slots = (size_t)(p + nu) * (q + nv);
    memset(MatrixN, 0, slots * 4 * k1 * k2 * sizeof(double));
    memset(valid, 0, slots);
    for (int i = p; i < nu; i++) {
        if (ku[i] == ku[i + 1]) continue;
        double Nu[OR_MAXK * OR_MAXK];
        if (use_cdb) or_basis_matrix_cdb(p, ku, i, Nu);
        else or_basis_matrix_nrc(p, ku, i, Nu);
        for (int j = q; j < nv; j++) {
            if (kv[j] == kv[j + 1]) continue;
            double Nv[OR_MAXK * OR_MAXK], NvT[OR_MAXK * OR_MAXK];
            if (use_cdb) or_basis_matrix_cdb(q, kv, j, Nv);
            else or_basis_matrix_nrc(q, kv, j, Nv);
            for (int m = 0; m < k2; m++)
                for (int n = 0; n < k2; n++) NvT[n * k2 + m] = Nv[m * k2 + n];
            for (int c = 0; c < 4; c++) {
                double mid[OR_MAXK * OR_MAXK], tmp[OR_MAXK * OR_MAXK];
                for (int s = 0; s < k1; s++)
                    for (int l = 0; l < k2; l++) {
                        size_t g = (size_t)(s + i - p) * nv + (l + j - q);
                        double wt = w[g];
                        mid[s * k2 + l] = (c == 3) ? wt : cp[g * 3 + c] * wt;
                    }
                /* tmp = mid * NvT (src/NURBS/Evaluate.cu:58-61) */
                for (int a = 0; a < k1; a++)
                    for (int b = 0; b < k2; b++) {
                        double sum = 0;
                        for (int l = 0; l < k2; l++) sum += mid[a * k2 + l] * NvT[l * k2 + b];
                        tmp[a * k2 + b] = sum;
                    }
                /* out = Nu * tmp (src/NURBS/Evaluate.cu:63-66) */
                double *out = MatrixN + or_matrix_n_position(i, j, p, q, nu, nv, c);
                for (int a = 0; a < k1; a++)
                    for (int b = 0; b < k2; b++) {
                        double sum = 0;
                        for (int l = 0; l < k1; l++) sum += Nu[a * k1 + l] * tmp[l * k2 + b];
                        out[a * k2 + b] = sum;
                    }
            }
            valid[(size_t)i * (q + nv) + j] = 1;
        }
    }
}

/* ------------------------------------------------------------------------------------------ */
/* a6-a9: evaluation -- MatrixSurfacePoint src/NURBS/Evaluate.cu:159-227, old-version rational   */
/* derivatives 535-930, EvaluateRandom 1023-1038, homogenousToCartesian src/util/ExtraMath.cu:3-5 */
/* ------------------------------------------------------------------------------------------ */
typedef struct {
    int p, q, nu, nv;
    const double *ku, *kv, *MatrixN;
} or_surf;

/* y[i] = sum_l a[i*n + l] * b[l]   (matrix_mult with k == 1, src/util/Matrix.cu:43-52) */
static void mv(const double *a, const double *b, double *y, int m, int n)
{
    for (int i = 0; i < m; i++) {
        double s = 0;
        for (int l = 0; l < n; l++) s += a[i * n + l] * b[l];
        y[i] = s;
    }
}
/* y[j] = sum_l a[l] * b[l*k + j]   (matrix_mult with m == 1) */
static void vm(const double *a, const double *b, double *y, int n, int k)
{
    for (int j = 0; j < k; j++) {
        double s = 0;
        for (int l = 0; l < n; l++) s += a[l] * b[l * k + j];
        y[j] = s;
    }
}
static double dot(const double *a, const double *b, int n)
{
    double s = 0;
    for (int l = 0; l < n; l++) s += a[l] * b[l];
    return s;
}

static void monomials(double t, int k, double *V, double *dV, double *ddV, double *dddV)
{
    double now = 1;
    for (int i = 0; i < k; i++) { V[i] = now; now *= t; }
    now = 1;
    dV[0] = 0;
    for (int i = 0; i < k - 1; i++) { dV[i + 1] = (i + 1) * now; now *= t; }
    now = 1;
    ddV[0] = 0;
    if (k > 1) ddV[1] = 0;
    for (int i = 0; i < k - 2; i++) { ddV[i + 2] = (i + 2) * (i + 1) * now; now *= t; }
    /* vectordUUU / vectordVVV of MatrixSurfaceRdUUU / RdVVV (src/NURBS/Evaluate.cu:408-502) */
    for (int i = 0; i < k && i < 3; i++) dddV[i] = 0;
    now = 1;
    for (int i = 0; i < k - 3; i++) { dddV[i + 3] = (i + 3) * (i + 2) * (i + 1) * now; now *= t; }
}

typedef struct { double x, y, z, w; } d4;

/* Homogeneous contraction R = U^T * M_c * V for the four components c (MatrixSurfacePoint). */
static d4 contract(const double *mid, int k1, int k2, const double *U, const double *V)
{
    double t0[OR_MAXK], t1[OR_MAXK], t2[OR_MAXK], t3[OR_MAXK];
    int kk = k1 * k2;
    mv(mid + 0 * kk, V, t0, k1, k2);
    mv(mid + 1 * kk, V, t1, k1, k2);
    mv(mid + 2 * kk, V, t2, k1, k2);
    mv(mid + 3 * kk, V, t3, k1, k2);
    d4 r = {dot(U, t0, k1), dot(U, t1, k1), dot(U, t2, k1), dot(U, t3, k1)};
    return r;
}
/* VV variant contracts U first (src/NURBS/Evaluate.cu:883-912) */
static d4 contract_ufirst(const double *mid, int k1, int k2, const double *U, const double *V)
{
    double t0[OR_MAXK], t1[OR_MAXK], t2[OR_MAXK], t3[OR_MAXK];
    int kk = k1 * k2;
    vm(U, mid + 0 * kk, t0, k1, k2);
    vm(U, mid + 1 * kk, t1, k1, k2);
    vm(U, mid + 2 * kk, t2, k1, k2);
    vm(U, mid + 3 * kk, t3, k1, k2);
    d4 r = {dot(t0, V, k2), dot(t1, V, k2), dot(t2, V, k2), dot(t3, V, k2)};
    return r;
}

typedef struct {
    double S[3], Su[3], Sv[3], Suu[3], Suv[3], Svv[3];
    double Suuu[3], Suuv[3], Suvv[3], Svvv[3];
    double R[4];
    int m, n;
} or_pt;

/* Point and Cartesian derivatives up to `order` (0..3) at global (u, v). */
static void eval_one(const or_surf *sf, double uu, double vv, int order, or_pt *o)
{
    int p = sf->p, q = sf->q, k1 = p + 1, k2 = q + 1;
    int m = or_find_span(p, sf->nu, sf->ku, uu);
    int n = or_find_span(q, sf->nv, sf->kv, vv);
    double hu = sf->ku[m + 1] - sf->ku[m], hv = sf->kv[n + 1] - sf->kv[n];
    double u = (uu - sf->ku[m]) / hu;
    double v = (vv - sf->kv[n]) / hv;
    double spanU = 1.0 / hu, spanV = 1.0 / hv;
    const double *mid = sf->MatrixN + or_matrix_n_position(m, n, p, q, sf->nu, sf->nv, 0);
    double U[OR_MAXK], dU[OR_MAXK], ddU[OR_MAXK], dddU[OR_MAXK];
    double V[OR_MAXK], dV[OR_MAXK], ddV[OR_MAXK], dddV[OR_MAXK];
    monomials(u, k1, U, dU, ddU, dddU);
    monomials(v, k2, V, dV, ddV, dddV);
    d4 R = contract(mid, k1, k2, U, V);
    o->m = m;
    o->n = n;
    o->R[0] = R.x; o->R[1] = R.y; o->R[2] = R.z; o->R[3] = R.w;
    double Rc[3] = {R.x, R.y, R.z};
    for (int c = 0; c < 3; c++) o->S[c] = Rc[c] / R.w;
    if (order < 1) return;
    double h = R.w;
    /* MatrixSurfaceDerivativeU / V, src/NURBS/Evaluate.cu:535-665 */
    d4 RdU = contract(mid, k1, k2, dU, V);
    d4 RdV = contract(mid, k1, k2, U, dV);
    double RdUc[3] = {RdU.x, RdU.y, RdU.z}, RdVc[3] = {RdV.x, RdV.y, RdV.z};
    for (int c = 0; c < 3; c++) {
        o->Su[c] = (Rc[c] * (-RdU.w / pow(h, 2)) + RdUc[c] / h) * spanU;
        o->Sv[c] = (Rc[c] * (-RdV.w / pow(h, 2)) + RdVc[c] / h) * spanV;
    }
    if (order < 2) return;
    /* MatrixSurfaceDerivativeUU / UV / VV, src/NURBS/Evaluate.cu:667-930 */
    d4 RdUU = contract(mid, k1, k2, ddU, V);
    d4 RdUV = contract(mid, k1, k2, dU, dV);
    d4 RdVV = contract_ufirst(mid, k1, k2, U, ddV);
    double RdUUc[3] = {RdUU.x, RdUU.y, RdUU.z}, RdUVc[3] = {RdUV.x, RdUV.y, RdUV.z},
           RdVVc[3] = {RdVV.x, RdVV.y, RdVV.z};
    for (int c = 0; c < 3; c++) {
        double huu = Rc[c] * ((2 * pow(RdU.w, 2) / pow(h, 3)) - (RdUU.w / pow(h, 2))) +
                     2 * RdUc[c] * (-RdU.w / pow(h, 2)) + RdUUc[c] / h;
        double hvv = Rc[c] * ((2 * pow(RdV.w, 2) / pow(h, 3)) - (RdVV.w / pow(h, 2))) +
                     2 * RdVc[c] * (-RdV.w / pow(h, 2)) + RdVVc[c] / h;
        double huv = (RdUVc[c] * h - RdV.w * RdUc[c] - RdVc[c] * RdU.w - Rc[c] * RdUV.w) / pow(h, 2) +
                     (2 * Rc[c] * RdU.w * RdV.w) / pow(h, 3);
        o->Suu[c] = huu * spanU * spanU;
        o->Svv[c] = hvv * spanV * spanV;
        o->Suv[c] = huv * spanU * spanV;
    }
    if (order < 3) return;
    /* Third order: homogeneous MatrixSurfaceRd{UUU,UUV,UVV,VVV} (src/NURBS/Evaluate.cu:408-502),
     * then the Leibniz quotient of MatrixSurfaceDerivative{UUU,UUV,UVV,VVV}
     * (src/NURBS/Evaluate.cu:960-1020) with the w_uv, w_uuv and w_uvv terms of UUV and UVV
     * subtracted (the reference adds them, which is exact only for constant weights).  All in the
     * span's local parameters, scaled by the chain factors at the end. */
    d4 RdUUU = contract(mid, k1, k2, dddU, V);
    d4 RdUUV = contract(mid, k1, k2, ddU, dV);
    d4 RdUVV = contract(mid, k1, k2, dU, ddV);
    d4 RdVVV = contract_ufirst(mid, k1, k2, U, dddV);
    const double A3[4][3] = {{RdUUU.x, RdUUU.y, RdUUU.z}, {RdUUV.x, RdUUV.y, RdUUV.z},
                             {RdUVV.x, RdUVV.y, RdUVV.z}, {RdVVV.x, RdVVV.y, RdVVV.z}};
    for (int c = 0; c < 3; c++) {
        /* local-parameter derivatives of orders 1 and 2 */
        double su = o->Su[c] / spanU, sv = o->Sv[c] / spanV;
        double suu = o->Suu[c] / (spanU * spanU), svv = o->Svv[c] / (spanV * spanV);
        double suv = o->Suv[c] / (spanU * spanV);
        double s0 = o->S[c];
        double uuu = (A3[0][c] - 3 * RdU.w * suu - 3 * RdUU.w * su - RdUUU.w * s0) / h;
        double uuv = (A3[1][c] - RdV.w * suu - 2 * RdU.w * suv - 2 * RdUV.w * su - RdUU.w * sv -
                      RdUUV.w * s0) / h;
        double uvv = (A3[2][c] - RdU.w * svv - 2 * RdV.w * suv - 2 * RdUV.w * sv - RdVV.w * su -
                      RdUVV.w * s0) / h;
        double vvv = (A3[3][c] - 3 * RdV.w * svv - 3 * RdVV.w * sv - RdVVV.w * s0) / h;
        o->Suuu[c] = uuu * spanU * spanU * spanU;
        o->Suuv[c] = uuv * spanU * spanU * spanV;
        o->Suvv[c] = uvv * spanU * spanV * spanV;
        o->Svvv[c] = vvv * spanV * spanV * spanV;
    }
}

/* Batch evaluation (EvaluateRandom, src/NURBS/Evaluate.cu:1023-1038, extended to order <= 3).
 * out is SoA: [S.x S.y S.z | Su.xyz | Sv.xyz | Suu.xyz | Suv.xyz | Svv.xyz | Suuu.xyz | Suuv.xyz |
 * Suvv.xyz | Svvv.xyz], each block n long; only the first 3, 9, 18 or 30 blocks (order 0..3) are
 * written. */
void or_eval(int p, int q, int nu, int nv, const double *ku, const double *kv,
             const double *MatrixN, const double *u, const double *v, long n, int order,
             double *out, int nthreads)
{
    or_surf sf = {p, q, nu, nv, ku, kv, MatrixN};
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(nthreads)
#else
    (void)nthreads;
#endif
    for (long i = 0; i < n; i++) {
        or_pt o;
        eval_one(&sf, u[i], v[i], order, &o);
        for (int c = 0; c < 3; c++) {
            out[(0 + c) * n + i] = o.S[c];
            if (order >= 1) {
                out[(3 + c) * n + i] = o.Su[c];
                out[(6 + c) * n + i] = o.Sv[c];
            }
            if (order >= 2) {
                out[(9 + c) * n + i] = o.Suu[c];
                out[(12 + c) * n + i] = o.Suv[c];
                out[(15 + c) * n + i] = o.Svv[c];
            }
            if (order >= 3) {
                out[(18 + c) * n + i] = o.Suuu[c];
                out[(21 + c) * n + i] = o.Suuv[c];
                out[(24 + c) * n + i] = o.Suvv[c];
                out[(27 + c) * n + i] = o.Svvv[c];
            }
        }
    }
}

/* Homogeneous point R (4 values per point, AoS) -- used by the tree build (GetPoints). */
void or_eval_homogeneous(int p, int q, int nu, int nv, const double *ku, const double *kv,
                         const double *MatrixN, const double *u, const double *v, long n,
                         double *R4)
{
    or_surf sf = {p, q, nu, nv, ku, kv, MatrixN};
    for (long i = 0; i < n; i++) {
        or_pt o;
        eval_one(&sf, u[i], v[i], 0, &o);
        for (int c = 0; c < 4; c++) R4[i * 4 + c] = o.R[c];
    }
}

/* ------------------------------------------------------------------------------------------ */
/* a20: Gauss-Newton -- src/kernel/kernel.cu:151-249                                            */
/* ------------------------------------------------------------------------------------------ */
enum { OR_CONV_STEP = 0, OR_CONV_DIST = 1, OR_MAXITER = 2, OR_SINGULAR = 3, OR_NONFINITE = 4,
       OR_NOSEED = 5 };

typedef struct {
    double tol_dist; /* stop when |S - P| < tol_dist, before the update (reference rule) */
    double tol_step; /* stop when max(|du|,|dv|) < tol_step, after the update (<= 0: off) */
    int max_iter;    /* reference: 10 */
    int clamp;       /* clamp (u,v) to the knot domain after each update */
    int newton;      /* 1: add the residual-curvature term to J^T J (exact Hessian) */
    int active;      /* 1: projected Newton -- a coordinate on its clamp bound whose descent
                        points outward is held and the other takes its one-dimensional step
                        (exact curvature where positive); a step after which the squared
                        distance grew by more than 1e-6 relative is halved (up to 8 times in a
                        row); the nearest accepted point is kept */
    double tol_quad; /* certificate of an exact-Hessian step (the library's rule, not in the
                        reference): stop when |H^-1| L |d|^2 / 2 <= tol_quad, L = |S''|
                        (3 |J| + |r| (k1/h_u + k2/h_v)), evaluated in the squared form
                        0.5 |S''|^2 (9 |J|^2 + |r|^2 m^2) |d|^4 (a+c)^2 <= tol_quad^2 det^2,
                        for a step that ends inside the span pair it was evaluated in
                        (<= 0: off) */
} or_gn_opts;

/* Returns status; (u,v) is updated in place; *iters gets the number of iterations begun. */
/* dom (optional): clamp rectangle {u0, u1, v0, v1} instead of the knot domain */
static int gn_one_in(const or_surf *sf, const double P[3], const or_gn_opts *op, double *pu,
                     double *pv, int *iters, const double *dom)
{
    double cu = *pu, cv = *pv;
    double ulo = sf->ku[sf->p], uhi = sf->ku[sf->nu], vlo = sf->kv[sf->q], vhi = sf->kv[sf->nv];
    if (dom) { ulo = dom[0]; uhi = dom[1]; vlo = dom[2]; vhi = dom[3]; }
    int st = OR_MAXITER, it, back = 0;
    double au = cu, av = cv, prr = INFINITY; /* active mode: the last accepted point */
    for (it = 0; it < op->max_iter; it++) {
        or_pt o;
        eval_one(sf, cu, cv, op->newton ? 2 : 1, &o);
        double d[3] = {o.S[0] - P[0], o.S[1] - P[1], o.S[2] - P[2]};
        double rr = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        double dist = sqrt(rr);
        if (op->active) { /* backtracking: a step after which the distance grew is halved */
            if (!(rr <= prr * (1.0 + 1e-6)) && back < 8) { /* grew, beyond rounding */
                cu = au + 0.5 * (cu - au);
                cv = av + 0.5 * (cv - av);
                back++;
                continue;
            }
            back = 0;
            au = cu;
            av = cv;
            prr = rr;
        }
        if (!(dist < INFINITY)) { st = OR_NONFINITE; break; } /* NaN or infinite residual */
        if (dist < op->tol_dist) { st = OR_CONV_DIST; break; }
        /* J = [Su Sv]; (J^T J)^-1 J^T (S - P), src/kernel/kernel.cu:195-239 */
        double a = 0, b = 0, c = 0, gu = 0, gv = 0;
        for (int k = 0; k < 3; k++) {
            a += o.Su[k] * o.Su[k];
            b += o.Su[k] * o.Sv[k];
            c += o.Sv[k] * o.Sv[k];
            gu += o.Su[k] * d[k];
            gv += o.Sv[k] * d[k];
        }
        double a1 = a, c1 = c; /* one-dimensional curvatures for a held coordinate */
        double j2 = a + c;
        int hexact = 0;
        if (op->newton) {
            double a2 = a, b2 = b, c2 = c;
            for (int k = 0; k < 3; k++) {
                a2 += d[k] * o.Suu[k];
                b2 += d[k] * o.Suv[k];
                c2 += d[k] * o.Svv[k];
            }
            if (a2 > 0) a1 = a2;
            if (c2 > 0) c1 = c2;
            if (a2 > 0 && a2 * c2 - b2 * b2 > 0) { a = a2; b = b2; c = c2; hexact = 1; }
        }
        double det = a * c - b * b;
        if (!(det == det)) { st = OR_NONFINITE; break; }
        if (det == 0) { st = OR_SINGULAR; break; }
        double du = (c * gu - b * gv) / det;
        double dv = (a * gv - b * gu) / det;
        if (op->active) {
            int hu = (cu <= ulo && gu > 0) || (cu >= uhi && gu < 0);
            int hv = (cv <= vlo && gv > 0) || (cv >= vhi && gv < 0);
            if (hu) { du = 0; dv = hv ? 0 : gv / c1; }
            else if (hv) { dv = 0; du = gu / a1; }
        }
        cu -= du;
        cv -= dv;
        if (op->clamp) {
            cu = cu < ulo ? ulo : (cu > uhi ? uhi : cu);
            cv = cv < vlo ? vlo : (cv > vhi ? vhi : cv);
        }
        if (!(cu == cu) || !(cv == cv)) { st = OR_NONFINITE; break; }
        if (op->tol_step > 0 && fabs(du) < op->tol_step && fabs(dv) < op->tol_step) {
            st = OR_CONV_STEP;
            it++;
            break;
        }
        /* the bound holds on one polynomial piece: a step that leaves the span pair it was
           evaluated in crosses a knot, where S'' may jump, and is not certified */
        if (op->newton && hexact && op->tol_quad > 0 && cu >= sf->ku[o.m] &&
            cu <= sf->ku[o.m + 1] && cv >= sf->kv[o.n] && cv <= sf->kv[o.n + 1]) {
            double h2 = 0;
            for (int k = 0; k < 3; k++)
                h2 += o.Suu[k] * o.Suu[k] + 2 * o.Suv[k] * o.Suv[k] + o.Svv[k] * o.Svv[k];
            double m = (sf->p + 1) / (sf->ku[o.m + 1] - sf->ku[o.m]) +
                       (sf->q + 1) / (sf->kv[o.n + 1] - sf->kv[o.n]);
            double d2 = du * du + dv * dv, tr = a + c;
            if (0.5 * h2 * (9 * j2 + rr * m * m) * d2 * d2 * tr * tr <=
                op->tol_quad * op->tol_quad * det * det) {
                st = OR_CONV_STEP;
                it++;
                break;
            }
        }
    }
    if (op->active) { /* the final step did not help: keep the last accepted point */
        or_pt o;
        eval_one(sf, cu, cv, 0, &o);
        double dx = o.S[0] - P[0], dy = o.S[1] - P[1], dz = o.S[2] - P[2];
        if (!(dx * dx + dy * dy + dz * dz <= prr)) { cu = au; cv = av; }
    }
    *pu = cu;
    *pv = cv;
    if (iters) *iters = it;
    return st;
}

static int gn_one(const or_surf *sf, const double P[3], const or_gn_opts *op, double *pu, double *pv,
                  int *iters)
{
    return gn_one_in(sf, P, op, pu, pv, iters, NULL);
}

/* Batch Gauss-Newton from given seeds.  xyz SoA (x[n], y[n], z[n]). */
void or_gauss_newton(int p, int q, int nu, int nv, const double *ku, const double *kv,
                     const double *MatrixN, const double *x, const double *y, const double *z,
                     long n, double tol_dist, double tol_step, int max_iter, int clamp, int newton,
                     double tol_quad, const double *seed_u, const double *seed_v, double *u_out,
                     double *v_out, unsigned char *status, int *iters, int nthreads)
{
    or_surf sf = {p, q, nu, nv, ku, kv, MatrixN};
    or_gn_opts op = {tol_dist, tol_step, max_iter, clamp, newton, 0, tol_quad};
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads)
#else
    (void)nthreads;
#endif
    for (long i = 0; i < n; i++) {
        double P[3] = {x[i], y[i], z[i]};
        double cu = seed_u[i], cv = seed_v[i];
        int it = 0;
        int st = gn_one(&sf, P, &op, &cu, &cv, &it);
        u_out[i] = cu;
        v_out[i] = cv;
        if (status) status[i] = (unsigned char)st;
        if (iters) iters[i] = it;
    }
}

/* ------------------------------------------------------------------------------------------ */
/* a13-a18: subdivision tree -- GetPoints src/kernel/kernel.cu:404-431, GetSecondDerivatives     */
/* 375-402, Max_Sequential_Addressing_Shared 707-733, TightBoundingBox 544-621, NeedDivide       */
/* 859-878, min/max(TreeNode) 885-901, constants src/kernel/kernel.cuh:13-39                     */
/* ------------------------------------------------------------------------------------------ */
#define OR_GAP 8
#define OR_PPE (OR_GAP + 1)
#define OR_FAN (OR_GAP * OR_GAP)
#define OR_THRESHOLD 0.5
#define OR_TREE_REF_PARENT 1

long or_tree_base(int level)
{
    long b = 0, g = 1;
    for (int l = 0; l < level; l++) { b += g; g *= OR_FAN; }
    return b;
}

static int need_divide(const double lb[3], const double rt[3])
{
    double dx = rt[0] - lb[0], dy = rt[1] - lb[1], dz = rt[2] - lb[2], t;
    if (dx > dy) { t = dx; dx = dy; dy = t; }
    if (dx > dz) { t = dx; dx = dz; dz = t; }
    if ((dx < OR_THRESHOLD * dy) && (dx < OR_THRESHOLD * dz)) return 0;
    return 1;
}
int or_need_divide(const double *lb, const double *rt) { return need_divide(lb, rt); }

/* Node record (13 doubles): uv_lb(2) uv_rt(2) lb(3) rt(3) centre point C(3).  Every node of
 * `levels` levels is computed (as if every parent were divided); valid[] marks the nodes the
 * reference itself would build (root, and children of divided nodes), divided[] = NeedDivide. */
void or_build_tree(int p, int q, int nu, int nv, const double *ku, const double *kv,
                   const double *MatrixN, int levels, int flags, double *nodes,
                   unsigned char *valid, unsigned char *divided)
{
    or_surf sf = {p, q, nu, nv, ku, kv, MatrixN};
    /* the reference lays the tree over [0,1]^2 (src/kernel/kernel.cu:383-387, 406-407); the
     * library maps that square onto the knot domain, u = u0 + Wu x, with cell sizes Wu/(8g),
     * Wv/(8g) in K -- identical to the reference on [0,1] knots */
    double u0 = ku[p], wu = ku[nu] - u0, v0 = kv[q], wv = kv[nv] - v0;
    long g = 1;
    for (int L = 0; L < levels; L++, g *= OR_GAP) {
        long base = or_tree_base(L);
        for (long bx = 0; bx < g; bx++)
            for (long by = 0; by < g; by++) {
                long idx = bx * g + by, ti = base + idx;
                double base_unit = 1.0 / g, unit = base_unit / OR_GAP;
                double mx = -INFINITY, lo[3] = {INFINITY, INFINITY, INFINITY},
                       hi[3] = {-INFINITY, -INFINITY, -INFINITY};
                double M1 = 0, M2 = 0, M3 = 0;
                for (int tx = 0; tx < OR_PPE; tx++)
                    for (int ty = 0; ty < OR_PPE; ty++) {
                        double su = u0 + wu * (base_unit * bx + tx * unit),
                               sv = v0 + wv * (base_unit * by + ty * unit);
                        or_pt o;
                        eval_one(&sf, su, sv, 2, &o);
                        for (int c = 0; c < 3; c++) {
                            double s = o.R[c] / o.R[3];
                            if (s > hi[c]) hi[c] = s;
                            if (s < lo[c]) lo[c] = s;
                        }
                        double e1 = fmax(fabs(o.Suu[0]), fmax(fabs(o.Suu[1]), fabs(o.Suu[2])));
                        double e2 = fmax(fabs(o.Suv[0]), fmax(fabs(o.Suv[1]), fabs(o.Suv[2])));
                        double e3 = fmax(fabs(o.Svv[0]), fmax(fabs(o.Svv[1]), fabs(o.Svv[2])));
                        if (e1 > M1) M1 = e1;
                        if (e2 > M2) M2 = e2;
                        if (e3 > M3) M3 = e3;
                    }
                (void)mx;
                double K = (M1 * wu * wu + 2 * M2 * wu * wv + M3 * wv * wv) * sqrt(3.0) /
                           (8.0 * OR_GAP * (double)g * OR_GAP * (double)g);
                double *nd = nodes + ti * 13;
                double unitg = 1.0 / g;
                nd[0] = u0 + wu * (bx * unitg);
                nd[1] = v0 + wv * (by * unitg);
                nd[2] = u0 + wu * ((bx + 1) * unitg);
                nd[3] = v0 + wv * ((by + 1) * unitg);
                for (int c = 0; c < 3; c++) {
                    nd[4 + c] = lo[c] - K;
                    nd[7 + c] = hi[c] + K;
                }
                long parent = -1;
                if (L > 0) {
                    if (flags & OR_TREE_REF_PARENT) parent = (ti - 1) / OR_FAN;
                    else parent = or_tree_base(L - 1) + (bx / OR_GAP) * (g / OR_GAP) + by / OR_GAP;
                }
                if (parent >= 0) {
                    const double *pn = nodes + parent * 13;
                    for (int c = 0; c < 3; c++) {
                        if (pn[7 + c] < nd[7 + c]) nd[7 + c] = pn[7 + c];
                        if (pn[4 + c] > nd[4 + c]) nd[4 + c] = pn[4 + c];
                    }
                }
                valid[ti] = (parent < 0) ? 1 : (unsigned char)(valid[parent] && divided[parent]);
                divided[ti] = (unsigned char)need_divide(nd + 4, nd + 7);
                /* centre point: the surface at the uv-rect centre used as the GN seed
                 * (src/kernel/kernel.cu:156) */
                double cu = (nd[0] + nd[2]) / 2, cv = (nd[1] + nd[3]) / 2;
                or_pt o;
                eval_one(&sf, cu, cv, 0, &o);
                for (int c = 0; c < 3; c++) nd[10 + c] = o.S[c];
            }
    }
}

/* ------------------------------------------------------------------------------------------ */
/* normal cones of the tree leaves -- NormalBoundingBox / NormalBoxToNormalTreeNode /             */
/* GetNd{UU,UV,VV}s, src/kernel/kernel.cu:623-704, 1486-1556, 1605-1636, with the normal          */
/* derivatives of src/NURBS/Evaluate.cu:1125-1135                                                 */
/* ------------------------------------------------------------------------------------------ */
static void cross3(const double *a, const double *b, double *o)
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

static double amax3(const double *a) { return fmax(fabs(a[0]), fmax(fabs(a[1]), fabs(a[2]))); }

/* angle between a and b, by atan2 of |a x b| and a . b */
static double angle3(const double *a, const double *b)
{
    double c[3];
    cross3(a, b, c);
    return atan2(sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]),
                 a[0] * b[0] + a[1] * b[1] + a[2] * b[2]);
}

/* N = S_u x S_v and its second derivatives (src/NURBS/Evaluate.cu:1125-1135; there N_uv also
 * carries S_uv x S_uv, which is zero) */
static void normal_ders(const or_pt *o, double *N, double *Nuu, double *Nuv, double *Nvv)
{
    double a[3], b[3], c[3];
    cross3(o->Su, o->Sv, N);
    cross3(o->Suuu, o->Sv, a);
    cross3(o->Suu, o->Suv, b);
    cross3(o->Su, o->Suuv, c);
    for (int k = 0; k < 3; k++) Nuu[k] = a[k] + 2.0 * b[k] + c[k];
    cross3(o->Suuv, o->Sv, a);
    cross3(o->Suu, o->Svv, b);
    cross3(o->Su, o->Suvv, c);
    for (int k = 0; k < 3; k++) Nuv[k] = a[k] + b[k] + c[k];
    cross3(o->Suvv, o->Sv, a);
    cross3(o->Suv, o->Svv, b);
    cross3(o->Su, o->Svvv, c);
    for (int k = 0; k < 3; k++) Nvv[k] = a[k] + 2.0 * b[k] + c[k];
}

/* out[12 i ..]: N, N_uu, N_uv, N_vv at (u[i], v[i]) */
void or_normal_ders(int p, int q, int nu, int nv, const double *ku, const double *kv,
                    const double *MatrixN, const double *u, const double *v, long n, double *out)
{
    or_surf sf = {p, q, nu, nv, ku, kv, MatrixN};
    for (long i = 0; i < n; i++) {
        or_pt o;
        eval_one(&sf, u[i], v[i], 3, &o);
        normal_ders(&o, out + 12 * i, out + 12 * i + 3, out + 12 * i + 6, out + 12 * i + 9);
    }
}

/* For every leaf (level levels-1, g = 8^(levels-1) per side, index bx*g + by): the box of the
 * normal N = S_u x S_v over the leaf's 9 x 9 samples (the tree's sample grid), enlarged by
 * K = (M1 Wu^2 + 2 M2 Wu Wv + M3 Wv^2) sqrt(3) / (8 (8g)^2) with M* the sample maxima of
 * |N_uu|inf, |N_uv|inf, |N_vv|inf (the position boxes' K, applied to N), then the cone around the
 * box centre m that holds the box: half-angle = the largest angle between m and a box corner.
 * out[4*leaf]: unit axis (3) and half-angle; half-angle 4 (> pi/2) when the box holds the origin
 * or a corner is 90 degrees or more from m (no bounded cone).  The reference stops at the box
 * and cone; its use in Project (kernel.cu:471-513) tests box corners only and is unfinished. */
void or_leaf_cones(int p, int q, int nu, int nv, const double *ku, const double *kv,
                   const double *MatrixN, int levels, double *out)
{
    or_surf sf = {p, q, nu, nv, ku, kv, MatrixN};
    double u0 = ku[p], wu = ku[nu] - u0, v0 = kv[q], wv = kv[nv] - v0;
    long g = 1;
    for (int L = 1; L < levels; L++) g *= OR_GAP;
    for (long bx = 0; bx < g; bx++)
        for (long by = 0; by < g; by++) {
            double base_unit = 1.0 / g, unit = base_unit / OR_GAP;
            double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
            double M1 = 0, M2 = 0, M3 = 0;
            for (int tx = 0; tx < OR_PPE; tx++)
                for (int ty = 0; ty < OR_PPE; ty++) {
                    double su = u0 + wu * (base_unit * bx + tx * unit),
                           sv = v0 + wv * (base_unit * by + ty * unit);
                    or_pt o;
                    eval_one(&sf, su, sv, 3, &o);
                    double N[3], Nuu[3], Nuv[3], Nvv[3];
                    normal_ders(&o, N, Nuu, Nuv, Nvv);
                    for (int k = 0; k < 3; k++) {
                        if (N[k] < lo[k]) lo[k] = N[k];
                        if (N[k] > hi[k]) hi[k] = N[k];
                    }
                    M1 = fmax(M1, amax3(Nuu));
                    M2 = fmax(M2, amax3(Nuv));
                    M3 = fmax(M3, amax3(Nvv));
                }
            double K = (M1 * wu * wu + 2 * M2 * wu * wv + M3 * wv * wv) * sqrt(3.0) /
                       (8.0 * OR_GAP * (double)g * OR_GAP * (double)g);
            double *o4 = out + (bx * g + by) * 4;
            int zero = 1;
            double m[3];
            for (int k = 0; k < 3; k++) {
                lo[k] -= K;
                hi[k] += K;
                if (!(lo[k] <= 0.0 && hi[k] >= 0.0)) zero = 0;
                m[k] = 0.5 * (lo[k] + hi[k]);
            }
            double th = 0.0;
            for (int cn = 0; cn < 8 && !zero; cn++) {
                double x[3] = {(cn & 4) ? hi[0] : lo[0], (cn & 2) ? hi[1] : lo[1], (cn & 1) ? hi[2] : lo[2]};
                th = fmax(th, angle3(m, x));
            }
            double mm = sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
            if (zero || !(th < 0.5 * M_PI) || !(mm > 0.0)) {
                o4[0] = o4[1] = o4[2] = 0.0;
                o4[3] = 4.0;
            } else {
                for (int k = 0; k < 3; k++) o4[k] = m[k] / mm;
                o4[3] = th;
            }
        }
}

/* ------------------------------------------------------------------------------------------ */
/* a19: seed search -- NewInverse src/kernel/kernel.cu:433-464, NewBoundingBoxCheck 466-469      */
/* ------------------------------------------------------------------------------------------ */
static int box_has(const double *nd, const double P[3])
{
    return (P[0] >= nd[4]) && (P[1] >= nd[5]) && (P[2] >= nd[6]) && (P[0] <= nd[7]) &&
           (P[1] <= nd[8]) && (P[2] <= nd[9]);
}

/* Reference inversion: scan valid nodes in index order; the first closed box holding P seeds
 * Gauss-Newton at the node's uv centre; stop at the first success (NewInverse). */
void or_invert_first_hit(int p, int q, int nu, int nv, const double *ku, const double *kv,
                         const double *MatrixN, const double *nodes, const unsigned char *valid,
                         long n_nodes, const double *x, const double *y, const double *z, long n,
                         double tol_dist, double tol_step, int max_iter, int clamp,
                         double *u_out, double *v_out, unsigned char *status, int *node_out,
                         int nthreads)
{
    or_surf sf = {p, q, nu, nv, ku, kv, MatrixN};
    or_gn_opts op = {tol_dist, tol_step, max_iter, clamp, 0, 0, 0.0};
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads)
#else
    (void)nthreads;
#endif
    for (long i = 0; i < n; i++) {
        double P[3] = {x[i], y[i], z[i]};
        int st = OR_NOSEED, hit = -1;
        double ru = NAN, rv = NAN;
        for (long t = 0; t < n_nodes; t++) {
            if (!valid[t]) continue;
            const double *nd = nodes + t * 13;
            if (!box_has(nd, P)) continue;
            double cu = (nd[0] + nd[2]) / 2, cv = (nd[1] + nd[3]) / 2;
            int s = gn_one(&sf, P, &op, &cu, &cv, NULL);
            hit = (int)t;
            ru = cu;
            rv = cv;
            st = s;
            if (s == OR_CONV_DIST || s == OR_CONV_STEP) break;
        }
        u_out[i] = ru;
        v_out[i] = rv;
        status[i] = (unsigned char)st;
        if (node_out) node_out[i] = hit;
    }
}

/* Nearest leaf: the leaf (level `levels-1` node) whose centre point is closest to P, ties to
 * the lower index; brute force over all leaves. */
void or_seed_nearest_leaf(const double *nodes, int levels, const double *x, const double *y,
                          const double *z, long n, int *leaf_out, double *seed_u, double *seed_v)
{
    long base = or_tree_base(levels - 1), cnt = or_tree_base(levels) - base;
    for (long i = 0; i < n; i++) {
        double best = INFINITY;
        long bi = -1;
        for (long t = 0; t < cnt; t++) {
            const double *nd = nodes + (base + t) * 13;
            double dx = x[i] - nd[10], dy = y[i] - nd[11], dz = z[i] - nd[12];
            double d2 = dx * dx + dy * dy + dz * dz;
            if (d2 < best) { best = d2; bi = t; }
        }
        leaf_out[i] = (int)(base + bi);
        const double *nd = nodes + (base + bi) * 13;
        seed_u[i] = (nd[0] + nd[2]) / 2;
        seed_v[i] = (nd[1] + nd[3]) / 2;
    }
}

/* Greedy-leaf seed (the library's GN_SEED_GREEDY_LEAF, not in the reference): quadtree over the
 * finest level with boxes = leaf box U leaf centre, unions upward; FP32 copies rounded outward and
 * widened by (1 + max|root box|) * 2^-20; greedy descent by FP32 squared box distance (ties to
 * the lower child, children ordered (x,y) = (0,0),(0,1),(1,0),(1,1)); then the nearest leaf
 * centre (FP32 centres and distances, ties to the lower reference index) among the 3 x 3 leaves
 * around it. */
/* refine: 1 = the 3 x 3 neighbour refinement of GN_SEED_GREEDY_LEAF; 0 = the descent's leaf */
static void greedy_leaf_impl(const double *nodes, int levels, const double *x, const double *y,
                             const double *z, long n, int *leaf_out, double *seed_u,
                             double *seed_v, int nthreads, int refine)
{
    int D = 3 * (levels - 1);
    long side = 1L << D, base = or_tree_base(levels - 1);
    /* boxes per quad level, row-major (ix*side_l + iy) */
    double **bx = malloc(sizeof(double *) * (D + 1));
    for (int l = 0; l <= D; l++) bx[l] = malloc(sizeof(double) * 6 * (1L << (2 * l)));
    for (long ix = 0; ix < side; ix++)
        for (long iy = 0; iy < side; iy++) {
            const double *nd = nodes + (base + ix * side + iy) * 13;
            double *b = bx[D] + (ix * side + iy) * 6;
            for (int k = 0; k < 3; k++) {
                b[k] = fmin(nd[4 + k], nd[10 + k]);
                b[3 + k] = fmax(nd[7 + k], nd[10 + k]);
            }
        }
    for (int l = D - 1; l >= 0; l--) {
        long s = 1L << l;
        for (long ix = 0; ix < s; ix++)
            for (long iy = 0; iy < s; iy++) {
                double *b = bx[l] + (ix * s + iy) * 6;
                for (int k = 0; k < 3; k++) { b[k] = INFINITY; b[3 + k] = -INFINITY; }
                for (int a = 0; a < 4; a++) {
                    long cx = 2 * ix + (a >> 1), cy = 2 * iy + (a & 1);
                    const double *cb = bx[l + 1] + (cx * (2 * s) + cy) * 6;
                    for (int k = 0; k < 3; k++) {
                        b[k] = fmin(b[k], cb[k]);
                        b[3 + k] = fmax(b[3 + k], cb[3 + k]);
                    }
                }
            }
    }
    double mag = 0;
    for (int k = 0; k < 6; k++) mag = fmax(mag, fabs(bx[0][k]));
    double margin = (1.0 + mag) * ldexp(1.0, -20);
    float **b32 = malloc(sizeof(float *) * (D + 1));
    for (int l = 0; l <= D; l++) {
        long cnt = 1L << (2 * l);
        b32[l] = malloc(sizeof(float) * 6 * cnt);
        for (long i = 0; i < cnt; i++)
            for (int k = 0; k < 3; k++) {
                double lo = bx[l][i * 6 + k] - margin, hi = bx[l][i * 6 + 3 + k] + margin;
                float lo32 = (float)lo, hi32 = (float)hi;
                if ((double)lo32 > lo) lo32 = nextafterf(lo32, -INFINITY);
                if ((double)hi32 < hi) hi32 = nextafterf(hi32, INFINITY);
                b32[l][i * 6 + k] = lo32;
                b32[l][i * 6 + 3 + k] = hi32;
            }
    }
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(nthreads)
#else
    (void)nthreads;
#endif
    for (long i = 0; i < n; i++) {
        double P[3] = {x[i], y[i], z[i]};
        float p[3] = {(float)P[0], (float)P[1], (float)P[2]};
        long ix = 0, iy = 0;
        for (int l = 0; l < D; l++) {
            float bd = INFINITY;
            long bxx = 2 * ix, byy = 2 * iy;
            long s = 1L << (l + 1);
            for (int a = 0; a < 4; a++) {
                long cx = 2 * ix + (a >> 1), cy = 2 * iy + (a & 1);
                const float *b = b32[l + 1] + (cx * s + cy) * 6;
                float sum = 0.f;
                for (int k = 0; k < 3; k++) {
                    float e = fmaxf(fmaxf(b[k] - p[k], p[k] - b[3 + k]), 0.f);
                    sum = fmaf(e, e, sum);
                }
                if (sum < bd) { bd = sum; bxx = cx; byy = cy; }
            }
            ix = bxx;
            iy = byy;
        }
        /* 3 x 3 leaves around the descent's leaf: FP32 centres and distances */
        float best = INFINITY;
        long bi = base + ix * side + iy;
        for (int dx = -1; dx <= 1 && refine; dx++)
            for (int dy = -1; dy <= 1; dy++) {
                long xx = ix + dx, yy = iy + dy;
                if (xx < 0 || yy < 0 || xx >= side || yy >= side) continue;
                long t = base + xx * side + yy;
                const double *q = nodes + t * 13 + 10;
                float ex = p[0] - (float)q[0], ey = p[1] - (float)q[1], ez = p[2] - (float)q[2];
                float dd = fmaf(ez, ez, fmaf(ey, ey, ex * ex));
                if (dd < best || (dd == best && t < bi)) { best = dd; bi = t; }
            }
        leaf_out[i] = (int)bi;
        const double *nd = nodes + bi * 13;
        seed_u[i] = (nd[0] + nd[2]) / 2;
        seed_v[i] = (nd[1] + nd[3]) / 2;
    }
    for (int l = 0; l <= D; l++) { free(bx[l]); free(b32[l]); }
    free(bx);
    free(b32);
}

void or_seed_greedy_leaf(const double *nodes, int levels, const double *x, const double *y,
                         const double *z, long n, int *leaf_out, double *seed_u, double *seed_v,
                         int nthreads)
{
    greedy_leaf_impl(nodes, levels, x, y, z, n, leaf_out, seed_u, seed_v, nthreads, 1);
}

/* The greedy descent's leaf alone (no neighbour refinement), as GN_SEED_GREEDY_STEP starts from. */
void or_seed_greedy_descent(const double *nodes, int levels, const double *x, const double *y,
                            const double *z, long n, int *leaf_out, double *seed_u,
                            double *seed_v, int nthreads)
{
    greedy_leaf_impl(nodes, levels, x, y, z, n, leaf_out, seed_u, seed_v, nthreads, 0);
}

/* The library's GN_SEED_GREEDY_STEP (no reference counterpart): the greedy descent's leaf (no
 * neighbour refinement: the step below moves up to a leaf cell anyway), then one
 * Gauss-Newton step from the leaf's seed (its uv centre, whose surface point is the node's
 * centre C) with J = [S_u S_v] evaluated there, (du, dv) = (J^T J)^-1 J^T (C - P), computed in
 * FP32 from C, P and J rounded to FP32, scaled down uniformly to at most one leaf cell per
 * direction, applied to the FP64 seed and clamped to the knot domain; a singular J^T J keeps the
 * leaf seed.  Written with the same explicit fmaf steps as the library. */
/* The step itself.  or_leafrec: what the library's 128-byte leaf record holds (FP32 centre C and
 * S_u, S_v, S_uu, S_uv, S_vv at the leaf's seed, the seed (u, v)); step_from_rec: the step from
 * it for one query P (leaves the seed when the linear step's system is singular). */
typedef struct {
    float C[3], ju[3], jv[3], huu[3], huv[3], hvv[3];
    double su, sv;
} or_leafrec;

static void leaf_rec(const or_surf *sf, const double *nodes, long t, or_leafrec *L)
{
    const double *nd = nodes + t * 13;
    L->su = (nd[0] + nd[2]) / 2;
    L->sv = (nd[1] + nd[3]) / 2;
    or_pt o;
    eval_one(sf, L->su, L->sv, 2, &o);
    for (int k = 0; k < 3; k++) {
        L->C[k] = (float)nd[10 + k];
        L->ju[k] = (float)o.Su[k];
        L->jv[k] = (float)o.Sv[k];
        L->huu[k] = (float)o.Suu[k];
        L->huv[k] = (float)o.Suv[k];
        L->hvv[k] = (float)o.Svv[k];
    }
}

typedef struct {
    float ihu, ihv, ih2u, ih2v;
    double u0, u1, v0, v1;
} or_stepdom;

static or_stepdom step_domain(const or_surf *sf, int levels)
{
    or_stepdom g;
    g.u0 = sf->ku[sf->p];
    g.u1 = sf->ku[sf->nu];
    g.v0 = sf->kv[sf->q];
    g.v1 = sf->kv[sf->nv];
    float cells = (float)(1L << (3 * (levels - 1)));
    float hu = (float)(g.u1 - g.u0) / cells, hv = (float)(g.v1 - g.v0) / cells;
    g.ihu = 1.f / hu;
    g.ihv = 1.f / hv;
    g.ih2u = 0.5f * g.ihu;
    g.ih2v = 0.5f * g.ihv;
    return g;
}

static void step_from_rec(const or_leafrec *L, const or_stepdom *g, const double P[3], double *su,
                          double *sv)
{
    const float *ju = L->ju, *jv = L->jv, *huu = L->huu, *huv = L->huv, *hvv = L->hvv;
    float r[3];
    *su = L->su;
    *sv = L->sv;
    for (int k = 0; k < 3; k++) r[k] = L->C[k] - (float)P[k];
    /* the linear step from the leaf's seed */
    float a = fmaf(ju[2], ju[2], fmaf(ju[1], ju[1], ju[0] * ju[0]));
    float b = fmaf(ju[2], jv[2], fmaf(ju[1], jv[1], ju[0] * jv[0]));
    float e = fmaf(jv[2], jv[2], fmaf(jv[1], jv[1], jv[0] * jv[0]));
    float gu = fmaf(ju[2], r[2], fmaf(ju[1], r[1], ju[0] * r[0]));
    float gv = fmaf(jv[2], r[2], fmaf(jv[1], r[1], jv[0] * r[0]));
    float det = fmaf(a, e, -(b * b));
    float idet = 1.f / det;
    float du = fmaf(e, gu, -(b * gv)) * idet;
    float dv = fmaf(a, gv, -(b * gu)) * idet;
    float sc = fmaxf(1.f, fmaxf(fabsf(du) * g->ihu, fabsf(dv) * g->ihv));
    if (!(det > 0.f && sc < INFINITY)) return;
    float isc = 1.f / sc;
    float tu = -(du * isc), tv = -(dv * isc);
    /* Gauss-Newton iterations on the leaf's quadratic model
     * m(t) = C + S_u tu + S_v tv + (S_uu tu^2 + 2 S_uv tu tv + S_vv tv^2) / 2 */
    for (int it = 0; it < 2; it++) {
        float tuu = tu * tu, tuv = tu * tv, tvv = tv * tv;
        float m[3], au[3], av[3];
        for (int k = 0; k < 3; k++) {
            float t = fmaf(ju[k], tu, fmaf(jv[k], tv, r[k]));
            float w = fmaf(huu[k], tuu, fmaf(2.f * huv[k], tuv, hvv[k] * tvv));
            m[k] = fmaf(0.5f, w, t);
            au[k] = fmaf(huu[k], tu, fmaf(huv[k], tv, ju[k]));
            av[k] = fmaf(huv[k], tu, fmaf(hvv[k], tv, jv[k]));
        }
        float A = fmaf(au[2], au[2], fmaf(au[1], au[1], au[0] * au[0]));
        float B = fmaf(au[2], av[2], fmaf(au[1], av[1], au[0] * av[0]));
        float E = fmaf(av[2], av[2], fmaf(av[1], av[1], av[0] * av[0]));
        float Gu = fmaf(au[2], m[2], fmaf(au[1], m[1], au[0] * m[0]));
        float Gv = fmaf(av[2], m[2], fmaf(av[1], m[1], av[0] * m[0]));
        float D = fmaf(A, E, -(B * B));
        float iD = 1.f / D;
        float nu_ = tu - fmaf(E, Gu, -(B * Gv)) * iD;
        float nv_ = tv - fmaf(A, Gv, -(B * Gu)) * iD;
        float s2 = fmaxf(1.f, fmaxf(fabsf(nu_) * g->ih2u, fabsf(nv_) * g->ih2v));
        if (D > 0.f && s2 < INFINITY) {
            float is2 = 1.f / s2;
            tu = nu_ * is2;
            tv = nv_ * is2;
        }
    }
    *su = fmin(fmax(*su + (double)tu, g->u0), g->u1);
    *sv = fmin(fmax(*sv + (double)tv, g->v0), g->v1);
}

/* The step for the points whose leaves (reference indices) are in leaf_out. */
static void seed_step_from_leaf(const or_surf *sfp, const double *nodes, int levels,
                                const double *x, const double *y, const double *z, long n,
                                const int *leaf_out, double *seed_u, double *seed_v, int nthreads)
{
    const or_stepdom g = step_domain(sfp, levels);
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(nthreads)
#endif
    for (long i = 0; i < n; i++) {
        or_leafrec L;
        leaf_rec(sfp, nodes, leaf_out[i], &L);
        const double P[3] = {x[i], y[i], z[i]};
        step_from_rec(&L, &g, P, &seed_u[i], &seed_v[i]);
    }
}

void or_seed_greedy_step(int p, int q, int nu, int nv, const double *ku, const double *kv,
                         const double *MatrixN, const double *nodes, int levels, const double *x,
                         const double *y, const double *z, long n, int *leaf_out, double *seed_u,
                         double *seed_v, int nthreads)
{
    or_surf sf = {p, q, nu, nv, ku, kv, MatrixN};
    or_seed_greedy_descent(nodes, levels, x, y, z, n, leaf_out, seed_u, seed_v, nthreads);
    seed_step_from_leaf(&sf, nodes, levels, x, y, z, n, leaf_out, seed_u, seed_v, nthreads);
}

/* The library's seed grid (gn_build_seed_grid; no reference counterpart).  Geometry from the
 * finest leaves' FP32 centres: lo / hi per axis, ext = the largest hi - lo (1 when it is not a
 * positive finite number), R = resolution, pad = max(1, R / 16), cell = ext / R,
 * origin = lo - pad * cell, dims = ceil((hi - lo) / cell) + 2 pad; every value a single FP32
 * operation.  A grid is too big above 2^25 cells or 2^36 cell-leaf pairs.  resolution 0 starts
 * at 2 * 2^D capped at 256 and halves while too big (down to 8).  Returns 0, or -1 when the
 * centres are not finite or the grid is too big (the library's GN_EINVAL cases). */
int or_seed_grid_geometry(const double *nodes, int levels, int resolution, int dims[3],
                          float org[3], float *cell_out)
{
    int D = 3 * (levels - 1);
    long side = 1L << D, base = or_tree_base(levels - 1);
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (long t = 0; t < side * side; t++)
        for (int k = 0; k < 3; k++) {
            float c = (float)nodes[(base + t) * 13 + 10 + k];
            lo[k] = fminf(lo[k], c);
            hi[k] = fmaxf(hi[k], c);
        }
    float ext = 0.f;
    for (int k = 0; k < 3; k++) {
        if (!isfinite(lo[k]) || !isfinite(hi[k])) return -1;
        float w = hi[k] - lo[k];
        ext = fmaxf(ext, w);
    }
    if (!(ext > 0.f) || !isfinite(ext)) ext = 1.f;
    long leaves = side * side, max_pairs = 1L << 36, cells = 0;
    float cell = 0.f;
    int R = resolution ? resolution : ((2 << D) < 256 ? (2 << D) : 256);
    for (;;) {
        int pad = R / 16 > 1 ? R / 16 : 1;
        cell = ext / (float)R;
        cells = 1;
        for (int k = 0; k < 3; k++) {
            float w = hi[k] - lo[k];
            float span = w / cell;
            float off = (float)pad * cell;
            org[k] = lo[k] - off;
            dims[k] = (int)ceilf(span) + 2 * pad;
            cells *= dims[k];
        }
        int big = cells > (1L << 25) || cells * leaves > max_pairs;
        if (!big) break;
        if (resolution || R <= 8) return -1;
        R /= 2;
    }
    (void)base;
    *cell_out = cell;
    return 0;
}

/* Each cell's leaf: the reference index t of least FP32 squared distance
 * fmaf(ez, ez, fmaf(ey, ey, ex * ex)) from the cell centre org + (i + 1/2) * cell (one fmaf per
 * axis) to the leaf's FP32 centre, the lowest t among equals.  grid[(i * dims[1] + j) * dims[2] + k]. */
void or_seed_grid_build(const double *nodes, int levels, const int dims[3], const float org[3],
                        float cell, int *grid, int nthreads)
{
    int D = 3 * (levels - 1);
    long side = 1L << D, base = or_tree_base(levels - 1), leaves = side * side;
    long cells = (long)dims[0] * dims[1] * dims[2];
    float *c32 = malloc(sizeof(float) * 3 * leaves);
    for (long t = 0; t < leaves; t++)
        for (int k = 0; k < 3; k++) c32[t * 3 + k] = (float)nodes[(base + t) * 13 + 10 + k];
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(nthreads)
#else
    (void)nthreads;
#endif
    for (long id = 0; id < cells; id++) {
        int k = (int)(id % dims[2]), j = (int)((id / dims[2]) % dims[1]),
            i = (int)(id / ((long)dims[2] * dims[1]));
        float cx = fmaf((float)i + 0.5f, cell, org[0]), cy = fmaf((float)j + 0.5f, cell, org[1]),
              cz = fmaf((float)k + 0.5f, cell, org[2]);
        float best = INFINITY;
        long bt = 0;
        for (long t = 0; t < leaves; t++) {
            float ex = cx - c32[t * 3], ey = cy - c32[t * 3 + 1], ez = cz - c32[t * 3 + 2];
            float dd = fmaf(ez, ez, fmaf(ey, ey, ex * ex));
            if (dd < best) { best = dd; bt = t; }
        }
        grid[id] = (int)(base + bt);
    }
    free(c32);
}

/* GN_SEED_GRID_STEP: the cell of P rounded to FP32, index floor(((float)P - org) * (1 / cell))
 * clamped to [0, dims - 1] (fmaxf takes a NaN to cell 0), its leaf, then the step of
 * or_seed_greedy_step from that leaf's uv centre. */
void or_seed_grid_step(int p, int q, int nu, int nv, const double *ku, const double *kv,
                       const double *MatrixN, const double *nodes, int levels, const int dims[3],
                       const float org[3], float cell, const int *grid, const double *x,
                       const double *y, const double *z, long n, int *leaf_out, double *seed_u,
                       double *seed_v, int nthreads)
{
    or_surf sf = {p, q, nu, nv, ku, kv, MatrixN};
    float inv = 1.f / cell;
    for (long i = 0; i < n; i++) {
        const double P[3] = {x[i], y[i], z[i]};
        int id[3];
        for (int k = 0; k < 3; k++) {
            float t = ((float)P[k] - org[k]) * inv;
            id[k] = (int)fminf(fmaxf(floorf(t), 0.f), (float)(dims[k] - 1));
        }
        int t = grid[((long)id[0] * dims[1] + id[1]) * dims[2] + id[2]];
        leaf_out[i] = t;
        const double *nd = nodes + (long)t * 13;
        seed_u[i] = (nd[0] + nd[2]) / 2;
        seed_v[i] = (nd[1] + nd[3]) / 2;
    }
    seed_step_from_leaf(&sf, nodes, levels, x, y, z, n, leaf_out, seed_u, seed_v, nthreads);
}

/* ------------------------------------------------------------------------------------------ */
/* Curves.  The reference has no curve code (README.md:10 names curves; src/ has none), so this  */
/* part restates textbook algorithms instead of reference lines: basis functions and their       */
/* derivatives by the Cox-de Boor triangle (Piegl & Tiller, The NURBS Book, A2.3), the rational  */
/* quotient C = A/w, C' = (A' - w'C)/w, C'' = (A'' - 2w'C' - w''C)/w (eq. 4.8 ff.), and 1-D       */
/* Gauss-Newton / Newton on f(t) = C'(t).(C(t) - P).  It shares no code with the matrix form of  */
/* the GPU path.  Pinned by: exact circle arcs (w = cos(pi/4) quarter arcs), scipy BSpline for   */
/* the non-rational value and derivatives, finite differences, round trips.                     */
/* Global projection (the library's GN_SEED_GLOBAL; the reference's normal-cone projection,
 * src/kernel/kernel.cu:471-513, 623-704, is unfinished).  Exhaustive form of the library's
 * branch and bound: exact-Hessian projected Newton (a coordinate held on its bound while the
 * descent points outward) from the greedy seed leaf's stepped seed (or_seed_greedy_leaf, then
 * seed_step_from_leaf) over the whole domain, then
 * from EVERY leaf's stepped seed for P (seed_step_from_leaf's step from that leaf, clamped into
 * its cell) with the iterate clamped to that leaf's uv cell (leaf_iter steps; the seed leaf
 * itself only if the first solve ended outside its cell), the nearest foot
 * point kept (ties to the earlier candidate), then a full-domain Newton polish from it that is
 * kept unless it is farther by more than rounding (1e-9 relative).  The tree is the reference's
 * (uv cells of [0,1]^2, nodes as or_build_tree); leaves in Morton order of their cell (x bits
 * odd). */
static double dist2_at(const or_surf *sf, const double P[3], double u, double v)
{
    or_pt o;
    eval_one(sf, u, v, 0, &o);
    double dx = o.S[0] - P[0], dy = o.S[1] - P[1], dz = o.S[2] - P[2];
    return dx * dx + dy * dy + dz * dz;
}

void or_project_global(int p, int q, int nu, int nv, const double *ku, const double *kv,
                       const double *MatrixN, const double *nodes, int levels, const double *x,
                       const double *y, const double *z, long n, int max_iter, int leaf_iter,
                       double tol_step, double *u_out, double *v_out, unsigned char *status,
                       int nthreads)
{
    or_surf sf = {p, q, nu, nv, ku, kv, MatrixN};
    long base = or_tree_base(levels - 1), side = 1L << (3 * (levels - 1));
    or_gn_opts full = {0.0, tol_step, max_iter, 1, 1, 1, 0.0},
               leaf = {0.0, tol_step, leaf_iter, 1, 1, 1, 0.0};
    int *gleaf = malloc(sizeof(int) * (n > 0 ? n : 1));
    double *gsu = malloc(sizeof(double) * (n > 0 ? n : 1)), *gsv = malloc(sizeof(double) * (n > 0 ? n : 1));
    greedy_leaf_impl(nodes, levels, x, y, z, n, gleaf, gsu, gsv, nthreads, 1);
    seed_step_from_leaf(&sf, nodes, levels, x, y, z, n, gleaf, gsu, gsv, nthreads);
    const or_stepdom g = step_domain(&sf, levels);
    or_leafrec *recs = malloc(sizeof(or_leafrec) * side * side);
    for (long t = 0; t < side * side; t++) leaf_rec(&sf, nodes, base + t, &recs[t]);
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
#else
    (void)nthreads;
#endif
    for (long i = 0; i < n; i++) {
        double P[3] = {x[i], y[i], z[i]};
        /* the greedy seed leaf (descent + 3 x 3, as or_seed_greedy_leaf) and its stepped seed
         * (seed_step_from_leaf) start the bound solve */
        long bl = gleaf[i] - base;
        const double *nd0 = nodes + (base + bl) * 13;
        double bu = gsu[i], bv = gsv[i];
        int bst = gn_one_in(&sf, P, &full, &bu, &bv, NULL, NULL);
        double best = dist2_at(&sf, P, bu, bv);
        /* the bound solve stands for the seed leaf's own solve only if it ended in that cell */
        int bl_done = bu >= nd0[0] && bu <= nd0[2] && bv >= nd0[1] && bv <= nd0[3];
        for (long t = 0; t < side * side; t++) {
            if (t == bl && bl_done) continue;
            const double *nd = nodes + (base + t) * 13;
            double dom[4] = {nd[0], nd[2], nd[1], nd[3]};
            /* the leaf's stepped seed for P, clamped into its cell */
            double cu, cv;
            step_from_rec(&recs[t], &g, P, &cu, &cv);
            cu = fmin(fmax(cu, nd[0]), nd[2]);
            cv = fmin(fmax(cv, nd[1]), nd[3]);
            int st = gn_one_in(&sf, P, &leaf, &cu, &cv, NULL, dom);
            double d2 = dist2_at(&sf, P, cu, cv);
            if (d2 < best) { best = d2; bu = cu; bv = cv; bst = st; }
        }
        double pu = bu, pv = bv;
        int pst = gn_one_in(&sf, P, &full, &pu, &pv, NULL, NULL);
        if (dist2_at(&sf, P, pu, pv) <= best * (1.0 + 1e-9)) { bu = pu; bv = pv; bst = pst; }
        u_out[i] = bu;
        v_out[i] = bv;
        status[i] = (unsigned char)bst;
    }
    free(recs);
    free(gleaf);
    free(gsu);
    free(gsv);
}

/* ------------------------------------------------------------------------------------------ */
#define OR_CURVE_SAMPLES 8

/* ders[k][j] = k-th derivative of N_{span-p+j,p} at u, k <= nd (A2.3) */
static void ders_basis(int span, double u, int p, int nd, const double *U, double ders[3][OR_MAXK])
{
    double ndu[OR_MAXK][OR_MAXK], left[OR_MAXK], right[OR_MAXK], a[2][OR_MAXK];
    ndu[0][0] = 1.0;
    for (int j = 1; j <= p; j++) {
        left[j] = u - U[span + 1 - j];
        right[j] = U[span + j] - u;
        double saved = 0.0;
        for (int r = 0; r < j; r++) {
            ndu[j][r] = right[r + 1] + left[j - r];
            double temp = ndu[r][j - 1] / ndu[j][r];
            ndu[r][j] = saved + right[r + 1] * temp;
            saved = left[j - r] * temp;
        }
        ndu[j][j] = saved;
    }
    for (int k = 0; k <= 2; k++)
        for (int j = 0; j <= p; j++) ders[k][j] = 0.0;
    for (int j = 0; j <= p; j++) ders[0][j] = ndu[j][p];
    for (int r = 0; r <= p; r++) {
        int s1 = 0, s2 = 1;
        a[0][0] = 1.0;
        for (int k = 1; k <= nd && k <= p; k++) {
            double d = 0.0;
            int rk = r - k, pk = p - k;
            if (r >= k) {
                a[s2][0] = a[s1][0] / ndu[pk + 1][rk];
                d = a[s2][0] * ndu[rk][pk];
            }
            int j1 = rk >= -1 ? 1 : -rk;
            int j2 = (r - 1 <= pk) ? k - 1 : p - r;
            for (int j = j1; j <= j2; j++) {
                a[s2][j] = (a[s1][j] - a[s1][j - 1]) / ndu[pk + 1][rk + j];
                d += a[s2][j] * ndu[rk + j][pk];
            }
            if (r <= pk) {
                a[s2][k] = -a[s1][k - 1] / ndu[pk + 1][r];
                d += a[s2][k] * ndu[r][pk];
            }
            ders[k][r] = d;
            int t = s1; s1 = s2; s2 = t;
        }
    }
    double f = p;
    for (int k = 1; k <= nd && k <= p; k++) {
        for (int j = 0; j <= p; j++) ders[k][j] *= f;
        f *= (p - k);
    }
}

typedef struct {
    int p, n;
    const double *k, *xyz, *w;
} or_curve;

/* C, C', C'' (xyz each) at t; order 0..2 */
static void curve_eval_one(const or_curve *c, double t, int order, double out[9])
{
    int span = or_find_span(c->p, c->n, c->k, t);
    double d[3][OR_MAXK];
    ders_basis(span, t, c->p, order, c->k, d);
    double A[3][4] = {{0}};
    for (int j = 0; j <= c->p; j++) {
        int g = span - c->p + j;
        double w = c->w[g];
        double hp[4] = {c->xyz[3 * g] * w, c->xyz[3 * g + 1] * w, c->xyz[3 * g + 2] * w, w};
        for (int kk = 0; kk <= order; kk++)
            for (int q = 0; q < 4; q++) A[kk][q] += d[kk][j] * hp[q];
    }
    double iw = 1.0 / A[0][3];
    for (int q = 0; q < 3; q++) out[q] = A[0][q] * iw;
    if (order >= 1)
        for (int q = 0; q < 3; q++) out[3 + q] = (A[1][q] - A[1][3] * out[q]) * iw;
    if (order >= 2)
        for (int q = 0; q < 3; q++)
            out[6 + q] = (A[2][q] - 2.0 * A[1][3] * out[3 + q] - A[2][3] * out[q]) * iw;
}

/* out: SoA blocks of n doubles -- C (3), then C' (3) if order >= 1, then C'' (3) if order 2 */
void or_curve_eval(int p, int cp, const double *knots, const double *xyz, const double *w,
                   const double *t, long n, int order, double *out, int nthreads)
{
    or_curve c = {p, cp, knots, xyz, w};
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(nthreads)
#else
    (void)nthreads;
#endif
    for (long i = 0; i < n; i++) {
        double o[9];
        curve_eval_one(&c, t[i], order, o);
        for (int q = 0; q < 3 * (order + 1); q++) out[q * n + i] = o[q];
    }
}

/* Seeds: parameter of the nearest (FP32 squared distance, first index on ties) of 8 samples per
 * domain span at t = k[m] + (j + 1/2) h / 8, samples rounded to FP32. */
void or_curve_seed(int p, int cp, const double *knots, const double *xyz, const double *w,
                   const double *x, const double *y, const double *z, long n, double *t_out)
{
    or_curve c = {p, cp, knots, xyz, w};
    long ns = (long)(cp - p) * OR_CURVE_SAMPLES;
    float *sp = (float *)malloc(sizeof(float) * 3 * ns);
    double *st = (double *)malloc(sizeof(double) * ns);
    for (long i = 0; i < ns; i++) {
        int slot = (int)(i / OR_CURVE_SAMPLES), j = (int)(i % OR_CURVE_SAMPLES);
        double lo = knots[slot + p], hi = knots[slot + p + 1];
        double t = lo + (j + 0.5) * (hi - lo) / OR_CURVE_SAMPLES;
        double o[9];
        curve_eval_one(&c, t, 0, o);
        for (int q = 0; q < 3; q++) sp[3 * i + q] = (float)o[q];
        st[i] = t;
    }
    for (long i = 0; i < n; i++) {
        float P[3] = {(float)x[i], (float)y[i], (float)z[i]};
        float best = INFINITY;
        long bj = 0;
        for (long j = 0; j < ns; j++) {
            float dx = sp[3 * j] - P[0], dy = sp[3 * j + 1] - P[1], dz = sp[3 * j + 2] - P[2];
            float d2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
            if (d2 < best) { best = d2; bj = j; }
        }
        t_out[i] = st[bj];
    }
    free(sp);
    free(st);
}

/* 1-D Gauss-Newton (newton: f' = C'.C' + C''.(C-P) when positive) from given seeds, with the
 * stopping rules of the surface iteration: |C-P| < tol_dist before the update, |dt| < tol_step
 * after it. */
void or_curve_invert(int p, int cp, const double *knots, const double *xyz, const double *w,
                     const double *x, const double *y, const double *z, long n, double tol_dist,
                     double tol_step, int max_iter, int clamp, int newton, const double *seed,
                     double *t_out, unsigned char *status, int *iters, int nthreads)
{
    or_curve c = {p, cp, knots, xyz, w};
    double lo = knots[p], hi = knots[cp];
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads)
#else
    (void)nthreads;
#endif
    for (long i = 0; i < n; i++) {
        double P[3] = {x[i], y[i], z[i]}, t = seed[i];
        int st = OR_MAXITER, it;
        for (it = 0; it < max_iter; it++) {
            double o[9];
            curve_eval_one(&c, t, newton ? 2 : 1, o);
            double r[3] = {o[0] - P[0], o[1] - P[1], o[2] - P[2]};
            double rr = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
            if (!(rr < INFINITY)) { st = OR_NONFINITE; break; } /* NaN or infinite residual */
            if (tol_dist > 0 && sqrt(rr) < tol_dist) { st = OR_CONV_DIST; break; }
            double f = o[3] * r[0] + o[4] * r[1] + o[5] * r[2];
            double fp = o[3] * o[3] + o[4] * o[4] + o[5] * o[5];
            if (newton) {
                double f2 = fp + o[6] * r[0] + o[7] * r[1] + o[8] * r[2];
                if (f2 > 0) fp = f2;
            }
            if (!(fp == fp)) { st = OR_NONFINITE; break; }
            if (fp == 0) { st = OR_SINGULAR; break; }
            double dt = f / fp;
            t -= dt;
            if (clamp) t = t < lo ? lo : (t > hi ? hi : t);
            if (!(t == t)) { st = OR_NONFINITE; it++; break; }
            if (tol_step > 0 && fabs(dt) < tol_step) { st = OR_CONV_STEP; it++; break; }
        }
        t_out[i] = t;
        if (status) status[i] = (unsigned char)st;
        if (iters) iters[i] = it;
    }
}
