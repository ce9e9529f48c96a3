// span.hip -- knot-span lookup and the per-span basis-matrix / span-table build
// (findSpan src/NURBS/Evaluate.cu:504-533, Nrc/Nabla 90-108, GetMatrixN 3-88).
//
// The reference builds one basis matrix entry at a time with a 4-way recursion (4^p calls per
// entry), allocates from the device heap and launches eight child grids per span.  Here one
// thread builds a whole k x k basis matrix with the triangular Cox-de Boor scheme carried on
// power-basis coefficients of the local parameter, and one workgroup per span slot forms the
// four products M_u * P_c * M_v^T with no dynamic parallelism and no device heap.
#include "gn_device.hpp"

namespace gn {

// One thread per span slot in one direction: basis matrix (k*k doubles) and 1/h.
__global__ void k_basis(int deg, int ncp, const double *__restrict__ kn, double *__restrict__ Mout,
                        double *__restrict__ ih) {
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    const int nslots = ncp - deg;
    if (slot >= nslots) return;
    const int span = slot + deg;
    const int k = deg + 1;
    double M[GN_MAXK * GN_MAXK];
    const double h = kn[span + 1] - kn[span];
    if (h == 0.0) {
        for (int e = 0; e < k * k; e++) Mout[slot * k * k + e] = 0.0;
        ih[slot] = 0.0;
        return;
    }
    basis_matrix(deg, kn, span, M);
    for (int e = 0; e < k * k; e++) Mout[slot * k * k + e] = M[e];
    ih[slot] = 1.0 / h;
}

// One workgroup per span slot (im, in); thread (r, s) writes the 4-vector patch[r][s][:] of
// M_u * P_w * M_v^T (homogeneous control points P_w = (x w, y w, z w, w)).
__global__ void k_span_table(int p, int q, int nu, int nv, const double *__restrict__ Mu,
                             const double *__restrict__ Mv, const double *__restrict__ ihu,
                             const double *__restrict__ ihv, const double *__restrict__ xyz,
                             const double *__restrict__ w, double *__restrict__ tab64,
                             float *__restrict__ tab32) {
    const int sv = nv - q;
    const int im = blockIdx.x / sv, in = blockIdx.x % sv;
    const int k1 = p + 1, k2 = q + 1;
    const int e = threadIdx.x;
    if (e >= k1 * k2) return;
    const int r = e / k2, s = e % k2;
    double acc[4] = {0, 0, 0, 0};
    const bool empty = (ihu[im] == 0.0) || (ihv[in] == 0.0);
    if (!empty) {
        for (int a = 0; a < k1; a++) {
            const double mu = Mu[im * k1 * k1 + r * k1 + a];
            double row[4] = {0, 0, 0, 0};
            for (int b = 0; b < k2; b++) {
                const double mv = Mv[in * k2 * k2 + s * k2 + b];
                const size_t g = (size_t)(im + a) * nv + (in + b);  // control point (m-p+a, n-q+b)
                const double wt = w[g];
                row[0] += mv * (xyz[g * 3 + 0] * wt);
                row[1] += mv * (xyz[g * 3 + 1] * wt);
                row[2] += mv * (xyz[g * 3 + 2] * wt);
                row[3] += mv * wt;
            }
            for (int c = 0; c < 4; c++) acc[c] += mu * row[c];
        }
    }
    const size_t base = ((size_t)blockIdx.x * k1 * k2 + e) * 4;
    for (int c = 0; c < 4; c++) {
        tab64[base + c] = acc[c];
        tab32[base + c] = (float)acc[c];
    }
}

__global__ void k_find_span(const double *__restrict__ kn, int deg, int ncp,
                            const double *__restrict__ u, size_t n, int32_t *__restrict__ out) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        out[i] = find_span(kn, deg, ncp, u[i]);
}

int launch_span_table(const gn_surface *s, const double *d_xyz, const double *d_w,
                      hipStream_t st) {
    const int su = s->nu - s->p, sv = s->nv - s->q;
    const int k1 = s->p + 1, k2 = s->q + 1;
    double *Mu = nullptr, *Mv = nullptr;
    GN_HIP(hipMalloc((void **)&Mu, sizeof(double) * su * k1 * k1));
    GN_HIP(hipMalloc((void **)&Mv, sizeof(double) * sv * k2 * k2));
    k_basis<<<grid_for(su, 64, 1u << 20), 64, 0, st>>>(s->p, s->nu, s->d_ku, Mu, s->d_ihu);
    k_basis<<<grid_for(sv, 64, 1u << 20), 64, 0, st>>>(s->q, s->nv, s->d_kv, Mv, s->d_ihv);
    k_span_table<<<su * sv, 64, 0, st>>>(s->p, s->q, s->nu, s->nv, Mu, Mv, s->d_ihu, s->d_ihv,
                                         d_xyz, d_w, s->d_tab64, s->d_tab32);
    GN_HIP(hipGetLastError());
    // one-time set-up: wait for the build before releasing the basis matrices
    GN_HIP(hipStreamSynchronize(st));
    GN_HIP(hipFree(Mu));
    GN_HIP(hipFree(Mv));
    return GN_OK;
}

int launch_find_span(const GnSurfDev &d, int dir, const double *u, size_t n, int32_t *span,
                     hipStream_t st) {
    if (n == 0) return GN_OK;
    const double *kn = dir == 0 ? d.ku : d.kv;
    const int deg = dir == 0 ? d.p : d.q, ncp = dir == 0 ? d.nu : d.nv;
    k_find_span<<<grid_for(n, 256, 8192), 256, 0, st>>>(kn, deg, ncp, u, n, span);
    GN_HIP(hipGetLastError());
    return GN_OK;
}

}  // namespace gn
