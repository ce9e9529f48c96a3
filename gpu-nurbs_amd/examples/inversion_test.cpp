// inversion_test.cpp -- the reference driver's INVERSION_TEST sequence (src/main.cu:277-327) on the
// C++ interface: readNu, span tables, a tree of `level` 8x8 levels, then NewInverse of n copies of
// one query point with the reference's distance tolerance and iteration count, first-hit seeding.
// Prints one JSON line with the evaluated point at (0.25, 0.25) and the inversion results.
//
// usage: inversion_test <file.nu> [x y z] [n] [level] [tol]
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "gnurbs.hpp"

#define HIPCHECK(x)                                                                  \
    do {                                                                             \
        hipError_t e_ = (x);                                                         \
        if (e_ != hipSuccess) {                                                      \
            std::fprintf(stderr, "HIP error %s at %s\n", hipGetErrorString(e_), #x); \
            return 2;                                                                \
        }                                                                            \
    } while (0)

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s file.nu [x y z] [n] [level] [tol]\n", argv[0]);
        return 1;
    }
    // defaults: the query hard-coded for center_bk.nu (src/kernel/kernel.cu:454), tree level 2
    // (src/main.cu:283), tolerance 0.01 and 10 iterations (src/main.cu:310, kernel.cu:157)
    double P[3] = {0.0, 0.200049, 0.616};
    if (argc >= 5) for (int k = 0; k < 3; k++) P[k] = std::atof(argv[2 + k]);
    const size_t n = argc >= 6 ? (size_t)std::atoll(argv[5]) : 4096;
    const int level = argc >= 7 ? std::atoi(argv[6]) : 2;
    const double tol = argc >= 8 ? std::atof(argv[7]) : 0.01;
    try {
        gnurbs::NurbsSurface s = gnurbs::NurbsSurface::readNu(argv[1]);
        s.BuildTree(level);
        double *d = nullptr;
        uint8_t *st = nullptr;
        HIPCHECK(hipMalloc(&d, sizeof(double) * 5 * n + sizeof(double) * 6));
        HIPCHECK(hipMalloc(&st, n));
        double *x = d, *y = d + n, *z = d + 2 * n, *u = d + 3 * n, *v = d + 4 * n, *q = d + 5 * n;
        std::vector<double> hx(n, P[0]), hy(n, P[1]), hz(n, P[2]);
        HIPCHECK(hipMemcpy(x, hx.data(), sizeof(double) * n, hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(y, hy.data(), sizeof(double) * n, hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(z, hz.data(), sizeof(double) * n, hipMemcpyHostToDevice));
        // the known-answer evaluation at (0.25, 0.25)
        const double quv[2] = {0.25, 0.25};
        HIPCHECK(hipMemcpy(q, quv, sizeof(quv), hipMemcpyHostToDevice));
        s.EvaluateRandom(q, q + 1, 1, 0, q + 2);
        gn_invert_opts o = gnurbs::default_opts();
        o.seed_mode = GN_SEED_REF_FIRST_HIT;
        o.tol_dist = tol;
        o.tol_step = 0.0;
        o.max_iter = 10;
        o.clamp = 0;
        hipEvent_t a, b;
        HIPCHECK(hipEventCreate(&a));
        HIPCHECK(hipEventCreate(&b));
        HIPCHECK(hipEventRecord(a, nullptr));
        s.NewInverse(x, y, z, n, u, v, st, o);
        HIPCHECK(hipEventRecord(b, nullptr));
        HIPCHECK(hipEventSynchronize(b));
        float ms = 0;
        HIPCHECK(hipEventElapsedTime(&ms, a, b));
        double ev[3], hu, hv;
        uint8_t hs;
        HIPCHECK(hipMemcpy(ev, q + 2, sizeof(ev), hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(&hu, u, sizeof(double), hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(&hv, v, sizeof(double), hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(&hs, st, 1, hipMemcpyDeviceToHost));
        std::vector<uint8_t> all(n);
        HIPCHECK(hipMemcpy(all.data(), st, n, hipMemcpyDeviceToHost));
        size_t same = 0;
        for (uint8_t c : all) same += c == hs;
        std::printf("{\"p=\": [%d, %d], \"S(0.25,0.25)\": [%.9f, %.9f, %.9f], \"u\": %.15g, \"v\": %.15g, "
                    "\"status\": %d, \"same_status\": %zu, \"n\": %zu, \"ms\": %.4f}\n",
                    s.degree_u(), s.degree_v(), ev[0], ev[1], ev[2], hu, hv, (int)hs, same, n, ms);
        HIPCHECK(hipFree(d));
        HIPCHECK(hipFree(st));
    } catch (const gnurbs::Error &e) {
        std::fprintf(stderr, "gnurbs error %d: %s\n", e.code, e.what());
        return 3;
    }
    return 0;
}
