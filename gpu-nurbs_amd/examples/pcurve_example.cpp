// pcurve_example.cpp -- the p-curve use (README.md:10) on the C++ interface: read a surface
// (.nu), take its v-iso-curve at u0 as a 3-D rational curve (homogeneous control points
// sum_i N_i(u0) (w_ij P_ij, w_ij) on the v knots), and fit its p-curve with NurbsSurface::PCurve.
// The exact p-curve of an iso-curve is u = u0, v = t.  Prints one JSON line.
//
// usage: pcurve_example <file.nu> [u0] [degree] [ncp] [samples]
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "gnurbs.hpp"

namespace {

struct NuFile {
    int p = -1, q = -1, nu = -1, nv = -1;
    std::vector<double> ku, kv, xyz, w;
};

bool read_nu(const char *path, NuFile &f) {  // the tagged NU / CP / KNu / KNv / DP format
    std::ifstream in(path);
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::string tag;
        if (!(ls >> tag)) continue;
        if (tag == "NU") ls >> f.p >> f.q;
        else if (tag == "CP") ls >> f.nu >> f.nv;
        else if (tag == "KNu" || tag == "KNv") {
            std::vector<double> &k = tag == "KNu" ? f.ku : f.kv;
            const int cnt = tag == "KNu" ? f.nu + f.p + 1 : f.nv + f.q + 1;
            k.resize(cnt);
            for (double &x : k) ls >> x;
        } else if (tag == "DP") {
            double x, y, z, ww;
            ls >> x >> y >> z >> ww;
            f.xyz.insert(f.xyz.end(), {x, y, z});
            f.w.push_back(ww);
        }
    }
    return f.p >= 0 && f.nu > 0 && (int)f.w.size() >= f.nu * f.nv;
}

// B-spline basis values N_{i,p}(u), i = 0..n-1 (Cox-de Boor, half-open spans; u = last knot
// takes the last span)
std::vector<double> basis(const std::vector<double> &k, int p, int n, double u) {
    std::vector<double> N(k.size() - 1, 0.0);
    int span = p;
    while (span < n - 1 && u >= k[span + 1]) span++;
    N[span] = 1.0;
    for (int d = 1; d <= p; d++)
        for (int i = 0; i + d < (int)k.size() - 1; i++) {
            double a = 0, b = 0;
            if (k[i + d] > k[i]) a = (u - k[i]) / (k[i + d] - k[i]) * N[i];
            if (k[i + d + 1] > k[i + 1]) b = (k[i + d + 1] - u) / (k[i + d + 1] - k[i + 1]) * N[i + 1];
            N[i] = a + b;
        }
    N.resize(n);
    return N;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s file.nu [u0] [degree] [ncp] [samples]\n", argv[0]);
        return 1;
    }
    const double u0 = argc > 2 ? std::atof(argv[2]) : 0.37;
    const int degree = argc > 3 ? std::atoi(argv[3]) : 3;
    const int ncp = argc > 4 ? std::atoi(argv[4]) : 12;
    const size_t samples = argc > 5 ? (size_t)std::atoll(argv[5]) : 20000;
    NuFile f;
    if (!read_nu(argv[1], f)) {
        std::fprintf(stderr, "cannot read %s\n", argv[1]);
        return 1;
    }
    // the v-iso-curve at u0
    const std::vector<double> Nu = basis(f.ku, f.p, f.nu, u0);
    std::vector<double> cxyz(3 * (size_t)f.nv, 0.0), cw(f.nv, 0.0);
    for (int j = 0; j < f.nv; j++) {
        double hx = 0, hy = 0, hz = 0, hw = 0;
        for (int i = 0; i < f.nu; i++) {
            const size_t e = (size_t)i * f.nv + j;
            const double c = Nu[i] * f.w[e];
            hx += c * f.xyz[3 * e];
            hy += c * f.xyz[3 * e + 1];
            hz += c * f.xyz[3 * e + 2];
            hw += c;
        }
        cxyz[3 * j] = hx / hw;
        cxyz[3 * j + 1] = hy / hw;
        cxyz[3 * j + 2] = hz / hw;
        cw[j] = hw;
    }
    try {
        gnurbs::NurbsSurface s = gnurbs::NurbsSurface::readNu(argv[1]);
        s.BuildTree(3);
        gnurbs::NurbsCurve c(f.q, f.kv, cxyz, cw);
        gn_invert_opts o = gnurbs::default_opts();
        o.tol_step = 1e-12;
        const gnurbs::NurbsSurface::PCurveFit fit = s.PCurve(c, samples, degree, ncp, o);
        std::printf("{\"u0\": %.17g, \"degree\": %d, \"ncp\": %d, \"samples\": %zu, \"max_dev\": %.6g, "
                    "\"failed\": %zu, \"cp_uv\": [",
                    u0, degree, ncp, samples, fit.max_dev, fit.failed);
        for (size_t i = 0; i < fit.cp_uv.size(); i++)
            std::printf("%s%.17g", i ? ", " : "", fit.cp_uv[i]);
        std::printf("], \"knots\": [");
        for (size_t i = 0; i < fit.knots.size(); i++)
            std::printf("%s%.17g", i ? ", " : "", fit.knots[i]);
        std::printf("]}\n");
    } catch (const gnurbs::Error &e) {
        std::fprintf(stderr, "gnurbs error %d: %s\n", e.code, e.what());
        return 2;
    }
    return 0;
}
