// Driver of the plug-in headers tests/cpp/plugin_ellipse.h and
// plugin_bspline.h, whose GPU and CPU engine subclasses sit side by side as in
// gcurval's samples. Built twice: by hipcc (both engines) and by a host
// compiler as C++ (the CPU engine only; with sanitizers in the CPU tests).
//
// Usage: plugin_engines <f32|f64> <ellipse|bspline> <cpu|gpu> <in.bin> <out.bin> [threads]
// in.bin (little endian):
//   ellipse: int32 n, float64 a[n], float64 b[n], int32 counts[n], int32 relative,
//            [float64 positions[sum counts] if relative == 1]
//   bspline: int32 n, int32 knot_off[n+1], float64 knots[knot_off[n]],
//            float64 ctrl[(knot_off[n] - 4 n) * 3], int32 counts[n]
//            (cubic 3-D; uniform positions j / (count - 1))
// out.bin: float64 lengths[n], float64 u[sum counts], float64 points[sum counts][dim]
#include "plugin_bspline.h"
#include "plugin_ellipse.h"

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

template <typename T>
static void rd(std::ifstream& f, T* p, size_t n) {
    f.read((char*)p, (std::streamsize)(sizeof(T) * n));
    if (!f) throw std::runtime_error("short input file");
}
template <typename T>
static void wr(std::ofstream& f, const T* p, size_t n) {
    f.write((const char*)p, (std::streamsize)(sizeof(T) * n));
}

static int g_threads = 1;

// The reference's call sequence on one engine: build, lengths (all, then one
// by id), parameters at the positions, then points at the same positions.
template <typename Dt, class Pt, int D, class Engine>
static void run(Engine& eng, int n, std::vector<int>& counts, Dt* positions, bool device_results,
                std::ofstream& out) {
    eng.createArclengthTable();
    Dt* L = eng.getAllCurveArclengths();
    std::vector<double> Ld(L, L + n);
    delete[] L;
    for (int c = 0; c < n; c++)
        if ((double)eng.getArclength(c) != Ld[c]) throw std::runtime_error("getArclength disagrees");
    wr(out, Ld.data(), Ld.size());
    const gcurval::GcurvalEnum where = device_results ? gcurval::gcurvalDevice : gcurval::gcurvalHost;
    auto us = eng.arclengthReverse(counts.data(), positions, true, where, gcurval::gcurvalReverseReturnUs);
    std::vector<Dt> hu(us.n_points);
#if defined(__HIP__)
    if (device_results) {
        if (us.n_points) gcurval::copyToHost(hu.data(), us.us, hu.size());
        gcurval::gFree(us.us);
    } else
#endif
    {
        std::copy(us.us, us.us + us.n_points, hu.begin());
        delete[] us.us;
    }
    std::vector<double> u(hu.begin(), hu.end());
    wr(out, u.data(), u.size());
    auto pts = eng.arclengthReverse(counts.data(), positions, true, gcurval::gcurvalHost,
                                    gcurval::gcurvalReverseReturnPoints);
    std::vector<double> flat;
    for (int i = 0; i < pts.n_points; i++)
        for (int d = 0; d < D; d++) flat.push_back((double)pts.points[i][d]);
    delete[] pts.points;
    wr(out, flat.data(), flat.size());
}

template <typename Dt>
static void drive(const std::string& family, const std::string& engine, std::ifstream& in, std::ofstream& out) {
    int n = 0;
    rd(in, &n, 1);
    const bool gpu = engine == "gpu";
#if !defined(__HIP__)
    if (gpu) throw std::runtime_error("this build has no GPU engine");
#endif
    if (family == "ellipse") {
        std::vector<double> ad(n), bd(n);
        rd(in, ad.data(), n);
        rd(in, bd.data(), n);
        std::vector<int> counts(n);
        rd(in, counts.data(), n);
        int relative = 0;
        rd(in, &relative, 1);
        std::vector<Dt> a(ad.begin(), ad.end()), b(bd.begin(), bd.end());
        std::vector<Dt> pos;
        if (relative) {
            size_t tot = 0;
            for (int c : counts) tot += (size_t)c;
            std::vector<double> pd(tot);
            rd(in, pd.data(), tot);
            pos.assign(pd.begin(), pd.end());
        }
        Dt* pp = relative ? pos.data() : nullptr;
        if (gpu) {
#if defined(__HIP__)
            plug_ellipse::ArclengthCalculateGpu<Dt> eng(n, a.data(), b.data());
            run<Dt, gcurval::Vector2X<Dt>, 2>(eng, n, counts, pp, true, out);
#endif
        } else {
            plug_ellipse::ArclengthCalculateCpu<Dt> eng(n, a.data(), b.data());
            eng.setThreads(g_threads);
            run<Dt, gcurval::Vector2X<Dt>, 2>(eng, n, counts, pp, false, out);
        }
    } else if (family == "bspline") {
        using Pt3 = gcurval::Vector3X<Dt>;
        std::vector<int> ko(n + 1);
        rd(in, ko.data(), n + 1);
        std::vector<double> kd(ko[n]);
        rd(in, kd.data(), kd.size());
        const size_t n_ctrl = (size_t)ko[n] - (size_t)4 * n;
        std::vector<double> cd(n_ctrl * 3);
        rd(in, cd.data(), cd.size());
        std::vector<int> counts(n);
        rd(in, counts.data(), n);
        std::vector<Dt> knots(kd.begin(), kd.end());
        std::vector<Pt3> ctrl(n_ctrl);
        for (size_t j = 0; j < n_ctrl; j++) ctrl[j] = Pt3(Dt(cd[3 * j]), Dt(cd[3 * j + 1]), Dt(cd[3 * j + 2]));
        if (gpu) {
#if defined(__HIP__)
            plug_bspline::ArclengthCalculateGpu<Dt, Pt3> eng(n, ko.data(), knots.data(), ctrl.data());
            run<Dt, Pt3, 3>(eng, n, counts, nullptr, true, out);
#endif
        } else {
            plug_bspline::ArclengthCalculateCpu<Dt, Pt3> eng(n, ko.data(), knots.data(), ctrl.data());
            eng.setThreads(g_threads);
            run<Dt, Pt3, 3>(eng, n, counts, nullptr, false, out);
        }
    } else {
        throw std::runtime_error("unknown family " + family);
    }
}

int main(int argc, char** argv) {
    if (argc != 6 && argc != 7) {
        std::fprintf(stderr, "usage: %s f32|f64 ellipse|bspline cpu|gpu in.bin out.bin [threads]\n", argv[0]);
        return 2;
    }
    if (argc == 7) g_threads = std::atoi(argv[6]);
    const std::string prec = argv[1], family = argv[2], engine = argv[3];
    try {
        if (engine != "cpu" && engine != "gpu") throw std::runtime_error("engine must be cpu or gpu");
        std::ifstream in(argv[4], std::ios::binary);
        std::ofstream out(argv[5], std::ios::binary);
        if (prec == "f32") drive<float>(family, engine, in, out);
        else if (prec == "f64") drive<double>(family, engine, in, out);
        else throw std::runtime_error("unknown precision " + prec);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "plugin_engines: %s\n", e.what());
        return 1;
    }
    return 0;
}
