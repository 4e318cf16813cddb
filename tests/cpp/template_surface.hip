// Driver of the header-only template surface for the GPU tests: curve
// plug-ins written the way gcurval's samples write theirs (an ellipse family
// with every length rule, a cubic B-spline family with a curve locator, a
// Cox-de Boor derivative and a per-thread scratch allocator), built with
// hipcc against include/gcurval/gcurval.h and libgcurval_amd.so.
//
// Usage: template_surface <precision> <family> <rule> <in.bin> <out.bin> [tolerance]
//   precision: f32 | f64 (the plug-ins' DataType);  family: ellipse | bspline;
//   rule: gl | simpson | vincent | polyline | param
// in.bin (little endian):
//   ellipse: int32 n, float64 a[n], float64 b[n], int32 counts[n], int32 relative,
//            [float64 positions[sum counts] if relative == 1]
//   bspline: int32 n, int32 knot_off[n+1], float64 knots[knot_off[n]],
//            float64 ctrl[(knot_off[n] - 4 n) * 3], int32 counts[n]
// out.bin: float64 lengths[n], float64 u[sum counts], float64 points[sum counts][dim]
#include <gcurval/gcurval.h>

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

// ---------------------------------------------------------------- ellipse
template <typename Dt>
struct CurvesEllipse : gcurval::CurvesBase<Dt, gcurval::Vector2X<Dt>> {
    using PointType = gcurval::Vector2X<Dt>;
    Dt* a = nullptr;
    Dt* b = nullptr;
};

template <typename Dt>
struct PackEllipse : gcurval::ParaPackBase<Dt> {
    Dt a, b;
};

template <typename Dt>
struct ParserEllipse {
    __host__ __device__ PackEllipse<Dt> operator()(const CurvesEllipse<Dt>& c, int cid, int sid, bool = false) const {
        PackEllipse<Dt> p;
        p.begin = Dt(0);
        p.end = Dt(2 * gcurval::kPi);
        p.cid = cid;
        p.sid = sid;
        p.a = c.a[cid];
        p.b = c.b[cid];
        return p;
    }
};

template <typename Dt>
struct PointEllipse {
    __host__ __device__ gcurval::Vector2X<Dt> operator()(Dt u, const PackEllipse<Dt>& p) const {
        return gcurval::Vector2X<Dt>(p.a * std::cos(u), p.b * std::sin(u));
    }
};

template <typename Dt>
struct DerivEllipse {
    __host__ __device__ gcurval::Vector2X<Dt> operator()(Dt u, const PackEllipse<Dt>& p) const {
        return gcurval::Vector2X<Dt>(-p.a * std::sin(u), p.b * std::cos(u));
    }
};

template <typename Dt, template <typename, typename, typename> class Calc, template <typename> class Getter>
struct TraitsEllipse : gcurval::CurveTraitsBase<CurvesEllipse<Dt>> {
    using ParaPackParser = ParserEllipse<Dt>;
    using ParaPack = PackEllipse<Dt>;
    using GetPoint = PointEllipse<Dt>;
    using GetDerivative = DerivEllipse<Dt>;
    using CalculateFunction = Calc<Dt, PackEllipse<Dt>, Getter<Dt>>;
};

template <typename Dt, class Traits>
class EllipseEngine : public gcurval::ArclengthCalculateGpu<CurvesEllipse<Dt>, Traits> {
    using Base = gcurval::ArclengthCalculateGpu<CurvesEllipse<Dt>, Traits>;

public:
    EllipseEngine(int n, const Dt* a, const Dt* b) : Base(n) {
        gcurval::gMalloc(&this->d_curve.a, sizeof(Dt) * n);
        gcurval::gMalloc(&this->d_curve.b, sizeof(Dt) * n);
        gcurval::copyToDevice(this->d_curve.a, a, n);
        gcurval::copyToDevice(this->d_curve.b, b, n);
    }
    ~EllipseEngine() override {
        gcurval::gFree_s(this->d_curve.a);
        gcurval::gFree_s(this->d_curve.b);
    }
};

// ---------------------------------------------------------------- cubic B-spline (3-D)
constexpr int kDeg = 3;

template <typename Dt>
struct CurvesBSpline : gcurval::CurvesBase<Dt, gcurval::Vector3X<Dt>> {
    using Pt3 = gcurval::Vector3X<Dt>;
    using PointType = Pt3;
    int* knots_offset = nullptr;  // [n+1]
    Dt* knot_vector = nullptr;
    Pt3* ctrl_ptrs = nullptr;     // control points of curve c start at knots_offset[c] - (deg + 1) c
};

template <typename Dt>
struct PackBSpline : gcurval::ParaPackBase<Dt> {
    int span;
    const Dt* knots;
    const gcurval::Vector3X<Dt>* ctrl;
};

// Per-thread scratch: basis values and the two knot-difference rows.
template <typename Dt>
struct ScratchBSpline {
    Dt* basis;
    Dt* left;
    Dt* right;
};
template <typename Dt>
struct AllocBSpline {
    static constexpr int kPerThread = 3 * (kDeg + 1);
    __host__ __device__ ScratchBSpline<Dt> operator()(uint8_t* base) const {
        Dt* t = (Dt*)base + kPerThread * threadIdx.x;
        return ScratchBSpline<Dt>{t, t + (kDeg + 1), t + 2 * (kDeg + 1)};
    }
    static int size(int block_dim) { return block_dim * kPerThread * (int)sizeof(Dt); }
};

template <typename Dt>
struct ParserBSpline {
    __host__ __device__ PackBSpline<Dt> operator()(const CurvesBSpline<Dt>& c, int cid, int sid, bool = false) const {
        const int ko = c.knots_offset[cid];
        PackBSpline<Dt> p;
        p.cid = cid;
        p.sid = sid;
        p.span = sid + kDeg;
        p.knots = c.knot_vector + ko;
        p.ctrl = c.ctrl_ptrs + (ko - (kDeg + 1) * cid);
        p.begin = p.knots[p.span];
        p.end = p.knots[p.span + 1];
        return p;
    }
};

template <typename Dt>
struct LocatorBSpline {
    __host__ __device__ gcurval::CurveIntervalPos operator()(const CurvesBSpline<Dt>& c, int cid) const {
        return gcurval::CurveIntervalPos{c.knots_offset[cid] - (2 * kDeg + 1) * cid,
                                         c.knots_offset[cid + 1] - (2 * kDeg + 1) * (cid + 1)};
    }
};

// Cox-de Boor basis functions of degree q on span s at u (The NURBS Book,
// algorithm A2.2), in the scratch rows.
template <typename Dt>
__host__ __device__ inline void basis_funs(Dt u, int s, int q, const Dt* T, ScratchBSpline<Dt>& w) {
    w.basis[0] = Dt(1);
    for (int j = 1; j <= q; j++) {
        w.left[j] = u - T[s + 1 - j];
        w.right[j] = T[s + j] - u;
        Dt saved = Dt(0);
        for (int r = 0; r < j; r++) {
            const Dt tmp = w.basis[r] / (w.right[r + 1] + w.left[j - r]);
            w.basis[r] = saved + w.right[r + 1] * tmp;
            saved = w.left[j - r] * tmp;
        }
        w.basis[j] = saved;
    }
}

template <typename Dt>
struct PointBSpline {
    using Pt3 = gcurval::Vector3X<Dt>;
    __host__ __device__ Pt3 operator()(Dt u, const PackBSpline<Dt>& p, ScratchBSpline<Dt>& w) const {
        basis_funs(u, p.span, kDeg, p.knots, w);
        Pt3 v = Pt3::Zero();
        for (int j = 0; j <= kDeg; j++) v += w.basis[j] * p.ctrl[p.span - kDeg + j];
        return v;
    }
};

// dC/du = sum_j N_{j,p-1}(u) p (P_{j+1} - P_j) / (t_{j+p+1} - t_{j+1}) over the span.
template <typename Dt>
struct DerivBSpline {
    using Pt3 = gcurval::Vector3X<Dt>;
    __host__ __device__ Pt3 operator()(Dt u, const PackBSpline<Dt>& p, ScratchBSpline<Dt>& w) const {
        const int s = p.span;
        basis_funs(u, s, kDeg - 1, p.knots, w);
        Pt3 v = Pt3::Zero();
        for (int j = 0; j <= kDeg - 1; j++) {
            const int i = s - kDeg + 1 + j;  // derivative control point Q_{i-1}
            const Dt den = p.knots[i + kDeg] - p.knots[i];
            if (den > Dt(0)) v += (w.basis[j] * Dt(kDeg) / den) * (p.ctrl[i] - p.ctrl[i - 1]);
        }
        return v;
    }
};

template <typename Dt>
struct TraitsBSpline : gcurval::CurveTraitsBase<CurvesBSpline<Dt>> {
    using ParaPackParser = ParserBSpline<Dt>;
    using ParaPack = PackBSpline<Dt>;
    using SharedMemAllocator = AllocBSpline<Dt>;
    using CurveLocator = LocatorBSpline<Dt>;
    using GetPoint = PointBSpline<Dt>;
    using GetDerivative = DerivBSpline<Dt>;
    using CalculateFunction =
        gcurval::CalculateFunctionGaussLegendreQuadratureImpl<Dt, PackBSpline<Dt>, DerivBSpline<Dt>>;
};

template <typename Dt>
class BSplineEngine : public gcurval::ArclengthCalculateGpu<CurvesBSpline<Dt>, TraitsBSpline<Dt>> {
    using Base = gcurval::ArclengthCalculateGpu<CurvesBSpline<Dt>, TraitsBSpline<Dt>>;
    using Pt3 = gcurval::Vector3X<Dt>;
    using Base::d_curve;

public:
    BSplineEngine(int n, int n_seg, const std::vector<int>& ko, const std::vector<Dt>& knots,
                  const std::vector<Pt3>& ctrl)
        : Base(n, n_seg) {
        gcurval::gMalloc(&d_curve.knots_offset, sizeof(int) * ko.size());
        gcurval::gMalloc(&d_curve.knot_vector, sizeof(Dt) * knots.size());
        gcurval::gMalloc(&d_curve.ctrl_ptrs, sizeof(Pt3) * ctrl.size());
        gcurval::copyToDevice(d_curve.knots_offset, ko.data(), ko.size());
        gcurval::copyToDevice(d_curve.knot_vector, knots.data(), knots.size());
        gcurval::copyToDevice(d_curve.ctrl_ptrs, ctrl.data(), ctrl.size());
    }
    ~BSplineEngine() override {
        gcurval::gFree_s(d_curve.knots_offset);
        gcurval::gFree_s(d_curve.knot_vector);
        gcurval::gFree_s(d_curve.ctrl_ptrs);
    }
};

// ---------------------------------------------------------------- driver
template <typename T>
static void rd(std::ifstream& f, T* p, size_t n) {
    f.read((char*)p, sizeof(T) * n);
    if (!f) throw std::runtime_error("short input file");
}
template <typename T>
static void wr(std::ofstream& f, const T* p, size_t n) {
    f.write((const char*)p, sizeof(T) * n);
}

static double g_tol = 0.0;  // 0: the engine's default

template <typename Dt, class Engine, class Pt, int D>
static void run(Engine& eng, int n, std::vector<int>& counts, std::vector<Dt>* pos, std::ofstream& out) {
    if (g_tol > 0.0) eng.setTolerance(g_tol);
    eng.createArclengthTable();
    eng.createArclengthTable();  // a second build in place must give the same table
    Dt* L = eng.getAllCurveArclengths();
    std::vector<double> Ld(L, L + n);
    if (n > 0 && (double)eng.getArclength(n - 1) != Ld[n - 1]) throw std::runtime_error("getArclength mismatch");
    delete[] L;
    eng.clearArclength();
    wr(out, Ld.data(), Ld.size());
    Dt* positions = pos ? pos->data() : nullptr;
    auto us = eng.arclengthReverse(counts.data(), positions, true, gcurval::gcurvalHost,
                                   gcurval::gcurvalReverseReturnUs);
    std::vector<double> u(us.us, us.us + us.n_points);
    delete[] us.us;
    wr(out, u.data(), u.size());
    auto pts = eng.arclengthReverse(counts.data(), positions, true, gcurval::gcurvalDevice,
                                    gcurval::gcurvalReverseReturnPoints);
    std::vector<Pt> hp(pts.n_points);
    if (pts.n_points) gcurval::copyToHost(hp.data(), pts.points, hp.size());
    gcurval::gFree(pts.points);
    std::vector<double> flat;
    for (const Pt& p : hp)
        for (int d = 0; d < D; d++) flat.push_back((double)p[d]);
    wr(out, flat.data(), flat.size());
}

template <typename D_, typename P_, typename G_>
using GL = gcurval::CalculateFunctionGaussLegendreQuadratureImpl<D_, P_, G_>;
template <typename D_, typename P_, typename G_>
using SIMP = gcurval::CalculateFunctionSimpsonMethodImpl<D_, P_, G_>;
template <typename D_, typename P_, typename G_>
using VF = gcurval::CalculateFunctionVincentForseyRuleImpl<D_, P_, G_>;
template <typename D_, typename P_, typename G_>
using POLY = gcurval::CalculateFunctionPolylineMethodImpl<D_, P_, G_>;
template <typename D_, typename P_, typename G_>
using PARAM = gcurval::CalculateFunctionParameterDebugMethodImpl<D_, P_, G_>;

template <typename Dt, template <typename, typename, typename> class Calc, template <typename> class Getter>
static void run_ellipse(int n, std::vector<Dt>& a, std::vector<Dt>& b, std::vector<int>& counts,
                        std::vector<Dt>* pos, std::ofstream& out) {
    EllipseEngine<Dt, TraitsEllipse<Dt, Calc, Getter>> eng(n, a.data(), b.data());
    run<Dt, decltype(eng), gcurval::Vector2X<Dt>, 2>(eng, n, counts, pos, out);
}

template <typename Dt>
static void drive(const std::string& family, const std::string& rule, std::ifstream& in, std::ofstream& out) {
    int n = 0;
    rd(in, &n, 1);
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
        std::vector<Dt>* pp = relative ? &pos : nullptr;
        if (rule == "gl") run_ellipse<Dt, GL, DerivEllipse>(n, a, b, counts, pp, out);
        else if (rule == "simpson") run_ellipse<Dt, SIMP, PointEllipse>(n, a, b, counts, pp, out);
        else if (rule == "vincent") run_ellipse<Dt, VF, PointEllipse>(n, a, b, counts, pp, out);
        else if (rule == "polyline") run_ellipse<Dt, POLY, PointEllipse>(n, a, b, counts, pp, out);
        else if (rule == "param") run_ellipse<Dt, PARAM, PointEllipse>(n, a, b, counts, pp, out);
        else throw std::runtime_error("unknown rule " + rule);
    } else if (family == "bspline") {
        using Pt3 = gcurval::Vector3X<Dt>;
        std::vector<int> ko(n + 1);
        rd(in, ko.data(), n + 1);
        std::vector<double> kd(ko[n]);
        rd(in, kd.data(), kd.size());
        const size_t n_ctrl = (size_t)ko[n] - (size_t)(kDeg + 1) * n;
        std::vector<double> cd(n_ctrl * 3);
        rd(in, cd.data(), cd.size());
        std::vector<int> counts(n);
        rd(in, counts.data(), n);
        std::vector<Dt> knots(kd.begin(), kd.end());
        std::vector<Pt3> ctrl(n_ctrl);
        for (size_t j = 0; j < n_ctrl; j++) ctrl[j] = Pt3(Dt(cd[3 * j]), Dt(cd[3 * j + 1]), Dt(cd[3 * j + 2]));
        const int n_seg = ko[n] - (2 * kDeg + 1) * n;
        BSplineEngine<Dt> eng(n, n_seg, ko, knots, ctrl);
        run<Dt, BSplineEngine<Dt>, Pt3, 3>(eng, n, counts, nullptr, out);
    } else {
        throw std::runtime_error("unknown family " + family);
    }
}

int main(int argc, char** argv) {
    if (argc != 6 && argc != 7) {
        std::fprintf(stderr,
                     "usage: %s f32|f64 ellipse|bspline gl|simpson|vincent|polyline|param in.bin out.bin [tol]\n",
                     argv[0]);
        return 2;
    }
    if (argc == 7) g_tol = std::atof(argv[6]);
    const std::string prec = argv[1], family = argv[2], rule = argv[3];
    try {
        std::ifstream in(argv[4], std::ios::binary);
        std::ofstream out(argv[5], std::ios::binary);
        if (prec == "f32") drive<float>(family, rule, in, out);
        else if (prec == "f64") drive<double>(family, rule, in, out);
        else throw std::runtime_error("unknown precision " + prec);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "template_surface: %s\n", e.what());
        return 1;
    }
    return 0;
}
