// An ellipse curve family written as gcurval's sample plug-ins are: curve
// arrays, parameter pack, parser, point and derivative functors, traits, and
// the GPU and CPU engine subclasses side by side in one header
// (the layout of samples/sample1_ellipse/ellipse.h:1-173). Against
// include/gcurval it needs only the substitutions INTEGRATION.md lists:
//   * the GPU subclass is guarded by __HIP__ (hipcc) instead of USE_CUDA;
//   * uploads go through gcurval::copyToDevice instead of cudaMemcpy;
//   * members of the dependent base are named this->d_curve / this->curve
//     (standard two-phase lookup);
//   * gcurval::kPi replaces EIGEN_PI.
#pragma once

#include <gcurval/gcurval.h>

#include <algorithm>
#include <cmath>

namespace plug_ellipse {

template <typename Dt, typename Vec = gcurval::Vector2X<Dt>>
GCURVAL_HD_FUNC Vec position(Dt t, Dt a, Dt b) {
    return Vec(a * std::cos(t), b * std::sin(t));
}

template <typename Dt, typename Vec = gcurval::Vector2X<Dt>>
GCURVAL_HD_FUNC Vec tangent(Dt t, Dt a, Dt b) {
    return Vec(-1.0 * a * std::sin(t), b * std::cos(t));
}

template <typename Dt>
struct Curves : gcurval::CurvesBase<Dt, gcurval::Vector2X<Dt>> {
    using PointType = gcurval::Vector2X<Dt>;
    Dt* a = nullptr;
    Dt* b = nullptr;
    GCURVAL_HD_FUNC explicit constexpr Curves(int n_curve = 0, Dt* a_ = nullptr, Dt* b_ = nullptr)
        : gcurval::CurvesBase<Dt, PointType>(n_curve), a(a_), b(b_) {}
};

template <typename Dt>
struct Pack : gcurval::ParaPackBase<Dt> {
    Dt a;
    Dt b;
    GCURVAL_HD_FUNC Pack() = default;
    GCURVAL_HD_FUNC constexpr Pack(Dt begin, Dt end, int cid, int sid, Dt a_ = Dt{}, Dt b_ = Dt{})
        : gcurval::ParaPackBase<Dt>{begin, end, cid, sid}, a(a_), b(b_) {}
};

// One knot interval per curve: t in [0, 2 pi].
template <typename Dt>
struct Parser : gcurval::ParaPackParserBase<Dt, Pack<Dt>> {
    GCURVAL_HD_FUNC Pack<Dt> operator()(const Curves<Dt>& c, int cid, int, bool = false) const {
        return Pack<Dt>{Dt(0), Dt(2 * gcurval::kPi), cid, 0, c.a[cid], c.b[cid]};
    }
};

template <typename Dt>
struct Point {
    GCURVAL_HD_FUNC gcurval::Vector2X<Dt> operator()(Dt t, const Pack<Dt>& p) const { return position(t, p.a, p.b); }
};

template <typename Dt>
struct Derivative {
    GCURVAL_HD_FUNC gcurval::Vector2X<Dt> operator()(Dt t, const Pack<Dt>& p) const { return tangent(t, p.a, p.b); }
};

template <typename Dt>
struct Traits : gcurval::CurveTraitsBase<Curves<Dt>> {
    using ParaPackParser = Parser<Dt>;
    using ParaPack = Pack<Dt>;
    using GetPoint = Point<Dt>;
    using GetDerivative = Derivative<Dt>;
    using CalculateFunction = gcurval::CalculateFunctionGaussLegendreQuadratureImpl<Dt, ParaPack, GetDerivative>;
};

#if defined(__HIP__)

template <typename Dt, int max_recursive_depth = gcurval::MaxRecursiveDepth<Dt>::value, int n_subdivided = 64>
class ArclengthCalculateGpu
    : public gcurval::ArclengthCalculateGpu<Curves<Dt>, Traits<Dt>, max_recursive_depth, n_subdivided> {
    using Base = gcurval::ArclengthCalculateGpu<Curves<Dt>, Traits<Dt>, max_recursive_depth, n_subdivided>;

public:
    explicit ArclengthCalculateGpu(int n_curve, Dt* a, Dt* b) : Base(n_curve, n_curve) {
        gcurval::gMalloc(&this->d_curve.a, sizeof(Dt) * n_curve);
        gcurval::gMalloc(&this->d_curve.b, sizeof(Dt) * n_curve);
        gcurval::copyToDevice(this->d_curve.a, a, n_curve);
        gcurval::copyToDevice(this->d_curve.b, b, n_curve);
    }
    ~ArclengthCalculateGpu() {
        gcurval::gFree_s(this->d_curve.a);
        gcurval::gFree_s(this->d_curve.b);
    }
};

#endif  // __HIP__

template <typename Dt, int max_recursive_depth = gcurval::MaxRecursiveDepth<Dt>::value>
class ArclengthCalculateCpu : public gcurval::ArclengthCalculateCpu<Curves<Dt>, Traits<Dt>, max_recursive_depth> {
    using Base = gcurval::ArclengthCalculateCpu<Curves<Dt>, Traits<Dt>, max_recursive_depth>;

public:
    explicit ArclengthCalculateCpu(int n_curve, Dt* a, Dt* b) : Base(n_curve, n_curve) {
        this->curve.a = new Dt[n_curve];
        this->curve.b = new Dt[n_curve];
        std::copy(a, a + n_curve, this->curve.a);
        std::copy(b, b + n_curve, this->curve.b);
    }
    ~ArclengthCalculateCpu() {
        gcurval::freeArray_s(this->curve.a);
        gcurval::freeArray_s(this->curve.b);
    }
};

}  // namespace plug_ellipse
