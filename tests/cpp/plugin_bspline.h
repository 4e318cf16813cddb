// A B-spline curve family written as gcurval's B-spline sample is: curves in
// a knot-offset layout, per-thread scratch rows for the Cox-de Boor triangle
// handed out by a scratch allocator, a curve locator, parser, point and
// derivative functors, traits, and the GPU and CPU engine subclasses side by
// side (the layout of samples/sample2_bspline/structs.h and
// samples/sample2_bspline/arclength_calculate.h:1-114). Against
// include/gcurval it needs the substitutions INTEGRATION.md lists, plus
// __HIP_DEVICE_COMPILE__ for __CUDA_ARCH__ in the scratch allocator.
//
// Storage of curve c (degree p, clamped): knots at knot_vector[knots_offset[c]
// ..], control points at ctrl_ptrs[knots_offset[c] - (p + 1) c ..], knot
// intervals [t_{s+p}, t_{s+p+1}] for s = 0 .. n_knots - 2p - 2.
#pragma once

#include <gcurval/gcurval.h>

#include <algorithm>

namespace plug_bspline {

// Cox-de Boor basis values of degree q on span s at u (The NURBS Book, A2.2)
// in the scratch rows N, left, right.
template <typename Dt>
GCURVAL_HD_FUNC void basis(Dt u, int s, int q, const Dt* T, Dt* N, Dt* left, Dt* right) {
    N[0] = Dt(1.0);
    for (int j = 1; j <= q; j++) {
        left[j] = u - T[s + 1 - j];
        right[j] = T[s + j] - u;
        Dt carry = 0.0;
        for (int r = 0; r < j; r++) {
            const Dt f = N[r] / (right[r + 1] + left[j - r]);
            N[r] = carry + right[r + 1] * f;
            carry = left[j - r] * f;
        }
        N[j] = carry;
    }
}

// First derivatives of the degree-p basis on span s (The NURBS Book, eq. 2.7
// on the degree p - 1 triangle): N'[r] = p (N_{r,p-1} / d_r - N_{r+1,p-1} / d_{r+1}).
template <int p, typename Dt>
GCURVAL_HD_FUNC void basis_derivative(Dt u, int s, const Dt* T, Dt* N, Dt* left, Dt* right) {
    basis(u, s, p - 1, T, N, left, right);
    left[p] = u - T[s + 1 - p];
    right[p] = T[s + p] - u;
    Dt carry = 0.0;
    for (int r = 0; r < p; r++) {
        const Dt f = p * N[r] / (right[r + 1] + left[p - r]);
        N[r] = carry - f;
        carry = f;
    }
    N[p] = carry;
}

// sum_i P_{s-p+i} w_i
template <int p, typename Dt, typename Pt>
GCURVAL_HD_FUNC Pt combine(int s, const Pt* P, const Dt* w) {
    Pt v = P[s - p] * w[0];
    for (int i = 1; i <= p; i++) v += P[s - p + i] * w[i];
    return v;
}

template <typename Dt, typename Pt, int p>
struct Curves : gcurval::CurvesBase<Dt, Pt> {
    using PointType = Pt;
    int* knots_offset = nullptr;
    Dt* knot_vector = nullptr;
    Pt* ctrl_ptrs = nullptr;
    GCURVAL_HD_FUNC constexpr Curves(int n_curve = 0, int* ko = nullptr, Dt* kv = nullptr, Pt* cp = nullptr)
        : gcurval::CurvesBase<Dt, Pt>(n_curve), knots_offset(ko), knot_vector(kv), ctrl_ptrs(cp) {}
};

template <typename Dt>
struct Scratch {
    Dt* N;
    Dt* left;
    Dt* right;
};

// 3 (p + 1) values per thread, indexed by the thread of the block on the GPU.
template <typename Dt, int p>
struct ScratchAllocator {
    static constexpr int kPerThread = 3 * (p + 1);
    GCURVAL_HD_FUNC Scratch<Dt> operator()(uint8_t* buffer) const {
#if defined(__HIP_DEVICE_COMPILE__)
        Dt* row = (Dt*)buffer + kPerThread * threadIdx.x;
#else
        Dt* row = (Dt*)buffer;
#endif
        return Scratch<Dt>{row, row + (p + 1), row + 2 * (p + 1)};
    }
    static int size(int block_dim = 1) { return block_dim * kPerThread * (int)sizeof(Dt); }
};

template <typename Dt, typename Pt>
struct Pack : gcurval::ParaPackBase<Dt> {
    int span;
    Dt* knot_vector;
    Pt* ctrl_ptrs;
    GCURVAL_HD_FUNC Pack() = default;
    GCURVAL_HD_FUNC constexpr Pack(Dt begin, Dt end, int cid, int sid, int span_ = 0, Dt* kv = nullptr,
                                   Pt* cp = nullptr)
        : gcurval::ParaPackBase<Dt>{begin, end, cid, sid}, span(span_), knot_vector(kv), ctrl_ptrs(cp) {}
};

template <typename Dt, typename Pt, int p>
struct Parser : gcurval::ParaPackParserBase<Dt, Pack<Dt, Pt>> {
    GCURVAL_HD_FUNC Pack<Dt, Pt> operator()(const Curves<Dt, Pt, p>& c, int cid, int sid,
                                            bool need_interval = false) const {
        const int ko = c.knots_offset[cid];
        Dt* T = c.knot_vector + ko;
        Pt* P = c.ctrl_ptrs + ko - (p + 1) * cid;
        const int s = sid + p;
        if (need_interval) return Pack<Dt, Pt>{T[s], T[s + 1], cid, sid, s, T, P};
        return Pack<Dt, Pt>{Dt(0), Dt(0), cid, sid, s, T, P};
    }
};

// Global knot-interval range of curve cid: n_knots - 2p - 1 intervals per curve.
template <typename Dt, typename Pt, int p>
struct Locator {
    GCURVAL_HD_FUNC constexpr gcurval::CurveIntervalPos operator()(const Curves<Dt, Pt, p>& c, int cid) const {
        return gcurval::CurveIntervalPos{c.knots_offset[cid] - (2 * p + 1) * cid,
                                         c.knots_offset[cid + 1] - (2 * p + 1) * (cid + 1)};
    }
};

template <typename Dt, typename Pt, int p>
struct Point {
    GCURVAL_HD_FUNC Pt operator()(Dt u, const Pack<Dt, Pt>& k, Scratch<Dt>& w) const {
        basis(u, k.span, p, k.knot_vector, w.N, w.left, w.right);
        return combine<p>(k.span, k.ctrl_ptrs, w.N);
    }
};

template <typename Dt, typename Pt, int p>
struct Derivative {
    GCURVAL_HD_FUNC Pt operator()(Dt u, const Pack<Dt, Pt>& k, Scratch<Dt>& w) const {
        basis_derivative<p>(u, k.span, k.knot_vector, w.N, w.left, w.right);
        return combine<p>(k.span, k.ctrl_ptrs, w.N);
    }
};

template <typename Dt, typename Pt, int p = 3>
struct Traits : gcurval::CurveTraitsBase<Curves<Dt, Pt, p>> {
    using ParaPack = Pack<Dt, Pt>;
    using ParaPackParser = Parser<Dt, Pt, p>;
    using SharedMemAllocator = ScratchAllocator<Dt, p>;
    using CurveLocator = Locator<Dt, Pt, p>;
    using GetPoint = Point<Dt, Pt, p>;
    using GetDerivative = Derivative<Dt, Pt, p>;
    using CalculateFunction = gcurval::CalculateFunctionGaussLegendreQuadratureImpl<Dt, ParaPack, GetDerivative>;
};

#if defined(__HIP__)

template <typename Dt, typename Pt, int p = 3, int max_recursive_depth = gcurval::MaxRecursiveDepth<Dt>::value,
          int n_subdivided = 64>
class ArclengthCalculateGpu
    : public gcurval::ArclengthCalculateGpu<Curves<Dt, Pt, p>, Traits<Dt, Pt, p>, max_recursive_depth, n_subdivided> {
    using Base = gcurval::ArclengthCalculateGpu<Curves<Dt, Pt, p>, Traits<Dt, Pt, p>, max_recursive_depth, n_subdivided>;

public:
    explicit ArclengthCalculateGpu(int n_curve, int* knots_offset, Dt* knot_vector, Pt* ctrl_ptrs)
        : Base(n_curve, knots_offset[n_curve] - (2 * p + 1) * n_curve) {
        const int n_knots = knots_offset[n_curve];
        const int n_ctrl = n_knots - (p + 1) * n_curve;
        gcurval::gMalloc(&this->d_curve.knot_vector, sizeof(Dt) * n_knots);
        gcurval::gMalloc(&this->d_curve.knots_offset, sizeof(int) * (n_curve + 1));
        gcurval::gMalloc(&this->d_curve.ctrl_ptrs, sizeof(Pt) * n_ctrl);
        gcurval::copyToDevice(this->d_curve.knot_vector, knot_vector, n_knots);
        gcurval::copyToDevice(this->d_curve.knots_offset, knots_offset, n_curve + 1);
        gcurval::copyToDevice(this->d_curve.ctrl_ptrs, ctrl_ptrs, n_ctrl);
    }
    ~ArclengthCalculateGpu() {
        gcurval::gFree_s(this->d_curve.knot_vector);
        gcurval::gFree_s(this->d_curve.knots_offset);
        gcurval::gFree_s(this->d_curve.ctrl_ptrs);
    }
};

#endif  // __HIP__

template <typename Dt, typename Pt, int p = 3, int max_recursive_depth = gcurval::MaxRecursiveDepth<Dt>::value>
class ArclengthCalculateCpu
    : public gcurval::ArclengthCalculateCpu<Curves<Dt, Pt, p>, Traits<Dt, Pt, p>, max_recursive_depth> {
    using Base = gcurval::ArclengthCalculateCpu<Curves<Dt, Pt, p>, Traits<Dt, Pt, p>, max_recursive_depth>;

public:
    explicit ArclengthCalculateCpu(int n_curve, int* knots_offset, Dt* knot_vector, Pt* ctrl_ptrs)
        : Base(n_curve, knots_offset[n_curve] - (2 * p + 1) * n_curve) {
        const int n_knots = knots_offset[n_curve];
        const int n_ctrl = n_knots - (p + 1) * n_curve;
        this->curve.knot_vector = new Dt[n_knots];
        this->curve.knots_offset = new int[n_curve + 1];
        this->curve.ctrl_ptrs = new Pt[n_ctrl];
        std::copy(knot_vector, knot_vector + n_knots, this->curve.knot_vector);
        std::copy(knots_offset, knots_offset + (n_curve + 1), this->curve.knots_offset);
        std::copy(ctrl_ptrs, ctrl_ptrs + n_ctrl, this->curve.ctrl_ptrs);
    }
    ~ArclengthCalculateCpu() {
        gcurval::freeArray_s(this->curve.knot_vector);
        gcurval::freeArray_s(this->curve.knots_offset);
        gcurval::freeArray_s(this->curve.ctrl_ptrs);
    }
};

}  // namespace plug_bspline
