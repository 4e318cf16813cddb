// Plug-in vocabulary of the header-only template surface: the types a curve
// family defines or derives from, with the names, members and call
// signatures of gcurval's own (gcurval/base/structs.inl.h:8-77,
// gcurval/base/functors.inl.h:9-342, gcurval/base/traits.inl.h:11-30,
// gcurval/base/util.h:14-104) so existing plug-ins compile unchanged apart
// from the CUDA/Eigen includes. Compiled by hipcc in the user's translation
// unit (device code targets gfx950); a host compiler (g++) may include it too,
// for the CPU engine alone (arclength_calculate_cpu.h).
#pragma once

#if defined(__HIP__)
#include <hip/hip_runtime.h>
#endif

#include <cmath>
#include <cstdint>
#include <type_traits>
#include <utility>

#ifndef GCURVAL_HD_FUNC
#if defined(__HIP__)
#define GCURVAL_HD_FUNC __host__ __device__
#else
#define GCURVAL_HD_FUNC
#endif
#endif

namespace gcurval {

// Where a result lives and what it holds (gcurval/base/util.h:71-77).
enum GcurvalEnum { gcurvalDevice, gcurvalHost, gcurvalReverseReturnUs, gcurvalReverseReturnPoints };

// Root sub-intervals per knot interval, the error bound and the depth cap of
// the adaptive test, with the reference's override macros GCURVAL_N_SUBDIVIDED,
// GCURVAL_MAX_ERROR and GCURVAL_MAX_RECURSIVE_DEPTH (gcurval/config.h:17-43).
#ifdef GCURVAL_N_SUBDIVIDED
constexpr int N_SUBDIVIDED = GCURVAL_N_SUBDIVIDED;
#else
constexpr int N_SUBDIVIDED = 64;
#endif
#ifdef GCURVAL_MAX_ERROR
constexpr float MAX_ERROR_FLOAT = GCURVAL_MAX_ERROR;
constexpr double MAX_ERROR_DOUBLE = GCURVAL_MAX_ERROR;
#else
constexpr float MAX_ERROR_FLOAT = 1e-6f;
constexpr double MAX_ERROR_DOUBLE = 1e-14;
#endif
template <typename Dt>
struct MaxRecursiveDepth {
#ifdef GCURVAL_MAX_RECURSIVE_DEPTH
    static constexpr int value = GCURVAL_MAX_RECURSIVE_DEPTH;
#else
    static constexpr int value = sizeof(Dt) <= 4 ? 14 : 23;
#endif
};
// Default absolute error bound of an engine in Dt (MAX_ERROR_FLOAT or
// MAX_ERROR_DOUBLE as a double).
template <typename Dt>
constexpr double default_max_error() {
    return sizeof(Dt) <= 4 ? (double)MAX_ERROR_FLOAT : MAX_ERROR_DOUBLE;
}

constexpr double kPi = 3.141592653589793238462643383279502884;

template <typename T>
GCURVAL_HD_FUNC inline T abs(T v) {
    return v < T(0) ? -v : v;
}

template <typename T>
inline void freeArray_s(T*& p) {
    if (p) {
        delete[] p;
        p = nullptr;
    }
}

// ---------------------------------------------------------------------------
// Fixed-size vectors standing in for the Eigen types plug-ins return
// (Vector2X / Vector3X / VectorX): element access, arithmetic with vectors
// and scalars, norm(), Zero(), rows().
// ---------------------------------------------------------------------------
template <typename T, int N>
struct VectorN {
    T v[N];

    VectorN() = default;  // VectorN{} is the zero vector
    template <typename... A, typename = std::enable_if_t<sizeof...(A) == N && (N > 1)>>
    GCURVAL_HD_FUNC VectorN(A... a) : v{T(a)...} {}
    GCURVAL_HD_FUNC explicit VectorN(T s) {
        for (int i = 0; i < N; i++) v[i] = s;
    }
    GCURVAL_HD_FUNC static VectorN Zero() { return VectorN(T(0)); }
    GCURVAL_HD_FUNC static constexpr int rows() { return N; }
    GCURVAL_HD_FUNC T& operator[](int i) { return v[i]; }
    GCURVAL_HD_FUNC const T& operator[](int i) const { return v[i]; }
    GCURVAL_HD_FUNC T& operator()(int i) { return v[i]; }
    GCURVAL_HD_FUNC const T& operator()(int i) const { return v[i]; }
    GCURVAL_HD_FUNC T x() const { return v[0]; }
    GCURVAL_HD_FUNC T y() const { return v[1]; }
    GCURVAL_HD_FUNC T squaredNorm() const {
        T s = T(0);
        for (int i = 0; i < N; i++) s += v[i] * v[i];
        return s;
    }
    GCURVAL_HD_FUNC T norm() const { return std::sqrt(squaredNorm()); }
    GCURVAL_HD_FUNC T dot(const VectorN& o) const {
        T s = T(0);
        for (int i = 0; i < N; i++) s += v[i] * o.v[i];
        return s;
    }
    GCURVAL_HD_FUNC VectorN& operator+=(const VectorN& o) {
        for (int i = 0; i < N; i++) v[i] += o.v[i];
        return *this;
    }
    GCURVAL_HD_FUNC VectorN& operator-=(const VectorN& o) {
        for (int i = 0; i < N; i++) v[i] -= o.v[i];
        return *this;
    }
    template <typename S, typename = std::enable_if_t<std::is_arithmetic<S>::value>>
    GCURVAL_HD_FUNC VectorN& operator*=(S s) {
        for (int i = 0; i < N; i++) v[i] *= T(s);
        return *this;
    }
    template <typename S, typename = std::enable_if_t<std::is_arithmetic<S>::value>>
    GCURVAL_HD_FUNC VectorN& operator/=(S s) {
        for (int i = 0; i < N; i++) v[i] /= T(s);
        return *this;
    }
};
template <typename T, int N>
GCURVAL_HD_FUNC inline VectorN<T, N> operator+(VectorN<T, N> a, const VectorN<T, N>& b) {
    return a += b;
}
template <typename T, int N>
GCURVAL_HD_FUNC inline VectorN<T, N> operator-(VectorN<T, N> a, const VectorN<T, N>& b) {
    return a -= b;
}
template <typename T, int N>
GCURVAL_HD_FUNC inline VectorN<T, N> operator-(VectorN<T, N> a) {
    for (int i = 0; i < N; i++) a.v[i] = -a.v[i];
    return a;
}
template <typename T, int N, typename S, typename = std::enable_if_t<std::is_arithmetic<S>::value>>
GCURVAL_HD_FUNC inline VectorN<T, N> operator*(VectorN<T, N> a, S s) {
    return a *= s;
}
template <typename T, int N, typename S, typename = std::enable_if_t<std::is_arithmetic<S>::value>>
GCURVAL_HD_FUNC inline VectorN<T, N> operator*(S s, VectorN<T, N> a) {
    return a *= s;
}
template <typename T, int N, typename S, typename = std::enable_if_t<std::is_arithmetic<S>::value>>
GCURVAL_HD_FUNC inline VectorN<T, N> operator/(VectorN<T, N> a, S s) {
    return a /= s;
}

template <typename T>
using Vector2X = VectorN<T, 2>;
template <typename T>
using Vector3X = VectorN<T, 3>;
template <typename T, int N>
using VectorX = VectorN<T, N>;

// ---------------------------------------------------------------------------
// Curve batches, parameter packs, interval ranges, results
// (gcurval/base/structs.inl.h:8-77).
// ---------------------------------------------------------------------------
template <typename Dt>
struct ForwardSpaceBase {
    Dt* begin = nullptr;
    Dt* end = nullptr;
    int* cid = nullptr;
    int* sid = nullptr;
};

template <typename Dt>
struct BackwardSpaceBase {
    int* ptid = nullptr;
    Dt* pal = nullptr;
};

template <typename Dt, typename Pt>
struct CurvesBase {
    using DataType = Dt;
    using PointType = Pt;
    int n_curve = 0;
    GCURVAL_HD_FUNC constexpr CurvesBase(int n = 0) : n_curve(n) {}
};

// Knot interval [begin, end] of curve cid, its segment index sid, and
// whatever else a family's parser adds.
template <typename Dt>
struct ParaPackBase {
    Dt begin;
    Dt end;
    int cid;
    int sid;
};

// Global knot-interval range [begin, end) of one curve.
template <typename T = void>
struct CurveIntervalPos_ {
    int begin;
    int end;
};
using CurveIntervalPos = CurveIntervalPos_<void>;

// No per-thread scratch: GetDerivative / GetPoint take (u, pack).
struct NoSharedMemAllocator {
    static constexpr int size(int) { return 0; }
};

template <typename DataType, typename PointType>
struct ArclengthReverseResult {
    int n_points = 0;
    GcurvalEnum store_where = gcurvalDevice;
    GcurvalEnum store_what = gcurvalReverseReturnUs;
    union {
        DataType* us = nullptr;
        PointType* points;
    };
};

// ---------------------------------------------------------------------------
// Plug-in functor defaults (gcurval/base/functors.inl.h:9-37, 262-269).
// ---------------------------------------------------------------------------
template <typename Dt, typename PointType, typename Rt = ParaPackBase<Dt>>
struct ParaPackParserBase {
    using result_type = Rt;
    GCURVAL_HD_FUNC Rt operator()(const CurvesBase<Dt, PointType>&, int cid, int sid,
                                  bool need_interval = false) const {
        (void)need_interval;
        return Rt{Dt(0), Dt(0), cid, sid};
    }
};

// Kept for traits that name it; the engine's task layout is internal.
template <typename Dt>
struct ParaPackSaverBase {};

// The fifth template parameter of the GPU engine
// (gcurval/detail/cuda/arclength_calculate.h:13-90, the per-level task
// buffers of the reference's BFS). Accepted so that engine declarations keep
// their signature; the engine ignores it: its refinement tasks live in a
// wave-local LDS stack, so there are no per-level buffers to manage.
template <typename Dt, int max_recursive_depth, typename ForwardSpace = ForwardSpaceBase<Dt>,
          typename BackwardSpace = BackwardSpaceBase<Dt>>
struct BufferManagerBase {
    static constexpr int N_FSPACE = 2;
    static constexpr int N_BSPACE = max_recursive_depth + 1;
};

// One knot interval per curve: global interval range {cid, cid + 1}.
template <typename Curve>
struct CurveLocatorBase {
    GCURVAL_HD_FUNC constexpr CurveIntervalPos operator()(const Curve&, int cid) const {
        return CurveIntervalPos{cid, cid + 1};
    }
};

namespace detail {
// Length of a functor result: scalars as they are, vectors through norm().
template <typename Dt, typename R>
GCURVAL_HD_FUNC inline Dt magnitude(const R& r) {
    if constexpr (std::is_arithmetic<R>::value) {
        return Dt(r);
    } else {
        return Dt(r.norm());
    }
}
// Call f(u, pack) or f(u, pack, scratch) depending on what the plug-in accepts.
template <typename F, typename Dt, typename Pack, typename View>
GCURVAL_HD_FUNC inline auto call(const F& f, Dt u, const Pack& pack, View* view) {
    if constexpr (std::is_same<View, void>::value) {
        (void)view;
        return f(u, pack);
    } else {
        return f(u, pack, *view);
    }
}
}  // namespace detail

// ---------------------------------------------------------------------------
// Length rules over a parameter range [begin, end]
// (gcurval/base/functors.inl.h:39-257). GaussLegendreQuadrature integrates
// the norm of the derivative with the 3-point rule (the rule the sample
// families select); the others take curve points.
// ---------------------------------------------------------------------------
template <typename Dt>
struct GaussLegendreQuadrature {
    static constexpr int N_GAUSS = 3;
    template <typename GetDerivative, typename Pack, typename View = void>
    GCURVAL_HD_FUNC Dt eval(Dt begin, Dt end, const GetDerivative& d, const Pack& pack, View* view = nullptr) const {
        const Dt k = (end - begin) / Dt(2);
        const Dt b = (end + begin) / Dt(2);
        const Dt x1 = Dt(0.7745966692414834), w0 = Dt(0.8888888888888889), w1 = Dt(0.5555555555555556);
        Dt acc = w0 * detail::magnitude<Dt>(detail::call(d, b, pack, view));
        acc += w1 * detail::magnitude<Dt>(detail::call(d, b - k * x1, pack, view));
        acc += w1 * detail::magnitude<Dt>(detail::call(d, b + k * x1, pack, view));
        return acc * k;
    }
};

template <typename Dt>
struct PolylineMethod {
    template <typename GetPoint, typename Pack, typename View = void>
    GCURVAL_HD_FUNC Dt eval(Dt begin, Dt end, const GetPoint& p, const Pack& pack, View* view = nullptr) const {
        const auto a = detail::call(p, begin, pack, view);
        const auto b = detail::call(p, end, pack, view);
        return detail::magnitude<Dt>(b - a);
    }
};

// Simpson's rule on the speed of the parabola through the end and middle
// points (end-point speeds from one-sided differences, middle from the chord).
template <typename Dt>
struct SimpsonMethod {
    template <typename GetPoint, typename Pack, typename View = void>
    GCURVAL_HD_FUNC Dt eval(Dt begin, Dt end, const GetPoint& p, const Pack& pack, View* view = nullptr) const {
        const Dt mid = (begin + end) / Dt(2);
        const auto v0 = detail::call(p, begin, pack, view);
        const auto v1 = detail::call(p, mid, pack, view);
        const auto v2 = detail::call(p, end, pack, view);
        const Dt s0 = detail::magnitude<Dt>(Dt(-3) * v0 + Dt(4) * v1 - v2);
        const Dt s1 = detail::magnitude<Dt>(v2 - v0);
        const Dt s2 = detail::magnitude<Dt>(v0 - Dt(4) * v1 + Dt(3) * v2);
        return Dt(1.0 / 6.0) * (s0 + Dt(4) * s1 + s2);
    }
};

// Vincent-Forsey: Richardson extrapolation of the one- and two-chord lengths.
template <typename Dt>
struct VincentForseyRule {
    template <typename GetPoint, typename Pack, typename View = void>
    GCURVAL_HD_FUNC Dt eval(Dt begin, Dt end, const GetPoint& p, const Pack& pack, View* view = nullptr) const {
        const Dt mid = (begin + end) / Dt(2);
        const auto v0 = detail::call(p, begin, pack, view);
        const auto v1 = detail::call(p, mid, pack, view);
        const auto v2 = detail::call(p, end, pack, view);
        const Dt two = detail::magnitude<Dt>(v0 - v1) + detail::magnitude<Dt>(v2 - v1);
        return Dt(4.0 / 3.0) * two - Dt(1.0 / 3.0) * detail::magnitude<Dt>(v2 - v0);
    }
};

// Parameter-width "length" (debugging aid of the reference).
template <typename Dt>
struct ParameterDebugMethod {
    template <typename GetPoint, typename Pack, typename View = void>
    GCURVAL_HD_FUNC Dt eval(Dt begin, Dt end, const GetPoint&, const Pack&, View* = nullptr) const {
        return abs(end - begin);
    }
};

// CalculateFunction<Rule>Impl<Dt, ParaPack, GetP>: the length rule a traits
// class selects, called as f(begin, end, pack) or f(begin, end, pack, scratch).
template <template <typename> class Rule, typename Dt, typename ParaPack, typename GetP>
struct CalculateFunctionImpl {
    using Getter = GetP;
    GCURVAL_HD_FUNC Dt operator()(Dt begin, Dt end, const ParaPack& pack) const {
        return Rule<Dt>{}.eval(begin, end, GetP{}, pack, (void*)nullptr);
    }
    template <typename View>
    GCURVAL_HD_FUNC Dt operator()(Dt begin, Dt end, const ParaPack& pack, View& view) const {
        return Rule<Dt>{}.eval(begin, end, GetP{}, pack, &view);
    }
};
template <typename Dt, typename ParaPack, typename GetP>
using CalculateFunctionGaussLegendreQuadratureImpl = CalculateFunctionImpl<GaussLegendreQuadrature, Dt, ParaPack, GetP>;
template <typename Dt, typename ParaPack, typename GetP>
using CalculateFunctionPolylineMethodImpl = CalculateFunctionImpl<PolylineMethod, Dt, ParaPack, GetP>;
template <typename Dt, typename ParaPack, typename GetP>
using CalculateFunctionSimpsonMethodImpl = CalculateFunctionImpl<SimpsonMethod, Dt, ParaPack, GetP>;
template <typename Dt, typename ParaPack, typename GetP>
using CalculateFunctionVincentForseyRuleImpl = CalculateFunctionImpl<VincentForseyRule, Dt, ParaPack, GetP>;
template <typename Dt, typename ParaPack, typename GetP>
using CalculateFunctionParameterDebugMethodImpl = CalculateFunctionImpl<ParameterDebugMethod, Dt, ParaPack, GetP>;

// ---------------------------------------------------------------------------
// Traits a curve family derives from (gcurval/base/traits.inl.h:11-30).
// ---------------------------------------------------------------------------
template <typename Curve>
struct CurveTraitsBase {
    using DataType = typename Curve::DataType;
    using PointType = typename Curve::PointType;
    using ForwardSpace = ForwardSpaceBase<DataType>;
    using BackwardSpace = BackwardSpaceBase<DataType>;
    using ParaPackParser = ParaPackParserBase<DataType, PointType>;
    using ParaPackSaver = ParaPackSaverBase<DataType>;
    using ParaPack = ParaPackBase<DataType>;
    using SharedMemAllocator = NoSharedMemAllocator;
    using CurveLocator = CurveLocatorBase<Curve>;
    using GetPoint = void;
    using GetDerivative = void;
    using CalculateFunction = void;
};

}  // namespace gcurval
