// Built-in curve families, evaluated in interval-local parameter x in [0, 1]:
//   * ellipse  C(t) = (a cos t, b sin t), t = 2 pi x, one knot interval per
//     curve (samples/sample1_ellipse/ellipse.h:5-97);
//   * clamped B-splines of degree 2..5 in 2-D or 3-D
//     (samples/sample2_bspline/BSpline.h:26-96, structs.h:6-185), whose
//     derivative on each knot span is converted once per interval into a
//     power-basis polynomial of degree p-1 in x, so a quadrature node costs a
//     Horner evaluation, a norm and a square root, with no divisions.
// Speeds are |dC/dx| = h |C'(u)| for u = u0 + h x, so integrals over x are arc
// lengths in the curve's (optionally normalised) units.
#pragma once

#include <stdint.h>
#include <math.h>
#include <type_traits>
#include <utility>

#if defined(__HIPCC__)
#include "device.h"
#endif

#ifndef GCURVAL_HD
#if defined(__HIPCC__)
#define GCURVAL_HD __host__ __device__
#else
#define GCURVAL_HD
#endif
#endif

namespace gcurval {
namespace dev {

constexpr double kTwoPi = 6.283185307179586476925286766559;

// Coefficient block stride per interval (p*D <= 15 used).
constexpr int kCoefStride = 16;

#if defined(__HIPCC__)
// 1 / v for finite nonzero normal v, within 1 ulp: v_rcp_f64 refined by two
// Newton steps (the IEEE-exact division adds scaling and fix-up steps that
// start values, scales and Newton updates do not need).
__device__ __forceinline__ double rcp_f64(double v) {
    double r = __builtin_amdgcn_rcp(v);
    double e = fma(-v, r, 1.0);
    r = fma(r, e, r);
    e = fma(-v, r, 1.0);
    return fma(r, e, r);
}
#endif

// 1 / x for the span polynomials: the refined hardware reciprocal on the GPU,
// IEEE division on the host (the two agree to 1 ulp).
GCURVAL_HD inline double span_rcp(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return rcp_f64(x);
#else
    return 1.0 / x;
#endif
}

// ---------------------------------------------------------------------------
// B-spline span -> power-basis derivative polynomial (double precision).
//
// For span s (knot T[s] <= u < T[s+1], h = T[s+1]-T[s]) the derivative curve is
// a degree q = p-1 B-spline with control points
//   Q_m = p (P_{m+1} - P_m) / (T[m+p+1] - T[m+1]),   m = s-p .. s-1,
// evaluated by de Boor's algorithm on the knot vector without its first and
// last knot. Running de Boor's recurrence on polynomials in x (alpha is linear
// in x) yields dC/du on the span as sum_j d_j x^j; multiplying by h gives
// dC/dx. Output layout: out[d*p + j] = coefficient of x^j of component d.
// Returns 0 on success, 1 if the span is degenerate (h <= 0; coefficients 0).
// ---------------------------------------------------------------------------
template <int P, int D>
GCURVAL_HD inline int span_poly(int s, const double* T, const double* Pc, int pstride, double scale_inv,
                                double* out, double* u0_out, double* h_out) {
    constexpr int q = P - 1;
    const double ts = T[s], h = T[s + 1] - T[s];
    if (u0_out) *u0_out = ts;
    if (h_out) *h_out = h;
    if (!(h > 0.0)) {
#pragma unroll
        for (int i = 0; i < P * D; i++) out[i] = 0.0;
        return 1;
    }
    // knot-only quantities: derivative-point factors and the linear alphas
    double fq[q + 1];
#pragma unroll
    for (int j = 0; j <= q; j++) {
        const int m = s - P + j;
        const double den = T[m + P + 1] - T[m + 1];
        fq[j] = den > 0.0 ? double(P) * span_rcp(den) : 0.0;
    }
    double a0[q + 1][q + 1], a1[q + 1][q + 1];
#pragma unroll
    for (int r = 1; r <= q; r++) {
#pragma unroll
        for (int j = q; j >= r; j--) {
            const int i = s - q + j;
            const double den = T[i + q - r + 1] - T[i];
            const double rd = den > 0.0 ? span_rcp(den) : 0.0;
            a0[r][j] = (ts - T[i]) * rd;
            a1[r][j] = h * rd;
        }
    }
    const double f = h * scale_inv;
#pragma unroll
    for (int c = 0; c < D; c++) {
        // d[j][e]: coefficient of x^e of the j-th de Boor point (component c)
        double d[q + 1][q + 1];
#pragma unroll
        for (int j = 0; j <= q; j++) {
            const int m = s - P + j;
#pragma unroll
            for (int e = 0; e <= q; e++) d[j][e] = 0.0;
            d[j][0] = fq[j] * (Pc[(m + 1) * pstride + c] - Pc[m * pstride + c]);
        }
        // step r raises the degree of each point by one: its inputs have
        // degree r - 1, so the coefficients above x^r stay zero and are not
        // touched (the same values as the full loops, with half the work)
#pragma unroll
        for (int r = 1; r <= q; r++) {
#pragma unroll
            for (int j = q; j >= r; j--) {
                // d_j <- (1 - a0 - a1 x) d_{j-1} + (a0 + a1 x) d_j
                double nw[q + 1];
#pragma unroll
                for (int e = 0; e < r; e++) nw[e] = (1.0 - a0[r][j]) * d[j - 1][e] + a0[r][j] * d[j][e];
                nw[r] = 0.0;
#pragma unroll
                for (int e = 0; e < r; e++) nw[e + 1] += a1[r][j] * (d[j][e] - d[j - 1][e]);
#pragma unroll
                for (int e = 0; e <= r; e++) d[j][e] = nw[e];
            }
        }
#pragma unroll
        for (int e = 0; e <= q; e++) out[c * P + e] = d[q][e] * f;
    }
    return 0;
}

// Runtime (degree, dimension) dispatch onto the fully unrolled span_poly.
GCURVAL_HD inline int bspline_span_poly(int p, int D, int s, const double* T, const double* P, int pstride,
                                        double scale_inv, double* out, double* u0_out, double* h_out) {
#define GCV_SPAN_CASE(PP, DD) \
    case 10 * PP + DD:        \
        return span_poly<PP, DD>(s, T, P, pstride, scale_inv, out, u0_out, h_out);
    switch (10 * p + D) {
        GCV_SPAN_CASE(1, 2)
        GCV_SPAN_CASE(1, 3)
        GCV_SPAN_CASE(2, 2)
        GCV_SPAN_CASE(2, 3)
        GCV_SPAN_CASE(3, 2)
        GCV_SPAN_CASE(3, 3)
        GCV_SPAN_CASE(4, 2)
        GCV_SPAN_CASE(4, 3)
        GCV_SPAN_CASE(5, 2)
        GCV_SPAN_CASE(5, 3)
        default:
            return 2;
    }
#undef GCV_SPAN_CASE
}

#if defined(__HIPCC__)

// sqrt(v) for v >= 0, within 1 ulp for v >= 2^-1000 (0 below): v_rsq_f64
// refined by one coupled Goldschmidt step on (g, h) ~ (sqrt v, 1/(2 sqrt v))
// and one Newton correction g += (v - g^2) h with the residual from an FMA.
// Speeds squared below 2^-1000 contribute nothing to an arc length, so the
// denormal scaling of a general square root is left out.
__device__ __forceinline__ double sqrt_pos(double v) {
    const double y = __builtin_amdgcn_rsq(v);
    double g = v * y, h = 0.5 * y;
    const double r = fma(-h, g, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    g = fma(fma(-g, g, v), h, g);
    return v >= 0x1p-1000 ? g : 0.0;
}

// ---------------------------------------------------------------------------
// sin^2(pi y) for y = 2x + shift, x in [0, 1] and shift 0 or -1/2, from one
// polynomial: q = rint(y) and e = y - q in [-1/2, 1/2] are exact, and with
// h = e^2 in [0, 1/4], sin^2(pi y) = sin^2(pi e) = h V(h) keeps full relative
// accuracy near its zeros. shift 0 gives sin^2(2 pi x), shift -1/2 gives
// cos^2(2 pi x). FP32: V is a degree-5 weighted least-squares fit (1.9e-7
// relative for sin^2, 2.3e-7 for cos^2 in FP32 evaluation); FP64: degree 9,
// the truncated Chebyshev expansion in powers of h (about 2 ulp).
// Derivation: tests/golden/fit_sincos.py.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float sin2_2pi(float x, float shift) {
    const float y = fmaf(2.0f, x, shift);
    const float e = y - rintf(y);
    const float h = e * e;
    float v = -3.336219310760498f;
    v = fmaf(v, h, 13.037117004394531f);
    v = fmaf(v, h, -30.098554611206055f);
    v = fmaf(v, h, 42.7269287109375f);
    v = fmaf(v, h, -32.469661712646484f);
    v = fmaf(v, h, 9.869604110717773f);
    return h * v;
}
__device__ __forceinline__ double sin2_2pi(double x, double shift) {
    const double y = fma(2.0, x, shift);
    const double e = y - rint(y);
    const double h = e * e;
    double v = -0.0016981312977267715;
    v = fma(v, h, 0.0180912904683608);
    v = fma(v, h, -0.1409737898673586);
    v = fma(v, h, 0.8571901937854678);
    v = fma(v, h, -3.9517676198423173);
    v = fma(v, h, 13.213128353891875);
    v = fma(v, h, -30.122320684487374);
    v = fma(v, h, 42.72840860331856);
    v = fma(v, h, -32.469697011333935);
    v = fma(v, h, 9.869604401089358);
    return h * v;
}

// sin^2(pi y) as sin2_2pi(double) with V of degree 7 (the Chebyshev expansion
// truncated after T_7, 1.9e-12 relative), for the final FP64 Newton step of
// FP32 builds, whose speeds need about 1e-10 relative accuracy (their
// quadrature carries the FP32 table's 1e-7).
__device__ __forceinline__ double sin2_2pi_polish(double x, double shift) {
    const double y = fma(2.0, x, shift);
    const double e = y - rint(y);
    const double h = e * e;
    double v = -0.1238973981824048;
    v = fma(v, h, 0.8503903438737084);
    v = fma(v, h, -3.9503437263639913);
    v = fma(v, h, 13.21296251345923);
    v = fma(v, h, -30.12231017436645);
    v = fma(v, h, 42.72840827584517);
    v = fma(v, h, -32.46969700744408);
    v = fma(v, h, 9.869604401081775);
    return h * v;
}

// sqrt(v) for v >= 0 to about 2^-49 relative (0 below 2^-1000): v_rsq_f64
// (about 2^-24) refined by one Newton correction g += (v - g^2) / (2 g) with
// the residual from an FMA; for the final FP64 Newton step of FP32 builds.
__device__ __forceinline__ double sqrt_polish(double v) {
    const double y = __builtin_amdgcn_rsq(v);
    const double g = v * y;
    const double g1 = fma(fma(-g, g, v), 0.5 * y, g);
    return v >= 0x1p-1000 ? g1 : 0.0;
}

// ---------------------------------------------------------------------------
// Ellipse evaluator. With A = 2 pi a', B = 2 pi b' and a', b' the optionally
// normalised semi-axes, speed(x)^2 = A^2 sin^2(2 pi x) + B^2 cos^2(2 pi x)
// = base + D S(x), base = min(A^2, B^2), D = |A^2 - B^2| and S the squared
// sine (A >= B) or cosine (A < B) that vanishes where the speed is smallest:
// both terms are non-negative, so the sum keeps full relative accuracy.
// ---------------------------------------------------------------------------
template <typename Dt>
struct EllipseEval {
    static constexpr bool kOneInterval = true;  // one knot interval per curve
    const double* a;
    const double* b;
    int normalize;

    struct Ctx {
        Dt base, D, shift;
    };

    __device__ __forceinline__ double scale(int i) const {
        // the loads are unconditional (ev.load reads the same two values), so
        // they issue with the caller's other per-interval loads
        const double ai = a[i], bi = b[i];
        const double s = fmax(fabs(ai), fabs(bi));
        return normalize && s > 0.0 ? s : 1.0;
    }
    __device__ __forceinline__ Ctx load(int i) const { return load(i, rcp_f64(scale(i))); }
    // with the reciprocal of scale(i) already at hand
    __device__ __forceinline__ Ctx load(int i, double inv_scale) const { return ctx_of(a[i], b[i], inv_scale); }
    __device__ __forceinline__ static Ctx ctx_of(double ai, double bi, double inv_scale) {
        const double A = kTwoPi * ai * inv_scale, B = kTwoPi * bi * inv_scale;
        const double A2 = A * A, B2 = B * B;
        const bool sine = A2 >= B2;
        return Ctx{Dt(sine ? B2 : A2), Dt(sine ? A2 - B2 : B2 - A2), Dt(sine ? 0.0 : -0.5)};
    }
    __device__ __forceinline__ double scale_of(double ai, double bi) const {
        const double s = fmax(fabs(ai), fabs(bi));
        return normalize && s > 0.0 ? s : 1.0;
    }
    // context and scale of wave-uniform interval i through scalar loads
    __device__ __forceinline__ void load_s(int i, Ctx& c, double& sc) const {
        const double ai = dev::load_s(a + i), bi = dev::load_s(b + i);
        sc = scale_of(ai, bi);
        c = ctx_of(ai, bi, rcp_f64(sc));
    }
    __device__ __forceinline__ float speed(const Ctx& c, float x) const {
        return __builtin_amdgcn_sqrtf(fmaf(float(c.D), sin2_2pi(x, float(c.shift)), float(c.base)));
    }
    __device__ __forceinline__ double speed(const Ctx& c, double x) const {
        return sqrt_pos(fma(double(c.D), sin2_2pi(x, double(c.shift)), double(c.base)));
    }
    // d speed / dx = D pi sin(2 pi y) / speed with y = 2x + shift (speed v > 0)
    __device__ __forceinline__ float speed_slope(const Ctx& c, float x, float v) const {
        return float(c.D) * 3.14159265358979f * sinpif(2.0f * fmaf(2.0f, x, float(c.shift))) / v;
    }
    __device__ __forceinline__ double speed_slope(const Ctx& c, double x, double v) const {
        return double(c.D) * 3.141592653589793 * sinpi(2.0 * fma(2.0, x, double(c.shift))) / v;
    }
    // (min(a', b') / max(a', b'))^2
    __device__ __forceinline__ double axis_ratio2(const Ctx& c) const {
        const double lo = double(c.base), hi = double(c.base) + double(c.D);
        return hi > 0.0 ? lo / hi : 1.0;
    }
    // interval-local x -> curve parameter t
    __device__ __forceinline__ double to_u(int, double x) const { return kTwoPi * x; }
};

// FP64 speed of the final Newton step of FP32 ellipse builds: the degree-7
// squared sine and the one-correction square root (speed to about 1e-12
// relative, 15 fewer FP64 instructions per sample than the full evaluator).
struct EllipsePolishEval : EllipseEval<double> {
    __device__ __forceinline__ double speed(const Ctx& c, double x) const {
        return sqrt_polish(fma(double(c.D), sin2_2pi_polish(x, double(c.shift)), double(c.base)));
    }
};

// ---------------------------------------------------------------------------
// B-spline evaluator for one (degree P, dimension D) class; coefficient
// blocks are precomputed per interval (cstride values each, stored as Ct and
// evaluated in Dt).
// ---------------------------------------------------------------------------
template <typename Dt, int P, int D, typename Ct = Dt>
struct BSplineEval {
    const Ct* coef;           // [n_int * cstride]
    int cstride;              // values per interval block (>= P * D)
    const double* int_scale;  // per interval normalisation scale (may be null -> 1)
    const double* u0;         // per interval knot T[s]
    const double* h;          // per interval span width

    struct Ctx {
        Dt c[P * D];
    };

    __device__ __forceinline__ double scale(int i) const { return int_scale ? int_scale[i] : 1.0; }
    __device__ __forceinline__ Ctx load(int i) const {
        Ctx k;
        const Ct* src = coef + (int64_t)i * cstride;
#pragma unroll
        for (int j = 0; j < P * D; j++) k.c[j] = Dt(src[j]);
        return k;
    }
    // context and scale of wave-uniform interval i: the coefficients through
    // vector loads (a block of up to 15 values in scalar registers, twice
    // over for the prefetch, spilled them), the scale through a scalar load
    __device__ __forceinline__ void load_s(int i, Ctx& k, double& sc) const {
        k = load(i);
        sc = int_scale ? dev::load_s(int_scale + i) : 1.0;
    }
    // coefficients are stored pre-normalised
    __device__ __forceinline__ Ctx load(int i, double) const { return load(i); }
    // the context of an evaluator over the same float coefficients, widened
    // (exact: each value is the stored float); only for Ct = float
    template <class C32, typename C = Ct,
              typename = typename std::enable_if<std::is_same<C, float>::value &&
                                                 sizeof(((C32*)nullptr)->c) == sizeof(float) * P * D>::type>
    __device__ __forceinline__ static Ctx widen(const C32& c32) {
        Ctx k;
#pragma unroll
        for (int j = 0; j < P * D; j++) k.c[j] = Dt(c32.c[j]);
        return k;
    }
    template <typename T>
    __device__ __forceinline__ T speed(const Ctx& k, T x) const {
        T s = T(0);
#pragma unroll
        for (int d = 0; d < D; d++) {
            T v = T(k.c[d * P + P - 1]);
#pragma unroll
            for (int j = P - 2; j >= 0; j--) v = v * x + T(k.c[d * P + j]);
            s += v * v;
        }
        return sq(s);
    }
    // speed and its slope d|C'|/dx = (C' . C'') / |C'| (0 where the speed
    // vanishes), C' and C'' from one Horner pass per component
    template <typename T>
    __device__ __forceinline__ void speed_and_slope(const Ctx& k, T x, T& v, T& vp) const {
        T s = T(0), sd = T(0);
#pragma unroll
        for (int d = 0; d < D; d++) {
            T c = T(k.c[d * P + P - 1]), dc = T(0);
#pragma unroll
            for (int j = P - 2; j >= 0; j--) {
                dc = dc * x + c;
                c = c * x + T(k.c[d * P + j]);
            }
            s += c * c;
            sd += c * dc;
        }
        v = sq(s);
        vp = v > T(0) ? sd * rcp(v) : T(0);
    }
    __device__ __forceinline__ static float rcp(float v) { return __builtin_amdgcn_rcpf(v); }
    __device__ __forceinline__ static double rcp(double v) { return rcp_f64(v); }
    __device__ __forceinline__ static float sq(float v) { return __builtin_amdgcn_sqrtf(v); }
    __device__ __forceinline__ static double sq(double v) { return sqrt_pos(v); }
    __device__ __forceinline__ double to_u(int i, double x) const { return u0[i] + h[i] * x; }
};

#endif  // __HIPCC__

}  // namespace dev
}  // namespace gcurval
