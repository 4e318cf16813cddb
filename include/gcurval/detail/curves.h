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
#pragma unroll
        for (int r = 1; r <= q; r++) {
#pragma unroll
            for (int j = q; j >= r; j--) {
                // d_j <- (1 - a0 - a1 x) d_{j-1} + (a0 + a1 x) d_j
                double nw[q + 1];
#pragma unroll
                for (int e = 0; e <= q; e++) nw[e] = (1.0 - a0[r][j]) * d[j - 1][e] + a0[r][j] * d[j][e];
#pragma unroll
                for (int e = 0; e < q; e++) nw[e + 1] += a1[r][j] * (d[j][e] - d[j - 1][e]);
#pragma unroll
                for (int e = 0; e <= q; e++) d[j][e] = nw[e];
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
// sin^2(2 pi x) and cos^2(2 pi x) for x in [0, 1] from one polynomial: r = 4x
// and its nearest integer q are exact, f = r - q lies in [-1/2, 1/2], and with
// g = f^2 the squared quarter-period sine sin^2(pi f / 2) = g U(g) <= 1/2;
// sin^2(2 pi x) and cos^2(2 pi x) are g U(g) and 1 - g U(g), swapped when q
// is odd, so both keep full relative accuracy near their zeros.
// FP32: U is a degree-4 weighted least-squares fit (max relative error
// 1.2e-7 in FP32 evaluation); FP64: degree 7, the truncated Chebyshev
// expansion in powers of g (2.2e-16, as the 11-term Taylor series). Derivation:
// gcurval_amd/fit_sincos.py.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void sincos2_2pi(float x, float& s2, float& c2) {
    const float r = 4.0f * x;
    const float q = rintf(r);
    const float f = r - q;
    const float g = f * f;
    const float u = fmaf(fmaf(fmaf(fmaf(0.012307871133089066f, g, -0.1175343245267868f), g, 0.667619526386261f), g,
                              -2.029355764389038f),
                         g, 2.4674010276794434f);
    const float lo = g * u, hi = 1.0f - lo;
    const bool odd = ((int)q) & 1;
    s2 = odd ? hi : lo;
    c2 = odd ? lo : hi;
}
__device__ __forceinline__ void sincos2_2pi(double x, double& s2, double& c2) {
    const double r = 4.0 * x;
    const double q = rint(r);
    const double f = r - q;
    const double g = f * f;
    double u = -2.0832065862542846e-06;
    u = fma(u, g, 5.229143784023563e-05);
    u = fma(u, g, -0.0009647813283021856);
    u = fma(u, g, 0.012903445013289263);
    u = fma(u, g, -0.11766531513610962);
    u = fma(u, g, 0.6676313844259415);
    u = fma(u, g, -2.029356063208368);
    u = fma(u, g, 2.4674011002723395);
    const double lo = g * u, hi = 1.0 - lo;
    const bool odd = ((int)q) & 1;  // |q| <= 4 on the domain
    s2 = odd ? hi : lo;
    c2 = odd ? lo : hi;
}

// ---------------------------------------------------------------------------
// Ellipse evaluator. Ctx holds A^2 and B^2 with A = 2 pi a', B = 2 pi b' and
// a', b' the optionally normalised semi-axes, so
// speed(x) = sqrt(A^2 sin^2(2 pi x) + B^2 cos^2(2 pi x)).
// ---------------------------------------------------------------------------
template <typename Dt>
struct EllipseEval {
    const double* a;
    const double* b;
    int normalize;

    struct Ctx {
        Dt A2, B2;
    };

    __device__ __forceinline__ double scale(int i) const {
        if (!normalize) return 1.0;
        double s = fmax(fabs(a[i]), fabs(b[i]));
        return s > 0.0 ? s : 1.0;
    }
    __device__ __forceinline__ Ctx load(int i) const { return load(i, rcp_f64(scale(i))); }
    // with the reciprocal of scale(i) already at hand
    __device__ __forceinline__ Ctx load(int i, double inv_scale) const {
        const double A = kTwoPi * a[i] * inv_scale, B = kTwoPi * b[i] * inv_scale;
        return Ctx{Dt(A * A), Dt(B * B)};
    }
    __device__ __forceinline__ float speed(const Ctx& c, float x) const {
        float s2, c2;
        sincos2_2pi(x, s2, c2);
        return __builtin_amdgcn_sqrtf(fmaf(float(c.A2), s2, float(c.B2) * c2));
    }
    __device__ __forceinline__ double speed(const Ctx& c, double x) const {
        double s2, c2;
        sincos2_2pi(x, s2, c2);
        return sqrt_pos(fma(double(c.A2), s2, double(c.B2) * c2));
    }
    // interval-local x -> curve parameter t
    __device__ __forceinline__ double to_u(int, double x) const { return kTwoPi * x; }
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
    __device__ __forceinline__ static float sq(float v) { return __builtin_amdgcn_sqrtf(v); }
    __device__ __forceinline__ static double sq(double v) { return sqrt_pos(v); }
    __device__ __forceinline__ double to_u(int i, double x) const { return u0[i] + h[i] * x; }
};

#endif  // __HIPCC__

}  // namespace dev
}  // namespace gcurval
