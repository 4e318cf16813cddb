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
        fq[j] = den > 0.0 ? double(P) / den : 0.0;
    }
    double a0[q + 1][q + 1], a1[q + 1][q + 1];
#pragma unroll
    for (int r = 1; r <= q; r++) {
#pragma unroll
        for (int j = q; j >= r; j--) {
            const int i = s - q + j;
            const double den = T[i + q - r + 1] - T[i];
            a0[r][j] = den > 0.0 ? (ts - T[i]) / den : 0.0;
            a1[r][j] = den > 0.0 ? h / den : 0.0;
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

// ---------------------------------------------------------------------------
// |sin 2 pi x| and |cos 2 pi x| for x in [0, 1]: r = 4x and its nearest
// integer q are exact, f = r - q lies in [-1/2, 1/2], and
//   sin(pi f / 2) = f S(f^2),  cos(pi f / 2) = C(f^2),
// swapped when q is odd (signs are not needed: only squares are used).
// FP32: S, C are weighted least-squares fits on f^2 in [0, 1/4]
// (gcurval_amd/fit_sincos.py; max abs error 8.6e-8 over x in [0, 1]).
// FP64: S, C are the Taylor series of sin(pi f/2)/f and cos(pi f/2) to
// f^16 / f^18 (max abs error 1.1e-16).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void abs_sincos_2pi(float x, float& as, float& ac) {
    const float r = 4.0f * x;
    const float q = rintf(r);
    const float f = r - q;
    const float g = f * f;
    const float sp = f * fmaf(fmaf(fmaf(-0.004601657390594482f, g, 0.07968003302812576f), g, -0.6459634900093079f), g,
                              1.5707963705062866f);
    const float cp = fmaf(fmaf(fmaf(fmaf(0.0009033629903569818f, g, -0.020859936252236366f), g, 0.2536691725254059f), g,
                               -1.2337005138397217f), g, 1.0f);
    const bool odd = ((int)q) & 1;
    as = odd ? cp : sp;
    ac = odd ? sp : cp;
}
__device__ __forceinline__ void abs_sincos_2pi(double x, double& as, double& ac) {
    const double r = 4.0 * x;
    const double q = rint(r);
    const double f = r - q;
    const double g = f * f;
    double sp = 6.0669357311061955e-12;
    sp = fma(sp, g, -6.688035109811468e-10);
    sp = fma(sp, g, 5.692172921967927e-08);
    sp = fma(sp, g, -3.598843235212085e-06);
    sp = fma(sp, g, 0.00016044118478735983);
    sp = fma(sp, g, -0.004681754135318688);
    sp = fma(sp, g, 0.07969262624616705);
    sp = fma(sp, g, -0.6459640975062463);
    sp = fma(sp, g, 1.5707963267948966);
    sp *= f;
    double cp = -5.294400200734623e-13;
    cp = fma(cp, g, 6.565963114979473e-11);
    cp = fma(cp, g, -6.386603083791852e-09);
    cp = fma(cp, g, 4.710874778818172e-07);
    cp = fma(cp, g, -2.5202042373060607e-05);
    cp = fma(cp, g, 0.0009192602748394266);
    cp = fma(cp, g, -0.02086348076335296);
    cp = fma(cp, g, 0.25366950790104803);
    cp = fma(cp, g, -1.2337005501361697);
    cp = fma(cp, g, 1.0);
    const bool odd = ((long long)q) & 1;
    as = odd ? cp : sp;
    ac = odd ? sp : cp;
}

// ---------------------------------------------------------------------------
// Ellipse evaluator. Ctx holds 2 pi a' and 2 pi b' with a', b' the optionally
// normalised semi-axes, so speed(x) = sqrt((A sin 2pi x)^2 + (B cos 2pi x)^2).
// ---------------------------------------------------------------------------
template <typename Dt>
struct EllipseEval {
    const double* a;
    const double* b;
    int normalize;

    struct Ctx {
        Dt A, B;
    };

    __device__ __forceinline__ double scale(int i) const {
        if (!normalize) return 1.0;
        double s = fmax(fabs(a[i]), fabs(b[i]));
        return s > 0.0 ? s : 1.0;
    }
    __device__ __forceinline__ Ctx load(int i) const { return load(i, 1.0 / scale(i)); }
    // with the reciprocal of scale(i) already at hand
    __device__ __forceinline__ Ctx load(int i, double inv_scale) const {
        return Ctx{Dt(kTwoPi * a[i] * inv_scale), Dt(kTwoPi * b[i] * inv_scale)};
    }
    __device__ __forceinline__ float speed(const Ctx& c, float x) const {
        float sn, cs;
        abs_sincos_2pi(x, sn, cs);
        const float u = float(c.A) * sn, v = float(c.B) * cs;
        return __builtin_amdgcn_sqrtf(fmaf(u, u, v * v));
    }
    __device__ __forceinline__ double speed(const Ctx& c, double x) const {
        double sn, cs;
        abs_sincos_2pi(x, sn, cs);
        const double u = double(c.A) * sn, v = double(c.B) * cs;
        return sqrt(fma(u, u, v * v));
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
    __device__ __forceinline__ static double sq(double v) { return sqrt(v); }
    __device__ __forceinline__ double to_u(int i, double x) const { return u0[i] + h[i] * x; }
};

#endif  // __HIPCC__

}  // namespace dev
}  // namespace gcurval
