// Per-interval input features of the learned seed, and their computation in
// the B-spline setup pass (which already holds each span's control points and
// derivative polynomial).
#pragma once

#include "device.h"
#include "curves.h"
#include "seeded.h"

namespace gcurval {
namespace dev {

constexpr int kFeat = 16;

// float -> bf16 bits, round to nearest even (finite inputs)
__host__ __device__ inline uint16_t f2bf(float f) {
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t b) {
    uint32_t u = (uint32_t)b << 16;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

#if defined(__HIPCC__)
// Hardware log2 / square root / reciprocal (about 1 ulp): the features are
// rounded to BF16 (8 significant bits) before they reach the network.
__device__ __forceinline__ float flog2(float x) { return __builtin_amdgcn_logf(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float fdiv(float x, float y) { return x * __builtin_amdgcn_rcpf(y); }
#endif

// ---------------------------------------------------------------------------
// Per-interval features (16, FP32, before standardisation), all in the
// normalised units the error test sees:
//   0-4   log2 |dC/dx| at x = 0, 1/4, 1/2, 3/4, 1
//   5     log2 of max / min of those speeds
//   6-8   cosines between dC/dx at (0, 1), (0, 1/2), (1/2, 1)
//   9     log2 (level-0 tolerance / Simpson length estimate)
//   10    log2 (|leading coefficient| / length estimate)
//   11    log2 of longest / shortest edge of the span's control polygon
//   12    smallest cosine between consecutive control-polygon edges
//   13    log2 (knot span x intervals of the curve)
//   14    class code (degree / 5 + dim / 10; 0 for ellipses)
//   15    min / max speed over 17 samples (the safety-floor quantity)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void features_common(const float sp[5], const float dv[5][3], float tol0, float* f) {
    float mn = sp[0], mx = sp[0];
#pragma unroll
    for (int q = 0; q < 5; q++) {
        f[q] = flog2(sp[q] + 1e-12f);
        mn = fminf(mn, sp[q]);
        mx = fmaxf(mx, sp[q]);
    }
    f[5] = flog2(fdiv(mx + 1e-12f, mn + 1e-12f));
    const int pa[3] = {0, 0, 2}, pb[3] = {4, 2, 4};
#pragma unroll
    for (int t = 0; t < 3; t++) {
        const int a = pa[t], b = pb[t];
        const float dot = dv[a][0] * dv[b][0] + dv[a][1] * dv[b][1] + dv[a][2] * dv[b][2];
        f[6 + t] = fdiv(dot, sp[a] * sp[b] + 1e-20f);
    }
    const float Lest = (sp[0] + 4.f * sp[2] + sp[4]) * (1.0f / 6.0f) + 1e-12f;
    f[9] = flog2(fdiv(tol0, Lest));
}

template <int P, int D>
__device__ void features_bspline(const float* c, double h, int n_int_curve, const double* ctrl, int pstride, int s,
                                 float tol0, float* f, float* r2_out) {
    // dC/dx at x = j / 16, j = 0..16: the min / max of |dC/dx|^2 over all
    // 17 (the safety-floor quantity) and the vectors at j = 0, 4, 8, 12, 16
    float sp[5], dv[5][3];
    float mn2 = 3.0e38f, mx2 = 0.0f;
#pragma unroll
    for (int j = 0; j <= 16; j++) {
        const float x = j * (1.0f / 16.0f);
        float s2 = 0.f, vd[3] = {0.f, 0.f, 0.f};
#pragma unroll
        for (int d = 0; d < D; d++) {
            float v = c[d * P + P - 1];
#pragma unroll
            for (int e = P - 2; e >= 0; e--) v = fmaf(v, x, c[d * P + e]);
            vd[d] = v;
            s2 = fmaf(v, v, s2);
        }
        mn2 = fminf(mn2, s2);
        mx2 = fmaxf(mx2, s2);
        if ((j & 3) == 0) {
            sp[j >> 2] = fsqrt(s2);
#pragma unroll
            for (int d = 0; d < 3; d++) dv[j >> 2][d] = vd[d];
        }
    }
    features_common(sp, dv, tol0, f);
    const float Lest = (sp[0] + 4.f * sp[2] + sp[4]) * (1.0f / 6.0f) + 1e-12f;
    float lead = 0.f;
#pragma unroll
    for (int d = 0; d < D; d++) lead += c[d * P + P - 1] * c[d * P + P - 1];
    f[10] = P > 1 ? flog2(fdiv(fsqrt(lead), Lest) + 1e-12f) : -40.f;
    // the span's control polygon P_{s-p} .. P_s
    float emin = 1e30f, emax = 0.f, cmin = 1.f;
    float pe[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int e = 0; e < P; e++) {
        float ev[3] = {0.f, 0.f, 0.f}, el = 0.f;
#pragma unroll
        for (int d = 0; d < D; d++) {
            ev[d] = (float)(ctrl[(s - P + e + 1) * pstride + d] - ctrl[(s - P + e) * pstride + d]);
            el += ev[d] * ev[d];
        }
        el = fsqrt(el);
        emin = fminf(emin, el);
        emax = fmaxf(emax, el);
        if (e > 0) {
            const float pl = fsqrt(pe[0] * pe[0] + pe[1] * pe[1] + pe[2] * pe[2]);
            const float dot = pe[0] * ev[0] + pe[1] * ev[1] + pe[2] * ev[2];
            cmin = fminf(cmin, fdiv(dot, pl * el + 1e-20f));
        }
        pe[0] = ev[0];
        pe[1] = ev[1];
        pe[2] = ev[2];
    }
    f[11] = flog2(fdiv(emax + 1e-12f, emin + 1e-12f));
    f[12] = cmin;
    f[13] = flog2((float)(h * n_int_curve) + 1e-12f);
    f[14] = P * 0.2f + D * 0.1f;
    // 17-sample min / max speed (the safety-floor quantity, as speed2_ratio)
    const float r2 = mx2 > 0.0f ? mn2 / mx2 : 1.0f;
    f[15] = fsqrt(r2);
    *r2_out = r2;
}

// Ellipse (normalised semi-axes a', b'): dC/dx = 2 pi (-a' sin 2 pi x, b' cos 2 pi x).
__device__ void features_ellipse(double a, double b, float tol0, float* f) {
    const float A = (float)(kTwoPi * a), B = (float)(kTwoPi * b);
    float sp[5], dv[5][3];
    const float sn[5] = {0.f, 1.f, 0.f, -1.f, 0.f}, cs[5] = {1.f, 0.f, -1.f, 0.f, 1.f};
#pragma unroll
    for (int q = 0; q < 5; q++) {
        dv[q][0] = -A * sn[q];
        dv[q][1] = B * cs[q];
        dv[q][2] = 0.f;
        sp[q] = fsqrt(dv[q][0] * dv[q][0] + dv[q][1] * dv[q][1]);
    }
    features_common(sp, dv, tol0, f);
    const float lo = fminf(fabsf(A), fabsf(B)), hi = fmaxf(fabsf(A), fabsf(B));
    f[10] = flog2(fdiv(hi + 1e-12f, lo + 1e-12f));
    f[11] = 0.f;
    f[12] = 1.f;
    f[13] = 0.f;
    f[14] = 0.f;
    f[15] = hi > 0.f ? fdiv(lo, hi) : 1.f;
}

}  // namespace dev
}  // namespace gcurval
