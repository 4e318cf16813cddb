// Curve points at sampled parameters for the prebuilt families (the
// point-returning variant of arclengthReverse: kn_toPoints of
// gcurval/detail/cuda/kernel/reverse.inl.cuh:344-363). One thread per sample;
// the curve comes from a search over the sample offsets, the knot interval
// from the sampler's local interval index. Raw (unnormalised) coordinates, FP64.
#pragma once

#include "device.h"

namespace gcurval {
namespace dev {

__device__ __forceinline__ int curve_of_sample(const int64_t* soff, int n_curve, int64_t g) {
    int lo = 0, hi = n_curve;  // last c with soff[c] <= g
    while (hi - lo > 1) {
        const int m = (lo + hi) >> 1;
        if (soff[m] <= g) lo = m; else hi = m;
    }
    return lo;
}

static __global__ __launch_bounds__(256) void k_points_ellipse(int n_curve, int64_t n_total, const int64_t* soff,
                                                               const double* a, const double* b, const double* u,
                                                               int out_dim, double* out) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n_total) return;
    const int c = curve_of_sample(soff, n_curve, g);
    double sn, cs;
    sincos(u[g], &sn, &cs);
    double* o = out + g * out_dim;
    o[0] = a[c] * cs;
    o[1] = b[c] * sn;
    for (int d = 2; d < out_dim; d++) o[d] = 0.0;
}

// de Boor's algorithm on span s = p + local interval (The NURBS Book, A3.1
// with the triangular scheme of A5.1), FP64.
static __global__ __launch_bounds__(256) void k_points_bspline(int n_curve, int64_t n_total, const int64_t* soff,
                                                               const int32_t* degree, const int32_t* dim,
                                                               const int64_t* knot_off, const int64_t* ctrl_off,
                                                               const double* knots, const double* ctrl,
                                                               int ctrl_stride, const double* u, const int32_t* seg,
                                                               int out_dim, double* out) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n_total) return;
    const int c = curve_of_sample(soff, n_curve, g);
    const int p = degree[c], D = dim[c];
    const double* T = knots + knot_off[c];
    const double* P = ctrl + ctrl_off[c] * ctrl_stride;
    // span of the sampler's interval index, clamped to the curve's spans
    const int np = (int)(ctrl_off[c + 1] - ctrl_off[c]);
    int s = seg[g] + p;
    s = s < p ? p : (s > np - 1 ? np - 1 : s);
    const double x = u[g];
    double d[6][3];
    for (int j = 0; j <= p; j++)
        for (int k = 0; k < 3; k++) d[j][k] = k < D ? P[(int64_t)(s - p + j) * ctrl_stride + k] : 0.0;
    for (int r = 1; r <= p; r++) {
        for (int j = p; j >= r; j--) {
            const int i = s - p + j;
            const double den = T[i + p - r + 1] - T[i];
            const double al = den > 0.0 ? (x - T[i]) / den : 0.0;
            for (int k = 0; k < 3; k++) d[j][k] = (1.0 - al) * d[j - 1][k] + al * d[j][k];
        }
    }
    double* o = out + g * out_dim;
    for (int k = 0; k < out_dim; k++) o[k] = k < 3 ? d[p][k] : 0.0;
}

}  // namespace dev
}  // namespace gcurval
