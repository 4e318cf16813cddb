// Kernels of the seeded table build: each knot interval i starts from
// 2^k_i roots (k_i in 0..6, predicted per interval) instead of a fixed
// n_sub. Intervals are bucketed by k (counting sort, wave-aggregated
// atomics) so every subdivision launch runs a uniform tile shape, and the
// root storage of interval i starts at root_off[i] (exclusive scan of 2^k_i).
#pragma once

#include "device.h"
#include "curves.h"

namespace gcurval {
namespace dev {

constexpr int kMaxSeedK = 6;
constexpr int kScanItems = 4;  // elements per thread in the scan kernels

constexpr int kSeedBins = kMaxSeedK + 1;  // bins per evaluator class
constexpr int kMaxBins = 80;               // 10 classes x 7 k values, rounded up

__device__ __forceinline__ int clamp_k(int k) { return k < 0 ? 0 : (k > kMaxSeedK ? kMaxSeedK : k); }

// Bucket of interval i: evaluator class index x 7 + k_i, where k_i is the
// seed raised to the interval's safety floor.
struct BinArgs {
    const int8_t* kk;         // [n_int] seed k
    const int8_t* kmin;       // [n_int] safety floor (k_seed_floor)
    const int8_t* cls_idx;    // optional [n_curve] class index of each curve
    const int32_t* int_curve; // interval -> curve (needed with cls_idx)
};
__device__ __forceinline__ int k_eff(const int8_t* kk, const int8_t* kmin, int64_t i) {
    const int k = clamp_k(kk[i]);
    const int m = kmin ? (int)kmin[i] : 0;
    return k > m ? k : m;
}
__device__ __forceinline__ int bin_of(const BinArgs& B, int64_t i) {
    int b = k_eff(B.kk, B.kmin, i);
    if (B.cls_idx) b += kSeedBins * B.cls_idx[B.int_curve ? B.int_curve[i] : i];
    return b;
}

// ---------------------------------------------------------------------------
// Safety floor of the seed. The halves-versus-whole test can pass by
// coincidence where |C'| nearly vanishes inside a wide root (a kink of the
// integrand that no node sees); the plain method's 64-way pre-split keeps
// such roots narrow. Sampling |dC/dx|^2 at 17 points of the interval, a
// minimum-to-maximum speed ratio below 0.05 forces the plain root count and
// below 0.2 at least 16 roots.
// ---------------------------------------------------------------------------
constexpr double kFloorPlain = 0.05;
constexpr double kFloorMid = 0.2;
constexpr int kFloorMidK = 4;

__device__ __forceinline__ int8_t floor_from_ratio(double r2, int ls_plain) {
    if (!(r2 >= kFloorPlain * kFloorPlain)) return (int8_t)ls_plain;
    if (r2 < kFloorMid * kFloorMid) return (int8_t)(kFloorMidK < ls_plain ? kFloorMidK : ls_plain);
    return 0;
}

template <int P, int D>
__device__ __forceinline__ double speed2_ratio(const double* c) {
    double mn = 1e300, mx = 0.0;
    for (int j = 0; j <= 16; j++) {
        const double x = j * (1.0 / 16.0);
        double s2 = 0.0;
#pragma unroll
        for (int d = 0; d < D; d++) {
            double v = c[d * P + P - 1];
#pragma unroll
            for (int e = P - 2; e >= 0; e--) v = v * x + c[d * P + e];
            s2 += v * v;
        }
        mn = s2 < mn ? s2 : mn;
        mx = s2 > mx ? s2 : mx;
    }
    return mx > 0.0 ? mn / mx : 1.0;
}

static __global__ __launch_bounds__(256) void k_seed_floor_bspline(int64_t n, const int32_t* int_curve,
                                                                   const int32_t* degree, const int32_t* dim,
                                                                   const double* coef64, int ls_plain,
                                                                   int8_t* kmin) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = int_curve[i];
    const double* cf = coef64 + i * kCoefStride;
    double r2 = 1.0;
#define GCV_FLOOR_CASE(PP, DD) \
    case 10 * PP + DD:         \
        r2 = speed2_ratio<PP, DD>(cf); \
        break;
    switch (10 * degree[c] + dim[c]) {
        GCV_FLOOR_CASE(1, 2)
        GCV_FLOOR_CASE(1, 3)
        GCV_FLOOR_CASE(2, 2)
        GCV_FLOOR_CASE(2, 3)
        GCV_FLOOR_CASE(3, 2)
        GCV_FLOOR_CASE(3, 3)
        GCV_FLOOR_CASE(4, 2)
        GCV_FLOOR_CASE(4, 3)
        GCV_FLOOR_CASE(5, 2)
        GCV_FLOOR_CASE(5, 3)
        default:
            break;
    }
#undef GCV_FLOOR_CASE
    kmin[i] = floor_from_ratio(r2, ls_plain);
}

// Ellipse: speed^2 = (2 pi)^2 (a^2 sin^2 + b^2 cos^2) spans [min(a,b)^2, max(a,b)^2].
static __global__ __launch_bounds__(256) void k_seed_floor_ellipse(int64_t n, const double* a, const double* b,
                                                                   int ls_plain, int8_t* kmin) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double aa = fabs(a[i]), bb = fabs(b[i]);
    const double mn = aa < bb ? aa : bb, mx = aa > bb ? aa : bb;
    kmin[i] = floor_from_ratio(mx > 0.0 ? (mn * mn) / (mx * mx) : 1.0, ls_plain);
}

// Histogram of bins over intervals into counts[0..n_bins) (zeroed beforehand).
static __global__ __launch_bounds__(256) void k_bucket_count(int64_t n, BinArgs B, int n_bins,
                                                             unsigned long long* counts) {
    __shared__ unsigned int h[kMaxBins];
    for (int j = threadIdx.x; j < kMaxBins; j += blockDim.x) h[j] = 0;
    __syncthreads();
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        atomicAdd(&h[bin_of(B, i)], 1u);
    __syncthreads();
    for (int j = threadIdx.x; j < n_bins; j += blockDim.x)
        if (h[j]) atomicAdd(&counts[j], (unsigned long long)h[j]);
}

// ranges[2b] = offset, ranges[2b+1] = count of bin b; cursor[b] = offset.
static __global__ void k_bucket_ranges(const unsigned long long* counts, int n_bins, int64_t* ranges,
                                       unsigned long long* cursor) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    unsigned long long off = 0;
    for (int b = 0; b < n_bins; b++) {
        ranges[2 * b] = (int64_t)off;
        ranges[2 * b + 1] = (int64_t)counts[b];
        cursor[b] = off;
        off += counts[b];
    }
}

// Scatter interval ids into their bins: one atomic per wave per distinct bin.
static __global__ __launch_bounds__(256) void k_bucket_scatter(int64_t n, BinArgs B, unsigned long long* cursor,
                                                               int32_t* list) {
    const int lane = lane_id();
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n; base += stride) {
        const int64_t i = base + threadIdx.x;
        const bool valid = i < n;
        const int b = valid ? bin_of(B, i) : 0;
        uint64_t todo = ballot(valid);
        while (todo) {
            const int leader = __ffsll((unsigned long long)todo) - 1;
            const int bl = __shfl(b, leader, 64);
            const bool mine = valid && b == bl;
            const uint64_t m = ballot(mine);
            unsigned long long pos = 0;
            if (lane == leader) pos = atomicAdd(&cursor[bl], (unsigned long long)popc64(m));
            pos = __shfl(pos, leader, 64);
            if (mine) list[pos + mbcnt(m)] = (int32_t)i;
            todo &= ~m;
        }
    }
}

// Exclusive scan of 2^k_i into root_off[0..n] in three passes: per-block
// sums of kScanItems * 256 elements, a single-block scan of those sums, and
// the per-block apply.
__device__ __forceinline__ int64_t block_scan_excl(int64_t v, int64_t* s_wave, int64_t* total) {
    // wave inclusive scan
    const int lane = lane_id();
    int64_t x = v;
    for (int d = 1; d < 64; d <<= 1) {
        const int64_t t = __shfl_up(x, (unsigned)d, 64);
        if (lane >= d) x += t;
    }
    const int w = threadIdx.x >> 6;
    if (lane == 63) s_wave[w] = x;
    __syncthreads();
    int64_t wbase = 0, tot = 0;
    const int nw = blockDim.x >> 6;
    for (int j = 0; j < nw; j++) {
        if (j < w) wbase += s_wave[j];
        tot += s_wave[j];
    }
    __syncthreads();
    *total = tot;
    return wbase + x - v;
}

static __global__ __launch_bounds__(256) void k_root_block_sums(int64_t n, const int8_t* kk, const int8_t* kmin,
                                                              int64_t* block_sums) {
    __shared__ int64_t s_wave[4];
    const int64_t base = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * kScanItems;
    int64_t v = 0;
    for (int j = 0; j < kScanItems; j++)
        if (base + j < n) v += (int64_t)1 << k_eff(kk, kmin, base + j);
    int64_t tot;
    block_scan_excl(v, s_wave, &tot);
    if (threadIdx.x == 0) block_sums[blockIdx.x] = tot;
}

// In place exclusive scan of nb block sums by one block.
static __global__ __launch_bounds__(256) void k_scan_block_sums(int64_t nb, int64_t* block_sums) {
    __shared__ int64_t s_wave[4];
    int64_t carry = 0;
    for (int64_t base = 0; base < nb; base += blockDim.x) {
        const int64_t i = base + threadIdx.x;
        const int64_t v = i < nb ? block_sums[i] : 0;
        int64_t tot;
        const int64_t ex = block_scan_excl(v, s_wave, &tot);
        if (i < nb) block_sums[i] = carry + ex;
        carry += tot;
        __syncthreads();
    }
}

static __global__ __launch_bounds__(256) void k_root_apply(int64_t n, const int8_t* kk, const int8_t* kmin,
                                                           const int64_t* block_offs, int64_t* root_off) {
    __shared__ int64_t s_wave[4];
    const int64_t base = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * kScanItems;
    int64_t vals[kScanItems];
    int64_t v = 0;
    for (int j = 0; j < kScanItems; j++) {
        vals[j] = base + j < n ? (int64_t)1 << k_eff(kk, kmin, base + j) : 0;
        v += vals[j];
    }
    int64_t tot;
    int64_t run = block_offs[blockIdx.x] + block_scan_excl(v, s_wave, &tot);
    for (int j = 0; j < kScanItems; j++) {
        if (base + j < n) root_off[base + j] = run;
        run += vals[j];
        if (base + j == n - 1) root_off[n] = run;
    }
}

// Constant k for every interval (uniform-k builds and tests).
static __global__ void k_fill_i8(int64_t n, int8_t* p, int8_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace dev
}  // namespace gcurval
