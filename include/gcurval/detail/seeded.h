// Kernels of the seeded table build, where each knot interval i starts from
// 2^k_i roots (k_i in 0..6, predicted per interval) instead of a fixed n_sub:
// the safety floor that raises a seed where the speed nearly vanishes, and
// the bucketing of intervals by (evaluator class, k) (a stable counting
// sort, wave-aggregated) so every subdivision tile has one shape. The roots
// are stored densely in interval order; the bucketing passes also write each
// interval's first root (root_off[i] = sum of 2^k_j over j < i).
#pragma once

#include "device.h"
#include "curves.h"

namespace gcurval {
namespace dev {

constexpr int kMaxSeedK = 6;

__device__ __forceinline__ int clamp_k(int k) { return k < 0 ? 0 : (k > kMaxSeedK ? kMaxSeedK : k); }

constexpr int kSeedBins = kMaxSeedK + 1;  // bins per evaluator class
constexpr int kMaxBins = 80;               // 10 classes x 7 k values, rounded up

// Bucket of interval i: evaluator class index x 7 + k_i, where k_i is the
// seed raised to the interval's safety floor.
struct BinArgs {
    const int8_t* kk;         // [n_int] seed k
    const int8_t* kmin;       // [n_int] safety floor (k_seed_floor)
    const int8_t* cls_idx;    // optional [n_curve] class index of each curve
    const int32_t* int_curve; // interval -> curve (needed with cls_idx)
    int8_t* root_k;           // out [n_int]: k_i, the log2 root count the build uses
};
__device__ __forceinline__ int k_eff(const int8_t* kk, const int8_t* kmin, int64_t i) {
    const int k = clamp_k(kk[i]);
    const int m = kmin ? (int)kmin[i] : 0;
    return k > m ? k : m;
}

// ---------------------------------------------------------------------------
// Safety floor of the seed. The halves-versus-whole test can pass by
// coincidence where |C'| nearly vanishes inside a wide root (a kink of the
// integrand that no node sees); the plain method's 64-way pre-split keeps
// such roots narrow. The seeded subdivision refines every segment wider than
// a plain root whose halves' node speeds dip (dip_forces), so the floor only
// keeps the roots of an interval whose 17-sample min / max speed ratio is
// below 0.2 narrow enough (4 roots) for those nodes to sample it at least
// as densely as the 17 points do (spacing <= 0.028 of the interval against
// 1/16); a non-finite ratio forces the plain root count.
// ---------------------------------------------------------------------------
constexpr double kFloorDip = 0.2;
constexpr int kFloorDipK = 2;

__device__ __forceinline__ int8_t floor_from_ratio(float r2, int ls_plain) {
    if (!(r2 == r2)) return (int8_t)ls_plain;
    if (r2 < float(kFloorDip * kFloorDip)) return (int8_t)(kFloorDipK < ls_plain ? kFloorDipK : ls_plain);
    return 0;
}

// min / max of |dC/dx|^2 over x = j / 16, j = 0..16 (power-basis coefficients
// c[d * P + e] of the derivative, FP32).
template <int P, int D>
__device__ __forceinline__ float speed2_ratio(const float* c) {
    float mn = 3.0e38f, mx = 0.0f;
    for (int j = 0; j <= 16; j++) {
        const float x = j * (1.0f / 16.0f);
        float s2 = 0.0f;
#pragma unroll
        for (int d = 0; d < D; d++) {
            float v = c[d * P + P - 1];
#pragma unroll
            for (int e = P - 2; e >= 0; e--) v = fmaf(v, x, c[d * P + e]);
            s2 = fmaf(v, v, s2);
        }
        mn = fminf(mn, s2);
        mx = fmaxf(mx, s2);
    }
    return mx > 0.0f ? mn / mx : 1.0f;
}

// FP32 copy of an interval's derivative coefficients (from the FP32 table when
// the build keeps one, else rounded from the FP64 table).
__device__ __forceinline__ void load_coef_f32(const float* coef32, const double* coef64, int cstride, int64_t i,
                                              int n, float* cf) {
    if (coef32) {
        for (int j = 0; j < n; j++) cf[j] = coef32[i * cstride + j];
    } else {
        for (int j = 0; j < n; j++) cf[j] = (float)coef64[i * cstride + j];
    }
}

__device__ __forceinline__ float bspline_speed2_ratio(int p, int D, const float* cf) {
#define GCV_FLOOR_CASE(PP, DD) \
    case 10 * PP + DD:         \
        return speed2_ratio<PP, DD>(cf);
    switch (10 * p + D) {
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
            return 1.0f;
    }
#undef GCV_FLOOR_CASE
}

static __global__ __launch_bounds__(256) void k_seed_floor_bspline(int64_t n, const int32_t* int_curve,
                                                                   const int32_t* degree, const int32_t* dim,
                                                                   const float* coef32, const double* coef64,
                                                                   int cstride, int ls_plain, int8_t* kmin) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = int_curve[i];
    const int p = degree[c], D = dim[c];
    float cf[kCoefStride];
    load_coef_f32(coef32, coef64, cstride, i, p * D, cf);
    kmin[i] = floor_from_ratio(bspline_speed2_ratio(p, D, cf), ls_plain);
}

// Ellipse: speed^2 = (2 pi)^2 (a^2 sin^2 + b^2 cos^2) spans [min(a,b)^2, max(a,b)^2].
__device__ __forceinline__ float ellipse_speed2_ratio(double a, double b) {
    const double aa = fabs(a), bb = fabs(b);
    const double mn = aa < bb ? aa : bb, mx = aa > bb ? aa : bb;
    return (float)(mx > 0.0 ? (mn * mn) / (mx * mx) : 1.0);
}

static __global__ __launch_bounds__(256) void k_seed_floor_ellipse(int64_t n, const double* a, const double* b,
                                                                   int ls_plain, int8_t* kmin) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    kmin[i] = floor_from_ratio(ellipse_speed2_ratio(a[i], b[i]), ls_plain);
}

// Block-wide exclusive scan (blockDim.x / 64 <= 16 waves); *total gets the
// block sum.
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

// Stable counting sort of intervals by bin (one radix pass). Wave w owns the
// contiguous chunk [w wchunk, (w + 1) wchunk) of intervals (wchunk a multiple
// of 64); bin counters live in registers, lane j holding bin j (bins 64..127
// in a second register), and each 64-interval group is ranked with one
// ballot per distinct bin, so there are no LDS atomics and no block barriers.
//   k_bucket_hist   bins8[i] = bin of i, root_k[i] = k_i; per-wave counts, bin-major:
//                   hist[b * nw + w]; roots of the chunk: root_chunk[w]
//   k_bucket_totals block per bin: total count of the bin
//   k_bucket_rows   block per bin: exclusive scan of the bin's row in place,
//                   offset by the totals of the bins before it, and
//                   ranges[2b] = first slot, ranges[2b + 1] = count of bin b;
//                   one more block: exclusive scan of root_chunk in place
//   k_bucket_place  list[hist[b * nw + w] + rank] = i, rank = order of i
//                   among the wave chunk's intervals of bin b, and
//                   root_off[i] = root_chunk[w] + roots of the chunk before i
// The list is ordered by (bin, interval), so a wave tile takes neighbouring
// intervals of one curve.
constexpr int kBucketChunk = 1024;      // intervals per wave chunk (more when capped)
constexpr int kBucketMaxWaves = 16384; // most wave chunks of the bucketing passes
constexpr int kBucketUnroll = 4;       // 64-interval groups in flight per wave

struct BinCounts {
    int lo = 0, hi = 0;  // lane j: count of bin j, and of bin 64 + j
    __device__ __forceinline__ int get(int bin) const {  // bin wave-uniform
        return bin < 64 ? __builtin_amdgcn_readlane(lo, bin) : __builtin_amdgcn_readlane(hi, bin - 64);
    }
    __device__ __forceinline__ void add(int bin, int v) {
        const int lane = lane_id();
        if (bin < 64) {
            if (lane == bin) lo += v;
        } else if (lane == bin - 64) {
            hi += v;
        }
    }
};

// Rank of each valid lane among the group's lanes of its bin, offset by the
// running count of that bin; the counts then include the group.
__device__ __forceinline__ int group_rank(bool valid, int b, BinCounts& cnt) {
    int rank = 0;
    uint64_t todo = ballot(valid);
    while (todo) {
        const int leader = __ffsll((unsigned long long)todo) - 1;
        const int bl = __builtin_amdgcn_readlane(b, leader);
        const uint64_t mm = ballot(valid && b == bl);
        if (valid && b == bl) rank = cnt.get(bl) + mbcnt(mm);
        cnt.add(bl, popc64(mm));
        todo &= ~mm;
    }
    return rank;
}

static __global__ __launch_bounds__(256) void k_bucket_hist(int64_t n, BinArgs B, int n_bins, int64_t wchunk, int nw,
                                                            int8_t* bins8, int32_t* hist, int64_t* root_chunk) {
    const int lane = lane_id();
    const int64_t w = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (w >= nw) return;
    const int64_t b0 = w * wchunk;
    const int64_t b1 = b0 + wchunk < n ? b0 + wchunk : n;
    BinCounts cnt;
    int64_t roots = 0;  // this lane's intervals' roots
    for (int64_t g0 = b0; g0 < b1; g0 += 64 * kBucketUnroll) {
        int bb[kBucketUnroll];
#pragma unroll
        for (int u = 0; u < kBucketUnroll; u++) {
            const int64_t i = g0 + 64 * u + lane;
            int k = 0, c = 0;
            if (i < b1) {
                k = k_eff(B.kk, B.kmin, i);
                if (B.cls_idx) c = B.cls_idx[B.int_curve ? B.int_curve[i] : i];
                B.root_k[i] = (int8_t)k;
                roots += 1 << k;
            }
            bb[u] = kSeedBins * c + k;
        }
#pragma unroll
        for (int u = 0; u < kBucketUnroll; u++) {
            const int64_t i = g0 + 64 * u + lane;
            const bool valid = i < b1;
            if (valid) bins8[i] = (int8_t)bb[u];
            group_rank(valid, bb[u], cnt);
        }
    }
    if (lane < n_bins) hist[(int64_t)lane * nw + w] = cnt.lo;
    if (lane + 64 < n_bins) hist[(int64_t)(lane + 64) * nw + w] = cnt.hi;
    for (int o = 32; o > 0; o >>= 1) roots += __shfl_xor(roots, o, 64);
    if (lane == 0) root_chunk[w] = roots;
}

static __global__ __launch_bounds__(1024) void k_bucket_totals(int nw, const int32_t* hist, int32_t* totals) {
    __shared__ int32_t s_wave[16];
    const int32_t* row = hist + (int64_t)blockIdx.x * nw;
    int32_t v = 0;
    for (int j = threadIdx.x; j < nw; j += blockDim.x) v += row[j];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane_id() == 0) s_wave[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        int32_t t = 0;
        for (int j = 0; j < (int)(blockDim.x >> 6); j++) t += s_wave[j];
        totals[blockIdx.x] = t;
    }
}

static __global__ __launch_bounds__(1024) void k_bucket_rows(int nw, int n_bins, int32_t* hist,
                                                             const int32_t* totals, int64_t* ranges,
                                                             int64_t* root_chunk) {
    __shared__ int64_t s_wave[16];
    __shared__ int64_t s_start;
    const int b = blockIdx.x;
    if (b == n_bins) {  // the last block: chunk root counts -> chunk root offsets
        int64_t carry = 0;
        for (int base = 0; base < nw; base += blockDim.x) {
            const int j = base + (int)threadIdx.x;
            const int64_t v = j < nw ? root_chunk[j] : 0;
            int64_t tot;
            const int64_t ex = block_scan_excl(v, s_wave, &tot);
            if (j < nw) root_chunk[j] = carry + ex;
            carry += tot;
        }
        return;
    }
    if (threadIdx.x < 64) {  // start of bin b: totals of bins 0 .. b-1 (b < kMaxBins <= 128)
        const int l = (int)threadIdx.x;
        int64_t v = (l < b ? totals[l] : 0) + (l + 64 < b ? totals[l + 64] : 0);
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        if (l == 0) s_start = v;
    }
    __syncthreads();
    const int64_t start = s_start;
    int32_t* row = hist + (int64_t)b * nw;
    int64_t carry = start;
    for (int base = 0; base < nw; base += blockDim.x) {
        const int j = base + (int)threadIdx.x;
        const int64_t v = j < nw ? row[j] : 0;
        int64_t tot;
        const int64_t ex = block_scan_excl(v, s_wave, &tot);
        if (j < nw) row[j] = (int32_t)(carry + ex);
        carry += tot;
    }
    if (threadIdx.x == 0) {
        ranges[2 * b] = start;
        ranges[2 * b + 1] = totals[b];
    }
}

static __global__ __launch_bounds__(256) void k_bucket_place(int64_t n, int n_bins, int64_t wchunk, int nw,
                                                             const int8_t* bins8, const int32_t* hist,
                                                             const int64_t* root_chunk, int32_t* list,
                                                             int64_t* root_off) {
    const int lane = lane_id();
    const int64_t w = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (w >= nw) return;
    const int64_t b0 = w * wchunk;
    const int64_t b1 = b0 + wchunk < n ? b0 + wchunk : n;
    BinCounts base;
    base.lo = lane < n_bins ? hist[(int64_t)lane * nw + w] : 0;
    base.hi = lane + 64 < n_bins ? hist[(int64_t)(lane + 64) * nw + w] : 0;
    int64_t roots = root_chunk[w];  // first root of the group (wave-uniform)
    for (int64_t g0 = b0; g0 < b1; g0 += 64 * kBucketUnroll) {
        int bb[kBucketUnroll];
#pragma unroll
        for (int u = 0; u < kBucketUnroll; u++) {
            const int64_t i = g0 + 64 * u + lane;
            bb[u] = i < b1 ? (int)bins8[i] : 0;
        }
#pragma unroll
        for (int u = 0; u < kBucketUnroll; u++) {
            const int64_t i = g0 + 64 * u + lane;
            const bool valid = i < b1;
            const int pos = group_rank(valid, bb[u], base);
            // roots of the group's intervals in interval order (<= 64 x 64: int)
            const int r = valid ? 1 << (bb[u] % kSeedBins) : 0;
            int incl = r;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const int y = __shfl_up(incl, o, 64);
                if (lane >= o) incl += y;
            }
            if (valid) {
                list[pos] = (int32_t)i;
                root_off[i] = roots + (incl - r);
            }
            roots += __shfl(incl, 63, 64);
        }
    }
}

}  // namespace dev
}  // namespace gcurval
