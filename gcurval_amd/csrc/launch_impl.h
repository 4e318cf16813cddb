// Definitions of the launchers; included once per instantiation unit.
#pragma once

#include <atomic>

#include "launch.h"

namespace gcurval {
namespace rt {

using namespace gcurval::dev;

template <typename Dt, int NGL>
void Launch<Dt, NGL>::subdivide_shape(int stack_cap, int want_wpb, int* waves_per_block, size_t* lds_bytes) {
    const int per_wave = subdiv_lds_per_wave<Dt>(stack_cap);
    int w = (want_wpb == 1 || want_wpb == 2 || want_wpb == 4) ? want_wpb : 4;
    while (w > 1 && w * per_wave > 64 * 1024) w >>= 1;
    *waves_per_block = w;
    *lds_bytes = (size_t)w * per_wave;
}

constexpr int kMaxLdsKey = 257;  // LDS sizes in 256-byte units, up to 64 KiB per block

// Resident blocks per CU of `kernel` for (waves per block, LDS bytes),
// queried once per shape and kernel; the cache is shared by every context and
// host thread (relaxed atomics: a race only repeats the query).
template <class K>
static hipError_t resident_blocks(K kernel, std::atomic<int> (&cache)[4][kMaxLdsKey], int wpb, size_t lds,
                                  int* per_cu) {
    const int lds_key = (lds & 255) ? kMaxLdsKey : (int)(lds >> 8);  // sizes here are multiples of 512
    const bool cached = wpb >= 1 && wpb <= 4 && lds_key < kMaxLdsKey;
    int n = cached ? cache[wpb - 1][lds_key].load(std::memory_order_relaxed) : 0;
    if (n == 0) {
        const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, 64 * wpb, lds);
        if (e != hipSuccess) return e;
        if (n < 1) n = 1;
        if (cached) cache[wpb - 1][lds_key].store(n, std::memory_order_relaxed);
    }
    *per_cu = n;
    return hipSuccess;
}

template <typename Dt, int NGL, class Ev, bool kWhole>
static hipError_t run_subdivide_as(const LaunchCfg& c, const Ev& ev, const SubdivArgs<Dt>& A, const char* name) {
    int wpb;
    size_t lds;
    Launch<Dt, NGL>::subdivide_shape(A.stack_cap, c.sub_wpb, &wpb, &lds);
    const int threads = 64 * wpb;
    const int ipw = 64 >> A.log2_nsub;
    const int64_t n_tiles = (A.n_slots + ipw - 1) / ipw;
    static std::atomic<int> cache[4][kMaxLdsKey] = {};
    int per_cu = 0;
    hipError_t e = resident_blocks(k_subdivide<Dt, NGL, Ev, kWhole>, cache, wpb, lds, &per_cu);
    if (e != hipSuccess) return e;
    // uniform builds launch 4 rounds of resident blocks: tiles differ in
    // refinement work, and blocks that retire early are replaced instead of
    // leaving their SIMDs idle (1M ellipses 0.438 -> 0.412 ms, 4M cubic
    // intervals 1.734 -> 1.663 ms); a seeded build's per-bin launches are
    // mostly small and measured slower that way, so they keep one round
    const int rounds = A.slot_range ? 1 : 4;
    int64_t blocks = (int64_t)c.n_cu * per_cu * rounds * (c.sub_grid_pct > 0 ? c.sub_grid_pct : 100) / 100;
    if (blocks < 1) blocks = 1;
    const int64_t need = (n_tiles + wpb - 1) / wpb;
    if (blocks > need) blocks = need;
    prof_begin(c.owner, name, c.stream);
    hipLaunchKernelGGL((k_subdivide<Dt, NGL, Ev, kWhole>), dim3((unsigned)blocks), dim3(threads), lds, c.stream, ev,
                       A);
    e = hipGetLastError();
    prof_end(c.owner, c.stream);
    if (e != hipSuccess || !A.rerun || !A.rerun_self) return e;
    // tiles that outgrew the small stack, again with the full one (the list
    // is usually empty and the waves return at once; its length is read on
    // the device)
    SubdivArgs<Dt> F = A;
    F.stack_cap = 64 * (A.max_depth + 2);
    const size_t flds = (size_t)subdiv_lds_per_wave<Dt>(F.stack_cap);
    const int64_t fblocks = std::min<int64_t>((int64_t)c.n_cu, A.n_slots);
    prof_begin(c.owner, "k_subdivide_rerun", c.stream);
    hipLaunchKernelGGL((k_subdivide_rerun<Dt, NGL, Ev, false>), dim3((unsigned)std::max<int64_t>(fblocks, 1)), dim3(64),
                       flds, c.stream, ev, F, (const int32_t*)A.rerun, (const unsigned long long*)A.rerun_count,
                       (const int8_t*)nullptr, A.ls_plain, A.base_depth);
    e = hipGetLastError();
    prof_end(c.owner, c.stream);
    return e;
}

template <typename Dt, int NGL, class Ev>
static hipError_t run_subdivide(const LaunchCfg& c, const Ev& ev, const SubdivArgs<Dt>& A, const char* name);

// Seeded subdivision of one class: the bins kernel over bins k = 0..5 (4
// rounds of resident blocks, as uniform builds), the whole-interval kernel
// over bin 6, and the re-run kernel (one wave per listed interval; the list
// is usually empty and its waves return at once).
template <typename Dt, int NGL, class Ev>
static hipError_t run_seeded(const LaunchCfg& c, const Ev& ev, const SeededLaunch<Dt>& S, const char* name) {
    const SubdivArgs<Dt>& A = S.bins;
    if (A.n_slots <= 0) return hipSuccess;
    hipError_t e;
    {
        int wpb;
        size_t lds;
        Launch<Dt, NGL>::subdivide_shape(A.stack_cap, c.sub_wpb, &wpb, &lds);
        static std::atomic<int> cache[4][kMaxLdsKey] = {};
        int per_cu = 0;
        if ((e = resident_blocks(k_subdivide_bins<Dt, NGL, Ev>, cache, wpb, lds, &per_cu)) != hipSuccess) return e;
        int64_t blocks = (int64_t)c.n_cu * per_cu * 4 * (c.sub_grid_pct > 0 ? c.sub_grid_pct : 100) / 100;
        const int64_t need = (A.n_slots + wpb - 1) / wpb;  // at most one tile per interval
        if (blocks > need) blocks = need;
        if (blocks < 1) blocks = 1;
        prof_begin(c.owner, name, c.stream);
        hipLaunchKernelGGL((k_subdivide_bins<Dt, NGL, Ev>), dim3((unsigned)blocks), dim3(64 * wpb), lds, c.stream,
                           ev, A, S.run);
        e = hipGetLastError();
        prof_end(c.owner, c.stream);
        if (e != hipSuccess) return e;
    }
    if ((e = run_subdivide<Dt, NGL>(c, ev, S.whole, name)) != hipSuccess) return e;
    {
        const size_t lds = (size_t)subdiv_lds_per_wave<Dt>(S.rerun.stack_cap);
        const int64_t blocks = std::min<int64_t>((int64_t)c.n_cu, A.n_slots);
        hipLaunchKernelGGL((k_subdivide_rerun<Dt, NGL, Ev>), dim3((unsigned)blocks), dim3(64), lds, c.stream, ev,
                           S.rerun, (const int32_t*)S.run.rerun, (const unsigned long long*)S.run.rerun_count,
                           S.root_k, S.run.ls_plain, S.run.base_depth);
        e = hipGetLastError();
    }
    return e;
}

template <typename Dt, int NGL>
hipError_t Launch<Dt, NGL>::seeded_ellipse(const LaunchCfg& c, const double* a, const double* b, int normalize,
                                           const SeededLaunch<Dt>& S) {
    EllipseEval<Dt> ev{a, b, normalize};
    return run_seeded<Dt, NGL>(c, ev, S, "k_subdivide_seeded_ellipse");
}

// Tiles of one whole 64-root interval (the plain method, 64-root bins) take
// the wave-uniform kernel; other root counts the per-lane one.
template <typename Dt, int NGL, class Ev>
static hipError_t run_subdivide(const LaunchCfg& c, const Ev& ev, const SubdivArgs<Dt>& A, const char* name) {
    if (A.n_slots <= 0) return hipSuccess;
    if (A.log2_nsub == 6) return run_subdivide_as<Dt, NGL, Ev, true>(c, ev, A, name);
    return run_subdivide_as<Dt, NGL, Ev, false>(c, ev, A, name);
}

template <typename Dt, int NGL>
hipError_t Launch<Dt, NGL>::subdivide_ellipse(const LaunchCfg& c, const double* a, const double* b,
                                              int normalize, SubdivArgs<Dt> A) {
    EllipseEval<Dt> ev{a, b, normalize};
    return run_subdivide<Dt, NGL>(c, ev, A, "k_subdivide_ellipse");
}

template <typename Dt, int P, int D>
struct MakeEv;
template <int P, int D>
struct MakeEv<float, P, D> {
    __host__ __device__ static BSplineEval<float, P, D> get(const BSplineTableDev& T) {
        return BSplineEval<float, P, D>{T.coef32, T.cstride, T.int_scale, T.u0, T.h};
    }
};
template <int P, int D>
struct MakeEv<double, P, D> {
    __host__ __device__ static BSplineEval<double, P, D> get(const BSplineTableDev& T) {
        return BSplineEval<double, P, D>{T.coef64, T.cstride, T.int_scale, T.u0, T.h};
    }
};
// FP64 evaluator of the sampler's final Newton step: an FP32 build keeps only
// FP32 coefficients and evaluates them in FP64 arithmetic.
template <typename Dt, int P, int D>
struct MakeEv64;
template <int P, int D>
struct MakeEv64<float, P, D> {
    __host__ __device__ static BSplineEval<double, P, D, float> get(const BSplineTableDev& T) {
        return BSplineEval<double, P, D, float>{T.coef32, T.cstride, T.int_scale, T.u0, T.h};
    }
};
template <int P, int D>
struct MakeEv64<double, P, D> {
    __host__ __device__ static BSplineEval<double, P, D> get(const BSplineTableDev& T) {
        return MakeEv<double, P, D>::get(T);
    }
};

template <typename Dt, int NGL>
hipError_t Launch<Dt, NGL>::subdivide_bspline(const LaunchCfg& c, int p, int D, const BSplineTableDev& T,
                                              SubdivArgs<Dt> A) {
#define GCV_SUBDIV_CASE(PP, DD) \
    case 10 * PP + DD:          \
        return run_subdivide<Dt, NGL>(c, MakeEv<Dt, PP, DD>::get(T), A, "k_subdivide_bspline");
    switch (bspline_class(p, D)) {
        GCV_SUBDIV_CASE(1, 2)
        GCV_SUBDIV_CASE(1, 3)
        GCV_SUBDIV_CASE(2, 2)
        GCV_SUBDIV_CASE(2, 3)
        GCV_SUBDIV_CASE(3, 2)
        GCV_SUBDIV_CASE(3, 3)
        GCV_SUBDIV_CASE(4, 2)
        GCV_SUBDIV_CASE(4, 3)
        GCV_SUBDIV_CASE(5, 2)
        GCV_SUBDIV_CASE(5, 3)
        default:
            return hipErrorInvalidValue;
    }
#undef GCV_SUBDIV_CASE
}

template <typename Dt, int NGL>
hipError_t Launch<Dt, NGL>::seeded_bspline(const LaunchCfg& c, int p, int D, const BSplineTableDev& T,
                                           const SeededLaunch<Dt>& S) {
#define GCV_SEEDED_CASE(PP, DD) \
    case 10 * PP + DD:          \
        return run_seeded<Dt, NGL>(c, MakeEv<Dt, PP, DD>::get(T), S, "k_subdivide_seeded_bspline");
    switch (bspline_class(p, D)) {
        GCV_SEEDED_CASE(1, 2)
        GCV_SEEDED_CASE(1, 3)
        GCV_SEEDED_CASE(2, 2)
        GCV_SEEDED_CASE(2, 3)
        GCV_SEEDED_CASE(3, 2)
        GCV_SEEDED_CASE(3, 3)
        GCV_SEEDED_CASE(4, 2)
        GCV_SEEDED_CASE(4, 3)
        GCV_SEEDED_CASE(5, 2)
        GCV_SEEDED_CASE(5, 3)
        default:
            return hipErrorInvalidValue;
    }
#undef GCV_SEEDED_CASE
}

// FP64 evaluator of the final Newton step: the polish evaluator for FP32
// builds (FP64 builds take no such step)
template <typename Dt>
using EllipsePolish = std::conditional_t<(sizeof(Dt) < sizeof(double)), EllipsePolishEval, EllipseEval<double>>;
template <typename Dt>
static EllipsePolish<Dt> make_polish(const double* a, const double* b, int normalize) {
    EllipsePolish<Dt> e;
    static_cast<EllipseEval<double>&>(e) = EllipseEval<double>{a, b, normalize};
    return e;
}

template <typename Dt, int NGL, bool kSeeded>
static void run_sample_runs(const LaunchCfg& c, const EllipseEval<Dt>& ev, const EllipsePolish<Dt>& ev64,
                            const SampleArgs& A, const int32_t* blk_curve) {
    // resident blocks per CU, queried once (shared by every context and host
    // thread; a race only repeats the query)
    static std::atomic<int> cached{0};
    int per_cu = cached.load(std::memory_order_relaxed);
    if (per_cu == 0) {
        int n = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &n, k_sample_runs<Dt, NGL, EllipseEval<Dt>, EllipsePolish<Dt>, kSeeded>, kSampleBlock, 0) !=
                hipSuccess || n < 1)
            n = 1;
        per_cu = n;
        cached.store(n, std::memory_order_relaxed);
    }
    const int64_t n_tiles = (A.n_total + 63) / 64;
    const int64_t wpb = kSampleBlock / 64;
    // 16 rounds of resident blocks: curves differ in cost (refined roots, the
    // cubic or quintic start), so one run per resident wave left SIMDs idle
    // while the slowest runs finished; shorter runs are handed out by the
    // dispatcher as blocks retire (1M ellipses, 970M samples: 9.14 -> 7.64 ms;
    // 32 rounds the same, 64 and 128 slower as runs start more often)
    constexpr int kRunRounds = 16;
    int64_t n_blocks = (int64_t)c.n_cu * per_cu * kRunRounds;
    if (n_blocks * wpb > n_tiles) n_blocks = (n_tiles + wpb - 1) / wpb;
    const int64_t tpw = (n_tiles + n_blocks * wpb - 1) / (n_blocks * wpb);
    hipLaunchKernelGGL((k_sample_runs<Dt, NGL, EllipseEval<Dt>, EllipsePolish<Dt>, kSeeded>),
                       dim3((unsigned)n_blocks), dim3(kSampleBlock), 0, c.stream, ev, ev64, A, blk_curve, tpw);
}

template <typename Dt, int NGL>
hipError_t Launch<Dt, NGL>::sample_ellipse(const LaunchCfg& c, const double* a, const double* b, int normalize,
                                           const SampleArgs& A, const int32_t* blk_curve) {
    if (A.n_total <= 0) return hipSuccess;
    EllipseEval<Dt> ev{a, b, normalize};
    const EllipsePolish<Dt> ev64 = make_polish<Dt>(a, b, normalize);
    const int64_t blocks = (A.n_total + kSampleBlock - 1) / kSampleBlock;
    prof_begin(c.owner, "k_sample_ellipse");
    // the common modes (uniform or explicit positions, plain or seeded roots)
    // are compiled in, so the kernel has no branch between its loads
    const bool pos = A.pos != nullptr, seeded = A.root_k != nullptr;
#define GCV_ELLIPSE_SAMPLE(P, S)                                                                          \
    hipLaunchKernelGGL((k_sample<Dt, NGL, EllipseEval<Dt>, EllipsePolish<Dt>, StaticMode<P, S>>),         \
                       dim3((unsigned)blocks), dim3(kSampleBlock), 0, c.stream, ev, ev64, A, blk_curve)
    if (pos && seeded) {
        GCV_ELLIPSE_SAMPLE(true, true);
    } else if (pos) {
        GCV_ELLIPSE_SAMPLE(true, false);
    } else {
        // uniform positions: persistent waves over contiguous runs of 64-sample
        // tiles, as many blocks as are resident at once (queried once)
        if (seeded)
            run_sample_runs<Dt, NGL, true>(c, ev, ev64, A, blk_curve);
        else
            run_sample_runs<Dt, NGL, false>(c, ev, ev64, A, blk_curve);
    }
#undef GCV_ELLIPSE_SAMPLE
    hipError_t e = hipGetLastError();
    prof_end(c.owner);
    return e;
}

// Mixed-class B-spline sampler: the class is uniform within a curve, so a
// wave diverges only where its samples straddle curves of different classes.
template <typename Dt, int NGL, class M>
__global__ __launch_bounds__(256, 8) void k_sample_bspline(BSplineTableDev T, int uniform_class, SampleArgs A,
                                                        const int32_t* blk_curve) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= A.n_total) return;
    __shared__ double s_win[kSampleBlock / 64][kRootWin];
    __shared__ double s_iwin[kSampleBlock / 64][64];
    const SampleLoc L = locate_sample(A, blk_curve, blockIdx.x, g);
    const int c = L.c;
    const int cls = A.class_of_curve ? A.class_of_curve[c] : uniform_class;
    int32_t seg = 0;
    double u = 0.0;
#define GCV_SAMPLE_CASE(PP, DD)                                                                     \
    case 10 * PP + DD:                                                                             \
        u = sample_one<Dt, NGL, M>(MakeEv<Dt, PP, DD>::get(T), MakeEv64<Dt, PP, DD>::get(T), A, c, g, \
                                   &seg, L.lo, L.hi, (lds_double*)s_win[threadIdx.x >> 6],          \
                                   (lds_double*)s_iwin[threadIdx.x >> 6]);                          \
        break;
    switch (cls) {
        GCV_SAMPLE_CASE(1, 2)
        GCV_SAMPLE_CASE(1, 3)
        GCV_SAMPLE_CASE(2, 2)
        GCV_SAMPLE_CASE(2, 3)
        GCV_SAMPLE_CASE(3, 2)
        GCV_SAMPLE_CASE(3, 3)
        GCV_SAMPLE_CASE(4, 2)
        GCV_SAMPLE_CASE(4, 3)
        GCV_SAMPLE_CASE(5, 2)
        GCV_SAMPLE_CASE(5, 3)
        default:
            break;
    }
#undef GCV_SAMPLE_CASE
    A.out_u[g] = u;
    if (A.out_seg) A.out_seg[g] = seg;
}

template <typename Dt, int NGL>
hipError_t Launch<Dt, NGL>::sample_bspline(const LaunchCfg& c, int uniform_class, const BSplineTableDev& T,
                                           const SampleArgs& A, const int32_t* blk_curve) {
    if (A.n_total <= 0) return hipSuccess;
    const int64_t blocks = (A.n_total + kSampleBlock - 1) / kSampleBlock;
    prof_begin(c.owner, "k_sample_bspline");
    const bool pos = A.pos != nullptr, seeded = A.root_k != nullptr;
#define GCV_BSPLINE_SAMPLE(P, S, ST)                                                                    \
    hipLaunchKernelGGL((k_sample_bspline<Dt, NGL, StaticMode<P, S, ST>>), dim3((unsigned)blocks),      \
                       dim3(kSampleBlock), 0, c.stream, T, uniform_class, A, blk_curve)
    if constexpr (sizeof(Dt) == sizeof(float)) {
        if (A.start_mode != kStartPlain) {
            // the learned / quintic Newton start has kernels of its own
            if (pos && seeded) {
                GCV_BSPLINE_SAMPLE(true, true, true);
            } else if (pos) {
                GCV_BSPLINE_SAMPLE(true, false, true);
            } else if (seeded) {
                GCV_BSPLINE_SAMPLE(false, true, true);
            } else {
                GCV_BSPLINE_SAMPLE(false, false, true);
            }
            hipError_t e = hipGetLastError();
            prof_end(c.owner);
            return e;
        }
    }
    if (pos && seeded) {
        GCV_BSPLINE_SAMPLE(true, true, false);
    } else if (pos) {
        GCV_BSPLINE_SAMPLE(true, false, false);
    } else if (seeded) {
        GCV_BSPLINE_SAMPLE(false, true, false);
    } else {
        GCV_BSPLINE_SAMPLE(false, false, false);
    }
#undef GCV_BSPLINE_SAMPLE
    hipError_t e = hipGetLastError();
    prof_end(c.owner);
    return e;
}

// Training labels of the learned Newton start: per knot interval, log2 of
// the worst error of the quintic Hermite start (relative to root width) at
// the targets 1/4, 1/2 and 3/4 of each root's length, against the converged
// FP64 Newton solution of the same single-rule partial integral the
// sampler's FP64 step evaluates; a refined root counts as error 1 (its
// samples never take the quintic start).
template <int NGL, class Ev, class Ev64>
__device__ float start_label(const Ev& ev, const Ev64& ev64, const SampleArgs& A, int64_t i) {
    int ls = A.log2_nsub;
    int64_t ro = i << ls;
    if (A.root_k) {
        ls = A.root_k[i];
        ro = A.root_off[i];
    }
    const uint64_t deep = A.deep[i];
    const typename Ev::Ctx ctx = ev.load((int)i, 1.0);
    const typename Ev64::Ctx cw = Ev64::widen(ctx);
    float worst = 0.0f;
    double prev = 0.0;
    for (int k = 0; k < (1 << ls); k++) {
        const double pk = A.root_prefix[ro + k];
        const double len = pk - prev;
        prev = pk;
        if ((deep >> k) & 1ull) {
            worst = 1.0f;
            continue;
        }
        float a, xm, b;
        task_bounds<float>(k, 0, 0u, ls, a, xm, b);
        const LeafEnds<float> E = leaf_ends_slopes(ev, ctx, a, b);
        const double wd = (double)b - (double)a;
        for (int j = 1; j <= 3; j++) {
            const double tau = 0.25 * j * len;
            float vs;
            const float d0 = invert_leaf<NGL>(ev, ctx, a, b, (float)tau, (float)len, 0, 1, &vs, true, E, true);
            double d = d0, lo = 0.0, hi = wd;
            for (int it = 0; it < 16; it++) {
                const double F = quad_from<NGL>(ev64, cw, (double)a, d) - tau;
                if (F > 0.0) hi = d; else lo = d;
                const double v = ev64.speed(cw, (double)a + d);
                double dn = v > 0.0 ? d - F / v : 0.5 * (lo + hi);
                if (!(dn >= lo && dn <= hi)) dn = 0.5 * (lo + hi);
                const bool done = fabs(dn - d) <= wd * 0x1p-45;
                d = dn;
                if (done) break;
            }
            const float e = (float)(fabs((double)d0 - d) / wd);
            if (!(e <= worst)) worst = e == e ? e : 1.0f;  // a NaN start counts as error 1
        }
    }
    return log2f(worst > 0x1p-40f ? worst : 0x1p-40f);
}

static __global__ __launch_bounds__(256) void k_start_labels(BSplineTableDev T, int uniform_class, const int32_t* int_curve,
                                                      SampleArgs A, float* out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n_int) return;
    const int c = int_curve[i];
    const int cls = A.class_of_curve ? A.class_of_curve[c] : uniform_class;
    float r = 0.0f;
#define GCV_LABEL_CASE(PP, DD)                                                                                \
    case 10 * PP + DD:                                                                                        \
        r = start_label<3>(MakeEv<float, PP, DD>::get(T), MakeEv64<float, PP, DD>::get(T), A, i);             \
        break;
    switch (cls) {
        GCV_LABEL_CASE(1, 2)
        GCV_LABEL_CASE(1, 3)
        GCV_LABEL_CASE(2, 2)
        GCV_LABEL_CASE(2, 3)
        GCV_LABEL_CASE(3, 2)
        GCV_LABEL_CASE(3, 3)
        GCV_LABEL_CASE(4, 2)
        GCV_LABEL_CASE(4, 3)
        GCV_LABEL_CASE(5, 2)
        GCV_LABEL_CASE(5, 3)
        default:
            break;
    }
#undef GCV_LABEL_CASE
    out[i] = r;
}

template <typename Dt, int NGL>
hipError_t Launch<Dt, NGL>::start_labels(const LaunchCfg& c, int uniform_class, const BSplineTableDev& T,
                                         const int32_t* int_curve, const SampleArgs& A, float* out) {
    if (A.n_int <= 0) return hipSuccess;
    if constexpr (sizeof(Dt) != sizeof(float) || NGL != 3) {
        return hipErrorInvalidValue;  // the quintic start serves FP32 GL-3 builds
    } else {
        prof_begin(c.owner, "k_start_labels");
        hipLaunchKernelGGL(k_start_labels, dim3((unsigned)((A.n_int + 255) / 256)), dim3(256), 0, c.stream, T,
                           uniform_class, int_curve, A, out);
        hipError_t e = hipGetLastError();
        prof_end(c.owner);
        return e;
    }
}

}  // namespace rt
}  // namespace gcurval
