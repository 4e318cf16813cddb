// Device-side building blocks shared by every gcurval kernel on gfx950:
// Gauss-Legendre rules, wave64 primitives, and the task-key encoding of the
// wave-local refinement stack.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gcurval {
namespace dev {

// ---------------------------------------------------------------------------
// Gauss-Legendre rules of order 1..5 on [-1, 1] in the node order of
// gcurval/base/constants.inl.h:74-92 (textbook values). The order of the
// node sum is kept so float rounding follows the reference's evaluation.
// ---------------------------------------------------------------------------
template <int N> struct GL;
template <> struct GL<1> {
    static constexpr double x[1] = {0.0};
    static constexpr double w[1] = {2.0};
};
template <> struct GL<2> {
    static constexpr double x[2] = {-0.5773502691896258, 0.5773502691896258};
    static constexpr double w[2] = {1.0, 1.0};
};
template <> struct GL<3> {
    static constexpr double x[3] = {0.0, -0.7745966692414834, 0.7745966692414834};
    static constexpr double w[3] = {0.8888888888888889, 0.5555555555555556, 0.5555555555555556};
};
template <> struct GL<4> {
    static constexpr double x[4] = {-0.33998104358485626, 0.33998104358485626, -0.8611363115940526,
                                    0.8611363115940526};
    static constexpr double w[4] = {0.6521451548625462, 0.6521451548625462, 0.34785484513745385,
                                    0.34785484513745385};
};
template <> struct GL<5> {
    static constexpr double x[5] = {0.0, -0.5384693101056831, 0.5384693101056831, -0.906179845938664,
                                    0.906179845938664};
    static constexpr double w[5] = {0.5688888888888889, 0.47862867049936647, 0.47862867049936647,
                                    0.23692688505618908, 0.23692688505618908};
};

// 8 nodes: the cross-check rule of GL-5 seeded builds (below)
template <> struct GL<8> {
    static constexpr double x[8] = {-0.18343464249564978, 0.18343464249564978, -0.525532409916329,
                                    0.525532409916329,    -0.7966664774136267, 0.7966664774136267,
                                    -0.9602898564975362,  0.9602898564975362};
    static constexpr double w[8] = {0.36268378337836177, 0.36268378337836177, 0.31370664587788705,
                                    0.31370664587788705, 0.22238103445337434, 0.22238103445337434,
                                    0.10122853629037669, 0.10122853629037669};
};

// An evaluator with `static constexpr bool kRule = true` supplies its own
// length rule `length(ctx, x0, x1)` (user plug-ins: the traits'
// CalculateFunction); the built-in curve families are integrated here.
template <class Ev, class = void>
struct HasLengthRule {
    static constexpr bool value = false;
};
template <class Ev>
struct HasLengthRule<Ev, decltype((void)Ev::kRule)> {
    static constexpr bool value = Ev::kRule;
};

// Node j of the N-point rule on [m - k, m + k]: k x_j + m, and m itself for
// the centre node of odd rules (no multiply by zero).
template <int N, int J, typename Dt>
__device__ __forceinline__ Dt gl_node(Dt m, Dt k) {
    if constexpr (GL<N>::x[J] == 0.0)
        return m;
    else
        return k * Dt(GL<N>::x[J]) + m;
}

// sum_j w_j speed(node j), started from the first term rather than 0 + term
template <int N, int J = 0, class Ev, typename Dt>
__device__ __forceinline__ Dt gl_sum(const Ev& ev, const typename Ev::Ctx& c, Dt m, Dt k, Dt acc) {
    if constexpr (J == N) {
        return acc;
    } else {
        const Dt t = Dt(GL<N>::w[J]) * ev.speed(c, gl_node<N, J>(m, k));
        return gl_sum<N, J + 1>(ev, c, m, k, J == 0 ? t : acc + t);
    }
}

// n-point quadrature of the evaluator's speed over [x0, x1] (interval-local
// parameter), in the arithmetic of Dt.
template <int N, class Ev, typename Dt>
__device__ __forceinline__ Dt quad(const Ev& ev, const typename Ev::Ctx& c, Dt x0, Dt x1) {
    if constexpr (HasLengthRule<Ev>::value) return ev.length(c, x0, x1);
    const Dt k = (x1 - x0) * Dt(0.5);
    const Dt m = (x1 + x0) * Dt(0.5);
    return gl_sum<N>(ev, c, m, k, Dt(0)) * k;
}

// The same sum, also folding the node speeds into [mn, mx] (evaluators with
// their own length rule expose no node speeds: mn and mx are left alone).
template <int N, int J = 0, class Ev, typename Dt>
__device__ __forceinline__ Dt gl_sum_mm(const Ev& ev, const typename Ev::Ctx& c, Dt m, Dt k, Dt acc, Dt& mn,
                                        Dt& mx) {
    if constexpr (J == N) {
        return acc;
    } else {
        const Dt v = ev.speed(c, gl_node<N, J>(m, k));
        mn = v < mn ? v : mn;
        mx = v > mx ? v : mx;
        const Dt t = Dt(GL<N>::w[J]) * v;
        return gl_sum_mm<N, J + 1>(ev, c, m, k, J == 0 ? t : acc + t, mn, mx);
    }
}
template <int N, class Ev, typename Dt>
__device__ __forceinline__ Dt quad_mm(const Ev& ev, const typename Ev::Ctx& c, Dt x0, Dt x1, Dt& mn, Dt& mx) {
    if constexpr (HasLengthRule<Ev>::value) return ev.length(c, x0, x1);
    const Dt k = (x1 - x0) * Dt(0.5);
    const Dt m = (x1 + x0) * Dt(0.5);
    return gl_sum_mm<N>(ev, c, m, k, Dt(0), mn, mx) * k;
}

// Dip guard of seeded builds: a segment of 2^-m of its knot interval that is
// wider than a plain root (m < ls_plain) is not accepted as a leaf while the
// speeds at its halves' nodes dip, min / max < 0.05, nor while
// m < min(4, ls_plain) and min / max < 0.2: the halves-versus-whole test can
// pass by coincidence where |C'| nearly vanishes inside a wide segment (a
// kink of the integrand between nodes), and the plain method's 64-way
// pre-split keeps such segments narrow.
template <typename Dt>
__device__ __forceinline__ bool dip_forces(int m, int ls_plain, Dt mn, Dt mx) {
    const int mid = ls_plain < 4 ? ls_plain : 4;
    return (m < ls_plain && mn < Dt(0.05) * mx) || (m < mid && mn < Dt(0.2) * mx);
}

// Cross-check of wide segments in seeded builds. The halves-versus-whole test
// compares two estimates of one rule; on a segment much wider than the
// distance to the integrand's nearest complex singularity (|C'| small next
// to |C''| times the width, not necessarily a dip) the two rule errors can
// agree, so the test passes with both estimates wrong (one dataset interval:
// 1.3e-6 of its curve's length from two roots). A segment wider than
// 2^-kCrossBelow of its interval is therefore accepted as a leaf only when
// L + R also agrees with a higher-order rule over the whole segment (GL-5
// for GL-3 builds, GL-8 for GL-5 builds) to the same tolerance. Narrower
// segments keep the plain method's test alone.
constexpr int kCrossBelow = 3;
// Segments wider than a plain root whose six node speeds vary by 2x or more
// (min < kCrossRatio max) take the check too: on the 100,000 dataset curves
// such segments of 1/8 to 1/32 of their interval, near a slow point (speed
// ratio 0.09-0.31), passed the halves test up to 1e-5 of their curve's
// length off, and the test costs little there (the segments are few).
constexpr float kCrossRatio = 0.5f;
template <typename Dt>
__device__ __forceinline__ bool cross_wanted(int m, int ls_plain, Dt mn, Dt mx) {
    return m < kCrossBelow || (m < ls_plain && mn < Dt(kCrossRatio) * mx);
}
template <int NGL> struct CrossRule;
template <> struct CrossRule<3> { static constexpr int n = 5; };
template <> struct CrossRule<5> { static constexpr int n = 8; };

template <int NGL, class Ev, typename Dt>
__device__ __forceinline__ bool cross_fails(const Ev& ev, const typename Ev::Ctx& c, Dt x0, Dt x1, Dt lr, Dt tol) {
    if constexpr (HasLengthRule<Ev>::value) {
        return false;
    } else {
        const Dt d = quad<CrossRule<NGL>::n>(ev, c, x0, x1) - lr;
        return (d < Dt(0) ? -d : d) > tol;
    }
}

// ---------------------------------------------------------------------------
// wave64 primitives
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63u); }

__device__ __forceinline__ uint64_t ballot(bool p) { return (uint64_t)__ballot(p); }

__device__ __forceinline__ int popc64(uint64_t m) { return __popcll(m); }

// Number of set bits of m in lanes below this lane.
__device__ __forceinline__ int mbcnt(uint64_t m) {
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Load through the scalar data cache (s_load) of read-only data at a
// wave-uniform address: a constant-address-space view of the same memory.
// Scalar loads are counted apart from vector memory operations, so waiting
// for them does not wait for the wave's outstanding vector stores.
template <typename T>
__device__ __forceinline__ T load_s(const T* p) {
    return *(const __attribute__((address_space(4))) T*)p;
}

// A wave-uniform copy of v (the first active lane's), kept in a scalar register.
__device__ __forceinline__ int uniform_int(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ double shfl_up_d(double v, int d) { return __shfl_up(v, (unsigned)d, 64); }
// Lane `src`'s v, for a wave-uniform src (v_readlane, no LDS round trip).
__device__ __forceinline__ double readlane_d(double v, int src) {
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, src);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), src);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ double shfl_d(double v, int src) { return __shfl(v, src, 64); }

// Lane `src`'s copy of a trivially copyable value, one 32-bit word at a time.
template <class T>
__device__ __forceinline__ T shfl_pod(const T& v, int src) {
    static_assert(sizeof(T) % 4 == 0, "shfl_pod needs a multiple of 4 bytes");
    constexpr int kWords = (int)(sizeof(T) / 4);
    int w[kWords];
    __builtin_memcpy(w, &v, sizeof(T));
#pragma unroll
    for (int j = 0; j < kWords; j++) w[j] = __shfl(w[j], src, 64);
    T out;
    __builtin_memcpy(&out, w, sizeof(T));
    return out;
}

// Orders one wave's LDS stores before its later LDS loads of other lanes'
// data. It emits no instruction (a wave's LDS operations complete in issue
// order); the fences and the barrier keep the compiler from moving LDS
// accesses across it.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

// Inclusive scan over aligned segments of 2^log2seg lanes, in double.
__device__ __forceinline__ double seg_scan_incl(double v, int log2seg) {
    const int k = lane_id() & ((1 << log2seg) - 1);
    for (int d = 1; d < (1 << log2seg); d <<= 1) {
        double t = shfl_up_d(v, d);
        if (k >= d) v += t;
    }
    return v;
}

// ---------------------------------------------------------------------------
// Refinement-task key: root lane (6 bits) | level (6 bits) | index (32 bits).
// A task at `level` with index `idx` below root k spans, in root-local units,
// [idx 2^-level, (idx+1) 2^-level]; in interval-local x it is
// [(k 2^level + idx) 2^-(level+log2nsub), ...], exact in float for
// log2nsub + level <= 24 and in double for <= 53.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t task_key(int root, int level, uint32_t idx) {
    return ((uint64_t)root << 38) | ((uint64_t)level << 32) | (uint64_t)idx;
}
__device__ __forceinline__ int key_root(uint64_t k) { return (int)(k >> 38) & 63; }
__device__ __forceinline__ int key_level(uint64_t k) { return (int)(k >> 32) & 63; }
__device__ __forceinline__ uint32_t key_idx(uint64_t k) { return (uint32_t)k; }

// 2^-e, exact (v_ldexp).
__device__ __forceinline__ float pow2m(float, int e) { return ldexpf(1.0f, -e); }
__device__ __forceinline__ double pow2m(double, int e) { return ldexp(1.0, -e); }

// num * 2^-shift, exact while num fits the mantissa.
template <typename Dt>
__device__ __forceinline__ Dt dyadic(uint64_t num, int shift) {
    return Dt(num) * pow2m(Dt(0), shift);
}

}  // namespace dev
}  // namespace gcurval
