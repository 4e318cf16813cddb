// gfx950 kernels of the arc-length engine (templates over the curve
// evaluator, the arithmetic type Dt and the Gauss-Legendre order).
//
//   k_subdivide   adaptive subdivision of every knot interval: one wave64 tile
//                 holds 64 root sub-intervals (n_sub per interval), tests
//                 |L + R - P| > eps(level) for each, refines failing roots
//                 cooperatively through a wave-local task stack with ballot /
//                 mbcnt compaction, and writes each root's inclusive in-interval
//                 prefix (FP64) plus the interval length.
//                 Replaces kn_curveLengthTableInit, kn_curveArclengthForward,
//                 kn_curveArclengthBackward_shfl and kn_mergeSubdividedIntoSeg
//                 (kernel/generic.inl.cuh:10-115, kernel/bfs.inl.cuh:9-125).
//   k_curve_scan  FP64 cumulative interval lengths per curve plus the curve
//                 totals. Replaces kn_mergeSeg and kn_getAllArclength
//                 (kernel/generic.inl.cuh:117-149).
//   k_sample      inverse arc-length map per sample: bracket the interval
//                 (FP64 cumulative) and root (FP64 prefix), re-walk the
//                 adaptive tree for refined roots, then Newton on
//                 int_{x0}^{x} |C'| = tau in Dt with a final FP64 step.
//                 Replaces kn_initSearchTask and kn_arclengthReverseDfsFast_Paper
//                 (kernel/reverse.inl.cuh:32-99, 216-341).
#pragma once

#include "device.h"
#include "curves.h"
#include "features.h"

namespace gcurval {
namespace dev {

// ---------------------------------------------------------------------------
// Subdivision
// ---------------------------------------------------------------------------
enum StatSlot { kStatQuad = 0, kStatDeepRoots = 1, kStatMaxStack = 2, kStatOverflow = 3, kStatTasks = 4,
                kStatDepthCap = 5, kStatRerun = 6, kStatTiles = 7, kStatPasses = 8, kStatCount = 9 };
// Build statistics live in kStatSlots rows of kStatStride counters (one
// 128-byte line each); block b adds into row b % kStatSlots, and readers sum
// (or take the maximum over) the rows. Spreading the rows keeps thousands of
// waves from serialising on one address.
constexpr int kStatSlots = 32;
constexpr int kStatStride = 16;
constexpr int kStatWords = kStatSlots * kStatStride;
// Counters of the seeded build after the statistics rows, zeroed with them
// at the start of every build: per evaluator class the number of intervals
// whose tile outgrew the small task stack (re-run with the full one).
constexpr int kQueueStride = 2;
constexpr int kQueueClasses = 16;
// then the cumulative scan's tile counter and its stall flag
constexpr int kScanTileWord = kStatWords + kQueueStride * kQueueClasses;
constexpr int kScanFailWord = kScanTileWord + 1;
constexpr int kStatBufWords = kScanFailWord + 1;

template <typename Dt>
struct SubdivArgs {
    int64_t n_slots;            // interval slots (n_int, or int_list length)
    const int32_t* int_list;    // optional slot -> interval map
    const int64_t* slot_range;  // optional device {offset, count} into int_list; overrides n_slots
    int log2_nsub;              // roots per interval = 2^log2_nsub (<= 64); root block of interval i at i << log2_nsub
    const int64_t* root_off;    // seeded builds: root block of interval i at root_off[i] (whole-interval,
                                // bins and re-run kernels); null: at i << log2_nsub
    Dt eps0;                    // level-0 tolerance per root (tol / n_sub)
    int max_depth;
    double* root_prefix;        // inclusive prefix of root lengths in each interval (normalised units)
    double* int_len;            // [n_int] interval length (curve units)
    double* curve_len;          // optional: one interval per curve -> also curve length
    uint64_t* deep;             // [n_int] bit k set if root k was refined
    int32_t* int_cost;          // optional: quadratures spent on each interval
    int stack_cap;              // task-stack entries per wave (LDS)
    unsigned long long* stats;
    // optional: a tile whose refinement outgrows stack_cap lists its
    // intervals here (count in *rerun_count) for a re-run with the full stack
    // of 64 (max_depth + 2) entries; null: the overflow flag is raised instead
    int32_t* rerun;
    unsigned long long* rerun_count;
    int rerun_self;             // the launcher runs the re-run kernel itself (uniform builds)
    int ls_plain;               // log2 roots of the plain build and its depth cap, for the re-run
    int base_depth;
};


// Task-stack key: float builds (depth <= 20) fit root | level | index in 32
// bits; double builds use 64 bits.
template <typename Dt> struct KeyOf { typedef uint64_t type; };
template <> struct KeyOf<float> { typedef uint32_t type; };

template <typename K> __device__ __forceinline__ K make_key(int root, int level, uint32_t idx);
template <> __device__ __forceinline__ uint32_t make_key<uint32_t>(int root, int level, uint32_t idx) {
    return ((uint32_t)root << 26) | ((uint32_t)level << 21) | idx;
}
template <> __device__ __forceinline__ uint64_t make_key<uint64_t>(int root, int level, uint32_t idx) {
    return task_key(root, level, idx);
}
__device__ __forceinline__ int key_root_of(uint32_t k) { return (int)(k >> 26) & 63; }
__device__ __forceinline__ int key_level_of(uint32_t k) { return (int)(k >> 21) & 31; }
__device__ __forceinline__ uint32_t key_idx_of(uint32_t k) { return k & 0x1FFFFFu; }
__device__ __forceinline__ int key_root_of(uint64_t k) { return key_root(k); }
__device__ __forceinline__ int key_level_of(uint64_t k) { return key_level(k); }
__device__ __forceinline__ uint32_t key_idx_of(uint64_t k) { return key_idx(k); }

constexpr int kScanSlots = 96;  // LDS slots of one wave's segmented scan

// Bytes of dynamic LDS per wave for a stack of `cap` entries: per-root FP64
// accumulators, the scan slots, per-root task counters, then the stack (keys,
// parent lengths).
template <typename Dt>
__host__ __device__ constexpr int subdiv_lds_per_wave(int cap) {
    return 64 * (int)sizeof(double) + kScanSlots * (int)sizeof(double) + 64 * (int)sizeof(int) +
           cap * (int)(sizeof(typename KeyOf<Dt>::type) + sizeof(Dt));
}

// Inclusive FP64 scan over aligned segments of S = 2^log2seg lanes through
// the wave's kScanSlots LDS slots. Segment s owns slots [s 3S/2, (s+1) 3S/2):
// S/2 zeros, then one slot per lane. A doubling step (d = 1, 2, .., S/2)
// stores every lane's value and adds the slot d below its own, which is one
// of the zeros for the segment's first d lanes, so a step is a store, a load
// and one add, with no lane mask or clamped index. The sums are those of
// seg_scan_incl, in the same order.
struct LdsSegScan {
    __attribute__((address_space(3))) double* slot;
    int steps;
    __device__ __forceinline__ LdsSegScan(void* slots, int log2seg) {
        typedef __attribute__((address_space(3))) double lds_f64;
        lds_f64* base = (lds_f64*)slots;
        const int lane = lane_id();
        const int S = 1 << log2seg, P = S >> 1;
        const int seg = lane >> log2seg, k = lane & (S - 1);
        slot = base + seg * (S + P) + P + k;
        if (k < P) base[seg * (S + P) + k] = 0.0;
        steps = log2seg;
        wave_lds_sync();
    }
    __device__ __forceinline__ double operator()(double v) const {
#pragma unroll
        for (int j = 0; j < 6; j++) {
            if (j < steps) {
                *slot = v;
                wave_lds_sync();
                v += slot[-(1 << j)];
                wave_lds_sync();
            }
        }
        return v;
    }
};

// Depth cap of a root when an interval starts from 2^ls roots instead of the
// plain 2^ls_plain: the finest width (interval / 2^(ls_plain + max_depth)) is
// kept, within the float key / exact-coordinate limits (level <= 20 and
// ls + level <= 24 in float; ls + level <= 53 and level <= 31 in double).
__host__ __device__ inline int effective_depth(int max_depth, int ls_plain, int ls, bool fp32) {
    int d = max_depth + (ls_plain > ls ? ls_plain - ls : 0);
    const int lim = fp32 ? (20 < 24 - ls ? 20 : 24 - ls) : (31 < 53 - ls ? 31 : 53 - ls);
    return d < lim ? d : lim;
}

// Bounds of task (root kr, level l, index idx) in interval-local x; exact.
template <typename Dt>
__device__ __forceinline__ void task_bounds(int kr, int l, uint32_t idx, int ls, Dt& x0, Dt& xm, Dt& x1) {
    if constexpr (sizeof(Dt) == 4) {
        // FP32 depths are capped at ls + l <= 24 (effective_depth), so the
        // numerators fit 32 bits and convert in one instruction
        const uint32_t base = ((uint32_t)kr << l) + idx;
        x0 = Dt(base) * pow2m(Dt(0), ls + l);
        x1 = Dt(base + 1u) * pow2m(Dt(0), ls + l);
        xm = Dt(2u * base + 1u) * pow2m(Dt(0), ls + l + 1);
        return;
    }
    const uint64_t base = ((uint64_t)kr << l) + idx;
    x0 = dyadic<Dt>(base, ls + l);
    x1 = dyadic<Dt>(base + 1, ls + l);
    xm = dyadic<Dt>(2 * base + 1, ls + l + 1);
}

template <typename Dt>
__device__ __forceinline__ Dt dabs(Dt v) { return v < Dt(0) ? -v : v; }

// Tolerance floor of segments no wider than a plain root (ls + l >= ls_plain)
// inside roots wider than the plain method's (ls < ls_plain: seeded and
// uniform-k builds). Their tolerance per width is 2^(ls_plain - ls) times
// tighter than the plain method's, and |L + R - P| cannot fall below the
// rounding of the sums (FP32, one root per interval: below the tolerance only
// where the speed is under about 0.2 in normalised units), so without a floor
// the test chases rounding noise down to the depth cap (a float32 simulation
// on 49,657 dataset intervals from one root each: 298M quadratures, 7.5M
// with the floor, 9.5M for the plain method; lengths unchanged). Such a
// segment is accepted once |L + R - P| is within 4 units of rounding of L + R
// (2^-21 relative in FP32, 2^-50 in FP64). Wider segments keep the
// tolerance alone: there |L + R - P| is no reliable bound on the error of
// L + R, which the sampler's inversion inside the leaf carries.
template <typename Dt>
__device__ __forceinline__ Dt rounding_floor(int ls, int l, int ls_plain) {
    return (ls < ls_plain && ls + l >= ls_plain) ? (sizeof(Dt) == 4 ? Dt(0x1p-21) : Dt(0x1p-50)) : Dt(0);
}
template <typename Dt>
__device__ __forceinline__ Dt tol_of(Dt eps, Dt rel, Dt sum) {
    const Dt f = rel * dabs(sum);
    return f > eps ? f : eps;
}

__device__ __forceinline__ int clamp_k_seed(int k) { return k < 0 ? 0 : (k > 6 ? 6 : k); }

__device__ __forceinline__ int64_t readfirstlane64(int64_t v) {
    return (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)((uint64_t)v >> 32)) << 32) |
                     (uint32_t)__builtin_amdgcn_readfirstlane((int)v));
}

// List the intervals of a tile whose refinement outgrew the stack (one lane
// each: the last root of every interval of the tile) for the full-stack
// re-run; v: the lane's slot holds an interval.
__device__ __forceinline__ void list_for_rerun(bool v, int i, int ls, int32_t* rerun,
                                               unsigned long long* rerun_count) {
    const int lane = lane_id();
    const bool mine = v && ((lane & ((1 << ls) - 1)) == (1 << ls) - 1);
    const uint64_t m = ballot(mine);
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(rerun_count, (unsigned long long)popc64(m));
    base = (unsigned long long)readfirstlane64((int64_t)__shfl((long long)base, 0, 64));
    if (mine) rerun[base + mbcnt(m)] = i;
}

// A wave's refinement-task stack in LDS: keys and parent lengths.
template <typename Dt>
struct TaskStack {
    typedef typename KeyOf<Dt>::type Key;
    Key* key;
    Dt* par;
    int cap;
};

// Per-wave counters of a subdivision launch, added into the statistics rows
// when the wave ends.
struct WaveCounters {
    unsigned long long n_quad = 0, n_deep = 0, n_tasks = 0, n_cap = 0, n_rerun = 0, n_tiles = 0, n_pass = 0;
    int max_stack = 0, overflow = 0;
    __device__ __forceinline__ void flush(unsigned long long* stats) const {
        if (lane_id() != 0) return;
        unsigned long long* st = stats + (blockIdx.x & (kStatSlots - 1)) * kStatStride;
        if (n_quad) atomicAdd(&st[kStatQuad], n_quad);
        if (n_deep) atomicAdd(&st[kStatDeepRoots], n_deep);
        if (n_tasks) atomicAdd(&st[kStatTasks], n_tasks);
        if (n_cap) atomicAdd(&st[kStatDepthCap], n_cap);
        if (n_rerun) atomicAdd(&st[kStatRerun], n_rerun);
        if (n_tiles) atomicAdd(&st[kStatTiles], n_tiles);
        if (n_pass) atomicAdd(&st[kStatPasses], n_pass);
        if (max_stack) atomicMax(&st[kStatMaxStack], (unsigned long long)max_stack);
        if (overflow) atomicMax(&st[kStatOverflow], 1ull);
    }
};

// The wave's LDS of a subdivision kernel: per-root FP64 accumulators, the
// scan slots, per-root task counters, then the stack.
template <typename Dt>
struct SubdivLds {
    double* acc;
    unsigned char* scan;
    int* cnt;
    TaskStack<Dt> stack;
    __device__ __forceinline__ SubdivLds(unsigned char* wbase, const SubdivArgs<Dt>& A) {
        typedef typename KeyOf<Dt>::type Key;
        acc = (double*)wbase;
        scan = wbase + 64 * sizeof(double);
        unsigned char* rest = scan + kScanSlots * sizeof(double);
        cnt = (int*)rest;
        stack.key = (Key*)(rest + 64 * sizeof(int));
        stack.par = (Dt*)(rest + 64 * sizeof(int) + A.stack_cap * sizeof(Key));
        stack.cap = A.stack_cap;
    }
};

// One wave tile: 64 roots, 2^ls per interval (64 / 2^ls intervals; kWhole:
// one interval, ls = 6). Each lane evaluates its root's whole (P) and halves
// (L, R) and tests |L + R - P| > eps0; failing roots are refined through the
// task stack (pop the top <= 64 tasks, push 2 children per failing task at
// ballot / mbcnt offsets, leaves add into per-root FP64 accumulators); then
// the FP64 in-interval prefix of the root lengths, the interval length and
// the refined-root mask are stored. `ctx` is lane r's interval context for
// root r; `rb` the lane's root block, root k of it at rb + k. Returns true
// (wave-uniform) if the refinement outgrew the stack: the tile's results are
// then incomplete and the caller has it re-run with a larger stack.
// kGuard (seeded builds): segments wider than a plain root (2^-ls_plain of
// the interval) whose halves' node speeds dip are refined whatever the test
// says (dip_forces).
template <typename Dt, int NGL, class Ev, bool kWhole, bool kGuard = false>
__device__ __forceinline__ bool subdivide_tile(const Ev& ev, const SubdivArgs<Dt>& A, const typename Ev::Ctx& ctx,
                                               int i, bool valid, int64_t rb, double scale_i, int ls,
                                               int max_depth, const SubdivLds<Dt>& W, const LdsSegScan& scan,
                                               WaveCounters& wc, int ls_plain = 6) {
    typedef typename KeyOf<Dt>::type Key;
    const int lane = lane_id();
    const int nsub = 1 << ls;
    const int ipw_log2 = 6 - ls;
    const int k = lane & (nsub - 1);
    const TaskStack<Dt>& stk = W.stack;
    // tolerance of a task at level l (whole-interval tiles hold plain roots:
    // no floor, and no code for it)
    auto tol = [&](int l, Dt sum) -> Dt {
        if constexpr (kWhole)
            return A.eps0 * pow2m(Dt(0), l);
        else
            return tol_of(A.eps0 * pow2m(Dt(0), l), rounding_floor<Dt>(ls, l, ls_plain), sum);
    };

    Dt x0, xm, x1;
    task_bounds<Dt>(k, 0, 0u, ls, x0, xm, x1);
    const Dt P = quad<NGL>(ev, ctx, x0, x1);
    Dt L, R;
    bool dip = false;
    if constexpr (kGuard) {
        Dt mn = Dt(3.0e38f), mx = Dt(0);
        L = quad_mm<NGL>(ev, ctx, x0, xm, mn, mx);
        R = quad_mm<NGL>(ev, ctx, xm, x1, mn, mx);
        dip = dip_forces(ls, ls_plain, mn, mx);
        if (ls < ls_plain) {
            // only where the other tests accept the root (the same decision)
            const bool need =
                valid && !dip && !(dabs((L + R) - P) > A.eps0) && cross_wanted(ls, ls_plain, mn, mx);
            if (ballot(need)) {
                // bounds through an empty asm: the check's node positions are
                // formed here per tile, not hoisted into registers held for
                // the whole kernel
                Dt cx0 = x0, cx1 = x1;
                asm volatile("" : "+v"(cx0), "+v"(cx1));
                if (need) dip = cross_fails<NGL>(ev, ctx, cx0, cx1, L + R, A.eps0);
            }
        }
    } else {
        L = quad<NGL>(ev, ctx, x0, xm);
        R = quad<NGL>(ev, ctx, xm, x1);
    }
    Dt len = L + R;
    const bool fail = valid && max_depth > 0 && (dabs(len - P) > A.eps0 || dip);
    wc.n_quad += 3ull * popc64(ballot(valid));
    wc.n_tiles++;
    const uint64_t fmask = ballot(fail);
    bool overflowed = false;

    if (fmask) {
        wc.n_deep += popc64(fmask);
        W.acc[lane] = 0.0;
        W.cnt[lane] = 0;
        const int pos = 2 * mbcnt(fmask);
        if (fail) {  // <= 128 entries: always in LDS (stack_cap >= 128)
            stk.key[pos] = make_key<Key>(lane, 1, 0u);
            stk.par[pos] = L;
            stk.key[pos + 1] = make_key<Key>(lane, 1, 1u);
            stk.par[pos + 1] = R;
        }
        int size = 2 * popc64(fmask);
        if (size > wc.max_stack) wc.max_stack = size;
        while (size > 0) {
            wc.n_pass++;
            const int n = size < 64 ? size : 64;
            const bool take = lane < n;
            Key kk = 0;
            Dt par = Dt(0);
            if (take) {
                kk = stk.key[size - 1 - lane];
                par = stk.par[size - 1 - lane];
            }
            size -= n;
            const int r = key_root_of(kk);
            const int l = key_level_of(kk);
            const uint32_t idx = key_idx_of(kk);
            // lane r holds the context of root r's interval
            const typename Ev::Ctx cr = (kWhole || ipw_log2 == 0) ? ctx : shfl_pod(ctx, r);
            Dt a0, am, a1;
            task_bounds<Dt>(r & (nsub - 1), l, idx, ls, a0, am, a1);
            Dt L2, R2;
            bool dip2 = false;
            Dt mn2 = Dt(3.0e38f), mx2 = Dt(0);
            if constexpr (kGuard) {
                L2 = quad_mm<NGL>(ev, cr, a0, am, mn2, mx2);
                R2 = quad_mm<NGL>(ev, cr, am, a1, mn2, mx2);
                dip2 = dip_forces(ls + l, ls_plain, mn2, mx2);
            } else {
                L2 = quad<NGL>(ev, cr, a0, am);
                R2 = quad<NGL>(ev, cr, am, a1);
            }
            const Dt lr = L2 + R2;
            const Dt tol2 = tol(l, lr);
            const bool over = dabs(lr - par) > tol2;
            if constexpr (kGuard) {
                const bool wide = take && !over && !dip2 && cross_wanted(ls + l, ls_plain, mn2, mx2);
                if (ballot(wide)) {
                    if (wide) dip2 = cross_fails<NGL>(ev, cr, a0, a1, lr, tol2);
                }
            }
            bool f2 = take && l < max_depth && (over || dip2);
            if (take && l >= max_depth && over) wc.n_cap++;  // a leaf that still fails: the depth cap
            const uint64_t tmask = ballot(take);
            wc.n_quad += 2ull * popc64(tmask);
            wc.n_tasks += popc64(tmask);
            if (A.int_cost && take) atomicAdd(&W.cnt[r], 1);
            uint64_t m2 = ballot(f2);
            if (size + 2 * popc64(m2) > stk.cap) {
                // The children do not fit: the tile stops refining (its
                // results are marked incomplete) and never writes past the
                // stack. A stack of 64 (max_depth + 1) entries always fits.
                overflowed = true;
                f2 = false;
                m2 = 0;
            }
            if (f2) {
                const int p2 = size + 2 * mbcnt(m2);
                stk.key[p2] = make_key<Key>(r, l + 1, 2u * idx);
                stk.par[p2] = L2;
                stk.key[p2 + 1] = make_key<Key>(r, l + 1, 2u * idx + 1u);
                stk.par[p2 + 1] = R2;
            } else if (take) {
                atomicAdd(&W.acc[r], (double)lr);
            }
            size += 2 * popc64(m2);
            if (size > wc.max_stack) wc.max_stack = size;
        }
        if (fail) len = (Dt)W.acc[lane];
    }

    // in-interval inclusive prefix over the n_sub roots (FP64)
    const double pre = scan(valid ? (double)len : 0.0);
    double cost = 0.0;
    if (A.int_cost) cost = scan(valid ? 3.0 + (fail ? 2.0 * W.cnt[lane] : 0.0) : 0.0);
    // Stores without branches: every lane stores, a lane with nothing of
    // its own to write repeats the store of a lane that has (same address,
    // same value). A fixed number of vector memory operations per tile
    // lets the compiler wait for the prefetched loads alone (it waits for
    // every outstanding operation where the count depends on a branch).
    // Interval values come from the interval's last root; a lane past the
    // end of the slot list takes lane 0's interval (always valid).
    const int own = valid ? lane : 0;                  // lane whose root prefix this lane stores
    const int last = (own & ~(nsub - 1)) + nsub - 1;   // last root of that lane's interval
    const double pre_o = kWhole ? pre : shfl_d(pre, own);
    const int64_t rb_o = kWhole ? rb : (int64_t)__shfl((long long)rb, own, 64);
    A.root_prefix[rb_o + (own & (nsub - 1))] = pre_o;
    const double pre_l = kWhole ? readlane_d(pre, 63) : shfl_d(pre, last);
    const double cost_l = kWhole ? readlane_d(cost, 63) : shfl_d(cost, last);
    const int i_l = kWhole ? i : __shfl(i, last, 64);
    const double L_int = pre_l * (kWhole ? scale_i : shfl_d(scale_i, last));
    A.int_len[i_l] = L_int;
    // optional outputs: a duplicate store into the same word of int_len
    // (curve lengths) or into the low word of deep (quadrature counts,
    // overwritten below). deep is written as two 32-bit words, a type the
    // count's store may alias, so the compiler keeps the two stores in order.
    double* const cl = A.curve_len ? A.curve_len : A.int_len;
    cl[i_l] = L_int;
    uint32_t* const dw = (uint32_t*)(A.deep + i_l);
    int32_t* const ic = A.int_cost ? A.int_cost + i_l : (int32_t*)dw;
    *ic = (int32_t)cost_l;
    const uint64_t bits = (ls == 6) ? fmask : ((fmask >> ((last >> ls) << ls)) & ((1ull << nsub) - 1ull));
    dw[0] = (uint32_t)bits;
    dw[1] = (uint32_t)(bits >> 32);
    return overflowed;
}

// Evaluators that can load the context and scale of a wave-uniform interval
// through scalar loads (load_s(i, ctx, scale)).
template <class Ev, class = void>
struct HasScalarLoad : std::false_type {};
template <class Ev>
struct HasScalarLoad<Ev, decltype((void)std::declval<const Ev&>().load_s(0, std::declval<typename Ev::Ctx&>(),
                                                                        std::declval<double&>()))>
    : std::true_type {};

// The wave-local stack holds at most 64 (max_depth + 1) tasks: each pop takes
// the top 64 entries and pushes their children on top, so the blocks left
// below the top hold <= 64 entries each and have strictly increasing minimum
// levels (1 .. max_depth), and the top block holds <= 128.
//
// Tiles of 2^log2_nsub roots per interval over the slots [0, n_slots) (of
// int_list when given), each wave striding over tiles. kWhole: every tile is
// one whole interval of 64 roots (log2_nsub = 6), so the interval, its
// context, scale and root base are wave-uniform: they come through scalar
// loads where the evaluator has them, and the tile loop issues no vector
// loads but the coefficients' (B-splines). Otherwise a tile holds
// 64 / 2^log2_nsub intervals and each lane loads its own.
template <typename Dt, int NGL, class Ev, bool kWhole>
__global__ __launch_bounds__(256) void k_subdivide(Ev ev, SubdivArgs<Dt> A) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = lane_id();
    const int wib = uniform_int((int)(threadIdx.x >> 6));
    const int64_t gwave = (int64_t)blockIdx.x * (blockDim.x >> 6) + wib;
    const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x >> 6);
    const SubdivLds<Dt> W(smem + (size_t)wib * subdiv_lds_per_wave<Dt>(A.stack_cap), A);
    const int32_t* list = A.int_list;
    int64_t n_slots = A.n_slots;
    if (A.slot_range) {
        list = A.int_list + load_s(A.slot_range);
        n_slots = load_s(A.slot_range + 1);
    }
    const int ls = kWhole ? 6 : A.log2_nsub;
    const int ipw_log2 = 6 - ls;  // interval slots per wave tile (log2)
    const int jslot = lane >> ls;
    const int64_t n_tiles = (n_slots + (1 << ipw_log2) - 1) >> ipw_log2;
    const LdsSegScan scan(W.scan, ls);
    WaveCounters wc;

    // Software pipeline over this wave's tiles: the interval id of tile t + 2
    // and the evaluator context, root base and scale of tile t + 1 are in
    // flight while tile t computes, so the dependent loads (list ->
    // coefficients) overlap the quadratures instead of stalling every tile.
    // (one unconditional-address load per form: a conditional load whose
    // register another path writes directly makes the compiler wait for every
    // load in flight, the prefetches included)
    auto interval_of = [&](int64_t tl, bool& ok) -> int {
        const int64_t slot = (tl << ipw_log2) + jslot;
        ok = slot < n_slots;
        const int64_t s = ok ? slot : 0;
        int r = (int)s;
        if (list) r = kWhole ? load_s(list + s) : list[s];
        return r;
    };
    // A stage: what tile t needs from memory (interval, context, scale,
    // root block).
    struct Stage {
        typename Ev::Ctx ctx;
        double sc;
        int64_t rb;
        int i;
        bool v;
    };
    // unconditional loads: past the last tile iv is still an interval of the
    // list (its first slot), so they stay in range and are simply unused
    auto fetch = [&](int iv, bool ok) -> Stage {
        Stage st;
        if constexpr (kWhole && HasScalarLoad<Ev>::value) {
            ev.load_s(iv, st.ctx, st.sc);
        } else {
            st.ctx = ev.load(iv);
            st.sc = ev.scale(iv);
        }
        st.rb = (int64_t)iv << ls;
        if constexpr (kWhole) {  // wave-uniform: a scalar load (a vector one would need a wave-wide wait)
            if (A.root_off) st.rb = load_s(A.root_off + iv);
        }
        st.i = iv;
        st.v = ok;
        return st;
    };

    // one copy of the tile body: results must not depend on which tiles a
    // wave takes, and two inlined copies may contract their FMAs differently
    Stage cur{};
    bool v_nxt = false;
    int i_nxt = 0;
    if (gwave < n_tiles) {
        bool ok;
        const int iv = interval_of(gwave, ok);
        cur = fetch(iv, ok);
        i_nxt = interval_of(gwave + nwaves, v_nxt);
    }
    for (int64_t tile = gwave; tile < n_tiles; tile += nwaves) {
        const Stage nxt = fetch(i_nxt, v_nxt);
        i_nxt = interval_of(tile + 2 * nwaves, v_nxt);
        // a tile that outgrows a small stack is listed for the full-stack
        // re-run (the statistics count the work of both passes); with the
        // full stack of 64 (max_depth + 2) entries it cannot overflow
        if (subdivide_tile<Dt, NGL, Ev, kWhole>(ev, A, cur.ctx, cur.i, cur.v, cur.rb, cur.sc, ls, A.max_depth, W,
                                                scan, wc, A.ls_plain)) {
            if (A.rerun) {
                list_for_rerun(cur.v, cur.i, ls, A.rerun, A.rerun_count);
                wc.n_rerun++;
            } else {
                wc.overflow = 1;
            }
        }
        cur = nxt;
    }
    wc.flush(A.stats);
}

// ---------------------------------------------------------------------------
// Seeded subdivision: interval i starts from 2^k_i roots (k_i in root_k);
// the roots are stored densely in interval order, interval i's block at
// root_off[i] = sum of 2^k_j over j < i (written by the bucketing). The
// bucketing sorts a class's intervals by k; k_subdivide_bins
// runs the class's bins k = 0..5 as one tile sequence (bin k's tiles hold
// 64 / 2^k intervals, the last one possibly partial), and the whole-interval
// kernel runs bin 6. The bins kernel keeps a small task stack, so more waves
// fit in LDS; a tile whose refinement outgrows it lists its intervals, which
// k_subdivide_rerun then runs again with the full stack.
// ---------------------------------------------------------------------------
constexpr int kSeedPartBins = 6;  // k = 0..5

struct BinRun {
    const int64_t* ranges;  // [2 * kSeedPartBins] {first slot, count} of this class's bins k = 0..5 (device)
    int ls_plain;           // log2 roots of the plain build (depth caps)
    int base_depth;         // depth cap of the plain build
    int32_t* rerun;         // out: intervals of tiles that outgrew the stack
    unsigned long long* rerun_count;
};

// The tiles of bin K (2^K roots per interval, 64 / 2^K intervals per tile)
// that fall to this wave. The bins' tiles form one sequence, tile T going to
// wave T mod nwaves; `t_begin` is bin K's first tile in the sequence and is
// advanced past the bin. K is a compile-time constant, so the root count,
// masks and depth cap of the tile body are constants too. Software pipeline
// as in k_subdivide: interval ids of tile t + 2, contexts of tile t + 1 in
// flight while tile t computes.
template <int K, typename Dt, int NGL, class Ev>
__device__ __forceinline__ void subdivide_bin(const Ev& ev, const SubdivArgs<Dt>& A, const BinRun& B,
                                              const SubdivLds<Dt>& W, int64_t gwave, int64_t nwaves,
                                              int64_t& t_begin, WaveCounters& wc) {
    constexpr bool kF32 = sizeof(Dt) == 4;
    constexpr int ipw_log2 = 6 - K;  // intervals per tile (log2)
    const int lane = lane_id();
    const int64_t first = load_s(B.ranges + 2 * K);
    const int64_t count = load_s(B.ranges + 2 * K + 1);
    const int64_t n_tiles = (count + (1 << ipw_log2) - 1) >> ipw_log2;
    int64_t t0 = (gwave - t_begin) % nwaves;  // this wave's first tile of the bin
    if (t0 < 0) t0 += nwaves;
    t_begin += n_tiles;
    if (t0 >= n_tiles) return;
    const int jslot = lane >> K;
    const LdsSegScan scan(W.scan, K);
    const int max_depth = effective_depth(B.base_depth, B.ls_plain, K, kF32);
    const int32_t* list = A.int_list + first;
    auto interval_of = [&](int64_t tl, bool& ok) -> int {
        const int64_t slot = (tl << ipw_log2) + jslot;
        ok = slot < count;
        return list[ok ? slot : 0];
    };
    struct Stage {
        typename Ev::Ctx ctx;
        double sc;
        int64_t rb;
        int i;
        bool v;
    };
    auto fetch = [&](int iv, bool ok) -> Stage {
        Stage st;
        st.ctx = ev.load(iv);
        st.sc = ev.scale(iv);
        st.rb = A.root_off[iv];
        st.i = iv;
        st.v = ok;
        return st;
    };
    bool ok0;
    Stage cur = fetch(interval_of(t0, ok0), ok0);
    bool v_nxt;
    int i_nxt = interval_of(t0 + nwaves, v_nxt);
    for (int64_t tile = t0; tile < n_tiles; tile += nwaves) {
        const Stage nxt = fetch(i_nxt, v_nxt);
        i_nxt = interval_of(tile + 2 * nwaves, v_nxt);
        if (subdivide_tile<Dt, NGL, Ev, false, true>(ev, A, cur.ctx, cur.i, cur.v, cur.rb, cur.sc, K, max_depth, W,
                                                     scan, wc, B.ls_plain)) {
            // list the tile's intervals (one lane each) for the full-stack re-run
            list_for_rerun(cur.v, cur.i, K, B.rerun, B.rerun_count);
            wc.n_rerun++;
        }
        cur = nxt;
    }
}

template <typename Dt, int NGL, class Ev>
__global__ __launch_bounds__(256) void k_subdivide_bins(Ev ev, SubdivArgs<Dt> A, BinRun B) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int wib = uniform_int((int)(threadIdx.x >> 6));
    const int64_t gwave = (int64_t)blockIdx.x * (blockDim.x >> 6) + wib;
    const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x >> 6);
    const SubdivLds<Dt> W(smem + (size_t)wib * subdiv_lds_per_wave<Dt>(A.stack_cap), A);
    WaveCounters wc;
    int64_t t_begin = 0;
    subdivide_bin<0, Dt, NGL>(ev, A, B, W, gwave, nwaves, t_begin, wc);
    subdivide_bin<1, Dt, NGL>(ev, A, B, W, gwave, nwaves, t_begin, wc);
    subdivide_bin<2, Dt, NGL>(ev, A, B, W, gwave, nwaves, t_begin, wc);
    subdivide_bin<3, Dt, NGL>(ev, A, B, W, gwave, nwaves, t_begin, wc);
    subdivide_bin<4, Dt, NGL>(ev, A, B, W, gwave, nwaves, t_begin, wc);
    subdivide_bin<5, Dt, NGL>(ev, A, B, W, gwave, nwaves, t_begin, wc);
    wc.flush(A.stats);
}

// Everything one class's seeded subdivision needs: the bins launch (small
// stack), the whole-interval launch over bin 6, and the re-run launch.
template <typename Dt>
struct SeededLaunch {
    SubdivArgs<Dt> bins;     // int_list = bucket list, n_slots = the class's interval count
    BinRun run;
    SubdivArgs<Dt> whole;    // int_list = bucket list, slot_range = bin 6's range
    SubdivArgs<Dt> rerun;    // full stack
    const int8_t* root_k;
};

// The intervals listed by k_subdivide_bins or k_subdivide, one wave tile
// each (root count 2^root_k[i], or 2^A.log2_nsub without root_k; the lanes
// past it idle), with the full task stack in LDS. kGuard: seeded builds
// (dip_forces), as the kernels whose tiles are re-run.
template <typename Dt, int NGL, class Ev, bool kGuard = true>
__global__ __launch_bounds__(64) void k_subdivide_rerun(Ev ev, SubdivArgs<Dt> A, const int32_t* list,
                                                        const unsigned long long* count, const int8_t* root_k,
                                                        int ls_plain, int base_depth) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = lane_id();
    const SubdivLds<Dt> W(smem, A);
    constexpr bool kF32 = sizeof(Dt) == 4;
    const int64_t n = (int64_t)load_s(count);
    WaveCounters wc;
    for (int64_t j = blockIdx.x; j < n; j += gridDim.x) {
        const int i = load_s(list + j);
        const int k = root_k ? (int)load_s(root_k + i) : A.log2_nsub;
        const int64_t rb = A.root_off ? load_s(A.root_off + i) : (int64_t)i << k;
        const typename Ev::Ctx ctx = ev.load(i);
        const double sc = ev.scale(i);
        const LdsSegScan scan(W.scan, k);
        if (subdivide_tile<Dt, NGL, Ev, false, kGuard>(ev, A, ctx, i, lane < (1 << k), rb, sc, k,
                                                       effective_depth(base_depth, ls_plain, k, kF32), W, scan, wc,
                                                       ls_plain))
            wc.overflow = 1;  // unreachable: the stack holds 64 (depth + 2) entries
    }
    wc.flush(A.stats);
}

// ---------------------------------------------------------------------------
// Per-curve FP64 cumulative scan (wave per curve), in place on int_cum.
// ---------------------------------------------------------------------------
static __global__ __launch_bounds__(256) void k_curve_scan(int n_curve, const int64_t* ioff, double* int_cum,
                                                    double* curve_len) {
    const int lane = lane_id();
    const int64_t gwave = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t c = gwave; c < n_curve; c += nwaves) {
        const int64_t b = ioff[c], e = ioff[c + 1];
        double carry = 0.0, last = 0.0;
        // 4 x 64 lengths in flight per iteration, then scanned in order
        for (int64_t base = b; base < e; base += 256) {
            double v[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int64_t j = base + 64 * u + lane;
                v[u] = j < e ? int_cum[j] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int64_t j = base + 64 * u + lane;
                const double w = seg_scan_incl(v[u], 6) + carry;
                if (j < e) int_cum[j] = w;
                carry = shfl_d(w, 63);
                // the curve length is its last interval's value itself (lanes
                // past the end sum the same terms in another order)
                const int64_t k = e - 1 - (base + 64 * u);
                if (k >= 0 && k < 64) last = shfl_d(w, (int)k);
            }
        }
        if (lane == 0) curve_len[c] = last;
    }
}

// ---------------------------------------------------------------------------
// Cumulative interval lengths of every curve in one pass over all intervals
// (a segmented inclusive FP64 scan, curves as segments), in place on
// int_cum, and each curve's length from its last interval. Its time depends
// on the interval count only, not on how the intervals split into curves.
// Single pass with decoupled look-back: a block takes the next tile of
// kScanTile intervals from a counter (so every tile it waits for belongs to
// a block that is already running), scans it, publishes its aggregate (the
// sum since its last curve start) and then, once the carry from the tiles
// before it is known, its inclusive prefix; the carry is the prefix of the
// nearest tile that has one (a tile holding a curve start has it at once)
// plus the aggregates after it, added left to right (the prefixes are sums
// in that same order, so the result does not depend on which tiles had
// published yet: builds are bit-reproducible).
// Tile words are device-coherent atomics (relaxed, no cache maintenance):
// a value is stored, the store is waited for, then the tile's status word
// {epoch, kind, has-start} is stored; a reader loads the status, then the
// value. The epoch (one per build) tells this build's words from stale
// ones, so the status array needs no clearing. A wait gives up after
// kScanMaxSpin polls and raises the stall flag (the host then reports the
// build as failed) instead of hanging.
// The tile moves through LDS twice: coalesced loads in, each thread's
// kScanItems consecutive intervals out of LDS for its serial scan, and the
// same way back for coalesced stores.
// ---------------------------------------------------------------------------
constexpr int kScanItems = 8;                      // intervals per thread
constexpr int kScanThreads = 512;
constexpr int kScanTile = kScanThreads * kScanItems;  // intervals per block tile
constexpr unsigned kScanMaxSpin = 1u << 20;
enum : uint64_t { kScanAgg = 1, kScanPrefix = 2, kScanHead = 4 };

struct ScanTiles {
    unsigned long long* status;  // [n_tiles]: epoch << 32 | kind | has-start
    double* agg;                 // [n_tiles] sum since the tile's last curve start
    double* prefix;              // [n_tiles] inclusive prefix of the tile's open curve
};

// (has a curve start, value) pairs under the segmented sum
struct SegVal {
    int f;
    double v;
};
__device__ __forceinline__ SegVal seg_add(SegVal a, SegVal b) { return SegVal{a.f | b.f, b.f ? b.v : a.v + b.v}; }

__device__ __forceinline__ SegVal wave_seg_scan_incl(SegVal x) {
    const int lane = lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int f = __shfl_up(x.f, o, 64);
        const double v = shfl_up_d(x.v, o);
        if (lane >= o) x = seg_add(SegVal{f, v}, x);
    }
    return x;
}

__device__ __forceinline__ void scan_put(double* p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double scan_get(const double* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// the value stores before it have completed when the status store issues
__device__ __forceinline__ void scan_publish(unsigned long long* st, unsigned long long v) {
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
    __hip_atomic_store(st, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// LDS slot of tile element e: one pad word per 32 keeps the blocked reads
// (thread t, elements 8t .. 8t + 7) free of bank conflicts
__device__ __forceinline__ int scan_slot(int e) { return e + (e >> 5); }

static __global__ __launch_bounds__(kScanThreads) void k_seg_scan(int64_t n, const int32_t* int_curve, double* int_cum,
                                                         double* curve_len, ScanTiles T, unsigned epoch,
                                                         unsigned long long* words) {
    __shared__ double s_x[kScanTile + kScanTile / 32];
    __shared__ int32_t s_c[kScanTile + kScanTile / 32];
    __shared__ SegVal s_wave[kScanThreads / 64];
    __shared__ int64_t s_tile;
    __shared__ double s_carry;
    __shared__ int32_t s_edge[2];  // curves of the intervals before and after the tile
    const int tid = (int)threadIdx.x, lane = lane_id(), w = tid >> 6;
    if (tid == 0) s_tile = (int64_t)atomicAdd(words + kScanTileWord, 1ull);
    __syncthreads();
    const int64_t tile = s_tile;
    const int64_t t0 = tile * kScanTile;
    // coalesced loads into LDS
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        const int e = k * kScanThreads + tid;
        const int64_t i = t0 + e;
        s_x[scan_slot(e)] = i < n ? int_cum[i] : 0.0;
        s_c[scan_slot(e)] = i < n ? int_curve[i] : -1;
    }
    if (tid == 0) s_edge[0] = t0 > 0 ? int_curve[t0 - 1] : -1;
    if (tid == 1) s_edge[1] = t0 + kScanTile < n ? int_curve[t0 + kScanTile] : -1;
    __syncthreads();
    // this thread's intervals: lengths, curve starts, and curve ends
    double x[kScanItems];
    int head[kScanItems], tail[kScanItems];
    {
        const int e0 = tid * kScanItems;
        int32_t prev = e0 > 0 ? s_c[scan_slot(e0 - 1)] : s_edge[0];
#pragma unroll
        for (int k = 0; k < kScanItems; k++) {
            const int e = e0 + k;
            const int32_t c = s_c[scan_slot(e)];
            const int32_t nx = e + 1 < kScanTile ? s_c[scan_slot(e + 1)] : s_edge[1];
            x[k] = s_x[scan_slot(e)];
            head[k] = c >= 0 && c != prev;
            tail[k] = c >= 0 && c != nx;
            prev = c;
        }
    }
    // serial scan of the thread's items
    SegVal t{0, 0.0};
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        t = seg_add(t, SegVal{head[k], x[k]});
        x[k] = t.v;
    }
    // block scan of the thread totals
    const SegVal incl = wave_seg_scan_incl(t);
    const int fx = __shfl_up(incl.f, 1, 64);
    const double vx = shfl_up_d(incl.v, 1);
    SegVal excl = lane ? SegVal{fx, vx} : SegVal{0, 0.0};  // within the wave
    if (lane == 63) s_wave[w] = incl;
    __syncthreads();
    SegVal before{0, 0.0}, tile_agg{0, 0.0};
#pragma unroll
    for (int q = 0; q < kScanThreads / 64; q++) {
        if (q < w) before = seg_add(before, s_wave[q]);
        tile_agg = seg_add(tile_agg, s_wave[q]);
    }
    excl = seg_add(before, excl);  // from the tile start to this thread
    // publish, look back (wave 0: 64 predecessors per poll), publish the prefix
    if (w == 0) {
        const uint64_t ep = (uint64_t)epoch << 32, hf = tile_agg.f ? kScanHead : 0;
        if (lane == 0) {
            scan_put(T.agg + tile, tile_agg.v);
            if (tile_agg.f) scan_put(T.prefix + tile, tile_agg.v);
            scan_publish(T.status + tile, ep | hf | (tile_agg.f ? kScanPrefix : kScanAgg));
        }
        // back to the nearest tile whose prefix is out (tile 0 starts curve
        // 0, so there is one), with every tile after it aggregated ...
        int64_t hi = tile - 1, j = -1;  // window: tiles hi - 63 .. hi, lane l polls hi - l
        bool stalled = false;
        unsigned spin = 0;
        while (hi >= 0) {
            const int64_t q = hi - lane;
            unsigned long long st = 0;
            if (q >= 0) st = __hip_atomic_load(T.status + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const bool mine = (st >> 32) == epoch;
            const uint64_t mp = ballot(q >= 0 && mine && (st & kScanPrefix));
            const uint64_t mr = ballot(q < 0 || (mine && (st & (kScanAgg | kScanPrefix))));
            if (mp) {
                const int l = __ffsll((unsigned long long)mp) - 1;  // the nearest such tile
                const uint64_t after = (1ull << l) - 1ull;
                if ((mr & after) == after) {
                    j = hi - l;
                    break;
                }
            } else if (mr == ~0ull) {
                hi -= 64;  // all aggregated, no prefix yet: further back
                spin = 0;
                continue;
            }
            if (++spin >= kScanMaxSpin) {  // stalled: flag the build, leave the carry short
                if (lane == 0) atomicMax(words + kScanFailWord, 1ull);
                stalled = true;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        // ... then forward from it, left to right
        double carry = 0.0;
        if (!stalled && j >= 0) {
            carry = scan_get(T.prefix + j);
            for (int64_t b0 = j + 1; b0 < tile; b0 += 64) {
                const int64_t q = b0 + lane;
                const double v = q < tile ? scan_get(T.agg + q) : 0.0;
                const int cnt = (int)(tile - b0 < 64 ? tile - b0 : 64);
                for (int u = 0; u < cnt; u++) carry += readlane_d(v, u);
            }
        }
        if (lane == 0) {
            if (!tile_agg.f) {
                scan_put(T.prefix + tile, carry + tile_agg.v);
                scan_publish(T.status + tile, ep | kScanPrefix);
            }
            s_carry = carry;
        }
    }
    __syncthreads();
    const SegVal base = seg_add(SegVal{0, s_carry}, excl);
    int seen = 0;
    {
        const int e0 = tid * kScanItems;
#pragma unroll
        for (int k = 0; k < kScanItems; k++) {
            seen |= head[k];
            const double v = seen ? x[k] : base.v + x[k];
            s_x[scan_slot(e0 + k)] = v;
            if (tail[k]) curve_len[s_c[scan_slot(e0 + k)]] = v;  // last interval of its curve
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        const int e = k * kScanThreads + tid;
        const int64_t i = t0 + e;
        if (i < n) int_cum[i] = s_x[scan_slot(e)];
    }
}

// Streaming copy in 16-byte vectors (HBM ceiling / counter calibration).
static __global__ __launch_bounds__(256) void k_copy16(uint4* __restrict__ dst, const uint4* __restrict__ src,
                                                       int64_t n16) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}

// VALU issue-rate calibration: every lane runs 8 independent chains of one
// operation (kind 0: FP32 FMA, 1: FP64 FMA, 2: FP32 square root) for `iters`
// steps, so a full grid measures that operation's sustained rate.
template <int Kind>
static __global__ __launch_bounds__(256) void k_valu_chain(double* out, int iters) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if constexpr (Kind == 1) {
        double a[8];
#pragma unroll
        for (int j = 0; j < 8; j++) a[j] = (double)(threadIdx.x + j) * 1e-3;
        for (int i = 0; i < iters; i++) {
#pragma unroll
            for (int j = 0; j < 8; j++) a[j] = fma(a[j], 0.999999, 1e-7);
        }
        double t = 0.0;
#pragma unroll
        for (int j = 0; j < 8; j++) t += a[j];
        out[g] = t;
    } else {
        float a[8];
#pragma unroll
        for (int j = 0; j < 8; j++) a[j] = (float)(threadIdx.x + j) * 1e-3f + (Kind == 2 ? 1.0f : 0.0f);
        for (int i = 0; i < iters; i++) {
#pragma unroll
            for (int j = 0; j < 8; j++) a[j] = Kind == 0 ? fmaf(a[j], 0.999999f, 1e-7f) : __builtin_amdgcn_sqrtf(a[j]);
        }
        float t = 0.0f;
#pragma unroll
        for (int j = 0; j < 8; j++) t += a[j];
        out[g] = t;
    }
}

static __global__ void k_fill_u64(unsigned long long* p, int n, unsigned long long v) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

// ---------------------------------------------------------------------------
// B-spline setup: per-curve scale and interval -> curve map (wave per curve),
// then per-interval derivative polynomials (thread per interval).
// ---------------------------------------------------------------------------
struct BSplineDev {
    int n_curve;
    int ctrl_stride;
    const int32_t* degree;
    const int32_t* dim;
    const int64_t* knot_off;
    const int64_t* ctrl_off;
    const double* knots;
    const double* ctrl;
};

// Input check of a batch before any table is built: B-spline knots finite and
// non-decreasing within each curve, control points finite (the coordinates a
// curve uses); ellipse semi-axes finite. The first bad curve and the reason
// land in *bad = curve * 4 + reason (atomic minimum, *bad starts at ~0).
enum BadInput { kBadKnotValue = 1, kBadKnotOrder = 2, kBadCtrl = 3, kBadAxis = 1 };
static __global__ __launch_bounds__(256) void k_check_bspline(BSplineDev B, unsigned long long* bad) {
    const int lane = lane_id();
    const int64_t gwave = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t c = gwave; c < B.n_curve; c += nwaves) {
        const int64_t k0 = B.knot_off[c], k1 = B.knot_off[c + 1];
        int reason = 0;
        for (int64_t j = k0 + lane; j < k1; j += 64) {
            const double t = B.knots[j];
            if (!__builtin_isfinite(t)) reason = reason > kBadKnotValue ? reason : kBadKnotValue;
            else if (j + 1 < k1 && B.knots[j + 1] < t) reason = reason > kBadKnotOrder ? reason : kBadKnotOrder;
        }
        const int D = B.dim[c], st = B.ctrl_stride;
        const int64_t p0 = B.ctrl_off[c] * st, p1 = B.ctrl_off[c + 1] * st;
        for (int64_t j = p0 + lane; j < p1; j += 64)
            if ((int)((j - p0) % st) < D && !__builtin_isfinite(B.ctrl[j])) reason = kBadCtrl;
        for (int o = 32; o > 0; o >>= 1) reason = max(reason, __shfl_xor(reason, o, 64));
        if (lane == 0 && reason) atomicMin(bad, (unsigned long long)c * 4ull + (unsigned long long)reason);
    }
}
static __global__ __launch_bounds__(256) void k_check_ellipse(int n, const double* a, const double* b,
                                                              unsigned long long* bad) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && !(__builtin_isfinite(a[i]) && __builtin_isfinite(b[i])))
        atomicMin(bad, (unsigned long long)i * 4ull + kBadAxis);
}

static __global__ __launch_bounds__(256) void k_bspline_curve_setup(BSplineDev B, const int64_t* ioff, int normalize,
                                                             double* curve_scale, double* int_scale,
                                                             int32_t* int_curve) {
    const int lane = lane_id();
    const int64_t gwave = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t c = gwave; c < B.n_curve; c += nwaves) {
        double sc = 1.0;
        if (normalize) {
            // two passes (centroid, then max |coordinate - centroid|), each
            // with 4 x 64 control points in flight per wave iteration: a
            // 2048-point curve is then a few latency rounds, not 32
            const int D = B.dim[c];
            const int64_t p0 = B.ctrl_off[c], p1 = B.ctrl_off[c + 1];
            const int64_t st = B.ctrl_stride;
            double cen[3] = {0, 0, 0};
            for (int64_t j0 = p0 + lane; j0 < p1; j0 += 256) {
                double x[4][3];
#pragma unroll
                for (int u = 0; u < 4; u++)
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        const int64_t j = j0 + 64 * u;
                        x[u][d] = (j < p1 && d < D) ? B.ctrl[j * st + d] : 0.0;
                    }
#pragma unroll
                for (int u = 0; u < 4; u++)
#pragma unroll
                    for (int d = 0; d < 3; d++) cen[d] += x[u][d];
            }
#pragma unroll
            for (int d = 0; d < 3; d++) {
                double v = cen[d];
                for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
                cen[d] = v / (double)(p1 - p0);
            }
            double m = 0.0;
            for (int64_t j0 = p0 + lane; j0 < p1; j0 += 256) {
                double x[4][3];
#pragma unroll
                for (int u = 0; u < 4; u++)
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        const int64_t j = j0 + 64 * u;
                        x[u][d] = (j < p1 && d < D) ? B.ctrl[j * st + d] : cen[d];
                    }
#pragma unroll
                for (int u = 0; u < 4; u++)
#pragma unroll
                    for (int d = 0; d < 3; d++) m = fmax(m, fabs(x[u][d] - cen[d]));
            }
            for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
            sc = m > 0.0 ? m : 1.0;
        }
        if (lane == 0) curve_scale[c] = sc;
        for (int64_t i = ioff[c] + lane; i < ioff[c + 1]; i += 64) {
            int_scale[i] = sc;
            int_curve[i] = (int32_t)c;
        }
    }
}

// One interval's coefficient block (P * D values, into whichever of the FP64
// and FP32 tables the build keeps), its parameter origin and width; the FP32
// coefficients and the width are also returned (learned-seed features).
template <int P, int D>
__device__ __forceinline__ void coef_block(int s, const double* T, const double* Pc, int pstride, double isc,
                                           double* c64, float* c32, double* u0, double* h, float* cf,
                                           double* h_out) {
    double out[P * D];
    double uu, hh;
    span_poly<P, D>(s, T, Pc, pstride, isc, out, &uu, &hh);
    *u0 = uu;
    *h = hh;
    *h_out = hh;
#pragma unroll
    for (int j = 0; j < P * D; j++) {
        cf[j] = (float)out[j];
        if (c64) c64[j] = out[j];
        if (c32) c32[j] = cf[j];
    }
}

// Coefficient pass of one evaluator class (degree P, dimension D), one thread
// per interval; `list` holds the class's intervals in a mixed batch (null:
// every interval is of this class).
template <int P, int D>
__global__ __launch_bounds__(256) void k_bspline_coef(BSplineDev B, const int32_t* list, int64_t n_slots,
                                                      const int64_t* ioff, const int32_t* int_curve,
                                                      const double* int_scale, double* coef64, float* coef32,
                                                      int cstride, double* u0, double* h) {
    const int64_t slot = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= n_slots) return;
    const int64_t i = list ? (int64_t)list[slot] : slot;
    const int c = int_curve[i];
    const double* T = B.knots + B.knot_off[c];
    const double* Pc = B.ctrl + B.ctrl_off[c] * B.ctrl_stride;
    float cf[P * D];
    double hh;
    coef_block<P, D>((int)(i - ioff[c]) + P, T, Pc, B.ctrl_stride, rcp_f64(int_scale[i]),
                     coef64 ? coef64 + i * cstride : nullptr, coef32 ? coef32 + i * cstride : nullptr, &u0[i], &h[i],
                     cf, &hh);
}

// ---------------------------------------------------------------------------
// Sampling
// ---------------------------------------------------------------------------
struct SampleArgs {
    int n_curve;
    int64_t n_total;
    const int64_t* soff;        // [n_curve+1] sample offsets
    const double* pos;          // optional positions; null -> uniform
    int uniform_closed;         // uniform: 1 -> j/(N-1), 0 -> j/N
    int is_relative;
    const int64_t* ioff;        // [n_curve+1] interval offsets
    const double* int_cum;      // [n_int] cumulative interval lengths (curve units)
    const double* curve_len;    // [n_curve]
    const double* root_prefix;  // inclusive root prefix per interval
    const int8_t* root_k;       // seeded builds: log2 root count per interval, root block at root_off[i];
                                // else (null) 2^log2_nsub roots at i << log2_nsub
    const int64_t* root_off;    // seeded builds: first root of interval i
    const uint64_t* deep;       // [n_int]
    int log2_nsub;
    double eps0;                // level-0 tolerance per root (same value as the build)
    int ls_plain;               // log2 roots of the plain build (depth caps of seeded roots)
    int max_depth;
    int newton_iters;
    int fp64_step;              // FP32 mode: finish with one FP64 Newton step
    double* out_u;              // [n_total]
    int32_t* out_seg;           // optional local interval index per sample
    const int32_t* class_of_curve;  // optional (mixed batches): evaluator class per curve
    const int32_t* blk_int;     // optional (uniform positions): interval of each sample block's first sample
    int64_t n_int;              // knot intervals of the batch (0: no interval windows)
    // Newton start of B-spline samples in FP32 builds (kStart*): the plain
    // cubic start with FP32 steps, the learned per-interval choice of the
    // quintic start (start_ok[i] != 0) or the quintic start everywhere
    int start_mode;
    const int8_t* start_ok;     // [n_int] learned decisions (kStartLearned)
    int8_t* out_iters;          // optional per sample: FP32 steps | FP64 steps << 4
};
constexpr int kStartPlain = 0, kStartLearned = 1, kStartQuintic = 2;

// first index in [lo, hi) with a[idx] > v (hi if none)
__device__ __forceinline__ int64_t upper_bound_i64(const int64_t* a, int64_t lo, int64_t hi, int64_t v) {
    while (lo < hi) {
        const int64_t m = (lo + hi) >> 1;
        if (a[m] > v) hi = m; else lo = m + 1;
    }
    return lo;
}

// blk_curve[b] as a valid curve index: an entry no curve wrote (offsets that
// do not cover [0, n_total)) reads as some curve, never out of range.
__device__ __forceinline__ int64_t blk_curve_at(const SampleArgs& A, const int32_t* blk_curve, int64_t b) {
    const int64_t c = blk_curve[b];
    return c < 0 ? 0 : (c < A.n_curve ? c : A.n_curve - 1);
}

// The curve c in [c0, c1] (c0 <= c1) holding sample g: the last with
// soff[c] <= g, and c0 when none is (inconsistent offsets).
__device__ __forceinline__ int curve_in(const SampleArgs& A, int64_t c0, int64_t c1, int64_t g) {
    const int64_t c = upper_bound_i64(A.soff, c0, c1 + 1, g) - 1;
    return (int)(c > c0 ? c : c0);
}

// Number of entries of the non-decreasing a[0, n) below v, i.e. the first
// index with a[idx] >= v (n if none); 32-bit offsets from a 64-bit base.
template <class P>
__device__ __forceinline__ uint32_t count_below(P a, uint32_t n, double v) {
    uint32_t lo = 0;
    while (n > 0) {
        const uint32_t h = n >> 1;
        if (a[lo + h] < v) { lo += h + 1; n -= h + 1; } else n = h;
    }
    return lo;
}
// The same over a power-of-two count 2^ls, clamped to 2^ls - 1: one probe per
// bit of the answer, no data-dependent trip count (global or LDS pointer).
template <class P>
__device__ __forceinline__ uint32_t count_below_pow2(P a, int ls, double v) {
    uint32_t k = 0;
    for (uint32_t st = (1u << ls) >> 1; st > 0; st >>= 1)
        if (a[k + st - 1] < v) k += st;
    return k;
}

// Root prefixes one wave stages in LDS for its root searches.
constexpr int kRootWin = 256;
typedef __attribute__((address_space(3))) double lds_double;

// LDS window of root prefixes for one wave: when the whole wave is active and
// the root blocks of its lanes lie in the kRootWin prefixes that start at the
// first lane's block and end at the last lane's (targets grow along a wave
// for uniform positions), the wave reads that window once, coalesced
// (fetch_roots, into registers v), then writes it to its LDS slice
// (store_roots), where each lane's block starts at offset `off`.
// fetch_roots returns false (nothing read) otherwise.
__device__ __forceinline__ bool fetch_roots(const double* root_prefix, int64_t ro, int ls,
                                            double (&v)[kRootWin / 64], int& off, int& span) {
    const int64_t r0 = (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)((uint64_t)ro >> 32))
                                  << 32) |
                                 (uint32_t)__builtin_amdgcn_readfirstlane((int)ro));
    const int64_t end = ro + (int64_t(1) << ls) - r0;  // this lane's block end in the window
    span = __builtin_amdgcn_readlane((int)(end < kRootWin ? end : kRootWin + 1), 63);
    const bool fits = ro >= r0 && end <= span && span <= kRootWin;
    if (__ballot(1) != ~0ull || __ballot(fits) != ~0ull) return false;
    const int lane = lane_id();
#pragma unroll
    for (int q = 0; q < kRootWin / 64; q++)
        if (lane + 64 * q < span) v[q] = root_prefix[r0 + lane + 64 * q];
    off = (int)(ro - r0);
    return true;
}
__device__ __forceinline__ void store_roots(lds_double* win, const double (&v)[kRootWin / 64], int span) {
    const int lane = lane_id();
#pragma unroll
    for (int q = 0; q < kRootWin / 64; q++)
        if (lane + 64 * q < span) win[lane + 64 * q] = v[q];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

// Root k of root-local target sn in a 2^ls-root block of inclusive prefixes
// (global or LDS), with the prefix before it and its length.
template <class P>
__device__ __forceinline__ void find_root(P rp, int ls, double sn, int& k, double& pre, double& rlen) {
    k = (int)count_below_pow2(rp, ls, sn);
    pre = k > 0 ? rp[k - 1] : 0.0;
    rlen = rp[k] - pre;
}

// Reciprocal for Newton updates and start values: one v_rcp (1 ulp) in float,
// v_rcp_f64 with two Newton steps (1 ulp) in double.
__device__ __forceinline__ float nrcp(float v) { return __builtin_amdgcn_rcpf(v); }
__device__ __forceinline__ double nrcp(double v) { return rcp_f64(v); }

// Quadrature of speed over [a, a + d], with the width d carried exactly.
template <int N, class Ev, typename Dt>
__device__ __forceinline__ Dt quad_from(const Ev& ev, const typename Ev::Ctx& c, Dt a, Dt d) {
    if constexpr (HasLengthRule<Ev>::value) return ev.length(c, a, a + d);
    const Dt k = d * Dt(0.5);
    const Dt m = a + k;
    return gl_sum<N>(ev, c, m, k, Dt(0)) * k;
}

// Composite rule: `pieces` (1 or 2) equal sub-ranges of [a, a + d].
template <int N, class Ev, typename Dt>
__device__ __forceinline__ Dt quad_from_n(const Ev& ev, const typename Ev::Ctx& c, Dt a, Dt d, int pieces) {
    if (pieces <= 1) return quad_from<N>(ev, c, a, d);
    const Dt h = d * Dt(0.5);
    return quad_from<N>(ev, c, a, h) + quad_from<N>(ev, c, a + h, h);
}

// Inverse arc length on one leaf [a, b] of the adaptive tree: find the offset
// d in [0, b - a] with int_a^{a+d} speed = tau. The start is the cubic Hermite
// interpolant of the inverse map x(s) (x' = 1 / speed at both ends); the first
// Newton step evaluates the integral by quadrature, later steps update it by
// the trapezoid rule over the (quadratically shrinking) step.
// After the `iters` steps a lane whose last step still exceeded 2^-14 (FP32)
// or 2^-26 (FP64) of the leaf width is not yet in Newton's quadratic regime
// (a poor start next to a near-zero speed): it keeps iterating, with the
// residual recomputed by quadrature, up to kExtraNewton more steps.
// Returns d; *v_out receives the speed at the last evaluated point (a usable
// Newton derivative for a final step at d).
// `pieces` = 2 for leaves wider than 1/64 of their knot interval: the error
// test bounds a leaf's quadrature error by tol x (leaf width / interval
// width), so halving keeps the partial integrals within tol / 64 as for the
// plain method's 64-way pre-split.
constexpr int kExtraNewton = 16;
// Speeds at a leaf's ends and their reciprocals (as invert_leaf computes them),
// and for the quintic start the second derivatives of the inverse map x(s) at
// the ends, x'' = -v'(x) / v^3.
template <typename Dt>
struct LeafEnds {
    Dt va, vb, rva, rvb, xa2, xb2;
};
template <class Ev, typename Dt>
__device__ __forceinline__ LeafEnds<Dt> leaf_ends(const Ev& ev, const typename Ev::Ctx& c, Dt a, Dt b) {
    LeafEnds<Dt> e;
    e.va = ev.speed(c, a);
    e.vb = ev.speed(c, b);
    e.rva = nrcp(e.va);
    e.rvb = nrcp(e.vb);
    e.xa2 = Dt(0);
    e.xb2 = Dt(0);
    return e;
}
// The same with the second derivatives (evaluators with speed_slope).
template <class Ev, typename Dt>
__device__ __forceinline__ LeafEnds<Dt> leaf_ends_quintic(const Ev& ev, const typename Ev::Ctx& c, Dt a, Dt b) {
    LeafEnds<Dt> e = leaf_ends(ev, c, a, b);
    if (e.va > Dt(0) && e.vb > Dt(0)) {
        e.xa2 = -ev.speed_slope(c, a, e.va) * e.rva * e.rva * e.rva;
        e.xb2 = -ev.speed_slope(c, b, e.vb) * e.rvb * e.rvb * e.rvb;
    }
    return e;
}
// Evaluators that give the speed and its slope in one pass (B-splines).
template <class Ev, class = void>
struct HasSpeedAndSlope : std::false_type {};
template <class Ev>
struct HasSpeedAndSlope<Ev, decltype((void)std::declval<const Ev&>().speed_and_slope(
                                std::declval<const typename Ev::Ctx&>(), 0.0f, std::declval<float&>(),
                                std::declval<float&>()))> : std::true_type {};
template <class Ev, typename Dt>
__device__ __forceinline__ LeafEnds<Dt> leaf_ends_slopes(const Ev& ev, const typename Ev::Ctx& c, Dt a, Dt b) {
    if constexpr (HasSpeedAndSlope<Ev>::value) {
        LeafEnds<Dt> e;
        Dt sa, sb;
        ev.speed_and_slope(c, a, e.va, sa);
        ev.speed_and_slope(c, b, e.vb, sb);
        e.rva = nrcp(e.va);
        e.rvb = nrcp(e.vb);
        e.xa2 = Dt(0);
        e.xb2 = Dt(0);
        if (e.va > Dt(0) && e.vb > Dt(0)) {
            e.xa2 = -sa * e.rva * e.rva * e.rva;
            e.xb2 = -sb * e.rvb * e.rvb * e.rvb;
        }
        return e;
    } else {
        return leaf_ends_quintic(ev, c, a, b);
    }
}
// With `quintic` (ends known, second derivatives included) the start is the
// quintic Hermite interpolant of x(s) and no Newton step follows here: the
// caller takes the FP64 step from it (*v_out = the speed at the start).
template <int NGL, class Ev, typename Dt>
__device__ __forceinline__ Dt invert_leaf(const Ev& ev, const typename Ev::Ctx& c, Dt a, Dt b, Dt tau, Dt len,
                                          int iters, int pieces, Dt* v_out, bool ends_known = false,
                                          const LeafEnds<Dt>& ends_in = LeafEnds<Dt>{}, bool quintic = false,
                                          int* n_steps = nullptr) {
    const Dt w = b - a;
    Dt t = len > Dt(0) ? tau * nrcp(len) : Dt(0);
    t = t < Dt(0) ? Dt(0) : (t > Dt(1) ? Dt(1) : t);
    LeafEnds<Dt> E;
    if (ends_known)
        E = ends_in;
    else
        E = leaf_ends(ev, c, a, b);
    const Dt va = E.va, vb = E.vb;
    if (ends_known && quintic && va > Dt(0) && vb > Dt(0) && len > Dt(0)) {
        const Dt t2 = t * t, t3 = t2 * t, t4 = t3 * t, t5 = t4 * t;
        const Dt H5 = Dt(10) * t3 - Dt(15) * t4 + Dt(6) * t5;     // value at the right end
        const Dt H1 = t - Dt(6) * t3 + Dt(8) * t4 - Dt(3) * t5;    // slope at the left end
        const Dt H4 = -Dt(4) * t3 + Dt(7) * t4 - Dt(3) * t5;       // slope at the right end
        const Dt H2 = Dt(0.5) * t2 - Dt(1.5) * t3 + Dt(1.5) * t4 - Dt(0.5) * t5;  // curvature, left
        const Dt H3 = Dt(0.5) * t3 - t4 + Dt(0.5) * t5;                           // curvature, right
        Dt d = w * H5 + len * (H1 * E.rva + H4 * E.rvb) + len * len * (H2 * E.xa2 + H3 * E.xb2);
        d = d > Dt(0) ? (d < w ? d : w) : Dt(0);
        *v_out = ev.speed(c, a + d);
        return d;
    }
    Dt d;
    if (va > Dt(0) && vb > Dt(0) && len > Dt(0)) {
        const Dt t2 = t * t, t3 = t2 * t;
        const Dt h10 = t3 - Dt(2) * t2 + t, h01 = Dt(3) * t2 - Dt(2) * t3, h11 = t3 - t2;
        d = w * h01 + len * (h10 * E.rva + h11 * E.rvb);
    } else {
        d = w * t;
    }
    d = d > Dt(0) ? (d < w ? d : w) : Dt(0);  // also maps a NaN start to 0
    *v_out = va;
    if (iters <= 0) return d;
    Dt lo = Dt(0), hi = w;
    Dt F = quad_from_n<NGL>(ev, c, a, d, pieces) - tau;
    Dt v = ev.speed(c, a + d);
    *v_out = v;
    const Dt conv = w * (sizeof(Dt) == 4 ? Dt(0x1p-14) : Dt(0x1p-26));
    for (int it = 0;; it++) {
        if (F > Dt(0)) hi = d; else lo = d;
        Dt dn = v > Dt(0) ? d - F * nrcp(v) : Dt(0.5) * (lo + hi);
        if (!(dn >= lo && dn <= hi)) dn = Dt(0.5) * (lo + hi);
        const bool past = it + 1 >= iters;
        if (past && (dabs(dn - d) <= conv || it + 1 >= iters + kExtraNewton)) {
            if (n_steps) *n_steps = it + 1;
            return dn;
        }
        const Dt vn = ev.speed(c, a + dn);
        if (past)
            F = quad_from_n<NGL>(ev, c, a, dn, pieces) - tau;
        else
            F += (dn - d) * (v + vn) * Dt(0.5);
        d = dn;
        v = vn;
        *v_out = v;
    }
}

// Re-walk the adaptive tree of a refined root in the build's arithmetic and
// return the build leaf [a, b] holding tau (root-local target), with tau
// reduced to a leaf-local target and the leaf's length.
// guard: the table is a seeded build, whose segments wider than a plain root
// (2^-ls_plain) were also refined where their node speeds dip (dip_forces).
template <typename Dt, int NGL, class Ev>
__device__ void walk_to_leaf(const Ev& ev, const typename Ev::Ctx& c, int k, int ls, Dt eps0, int max_depth,
                             double tau, Dt& a, Dt& b, double& tau_leaf, Dt& leaf_len, bool guard = false,
                             int ls_plain = 6) {
    Dt stackR[64];
    int l = 0;
    uint32_t idx = 0;
    Dt x0, xm, x1;
    task_bounds<Dt>(k, 0, 0u, ls, x0, xm, x1);
    Dt P = quad<NGL>(ev, c, x0, x1);
    double acc = 0.0;
    for (int it = 0; it < (1 << 26); it++) {
        task_bounds<Dt>(k, l, idx, ls, x0, xm, x1);
        Dt L, R;
        bool dip = false;
        Dt mn = Dt(3.0e38f), mx = Dt(0);
        if (guard) {
            L = quad_mm<NGL>(ev, c, x0, xm, mn, mx);
            R = quad_mm<NGL>(ev, c, xm, x1, mn, mx);
            dip = dip_forces(ls + l, ls_plain, mn, mx);
        } else {
            L = quad<NGL>(ev, c, x0, xm);
            R = quad<NGL>(ev, c, xm, x1);
        }
        const Dt lr = L + R;
        const Dt tol = tol_of(eps0 * pow2m(Dt(0), l), rounding_floor<Dt>(ls, l, ls_plain), lr);
        bool split = dabs(lr - P) > tol || dip;
        // the build's cross-check of wide segments, where the other tests accept
        if (guard && !split && cross_wanted(ls + l, ls_plain, mn, mx))
            split = cross_fails<NGL>(ev, c, x0, x1, lr, tol);
        if (l < max_depth && split) {
            stackR[l + 1] = R;
            l++;
            idx = 2u * idx;
            P = L;
            continue;
        }
        // build leaf
        const double v = (double)lr;
        if (tau <= acc + v) {
            a = x0; b = x1; tau_leaf = tau - acc; leaf_len = lr;
            return;
        }
        acc += v;
        while (l > 0 && (idx & 1u)) { idx >>= 1; l--; }
        if (l == 0) {  // target beyond the root's last leaf (rounding): clamp to its end
            a = x0; b = x1; tau_leaf = v; leaf_len = lr;
            return;
        }
        idx += 1u;
        P = stackR[l];
    }
    a = x0; b = x1; tau_leaf = 0.0; leaf_len = Dt(0);
}

// Context of the FP64 Newton step: widened from the FP32 context where that
// is exact (B-spline coefficients stored as float), else loaded.
template <class Ev64, class Ctx32, class = void>
struct HasWiden : std::false_type {};
template <class Ev64, class Ctx32>
struct HasWiden<Ev64, Ctx32, decltype((void)Ev64::widen(std::declval<const Ctx32&>()))> : std::true_type {};
template <class Ev64, class Ctx32>
__device__ __forceinline__ typename Ev64::Ctx widen_or_load(const Ev64& ev64, const Ctx32& c32, int i, double inv_sc) {
    if constexpr (HasWiden<Ev64, Ctx32>::value)
        return Ev64::widen(c32);
    else
        return ev64.load(i, inv_sc);
}

// Samples per block of the sampling kernels (blockDim.x).
constexpr int kSampleBlock = 256;

// Sampling mode: explicit positions (pos) and seeded root counts (root_k)
// read from the arguments at run time, or fixed at compile time so that a
// kernel has no branch between its loads.
struct RuntimeMode {
    static constexpr bool kStatic = false, kPos = false, kSeeded = false, kStart = false;
};
// kStart: the kernel carries the learned / quintic Newton start (kept out of
// the plain kernels, whose registers and code it would otherwise grow)
template <bool P, bool S, bool St = false>
struct StaticMode {
    static constexpr bool kStatic = true, kPos = P, kSeeded = S, kStart = St;
};
template <class M>
__device__ __forceinline__ bool mode_pos(const SampleArgs& A) {
    if constexpr (M::kStatic) return M::kPos; else return A.pos != nullptr;
}
template <class M>
__device__ __forceinline__ bool mode_seeded(const SampleArgs& A) {
    if constexpr (M::kStatic) return M::kSeeded; else return A.root_k != nullptr;
}

// Target arc length of sample g (curve c) in curve units, in two parts: the
// curve's loads (target_in; unconditional, their addresses are always valid)
// and the arithmetic (target_of), so a caller can issue its other loads in
// between.
struct TargetIn {
    int64_t s0, s1;  // the curve's sample offsets
    double cl;       // its length
    double pv;       // explicit position of the sample
};
template <class M>
__device__ __forceinline__ TargetIn target_in(const SampleArgs& A, int c, int64_t g) {
    TargetIn t;
    t.s0 = A.soff[c];
    t.s1 = A.soff[c + 1];
    t.cl = A.curve_len[c];
    t.pv = mode_pos<M>(A) ? A.pos[g] : 0.0;
    return t;
}
template <class M>
__device__ __forceinline__ double target_of(const SampleArgs& A, const TargetIn& t, int64_t g) {
    double pos = t.pv;
    if (!mode_pos<M>(A)) {
        // j / (N - 1) (closed; 0 for N = 1) or j / N to within 1.5 ulp; a
        // position a rounding above 1 is clamped to the curve's end by the
        // searches below
        const int64_t N = t.s1 - t.s0, j = g - t.s0;
        const double den = (double)(A.uniform_closed ? N - 1 : N);
        const double r = rcp_f64(den > 0.0 ? den : 1.0);
        pos = den > 0.0 ? (double)j * r : 0.0;
    }
    return A.is_relative ? pos * t.cl : pos;
}
template <class M = RuntimeMode>
__device__ __forceinline__ double sample_target(const SampleArgs& A, int c, int64_t g) {
    return target_of<M>(A, target_in<M>(A, c, g), g);
}

// Evaluators whose curves are one knot interval each (interval i = curve i).
template <class Ev, class = void>
struct OneIntervalPerCurve : std::false_type {};
template <class Ev>
struct OneIntervalPerCurve<Ev, std::enable_if_t<Ev::kOneInterval>> : std::true_type {};

// Interval of target s on curve c: the first cumulative length >= s among the
// curve's first n_int - 1, so a target past the last boundary lands in the
// last interval; the search is limited to [lo, hi] (clamped to the curve).
__device__ __forceinline__ int64_t interval_of_target(const SampleArgs& A, int c, double s, int64_t lo, int64_t hi) {
    const int64_t ib = A.ioff[c], ie = A.ioff[c + 1];
    lo = lo > ib ? lo : ib;
    hi = hi < ie - 1 ? hi : ie - 1;
    if (lo > hi) lo = hi;  // a stale hint: still a valid interval of the curve
    return lo + count_below(A.int_cum + lo, (uint32_t)(hi - lo), s);
}

// Intervals of sample block blk (uniform positions): the sampler of a block's
// samples on curve blk_curve[blk] searches only between this block's first
// interval and the next block's (one interval of slack on each side).
static __global__ __launch_bounds__(256) void k_block_interval(SampleArgs A, const int32_t* blk_curve,
                                                               int32_t* blk_int) {
    const int64_t n_blk = (A.n_total + kSampleBlock - 1) / kSampleBlock;
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n_blk) return;
    const int c = (int)blk_curve_at(A, blk_curve, b);
    const double s = sample_target(A, c, b * kSampleBlock);
    blk_int[b] = (int32_t)interval_of_target(A, c, s, 0, INT64_MAX);
}

// Interval of sample g on curve c (several intervals per curve) and the
// cumulative length before it. The search range [hint_lo, hint_hi] comes from
// the block interval map; when it is the same for the whole (fully active)
// wave and holds at most 63 intervals, the wave reads int_cum over it (plus
// the entry before) once, coalesced, into its LDS slice `iwin` in the same
// round trip as the curve's offsets, and searches there. Same result as the
// global search.
template <class M>
__device__ __forceinline__ int64_t locate_interval(const SampleArgs& A, const TargetIn& tin, int c, int64_t g,
                                                   int64_t hint_lo, int64_t hint_hi, lds_double* iwin, double& s,
                                                   int64_t& ib, double& prev) {
    const int64_t b0 = A.ioff[c], ie = A.ioff[c + 1];
    int64_t w0 = 0;
    bool ok = false;
    double v = 0.0;
    const int lane = lane_id();
    if (iwin) {
        const int64_t lo_u = readfirstlane64(hint_lo), hi_u = readfirstlane64(hint_hi);
        w0 = lo_u - 1 > 0 ? lo_u - 1 : 0;
        const int64_t w1 = hi_u < A.n_int - 1 ? hi_u : A.n_int - 1;
        const int64_t wn = w1 - w0 + 1;
        ok = __ballot(1) == ~0ull && __ballot(hint_lo == lo_u && hint_hi == hi_u) == ~0ull && wn >= 1 && wn <= 64;
        if (ok) v = A.int_cum[w0 + (lane < wn ? lane : wn - 1)];
    }
    s = target_of<M>(A, tin, g);
    int64_t lo = hint_lo > b0 ? hint_lo : b0;
    const int64_t hi = hint_hi < ie - 1 ? hint_hi : ie - 1;
    if (lo > hi) lo = hi;  // a stale hint: still a valid interval of the curve
    int64_t i;
    if (ok) {
        iwin[lane] = v;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
        i = lo + count_below((const lds_double*)(iwin + (lo - w0)), (uint32_t)(hi - lo), s);
        prev = i > b0 ? iwin[i - 1 - w0] : 0.0;
    } else {
        i = lo + count_below(A.int_cum + lo, (uint32_t)(hi - lo), s);
        prev = i > b0 ? A.int_cum[i - 1] : 0.0;
    }
    ib = b0;
    return i;
}

// Tail of one sample once its root is known (root k of the 2^ls-root block of
// interval i, the prefix before it, its length, the root-local target sn):
// re-walk a refined root to its leaf, Newton in Dt from the Hermite start,
// then the FP64 step. c64 is the FP64 context for evaluators that load it
// (read before any wave barrier); evaluators that widen their Dt context do
// so here.
template <typename Dt, int NGL, class Ev, class Ev64, bool kStart = false>
__device__ __forceinline__ double finish_sample(const Ev& ev, const Ev64& ev64, const SampleArgs& A, int i,
                                                const typename Ev::Ctx& ctx, const typename Ev64::Ctx& c64,
                                                uint64_t deep, int ls, int k, double sn, double pre, double rlen,
                                                bool root_ends_known = false,
                                                const LeafEnds<Dt>& root_ends = LeafEnds<Dt>{},
                                                bool quintic = false, int* iters = nullptr) {
    const Dt eps0 = (Dt)A.eps0;
    double tau = sn - pre;
    tau = tau < 0.0 ? 0.0 : (tau > rlen ? rlen : tau);
    Dt a, b, xm;
    task_bounds<Dt>(k, 0, 0u, ls, a, xm, b);
    Dt len = (Dt)rlen;
    bool ends_known = root_ends_known;
    if ((deep >> k) & 1ull) {
        double tl;
        const int depth = effective_depth(A.max_depth, A.ls_plain, ls, sizeof(Dt) == 4);
        walk_to_leaf<Dt, NGL>(ev, ctx, k, ls, eps0, depth, tau, a, b, tl, len, A.root_k != nullptr, A.ls_plain);
        tau = tl;
        ends_known = false;  // the leaf is below the root
    }
    Dt v_last;
    const int pieces = 1;
    constexpr bool f32 = sizeof(Dt) < sizeof(double);
    const bool step64 = f32 && A.fp64_step;
    // one FP64 Newton step on the same leaf: the residual is evaluated in
    // FP64; the derivative is the last FP32 speed (its 1e-7 relative error
    // only scales a residual that is already small)
    auto newton64 = [&](double d0, Dt v) -> double {
        typename Ev64::Ctx cw;
        if constexpr (HasWiden<Ev64, typename Ev::Ctx>::value)
            cw = Ev64::widen(ctx);
        else
            cw = c64;
        const double ad = (double)a, wd = (double)b - (double)a;
        const double F = quad_from_n<NGL>(ev64, cw, ad, d0, pieces) - tau;
        if (!(v > Dt(0))) return d0;
        const double dn = d0 - F * (double)nrcp(v);
        return dn < 0.0 ? 0.0 : (dn > wd ? wd : dn);
    };
    int n32 = 0, n64 = 0;
    double dd;
    if constexpr (kStart && f32 && HasSpeedAndSlope<Ev>::value) {
        // learned (or forced) Newton start, decided per wave so that the
        // wave's lanes run one path: when every lane's interval is marked
        // (and no lane sits on a refined root), the quintic Hermite start of
        // the inverse map goes straight to the FP64 step, which a lane keeps
        // when it moved the start by at most 2^-12 of the leaf w (Newton's
        // error after it is then about (w |v'|/v) 2^-25 w); a lane whose step
        // moved more falls back to the plain start and steps below
        const bool want = step64 && !((deep >> k) & 1ull) &&
                          (A.start_mode == kStartQuintic || (A.start_mode == kStartLearned && A.start_ok[i] != 0));
        bool plain = true;
        dd = 0.0;
        LeafEnds<Dt> E{};
        if (__all(want)) {
            E = leaf_ends_slopes(ev, ctx, a, b);
            Dt vs;
            const Dt d0 = invert_leaf<NGL>(ev, ctx, a, b, (Dt)tau, len, 0, pieces, &vs, true, E, true);
            dd = newton64((double)d0, vs);
            n64 = 1;
            plain = !(dabs(dd - (double)d0) <= ((double)b - (double)a) * 0x1p-12);
            ends_known = true;
        }
        if (plain) {
            const Dt d = invert_leaf<NGL>(ev, ctx, a, b, (Dt)tau, len, A.newton_iters, pieces, &v_last, ends_known,
                                          ends_known ? E : root_ends, false, &n32);
            dd = (double)d;
            if (step64) {
                dd = newton64(dd, v_last);
                n64++;
            }
        }
    } else {
        // the quintic start replaces the Dt Newton steps only where the FP64 step follows
        const bool q = quintic && ends_known && step64;
        const Dt d = invert_leaf<NGL>(ev, ctx, a, b, (Dt)tau, len, A.newton_iters, pieces, &v_last, ends_known,
                                      root_ends, q, &n32);
        dd = (double)d;
        if (step64) {
            dd = newton64(dd, v_last);
            n64 = 1;
        }
    }
    if (iters) *iters = n32 | (n64 << 4);
    return ev.to_u(i, (double)a + dd);
}

// FP64 context read with the Dt one (before any wave barrier) for evaluators
// that cannot widen theirs.
template <class Ev64, class Ctx32>
__device__ __forceinline__ typename Ev64::Ctx ctx64_early(const Ev64& ev64, int i, double inv_sc) {
    if constexpr (HasWiden<Ev64, Ctx32>::value)
        return typename Ev64::Ctx{};
    else
        return ev64.load(i, inv_sc);
}

template <typename Dt, int NGL, class M = RuntimeMode, class Ev, class Ev64>
__device__ __forceinline__ double sample_one(const Ev& ev, const Ev64& ev64, const SampleArgs& A, int c,
                                             int64_t g, int32_t* seg_out, int64_t hint_lo = 0,
                                             int64_t hint_hi = INT64_MAX, lds_double* win = nullptr,
                                             lds_double* iwin = nullptr) {
    constexpr bool one = OneIntervalPerCurve<Ev>::value;
    const TargetIn tin = target_in<M>(A, c, g);
    // one-interval evaluators know the interval before the target: their
    // target arithmetic waits until the interval's loads below are issued
    double s = 0.0, prev = 0.0;
    int64_t ib = c, i = c;
    if constexpr (!one) i = locate_interval<M>(A, tin, c, g, hint_lo, hint_hi, iwin, s, ib, prev);
    if (seg_out) *seg_out = (int32_t)(i - ib);
    // every load that depends on the interval only is issued before the root
    // window is staged, so one wait covers them all
    const double sc = ev.scale((int)i);
    const uint64_t deep = A.deep[i];
    int ls = A.log2_nsub;
    int64_t ro = (int64_t)i << ls;
    if (mode_seeded<M>(A)) {
        ro = A.root_off[i];
        ls = A.root_k[i];
    }
    int off = 0, span = 0;
    double rv[kRootWin / 64];
    const bool staged = win && fetch_roots(A.root_prefix, ro, ls, rv, off, span);
    if constexpr (one) s = target_of<M>(A, tin, g);
    const double inv_sc = rcp_f64(sc);
    const double sn = (s - prev) * inv_sc;
    // the contexts' loads go out with the root window's; the window is
    // written to LDS after them (a wave barrier before them would hold them
    // back, or force loads above to be repeated)
    const typename Ev::Ctx ctx = ev.load((int)i, inv_sc);
    const typename Ev64::Ctx c64 = ctx64_early<Ev64, typename Ev::Ctx>(ev64, (int)i, inv_sc);
    if (staged) store_roots(win, rv, span);
    int k;
    double pre, rlen;
    if (staged)
        find_root((const lds_double*)(win + off), ls, sn, k, pre, rlen);
    else
        find_root(A.root_prefix + ro, ls, sn, k, pre, rlen);
    int its = 0;
    const double u = finish_sample<Dt, NGL, Ev, Ev64, M::kStart>(ev, ev64, A, (int)i, ctx, c64, deep, ls, k, sn,
                                                                 pre, rlen, false, LeafEnds<Dt>{}, false, &its);
    if (A.out_iters) A.out_iters[g] = (int8_t)its;
    return u;
}

// Curve of the first sample of every block of kSampleBlock samples: thread per
// curve writes the blocks whose first sample falls in its range, so the
// sampler finds a lane's curve with one load and a search over the few
// curves one block spans. The block range is clamped to [0, n_blk), so sample
// offsets that are not non-decreasing from 0 to n_total write nothing out of
// bounds (their results are unspecified, see blk_curve_at).

static __global__ __launch_bounds__(256) void k_block_curve(int n_curve, const int64_t* soff, int64_t n_blk,
                                                            int32_t* blk_curve) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n_curve) return;
    int64_t b0 = (soff[c] + kSampleBlock - 1) / kSampleBlock;
    int64_t b1 = (soff[c + 1] + kSampleBlock - 1) / kSampleBlock;
    b0 = b0 < 0 ? 0 : b0;
    b1 = b1 < n_blk ? b1 : n_blk;
    for (int64_t b = b0; b < b1; b++) blk_curve[b] = (int32_t)c;
}

// Curve of sample g of block `blk` (blockDim.x == kSampleBlock).
__device__ __forceinline__ int sample_curve(const SampleArgs& A, const int32_t* blk_curve, int64_t blk, int64_t g) {
    const int64_t n_blk = (A.n_total + kSampleBlock - 1) / kSampleBlock;
    const int64_t c0 = blk_curve_at(A, blk_curve, blk);
    int64_t c1 = blk + 1 < n_blk ? blk_curve_at(A, blk_curve, blk + 1) : A.n_curve - 1;
    c1 = c1 > c0 ? c1 : c0;
    return curve_in(A, c0, c1, g);
}

// Curve of sample g and, with the block interval map (uniform positions), the
// interval range its search may be limited to: targets are non-decreasing
// along a curve, so a sample lies between the intervals of its block's first
// sample and the next block's first sample when those are on its curve.
struct SampleLoc {
    int c;
    int64_t lo, hi;
};
__device__ __forceinline__ SampleLoc locate_sample(const SampleArgs& A, const int32_t* blk_curve, int64_t blk,
                                                   int64_t g) {
    const int64_t n_blk = (A.n_total + kSampleBlock - 1) / kSampleBlock;
    const int64_t c0 = blk_curve_at(A, blk_curve, blk);
    const bool has_next = blk + 1 < n_blk;
    int64_t c1 = has_next ? blk_curve_at(A, blk_curve, blk + 1) : A.n_curve - 1;
    c1 = c1 > c0 ? c1 : c0;
    // a block whose first sample and the next block's are on one curve lies
    // on that curve
    SampleLoc L{c0 == c1 ? (int)c0 : curve_in(A, c0, c1, g), 0, INT64_MAX};
    if (A.blk_int) {
        if (L.c == c0) L.lo = (int64_t)A.blk_int[blk] - 1;
        if (has_next && L.c == c1) L.hi = (int64_t)A.blk_int[blk + 1] + 1;
    }
    return L;
}

template <typename Dt, int NGL, class Ev, class Ev64, class M = RuntimeMode>
__global__ __launch_bounds__(256) void k_sample(Ev ev, Ev64 ev64, SampleArgs A, const int32_t* blk_curve) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= A.n_total) return;
    __shared__ double s_win[kSampleBlock / 64][kRootWin];
    lds_double* win = (lds_double*)s_win[threadIdx.x >> 6];
    lds_double* iwin = nullptr;
    if constexpr (!OneIntervalPerCurve<Ev>::value) {
        __shared__ double s_iwin[kSampleBlock / 64][64];
        iwin = (lds_double*)s_iwin[threadIdx.x >> 6];
    }
    int32_t seg;
    double u;
    const int64_t n_blk = (A.n_total + kSampleBlock - 1) / kSampleBlock;
    const int c0 = (int)blk_curve_at(A, blk_curve, blockIdx.x);
    if (!A.blk_int && blockIdx.x + 1 < n_blk && blk_curve_at(A, blk_curve, blockIdx.x + 1) == c0) {
        // the whole block lies on curve c0: the curve index is block-uniform,
        // so its per-curve data comes through scalar loads issued together
        u = sample_one<Dt, NGL, M>(ev, ev64, A, c0, g, &seg, 0, INT64_MAX, win, iwin);
    } else {
        const SampleLoc L = locate_sample(A, blk_curve, blockIdx.x, g);
        u = sample_one<Dt, NGL, M>(ev, ev64, A, L.c, g, &seg, L.lo, L.hi, win, iwin);
    }
    A.out_u[g] = u;
    if (A.out_seg) A.out_seg[g] = seg;
}

// Wave-uniform copy of a plain value (each 32-bit word from the first active
// lane), so the compiler keeps it in scalar registers.
template <class T>
__device__ __forceinline__ T uniform_pod(const T& v) {
    static_assert(sizeof(T) % 4 == 0, "32-bit words");
    uint32_t w[sizeof(T) / 4];
    __builtin_memcpy(w, &v, sizeof(T));
#pragma unroll
    for (int q = 0; q < (int)(sizeof(T) / 4); q++) w[q] = (uint32_t)__builtin_amdgcn_readfirstlane((int)w[q]);
    T r;
    __builtin_memcpy(&r, w, sizeof(T));
    return r;
}

// Persistent sampling of one-interval curves (ellipses) at uniform positions.
// Every wave takes a contiguous run of 64-sample tiles. Tiles of a run mostly
// lie on one curve, so the wave reads a curve's data (sample offsets, length,
// evaluator contexts, refined-root mask, and the root prefixes into its LDS
// slice) once and keeps it for every tile inside that curve. A tile that
// reaches into later curves is finished curve by curve: the lanes on the
// current curve, then the wave loads the next curve and takes its lanes. Same
// arithmetic as sample_one, so the same results.
template <typename Dt, int NGL, class Ev, class Ev64, bool kSeeded>
__global__ __launch_bounds__(256) void k_sample_runs(Ev ev, Ev64 ev64, SampleArgs A, const int32_t* blk_curve,
                                                     int64_t tiles_per_wave) {
    static_assert(OneIntervalPerCurve<Ev>::value, "run sampling needs one interval per curve");
    __shared__ double s_run[kSampleBlock / 64][64];
    __shared__ Dt s_ends[kSampleBlock / 64][64 * 6];  // per root: va, vb, 1/va, 1/vb, x''(a), x''(b)
    const int lane = lane_id();
    // the wave index is uniform, and so is the whole tile loop below
    const int wib = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    lds_double* rwin = (lds_double*)s_run[wib];
    typedef __attribute__((address_space(3))) Dt lds_dt;
    lds_dt* ends = (lds_dt*)s_ends[wib];
    const int64_t gw = (int64_t)blockIdx.x * (blockDim.x >> 6) + wib;
    const int64_t n_tiles = (A.n_total + 63) >> 6;
    const int64_t t0 = gw * tiles_per_wave;
    const int64_t t1 = t0 + tiles_per_wave < n_tiles ? t0 + tiles_per_wave : n_tiles;
    const int64_t n_blk = (A.n_total + kSampleBlock - 1) / kSampleBlock;

    // run state (wave-uniform): curve cr holds samples [s0, s1)
    struct State {
        typename Ev::Ctx ctx;
        typename Ev64::Ctx c64;
        double rN, cl, inv_sc;
        uint64_t deep;
        int64_t cr, s0, s1;
        int32_t ls, den_pos, quintic;
    };
    State S{};
    for (int64_t t = t0; t < t1; t++) {
        const int64_t gf = t << 6;
        const int64_t gl = gf + 63 < A.n_total - 1 ? gf + 63 : A.n_total - 1;
        const int64_t g = gf + lane;
        int64_t next = gf;  // first sample of the tile not yet taken (wave-uniform)
        for (;;) {
            if (next < S.s0 || next >= S.s1) {
                // load the curve holding sample `next`
                const int64_t blk = next / kSampleBlock;
                const int64_t ca = readfirstlane64(blk_curve_at(A, blk_curve, blk));
                int64_t cb = readfirstlane64(blk + 1 < n_blk ? blk_curve_at(A, blk_curve, blk + 1) : A.n_curve - 1);
                cb = cb > ca ? cb : ca;
                const int64_t cn = ca == cb ? ca : readfirstlane64(curve_in(A, ca, cb, next));
                State N;
                N.cr = cn;
                N.s0 = A.soff[cn];
                N.s1 = A.soff[cn + 1];
                const double clv = A.curve_len[cn];
                const double sc = ev.scale((int)cn);
                N.deep = A.deep[cn];
                int64_t ro = cn << A.log2_nsub;
                N.ls = A.log2_nsub;
                if constexpr (kSeeded) {
                    ro = A.root_off[cn];
                    N.ls = A.root_k[cn];
                }
                const double v = lane < (1 << N.ls) ? A.root_prefix[ro + lane] : 0.0;
                const int64_t Ns = N.s1 - N.s0;
                const double den = (double)(A.uniform_closed ? Ns - 1 : Ns);
                N.den_pos = den > 0.0;
                N.rN = rcp_f64(den > 0.0 ? den : 1.0);
                N.cl = clv;
                N.inv_sc = rcp_f64(sc);
                N.ctx = ev.load((int)cn, N.inv_sc);
                N.c64 = ctx64_early<Ev64, typename Ev::Ctx>(ev64, (int)cn, N.inv_sc);
                // axis ratio >= 0.2: the quintic start is within 3.2e-5 of a root's width, so one FP64
                // Newton step lands within 6e-11 of it (FP32 builds with the FP64 step only)
                N.quintic = sizeof(Dt) < sizeof(double) && A.fp64_step && ev.axis_ratio2(N.ctx) >= 0.04;
                S = uniform_pod(N);
                rwin[lane] = v;
                if (lane < (1 << S.ls)) {
                    // the Hermite start's end speeds of every root of the curve
                    Dt ra, rm, rb;
                    task_bounds<Dt>(lane, 0, 0u, S.ls, ra, rm, rb);
                    const LeafEnds<Dt> e = S.quintic ? leaf_ends_quintic(ev, S.ctx, ra, rb) : leaf_ends(ev, S.ctx, ra, rb);
                    ends[6 * lane] = e.va;
                    ends[6 * lane + 1] = e.vb;
                    ends[6 * lane + 2] = e.rva;
                    ends[6 * lane + 3] = e.rvb;
                    ends[6 * lane + 4] = e.xa2;
                    ends[6 * lane + 5] = e.xb2;
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
            }
            if (g >= next && g < S.s1 && g <= gl) {
                // target as target_of (the same FP64 operations, so samples
                // at root boundaries pick the same root as the per-sample path)
                const int64_t j = g - S.s0;
                const double pos = S.den_pos ? (double)j * S.rN : 0.0;
                const double sv = A.is_relative ? pos * S.cl : pos;
                const double sn = sv * S.inv_sc;
                int k;
                double pre, rlen;
                find_root((const lds_double*)rwin, S.ls, sn, k, pre, rlen);
                const LeafEnds<Dt> E{ends[6 * k],     ends[6 * k + 1], ends[6 * k + 2],
                                     ends[6 * k + 3], ends[6 * k + 4], ends[6 * k + 5]};
                A.out_u[g] = finish_sample<Dt, NGL>(ev, ev64, A, (int)S.cr, S.ctx, S.c64, S.deep, S.ls, k, sn, pre,
                                                    rlen, true, E, S.quintic != 0);
                if (A.out_seg) A.out_seg[g] = 0;
            }
            // S.s1 <= next only for offsets that do not cover [0, n_total):
            // the rest of the tile is left unwritten instead of looping
            if (gl < S.s1 || S.s1 <= next) break;
            next = S.s1;
        }
    }
}

}  // namespace dev
}  // namespace gcurval
