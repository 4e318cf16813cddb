// Header-only CPU engine for user curve plug-ins: the template
// gcurval::ArclengthCalculateCpu<Curve, CurveTraits, max_recursive_depth> with
// the constructors, protected members and methods of the reference's
// (gcurval/detail/cpu/arclength_calculate.h:10-440). Results are always host
// arrays the caller deletes; there is no clearArclength.
//
// The algorithm is the reference CPU algorithm, so that a plug-in's CPU
// subclass gives the reference's numbers:
//   * table build: each knot interval [begin, end] of a curve is bisected
//     recursively while |L + R - P| > tol 2^-level (level < max depth), with
//     the plug-in's length rule; every split adds the row (mid, cumulative
//     length, interval) and every interval adds (end, interval length,
//     interval), then the interval's rows are offset by the curve's
//     cumulative length (double accumulation for float engines, compensated
//     for double ones);
//   * sampling: lower_bound of the target in the curve's cumulative column,
//     then a target-directed descent of that row's parameter range that
//     returns the midpoint of the leaf half holding the target once the half
//     is narrower than the tolerance or the depth cap is reached.
// Curves are independent, so setThreads(n) builds and samples n curve shards
// at once with identical results (the reference runs one thread).
//
// Host code only: it compiles with hipcc or with a plain host compiler.
#pragma once

#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <thread>
#include <vector>

#include "base.h"

namespace gcurval {

namespace cpu_detail {

// Accumulator of the wider type: double for float and double engines
// (wider_real_type of gcurval/base/util.h:23-49).
using Wide = double;

// sum + a for a float (a wider accumulator is exact enough) and the
// compensated form sum + (a - carry) for a double
// (precisely_summation, gcurval/base/util.h:134-151).
inline Wide add_to(Wide sum, float a, Wide) { return sum + a; }
inline Wide add_to(Wide sum, double a, Wide carry) { return sum + (a - carry); }

// sum += a in place; Kahan's compensated step for double addends
// (precisely_summation_ref, gcurval/base/util.h:153-172).
inline void accumulate(Wide& sum, float a, Wide&) { sum += a; }
inline void accumulate(Wide& sum, double a, Wide& carry) {
    const Wide y = a - carry;
    const Wide t = sum + y;
    carry = (t - sum) - y;
    sum = t;
}

// Host error bound of recursion level l: tol 2^-l in Dt
// (h_eps tables of gcurval/base/constants.inl.h:22-29).
template <typename Dt>
inline Dt level_bound(Dt tol, int l) {
    return Dt(std::ldexp((double)tol, -l));
}

// Midpoint (begin + end) / 2 with the sum rounded in Dt first, as the
// reference writes it.
template <typename Dt>
inline Dt midpoint(Dt a, Dt b) {
    return Dt((a + b) / 2.0);
}

}  // namespace cpu_detail

template <typename Curve, typename CurveTraits,
          int max_recursive_depth = MaxRecursiveDepth<typename Curve::DataType>::value>
class ArclengthCalculateCpu {
public:
    using Dt = typename Curve::DataType;
    using PointType = typename Curve::PointType;
    using CalculateFunction = typename CurveTraits::CalculateFunction;
    using ParaPackParser = typename CurveTraits::ParaPackParser;
    using ParaPackSaver = typename CurveTraits::ParaPackSaver;
    using SharedMemAllocator = typename CurveTraits::SharedMemAllocator;
    using CurveLocator = typename CurveTraits::CurveLocator;
    using GetPoint = typename CurveTraits::GetPoint;
    using ParaPack = typename CurveTraits::ParaPack;

    static constexpr int MAX_RECURSIVE_DEPTH = max_recursive_depth;
    static_assert(max_recursive_depth >= 0 && max_recursive_depth < 64, "max_recursive_depth out of range");

    explicit ArclengthCalculateCpu(int n_curve, int n_seg) : n_seg(n_seg) {
        if (n_curve < 0 || n_seg < 0) throw std::invalid_argument("gcurval: negative curve or interval count");
        curve.n_curve = n_curve;
    }
    explicit ArclengthCalculateCpu(int n_curve) : ArclengthCalculateCpu(n_curve, n_curve) {}

    virtual ~ArclengthCalculateCpu() { freeArray_s(arclength); }
    ArclengthCalculateCpu(const ArclengthCalculateCpu&) = delete;
    ArclengthCalculateCpu& operator=(const ArclengthCalculateCpu&) = delete;

    // Absolute error bound (default MAX_ERROR_FLOAT / MAX_ERROR_DOUBLE): the
    // level-0 bound of every knot interval and the sampler's stop width.
    void setTolerance(double tol) {
        if (!(tol > 0.0)) throw std::invalid_argument("gcurval: tolerance must be positive");
        tol_ = Dt(tol);
    }
    // Host threads over curve shards (1: the reference's single thread; 0:
    // every hardware thread). Results do not depend on it.
    void setThreads(int n) {
        if (n < 0) throw std::invalid_argument("gcurval: negative thread count");
        threads_ = n;
    }

    // createArclengthTable (cpu/arclength_calculate.h:80-139).
    void createArclengthTable() {
        const int n_curve = curve.n_curve;
        rows_.assign((size_t)n_curve, Rows{});
        freeArray_s(arclength);
        for_shards([this](int c0, int c1) {
            std::vector<uint8_t> scratch(scratch_bytes());
            for (int c = c0; c < c1; c++) build_curve(c, scratch.data());
        });
        built_ = true;
    }

    // Per-curve lengths as a new array the caller delete[]s (:141-144).
    Dt* getAllCurveArclengths() {
        lengths();
        Dt* out = new Dt[(size_t)curve.n_curve];
        std::copy(arclength, arclength + curve.n_curve, out);
        return out;
    }

    Dt getArclength(int cid) {
        if (cid < 0 || cid >= curve.n_curve) throw std::out_of_range("gcurval: curve id out of range");
        lengths();
        return arclength[cid];
    }

    // arclengthReverse (:152-238): n_ptrs_for_curves samples per curve at the
    // concatenated sample_position (fractions of the curve length when
    // is_relative, else arc lengths; nullptr: uniform fractions j / (n - 1)).
    // Results are host arrays (delete[]) of parameters u or of curve points;
    // result_is_device is accepted and, as in the reference, ignored.
    ArclengthReverseResult<Dt, PointType> arclengthReverse(int* n_ptrs_for_curves, Dt* sample_position,
                                                           bool is_relative,
                                                           GcurvalEnum result_is_device = gcurvalHost,
                                                           GcurvalEnum result_u_or_point = gcurvalReverseReturnUs) {
        (void)result_is_device;
        lengths();
        const int n_curve = curve.n_curve;
        std::vector<int64_t> off((size_t)n_curve + 1, 0);
        for (int c = 0; c < n_curve; c++) {
            if (n_ptrs_for_curves[c] < 0) throw std::invalid_argument("gcurval: negative sample count");
            off[c + 1] = off[c] + n_ptrs_for_curves[c];
        }
        if (off[n_curve] > INT32_MAX) throw std::invalid_argument("gcurval: more than 2^31 - 1 samples");
        ArclengthReverseResult<Dt, PointType> res;
        res.n_points = (int)off[n_curve];
        res.store_where = gcurvalHost;
        res.store_what = result_u_or_point;
        const bool points = result_u_or_point == gcurvalReverseReturnPoints;
        if (points) {
            if constexpr (std::is_same<GetPoint, void>::value)
                throw std::invalid_argument("gcurval: points requested from a plug-in without GetPoint");
            else
                res.points = new PointType[(size_t)res.n_points];
        } else {
            res.us = new Dt[(size_t)res.n_points];
        }
        for_shards([&](int c0, int c1) {
            std::vector<uint8_t> scratch(scratch_bytes());
            for (int c = c0; c < c1; c++) {
                const int64_t n = off[c + 1] - off[c];
                for (int64_t j = 0; j < n; j++) {
                    const int64_t g = off[c] + j;
                    Dt pos;
                    if (sample_position) pos = sample_position[g];
                    else pos = n > 1 ? Dt((double)j / (double)(n - 1)) : Dt(0);
                    sample_curve(c, is_relative ? Dt(pos * arclength[c]) : pos, scratch.data(), res, g, points);
                }
            }
        });
        return res;
    }

protected:
    Curve curve;
    int n_seg = 0;
    // Per-curve length column (the reference's `arclength`), filled from the
    // tables on first use.
    Dt* arclength = nullptr;

    // Table of one curve: parameter, cumulative length along the curve and
    // knot-interval index of every row (u_table / arclength_table /
    // segid_table of the reference).
    struct Rows {
        std::vector<Dt> u, s;
        std::vector<int> seg;
    };
    std::vector<Rows> rows_;

private:
    Dt tol_ = Dt(default_max_error<Dt>());
    int threads_ = 1;
    bool built_ = false;

    static size_t scratch_bytes() {
        const int n = SharedMemAllocator::size(1);
        return (size_t)(n > 0 ? n : 1);
    }

    // The plug-in's length rule over [b, e], with the per-thread scratch view
    // when the plug-in has one (DoArclengthCalculateStep, functors.inl.h:271-292).
    Dt rule(Dt b, Dt e, const ParaPack& pack, uint8_t* scratch) const {
        if constexpr (std::is_same<SharedMemAllocator, NoSharedMemAllocator>::value) {
            (void)scratch;
            return CalculateFunction{}(b, e, pack);
        } else {
            auto view = SharedMemAllocator{}(scratch);
            return CalculateFunction{}(b, e, pack, view);
        }
    }

    template <typename F>
    void for_shards(F&& f) {
        const int n_curve = curve.n_curve;
        int n = threads_ > 0 ? threads_ : (int)std::max(1u, std::thread::hardware_concurrency());
        n = std::max(1, std::min(n, n_curve));
        if (n <= 1) {
            f(0, n_curve);
            return;
        }
        std::vector<std::thread> pool;
        for (int t = 0; t < n; t++) {
            const int c0 = (int)((int64_t)n_curve * t / n), c1 = (int)((int64_t)n_curve * (t + 1) / n);
            pool.emplace_back([&f, c0, c1]() { f(c0, c1); });
        }
        for (auto& th : pool) th.join();
    }

    void lengths() {
        if (!built_) throw std::logic_error("gcurval: createArclengthTable must run first");
        if (arclength) return;
        arclength = new Dt[(size_t)curve.n_curve];
        for (int c = 0; c < curve.n_curve; c++)
            arclength[c] = rows_[c].s.empty() ? Dt(0) : rows_[c].s.back();
    }

    // Bisection of [b, e] (length estimate `whole`) below level `level`:
    // appends a row at every split midpoint with the length accumulated
    // before it (`before`, interval-local) and returns the refined length
    // (arclengthRecursive, cpu/arclength_calculate.h:242-278).
    Dt refine(Dt b, Dt e, Dt whole, int level, const ParaPack& pack, uint8_t* scratch, cpu_detail::Wide before,
              int seg, Rows& r) const {
        const Dt m = cpu_detail::midpoint(b, e);
        Dt left = rule(b, m, pack, scratch);
        Dt right = rule(m, e, pack, scratch);
        if (level < max_recursive_depth &&
            std::abs(left + right - whole) > cpu_detail::level_bound(tol_, level)) {
            left = refine(b, m, left, level + 1, pack, scratch, before, seg, r);
            before += left;
            r.u.push_back(m);
            r.s.push_back(Dt(before));
            r.seg.push_back(seg);
            right = refine(m, e, right, level + 1, pack, scratch, before, seg, r);
        }
        return left + right;
    }

    void build_curve(int c, uint8_t* scratch) {
        Rows& r = rows_[c];
        const CurveIntervalPos range = CurveLocator{}(curve, c);
        const int n_int = range.end - range.begin;
        cpu_detail::Wide along = 0, carry = 0;  // curve length so far
        for (int seg = 0; seg < n_int; seg++) {
            const ParaPack pack = ParaPackParser{}(curve, c, seg, true);
            const Dt whole = rule(pack.begin, pack.end, pack, scratch);
            const size_t first = r.s.size();
            const cpu_detail::Wide len = refine(pack.begin, pack.end, whole, 0, pack, scratch, 0, seg, r);
            r.u.push_back(pack.end);
            r.s.push_back(Dt(len));
            r.seg.push_back(seg);
            for (size_t k = first; k < r.s.size(); k++) r.s[k] = Dt(cpu_detail::add_to(along, r.s[k], carry));
            // len is of the wide type, so the compensated step applies to
            // float engines too (cpu/arclength_calculate.h:123)
            cpu_detail::accumulate(along, len, carry);
        }
    }

    // Target-directed descent of [b, e] (arclengthRecursiveWithTarget_Paper,
    // cpu/arclength_calculate.h:378-439): `done` is the length accumulated
    // before b; *u receives the midpoint of the leaf half holding `target`.
    Dt descend(Dt b, Dt e, Dt whole, Dt target, cpu_detail::Wide done, cpu_detail::Wide carry, int level,
               const ParaPack& pack, uint8_t* scratch, Dt& u, bool& found) const {
        const Dt m = cpu_detail::midpoint(b, e);
        Dt left = rule(b, m, pack, scratch);
        Dt right = rule(m, e, pack, scratch);
        const bool at_cap = level == max_recursive_depth;
        if (level < max_recursive_depth && std::abs((left + right) - whole) > cpu_detail::level_bound(tol_, level)) {
            left = descend(b, m, left, target, done, carry, level + 1, pack, scratch, u, found);
            if (!found) {
                cpu_detail::accumulate(done, left, carry);
                right = descend(m, e, right, target, done, carry, level + 1, pack, scratch, u, found);
            }
        } else if (target <= cpu_detail::add_to(done, left, carry)) {
            if (m - b < tol_ || at_cap) {
                u = Dt((b + m) / Dt(2.0));
                found = true;
            } else {
                left = descend(b, m, left, target, done, carry, level + 1, pack, scratch, u, found);
            }
        } else if (target <= cpu_detail::add_to(done, Dt(left + right), carry)) {
            if (e - m < tol_ || at_cap) {
                u = Dt((m + e) / Dt(2.0));
                found = true;
            } else {
                cpu_detail::accumulate(done, left, carry);
                right = descend(m, e, right, target, done, carry, level + 1, pack, scratch, u, found);
            }
        }
        return left + right;
    }

    // One sample of curve c at curve-length target (arclengthReverse's loop
    // body, cpu/arclength_calculate.h:191-232). The table row before the
    // first bounds it by the start of the curve's first knot interval (the
    // reference writes 0, which that start is for its sample families).
    void sample_curve(int c, Dt target, uint8_t* scratch, ArclengthReverseResult<Dt, PointType>& res, int64_t g,
                      bool points) const {
        const Rows& r = rows_[c];
        const int n = (int)r.s.size();
        if (n == 0) {  // a curve without knot intervals
            if (points) res.points[g] = PointType{};
            else res.us[g] = Dt(0);
            return;
        }
        int it = (int)(std::lower_bound(r.s.begin(), r.s.end(), target) - r.s.begin());
        if (it == n) it = n - 1;
        Dt b;
        int seg = 0;
        if (it > 0) {
            target -= r.s[it - 1];
            b = r.u[it - 1];
            seg = r.seg[it];
        } else {
            b = ParaPackParser{}(curve, c, 0, true).begin;
        }
        const Dt e = r.u[it];
        const ParaPack pack = ParaPackParser{}(curve, c, seg);
        const Dt whole = rule(b, e, pack, scratch);
        Dt u = e;
        bool found = false;
        descend(b, e, whole, target, 0.0, 0.0, 0, pack, scratch, u, found);
        if constexpr (!std::is_same<GetPoint, void>::value) {
            if (points) {
                if constexpr (std::is_same<SharedMemAllocator, NoSharedMemAllocator>::value) {
                    res.points[g] = GetPoint{}(u, pack);
                } else {
                    auto view = SharedMemAllocator{}(scratch);
                    res.points[g] = GetPoint{}(u, pack, view);
                }
                return;
            }
        }
        res.us[g] = u;
    }
};

}  // namespace gcurval
