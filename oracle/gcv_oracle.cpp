// gcurval oracle — CPU restatement of gcurval's arc-length table build and
// arc-length sampling, used ONLY as test infrastructure.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load this library, and only as the reference results are compared with (or
// as the timed CPU baseline).
// The product path (gcurval_amd/) never links, loads or calls it.
//
// Three engines live here:
//   * RefCpu      — restates ArclengthCalculateCpu
//                   (gcurval/detail/cpu/arclength_calculate.h:10-440): no
//                   pre-split, host epsilon table, recursive bisection, table
//                   of (u, cumulative length, segid), recursive target search.
//   * RefGpuPlain — serial restatement of the GPU engine's semantics
//                   (gcurval/detail/cuda/arclength_calculate.h:185-505 and the
//                   kernels in kernel/generic.inl.cuh, bfs.inl.cuh,
//                   reverse.inl.cuh): n_sub-way pre-split, device epsilon
//                   table (tol / n_sub), pairwise tree sums, Brent-Kung prefix
//                   in Dt, double-accumulated per-curve merge, DFS sampler
//                   returning a leaf-half midpoint.
//   * Truth       — long-double adaptive 20-point Gauss-Legendre to ~1e-17
//                   relative, plus a leaf-bracketed Newton inverse map.
//
// Parity pins (see tests/test_oracle_*.py): GL tables vs numpy leggauss,
// ellipse lengths vs the closed form 4aE(m) (mpmath fixture), B-spline
// derivatives vs scipy BSpline.derivative, B-spline lengths vs mpmath quad,
// degree-1 polyline and collinear special cases, scale equivariance,
// additivity, brute-force tree enumeration on tiny depths.
//
// Built with -ffp-contract=off so that float rounding follows the written
// expression order.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <thread>
#include <vector>

namespace orc {

// ---------------------------------------------------------------------------
// Gauss-Legendre tables, orders 1..5, in the node order the reference uses
// (gcurval/base/constants.inl.h:74-92). The values are the textbook ones;
// the order matters for float rounding of the node sum.
// ---------------------------------------------------------------------------
static const double GLX[5][5] = {
    {0., 0., 0., 0., 0.},
    {-0.5773502691896258, 0.5773502691896258, 0., 0., 0.},
    {0., -0.7745966692414834, 0.7745966692414834, 0., 0.},
    {-0.33998104358485626, 0.33998104358485626, -0.8611363115940526, 0.8611363115940526, 0.},
    {0., -0.5384693101056831, 0.5384693101056831, -0.906179845938664, 0.906179845938664}};
static const double GLW[5][5] = {
    {2., 0., 0., 0., 0.},
    {1., 1., 0., 0., 0.},
    {0.8888888888888889, 0.5555555555555556, 0.5555555555555556, 0., 0.},
    {0.6521451548625462, 0.6521451548625462, 0.34785484513745385, 0.34785484513745385, 0.},
    {0.5688888888888889, 0.47862867049936647, 0.47862867049936647, 0.23692688505618908,
     0.23692688505618908}};

// ---------------------------------------------------------------------------
// Curve batch description shared with Python (ctypes mirror in tests).
// ---------------------------------------------------------------------------
extern "C" {
typedef struct orc_batch {
    int kind;                 // 0 = ellipse, 1 = B-spline
    int n_curve;
    const double* a;          // ellipse semi-axes (kind 0)
    const double* b;
    const int* degree;        // per curve (kind 1)
    const int* dim;           // per curve, 2 or 3 (kind 1)
    const int64_t* knot_off;  // n_curve+1 offsets into knots
    const int64_t* ctrl_off;  // n_curve+1 offsets (in points) into ctrl
    const double* knots;
    const double* ctrl;       // point j of curve c: ctrl[(ctrl_off[c]+j)*ctrl_stride + d]
    int ctrl_stride;          // 2 or 3
} orc_batch;

typedef struct orc_opts {
    int precision;      // 0 = float, 1 = double
    int gl_order;       // 1..5
    int n_sub;          // pre-split count (GPU-plain semantics)
    int max_depth;      // recursion cap
    double tol;         // absolute tolerance (reference semantics)
    int normalize;      // 1: evaluate each curve scaled to unit max-abs coordinate
} orc_opts;
}

static int n_intervals(const orc_batch& B, int c) {
    if (B.kind == 0) return 1;
    int p = B.degree[c];
    return (int)(B.knot_off[c + 1] - B.knot_off[c]) - 2 * p - 1;
}

// Per-curve normalisation scale: the max-abs coordinate after centring on
// the control-point centroid (samples/sample2_bspline/dataset_loader.h:70-96),
// or max(a, b) for an ellipse (samples/sample1_ellipse/main.cpp:32-37).
static double curve_scale(const orc_batch& B, int c) {
    if (B.kind == 0) return std::max(std::fabs(B.a[c]), std::fabs(B.b[c]));
    int D = B.dim[c];
    int64_t n = B.ctrl_off[c + 1] - B.ctrl_off[c];
    const double* P = B.ctrl + B.ctrl_off[c] * B.ctrl_stride;
    double cen[3] = {0, 0, 0};
    for (int64_t j = 0; j < n; j++)
        for (int d = 0; d < D; d++) cen[d] += P[j * B.ctrl_stride + d];
    for (int d = 0; d < D; d++) cen[d] /= (double)n;
    double m = 0.0;
    for (int64_t j = 0; j < n; j++)
        for (int d = 0; d < D; d++) m = std::max(m, std::fabs(P[j * B.ctrl_stride + d] - cen[d]));
    return m > 0.0 ? m : 1.0;
}

// ---------------------------------------------------------------------------
// Reference-semantics curve evaluators in data type Dt.
// ---------------------------------------------------------------------------
template <typename Dt>
struct Curve {
    // ellipse
    Dt a = 0, b = 0;
    // B-spline (curve-local arrays)
    int p = 0, D = 0;
    std::vector<Dt> T;      // knots
    std::vector<Dt> P;      // control points, D per point
    int n_int = 0;
    double scale = 1.0;     // normalisation scale (lengths in curve units = scale x evaluated)

    // interval [begin, end] in u of local interval segid
    // (ellipse: samples/sample1_ellipse/ellipse.h:60-64; B-spline:
    //  samples/sample2_bspline/structs.h:126-150).
    void interval(int segid, Dt& u0, Dt& u1) const {
        if (D == 0) {
            u0 = Dt(0);
            u1 = Dt(2 * 3.14159265358979323846264338327950288L);
        } else {
            u0 = T[segid + p];
            u1 = T[segid + p + 1];
        }
    }

    // |C'(u)| (ellipse.h:11-15; BSpline.h:50-96 with Eigen norm)
    Dt speed(Dt u, int segid) const {
        if (D == 0) {
            Dt dx = (Dt)(-1.0 * a * std::sin(u));
            Dt dy = (Dt)(b * std::cos(u));
            return std::sqrt(dx * dx + dy * dy);
        }
        Dt d[3];
        deriv(u, segid + p, d);
        Dt s = 0;
        for (int k = 0; k < D; k++) s += d[k] * d[k];
        return std::sqrt(s);
    }

    // First derivative on span s (local knot index): Cox-de Boor triangle of
    // degree p-1, then N'_{i,p} = p (N_{i,p-1}/(t_{i+p}-t_i) - N_{i+1,p-1}/(t_{i+p+1}-t_{i+1}))
    // written with the left/right differences (The NURBS Book, A2.2/eq. 2.7).
    void deriv(Dt u, int s, Dt* out) const {
        Dt N[8], left[8], right[8];
        N[0] = Dt(1);
        for (int j = 1; j < p; j++) {
            left[j] = u - T[s + 1 - j];
            right[j] = T[s + j] - u;
            Dt saved = 0;
            for (int r = 0; r < j; r++) {
                Dt tmp = N[r] / (right[r + 1] + left[j - r]);
                N[r] = saved + right[r + 1] * tmp;
                saved = left[j - r] * tmp;
            }
            N[j] = saved;
        }
        left[p] = u - T[s + 1 - p];
        right[p] = T[s + p] - u;
        Dt saved = 0;
        for (int r = 0; r < p; r++) {
            Dt tmp = p * N[r] / (right[r + 1] + left[p - r]);
            N[r] = saved - tmp;
            saved = tmp;
        }
        N[p] = saved;
        for (int k = 0; k < D; k++) {
            Dt acc = P[(s - p) * D + k] * N[0];
            for (int i = 1; i <= p; i++) acc += P[(s - p + i) * D + k] * N[i];
            out[k] = acc;
        }
    }
};

template <typename Dt>
static Curve<Dt> load_curve(const orc_batch& B, int c, bool normalize) {
    Curve<Dt> cv;
    double sc = normalize ? curve_scale(B, c) : 1.0;
    cv.scale = sc;
    if (B.kind == 0) {
        cv.a = (Dt)(B.a[c] / sc);
        cv.b = (Dt)(B.b[c] / sc);
        cv.n_int = 1;
        return cv;
    }
    cv.p = B.degree[c];
    cv.D = B.dim[c];
    int64_t nk = B.knot_off[c + 1] - B.knot_off[c];
    int64_t np = B.ctrl_off[c + 1] - B.ctrl_off[c];
    cv.T.resize(nk);
    for (int64_t i = 0; i < nk; i++) cv.T[i] = (Dt)B.knots[B.knot_off[c] + i];
    cv.P.resize(np * cv.D);
    const double* P = B.ctrl + B.ctrl_off[c] * B.ctrl_stride;
    for (int64_t j = 0; j < np; j++)
        for (int d = 0; d < cv.D; d++) cv.P[j * cv.D + d] = (Dt)(P[j * B.ctrl_stride + d] / sc);
    cv.n_int = (int)nk - 2 * cv.p - 1;
    return cv;
}

// Gauss-Legendre quadrature in the reference's arithmetic
// (gcurval/base/functors.inl.h:46-66).
template <typename Dt>
static Dt gl(const Curve<Dt>& cv, int segid, int order, Dt begin, Dt end) {
    Dt k = (Dt)((end - begin) / 2.0);
    Dt b = (Dt)((end + begin) / 2.0);
    Dt acc = 0;
    for (int j = 0; j < order; j++) {
        Dt u = k * (Dt)GLX[order - 1][j] + b;
        acc += (Dt)GLW[order - 1][j] * cv.speed(u, segid);
    }
    acc *= k;
    return acc;
}

// precisely_summation / precisely_summation_ref (gcurval/base/util.h:134-172)
static inline double psum(double sum, float a, double) { return sum + a; }
static inline double psum(double sum, double a, double residual) { return sum + (a - residual); }
static inline void psum_ref(double& sum, float a, double&) { sum += a; }
static inline void psum_ref(double& sum, double a, double& residual) {
    double y = a - residual;
    double t = sum + y;
    residual = (t - sum) - y;
    sum = t;
}

// Epsilon tables (gcurval/base/constants.inl.h:8-34): host eps(l) = tol 2^-l,
// device eps(l) = (tol / n_sub) 2^-l, both computed in Dt.
template <typename Dt>
static Dt eps_level(double tol, int n_div, int level) {
    Dt base = (Dt)((Dt)tol / (Dt)n_div);
    return (Dt)(std::ldexp((double)base, -level));
}

// ===========================================================================
// RefCpu — ArclengthCalculateCpu (gcurval/detail/cpu/arclength_calculate.h)
// ===========================================================================
template <typename Dt>
struct RefCpu {
    const orc_batch& B;
    orc_opts o;
    std::vector<Curve<Dt>> curves;
    std::vector<std::vector<Dt>> u_tab, s_tab;
    std::vector<std::vector<int>> seg_tab;
    std::vector<Dt> length;

    RefCpu(const orc_batch& b, const orc_opts& op) : B(b), o(op) {
        curves.reserve(B.n_curve);
        for (int c = 0; c < B.n_curve; c++) curves.push_back(load_curve<Dt>(B, c, o.normalize != 0));
        u_tab.resize(B.n_curve);
        s_tab.resize(B.n_curve);
        seg_tab.resize(B.n_curve);
        length.resize(B.n_curve);
    }

    Dt eps(int l) const { return eps_level<Dt>(o.tol, 1, l); }
    Dt reverse_width() const { return (Dt)o.tol; }

    // arclengthRecursive (cpu/arclength_calculate.h:242-278)
    Dt rec(const Curve<Dt>& cv, int segid, Dt begin, Dt end, Dt parent, int level, double total,
           std::vector<Dt>& us, std::vector<Dt>& ss, std::vector<int>& sg) {
        Dt mid = (Dt)((begin + end) / 2.0);
        Dt left = gl(cv, segid, o.gl_order, begin, mid);
        Dt right = gl(cv, segid, o.gl_order, mid, end);
        if (level < o.max_depth && std::fabs(left + right - parent) > eps(level)) {
            left = rec(cv, segid, begin, mid, left, level + 1, total, us, ss, sg);
            total += left;
            us.push_back(mid);
            ss.push_back((Dt)total);
            sg.push_back(segid);
            right = rec(cv, segid, mid, end, right, level + 1, total, us, ss, sg);
        }
        return left + right;
    }

    // createArclengthTable (cpu/arclength_calculate.h:80-139)
    void build_curve(int c) {
        const Curve<Dt>& cv = curves[c];
        std::vector<Dt>& us = u_tab[c];
        std::vector<Dt>& ss = s_tab[c];
        std::vector<int>& sg = seg_tab[c];
        us.clear();
        ss.clear();
        sg.clear();
        double acc_curve = 0, residual = 0;
        for (int segid = 0; segid < cv.n_int; segid++) {
            Dt b, e;
            cv.interval(segid, b, e);
            double acc_seg = 0;
            Dt parent = gl(cv, segid, o.gl_order, b, e);
            size_t start = ss.size();
            acc_seg = rec(cv, segid, b, e, parent, 0, acc_seg, us, ss, sg);
            us.push_back(e);
            ss.push_back((Dt)acc_seg);
            sg.push_back(segid);
            for (size_t i = start; i < ss.size(); i++) ss[i] = (Dt)psum(acc_curve, ss[i], residual);
            // acc_seg is the wide type, so the reference deduces the Kahan
            // overload here even in float mode (cpu/arclength_calculate.h:123).
            psum_ref(acc_curve, acc_seg, residual);
        }
        length[c] = ss.empty() ? Dt(0) : ss.back();
    }

    void build(int n_threads) {
        if (n_threads <= 1) {
            for (int c = 0; c < B.n_curve; c++) build_curve(c);
            return;
        }
        // Shards balanced by interval count; each curve is independent.
        std::vector<int64_t> w(B.n_curve + 1, 0);
        for (int c = 0; c < B.n_curve; c++) w[c + 1] = w[c] + curves[c].n_int + 1;
        std::vector<std::thread> th;
        for (int t = 0; t < n_threads; t++) {
            th.emplace_back([this, &w, t, n_threads]() {
                int64_t lo = w.back() * t / n_threads, hi = w.back() * (t + 1) / n_threads;
                int c0 = (int)(std::lower_bound(w.begin(), w.end(), lo) - w.begin());
                for (int c = c0; c < B.n_curve && w[c] < hi; c++)
                    if (w[c] >= lo) build_curve(c);
            });
        }
        for (auto& x : th) x.join();
    }

    // arclengthRecursiveWithTarget_Paper (cpu/arclength_calculate.h:378-439)
    Dt rec_target(const Curve<Dt>& cv, int segid, Dt begin, Dt end, Dt parent, Dt target, double acc,
                  double residual, int level, Dt& u, bool& found) {
        Dt mid = (Dt)((begin + end) / 2.0);
        Dt left = gl(cv, segid, o.gl_order, begin, mid);
        Dt right = gl(cv, segid, o.gl_order, mid, end);
        if (level < o.max_depth && std::fabs((left + right) - parent) > eps(level)) {
            left = rec_target(cv, segid, begin, mid, left, target, acc, residual, level + 1, u, found);
            if (!found) {
                psum_ref(acc, left, residual);
                right = rec_target(cv, segid, mid, end, right, target, acc, residual, level + 1, u, found);
            }
        } else if ((double)target <= psum(acc, left, residual)) {
            if (mid - begin < reverse_width() || level == o.max_depth) {
                u = (Dt)((begin + mid) / Dt(2.0));
                found = true;
            } else {
                left = rec_target(cv, segid, begin, mid, left, target, acc, residual, level + 1, u, found);
            }
        } else if ((double)target <= psum(acc, (Dt)(left + right), residual)) {
            if (end - mid < reverse_width() || level == o.max_depth) {
                u = (Dt)((mid + end) / Dt(2.0));
                found = true;
            } else {
                psum_ref(acc, left, residual);
                right = rec_target(cv, segid, mid, end, right, target, acc, residual, level + 1, u, found);
            }
        }
        return left + right;
    }

    // arclengthReverse (cpu/arclength_calculate.h:152-238), us output.
    // First-entry begin is the curve's first knot (the reference hard-codes
    // 0.0, which equals it for the ellipse and for clamped knots on [0, 1]).
    void reverse_range(const int64_t* off, const double* pos, int is_relative, double* out, int c0,
                       int c1) {
        for (int c = c0; c < c1; c++) {
            const Curve<Dt>& cv = curves[c];
            const std::vector<Dt>& ss = s_tab[c];
            const std::vector<Dt>& us = u_tab[c];
            const std::vector<int>& sg = seg_tab[c];
            int n = (int)ss.size();
            Dt first_u0, dummy;
            cv.interval(0, first_u0, dummy);
            for (int64_t i = off[c]; i < off[c + 1]; i++) {
                Dt target = is_relative ? (Dt)((Dt)pos[i] * length[c]) : (Dt)(pos[i] / cv.scale);
                int it = (int)(std::lower_bound(ss.begin(), ss.end(), target) - ss.begin());
                if (it == n) it = n - 1;
                Dt begin = first_u0, end;
                int segid = 0;
                if (it > 0) {
                    target -= ss[it - 1];
                    begin = us[it - 1];
                    segid = sg[it];
                }
                end = us[it];
                Dt parent = gl(cv, segid, o.gl_order, begin, end);
                Dt u = end;
                bool found = false;
                rec_target(cv, segid, begin, end, parent, target, 0.0, 0.0, 0, u, found);
                out[i] = (double)u;
            }
        }
    }
};

// ===========================================================================
// RefGpuPlain — serial restatement of ArclengthCalculateCuda semantics
// ===========================================================================
template <typename Dt>
struct RefGpuPlain {
    const orc_batch& B;
    orc_opts o;
    std::vector<Curve<Dt>> curves;
    std::vector<int64_t> ioff;      // n_curve+1 global interval offsets
    std::vector<Dt> seg_cum;        // per global interval, cumulative within curve (Dt)
    std::vector<Dt> sub_prefix;     // per root, inclusive prefix within interval (Dt)
    std::vector<int> root_depth;    // per root: deepest level reached (0 = passed at level 0)
    std::vector<Dt> length;
    int64_t n_tasks = 0;            // subdivision tasks evaluated (2 GL each, +1 GL at level 0)

    RefGpuPlain(const orc_batch& b, const orc_opts& op) : B(b), o(op) {
        curves.reserve(B.n_curve);
        ioff.assign(B.n_curve + 1, 0);
        for (int c = 0; c < B.n_curve; c++) {
            curves.push_back(load_curve<Dt>(B, c, o.normalize != 0));
            ioff[c + 1] = ioff[c] + curves[c].n_int;
        }
        seg_cum.assign(ioff.back(), Dt(0));
        sub_prefix.assign(ioff.back() * o.n_sub, Dt(0));
        root_depth.assign(ioff.back() * o.n_sub, 0);
        length.assign(B.n_curve, Dt(0));
    }

    Dt eps(int l) const { return eps_level<Dt>(o.tol, o.n_sub, l); }

    // One BFS task and its subtree, summed as the backward pass does
    // (kernel/bfs.inl.cuh:9-125: children (b,m),(m,e); parent = left + right).
    Dt task(const Curve<Dt>& cv, int segid, Dt begin, Dt end, Dt parent, int level, int& depth) {
        n_tasks++;
        if (level > depth) depth = level;
        Dt mid = (begin + end) / 2.0f;
        Dt left = gl(cv, segid, o.gl_order, begin, mid);
        Dt right = gl(cv, segid, o.gl_order, mid, end);
        if (level < o.max_depth && std::fabs(left + right - parent) > eps(level)) {
            Dt l = task(cv, segid, begin, mid, left, level + 1, depth);
            Dt r = task(cv, segid, mid, end, right, level + 1, depth);
            return l + r;
        }
        return left + right;
    }

    // Brent-Kung inclusive scan in Dt (kernel/generic.inl.cuh:52-80); the
    // updates within one stride touch disjoint elements, so running the
    // lanes in sequence gives the parallel result.
    static void brent_kung(Dt* buf, int n) {
        for (int stride = 1; stride < n; stride *= 2)
            for (int lane = 0; lane < n; lane++) {
                int index = (lane + 1) * 2 * stride - 1;
                if (index < n) buf[index] += buf[index - stride];
            }
        for (int stride = n / 4; stride > 0; stride /= 2)
            for (int lane = 0; lane < n; lane++) {
                int index = (lane + 1) * stride * 2 - 1;
                if (index + stride < n) buf[index + stride] += buf[index];
            }
    }

    void build() {
        const int ns = o.n_sub;
        std::vector<Dt> seg_len(ioff.back());
        for (int c = 0; c < B.n_curve; c++) {
            const Curve<Dt>& cv = curves[c];
            for (int segid = 0; segid < cv.n_int; segid++) {
                int64_t g = ioff[c] + segid;
                Dt pb, pe;
                cv.interval(segid, pb, pe);
                // kn_curveLengthTableInit (kernel/generic.inl.cuh:10-48)
                Dt delta = (pe - pb) / Dt(ns);
                Dt* buf = &sub_prefix[g * ns];
                for (int k = 0; k < ns; k++) {
                    Dt b = pb + delta * (Dt)k;
                    Dt e = pb + delta * (Dt)(k + 1);
                    Dt parent = gl(cv, segid, o.gl_order, b, e);
                    n_tasks++;  // level-0 task costs one extra quadrature
                    int depth = 0;
                    buf[k] = task(cv, segid, b, e, parent, 0, depth);
                    root_depth[g * ns + k] = depth;
                }
                brent_kung(buf, ns);
                seg_len[g] = buf[ns - 1];
            }
            // kn_mergeSeg (kernel/generic.inl.cuh:117-137)
            int64_t b0 = ioff[c], e0 = ioff[c + 1];
            if (e0 > b0) {
                double sum = seg_len[b0], residual = 0;
                seg_cum[b0] = seg_len[b0];
                for (int64_t i = b0 + 1; i < e0; i++) {
                    psum_ref(sum, seg_len[i], residual);
                    seg_cum[i] = (Dt)sum;
                }
                length[c] = seg_cum[e0 - 1];
            }
        }
    }

    // kn_initSearchTask + kn_arclengthReverseDfsFast_Paper
    // (kernel/reverse.inl.cuh:32-99, 216-341). Targets beyond the last
    // cumulative entry are clamped into the last interval / sub-interval
    // (the reference reads past the end there).
    void reverse_range(const int64_t* off, const double* pos, int is_relative, double* out,
                       int32_t* seg_out, int c0, int c1) {
        const int ns = o.n_sub;
        const Dt REV = (Dt)o.tol;
        const int MAXD = o.max_depth;
        std::vector<Dt> stack(MAXD + 2);
        for (int c = c0; c < c1; c++) {
            const Curve<Dt>& cv = curves[c];
            int64_t ib = ioff[c], ie = ioff[c + 1];
            for (int64_t i = off[c]; i < off[c + 1]; i++) {
                Dt al = is_relative ? (Dt)pos[i] : (Dt)(pos[i] / cv.scale);
                if (is_relative) al *= length[c];
                int segid = (int)(std::lower_bound(seg_cum.begin() + ib, seg_cum.begin() + ie, al) -
                                  (seg_cum.begin() + ib));
                if (segid >= (int)(ie - ib)) segid = (int)(ie - ib) - 1;
                if (segid != 0) al -= seg_cum[ib + segid - 1];
                int64_t so = (ib + segid) * ns;
                int subid = (int)(std::lower_bound(sub_prefix.begin() + so, sub_prefix.begin() + so + ns, al) -
                                  (sub_prefix.begin() + so));
                if (subid >= ns) subid = ns - 1;
                if (subid != 0) al -= sub_prefix[so + subid - 1];
                if (seg_out) seg_out[i] = segid;
                Dt pb, pe;
                cv.interval(segid, pb, pe);
                Dt delta = (pe - pb) / (Dt)ns;
                double begin = (double)(pb + delta * (Dt)subid);
                double end = (double)(pb + delta * (Dt)(subid + 1));
                Dt target = al;

                int level = 0, top = 0;
                uint32_t path = 1;
                Dt result_u = (Dt)end;
                bool found = false;
                double acc = 0, residual = 0;
                Dt parent = gl(cv, segid, o.gl_order, (Dt)begin, (Dt)end);
                while (!found && path != 0) {
                    double mid = (begin + end) / 2.0;
                    Dt left = gl(cv, segid, o.gl_order, (Dt)begin, (Dt)mid);
                    Dt right = gl(cv, segid, o.gl_order, (Dt)mid, (Dt)end);
                    Dt lr = left + right;
                    if (level < MAXD && std::fabs(lr - parent) > eps(level)) {
                        level++;
                        end = mid;
                        path <<= 1;
                        stack[++top] = right;
                        parent = left;
                    } else if ((double)target <= psum(acc, left, residual)) {
                        if (mid - begin < (double)REV || level == MAXD) {
                            found = true;
                            result_u = (Dt)((begin + mid) / 2.0);
                        } else {
                            level++;
                            end = mid;
                            path <<= 1;
                            stack[++top] = right;
                            parent = left;
                        }
                    } else if ((double)target <= psum(acc, lr, residual)) {
                        if (end - mid < (double)REV || level == MAXD) {
                            found = true;
                            result_u = (Dt)((mid + end) / 2.0);
                        } else {
                            level++;
                            begin = mid;
                            psum_ref(acc, left, residual);
                            path = (path << 1) | 1u;
                            parent = right;
                        }
                    } else {
                        uint32_t mul = path - (path & (path + 1));
                        int dr = __builtin_popcount(mul);
                        mul++;
                        level -= dr;
                        path >>= dr;
                        if (path != 0) {
                            double tmp = end;
                            end += (end - begin) * (double)mul;
                            begin = tmp;
                            psum_ref(acc, lr, residual);
                            path |= 1u;
                            parent = stack[top];
                        }
                        top--;
                    }
                }
                out[i] = (double)result_u;
            }
        }
    }
};

// ===========================================================================
// Truth — long-double adaptive 20-point Gauss-Legendre and Newton inverse
// ===========================================================================
typedef long double LD;

struct GL20 {
    LD x[20], w[20];
    GL20() {
        // Nodes by Newton iteration on P_20 from the Chebyshev-like start
        // cos(pi (i + 3/4) / (n + 1/2)); weights 2 / ((1-x^2) P'(x)^2).
        const int n = 20;
        for (int i = 0; i < n; i++) {
            LD z = std::cos(3.141592653589793238462643383279502884L * (i + 0.75L) / (n + 0.5L));
            for (int it = 0; it < 100; it++) {
                LD p1 = 1, p2 = 0;
                for (int j = 1; j <= n; j++) {
                    LD p3 = p2;
                    p2 = p1;
                    p1 = ((2 * j - 1) * z * p2 - (j - 1) * p3) / j;
                }
                LD pp = n * (z * p1 - p2) / (z * z - 1);
                LD dz = p1 / pp;
                z -= dz;
                if (std::fabs(dz) < 1e-19L) break;
            }
            LD p1 = 1, p2 = 0;
            for (int j = 1; j <= n; j++) {
                LD p3 = p2;
                p2 = p1;
                p1 = ((2 * j - 1) * z * p2 - (j - 1) * p3) / j;
            }
            LD pp = n * (z * p1 - p2) / (z * z - 1);
            x[i] = z;
            w[i] = 2 / ((1 - z * z) * pp * pp);
        }
    }
};
static const GL20& gl20() {
    static GL20 g;
    return g;
}

struct TruthCurve {
    int kind = 0;
    LD a = 0, b = 0;
    int p = 0, D = 0;
    std::vector<LD> T, P;
    int n_int = 0;
    LD scale = 1;

    void interval(int segid, LD& u0, LD& u1) const {
        if (kind == 0) {
            u0 = 0;
            u1 = 2 * 3.141592653589793238462643383279502884L;
        } else {
            u0 = T[segid + p];
            u1 = T[segid + p + 1];
        }
    }
    // derivative via derivative control points Q_i = p (P_{i+1}-P_i)/(t_{i+p+1}-t_{i+1})
    // and de Boor's algorithm on the degree p-1 curve.
    void deriv(LD u, int segid, LD* out) const {
        if (kind == 0) {
            out[0] = -a * std::sin(u);
            out[1] = b * std::cos(u);
            return;
        }
        int s = segid + p;  // span
        int q = p - 1;
        LD d[8][3];
        for (int j = 0; j <= q; j++) {
            int i = s - q + j - 1;  // derivative control index (Q_i uses P_i, P_{i+1})
            LD den = T[i + p + 1] - T[i + 1];
            for (int k = 0; k < D; k++)
                d[j][k] = den > 0 ? p * (P[(i + 1) * D + k] - P[i * D + k]) / den : 0;
        }
        // de Boor on knots T' = T[1..] (degree q): span s-1 in T' terms -> use T directly
        for (int r = 1; r <= q; r++)
            for (int j = q; j >= r; j--) {
                int i = s - q + j;  // knot index in T for the degree-q spline on T[1..]
                LD den = T[i + q - r + 1] - T[i];
                LD alpha = den > 0 ? (u - T[i]) / den : 0;
                for (int k = 0; k < D; k++) d[j][k] = (1 - alpha) * d[j - 1][k] + alpha * d[j][k];
            }
        for (int k = 0; k < D; k++) out[k] = d[q][k];
    }
    LD speed(LD u, int segid) const {
        LD d[3];
        deriv(u, segid, d);
        int DD = kind == 0 ? 2 : D;
        LD s = 0;
        for (int k = 0; k < DD; k++) s += d[k] * d[k];
        return std::sqrt(s);
    }
    LD gl(LD b0, LD b1, int segid) const {
        const GL20& g = gl20();
        LD k = (b1 - b0) / 2, m = (b1 + b0) / 2, acc = 0;
        for (int j = 0; j < 20; j++) acc += g.w[j] * speed(k * g.x[j] + m, segid);
        return acc * k;
    }
};

static TruthCurve load_truth(const orc_batch& B, int c, bool normalize) {
    TruthCurve t;
    t.kind = B.kind;
    t.scale = normalize ? (LD)curve_scale(B, c) : 1.0L;
    if (B.kind == 0) {
        t.a = (LD)B.a[c];
        t.b = (LD)B.b[c];
        t.n_int = 1;
        return t;
    }
    t.p = B.degree[c];
    t.D = B.dim[c];
    int64_t nk = B.knot_off[c + 1] - B.knot_off[c];
    int64_t np = B.ctrl_off[c + 1] - B.ctrl_off[c];
    t.T.resize(nk);
    for (int64_t i = 0; i < nk; i++) t.T[i] = B.knots[B.knot_off[c] + i];
    t.P.resize(np * t.D);
    const double* P = B.ctrl + B.ctrl_off[c] * B.ctrl_stride;
    for (int64_t j = 0; j < np; j++)
        for (int d = 0; d < t.D; d++) t.P[j * t.D + d] = P[j * B.ctrl_stride + d];
    t.n_int = (int)nk - 2 * t.p - 1;
    return t;
}

struct TruthLeaf {
    LD u0, u1, cum_before;
};

struct Truth {
    std::vector<TruthCurve> curves;
    std::vector<int64_t> ioff;
    std::vector<LD> int_cum;                   // per interval, cumulative within curve
    std::vector<std::vector<TruthLeaf>> leaves;  // per interval
    std::vector<LD> length;

    void adapt(const TruthCurve& cv, int segid, LD b0, LD b1, LD whole, int depth, LD tol,
               std::vector<TruthLeaf>& out, LD& acc) {
        LD m = (b0 + b1) / 2;
        LD l = cv.gl(b0, m, segid), r = cv.gl(m, b1, segid);
        if (depth < 48 && std::fabs(l + r - whole) > tol && b1 - b0 > 1e-15L * (std::fabs(b1) + 1e-300L)) {
            adapt(cv, segid, b0, m, l, depth + 1, tol / 2, out, acc);
            adapt(cv, segid, m, b1, r, depth + 1, tol / 2, out, acc);
            return;
        }
        out.push_back({b0, b1, acc});
        acc += l + r;
    }

    void build(const orc_batch& B) {
        curves.clear();
        ioff.assign(B.n_curve + 1, 0);
        for (int c = 0; c < B.n_curve; c++) {
            curves.push_back(load_truth(B, c, false));
            ioff[c + 1] = ioff[c] + curves[c].n_int;
        }
        int_cum.assign(ioff.back(), 0);
        leaves.assign(ioff.back(), {});
        length.assign(B.n_curve, 0);
        for (int c = 0; c < B.n_curve; c++) {
            const TruthCurve& cv = curves[c];
            LD cum = 0;
            for (int s = 0; s < cv.n_int; s++) {
                LD u0, u1;
                cv.interval(s, u0, u1);
                LD acc = 0;
                if (u1 > u0) {
                    // start from 8 pieces so that a spike cannot hide between nodes
                    for (int k = 0; k < 8; k++) {
                        LD a0 = u0 + (u1 - u0) * k / 8, a1 = u0 + (u1 - u0) * (k + 1) / 8;
                        LD whole = cv.gl(a0, a1, s);
                        adapt(cv, s, a0, a1, whole, 0, std::max(std::fabs(whole) * 1e-17L, 1e-30L),
                              leaves[ioff[c] + s], acc);
                    }
                }
                cum += acc;
                int_cum[ioff[c] + s] = cum;
            }
            length[c] = cum;
        }
    }

    // Inverse map: interval by lower_bound over cumulative lengths, leaf by
    // binary search, then safeguarded Newton within the leaf.
    LD invert(int c, LD target, int32_t* seg_out) const {
        const TruthCurve& cv = curves[c];
        int64_t ib = ioff[c], ie = ioff[c + 1];
        int64_t g = std::lower_bound(int_cum.begin() + ib, int_cum.begin() + ie, target) - int_cum.begin();
        if (g >= ie) g = ie - 1;
        int segid = (int)(g - ib);
        if (seg_out) *seg_out = segid;
        LD local = target - (g > ib ? int_cum[g - 1] : 0);
        const std::vector<TruthLeaf>& lv = leaves[g];
        LD u0, u1;
        cv.interval(segid, u0, u1);
        if (lv.empty()) return u0;
        size_t lo = 0, hi = lv.size();
        while (hi - lo > 1) {
            size_t m = (lo + hi) / 2;
            if (lv[m].cum_before < local) lo = m; else hi = m;
        }
        const TruthLeaf& L = lv[lo];
        LD tau = local - L.cum_before;
        LD a = L.u0, b = L.u1;
        LD x = a + (b - a) * 0.5L;
        for (int it = 0; it < 100; it++) {
            LD f = cv.gl(L.u0, x, segid) - tau;
            if (f > 0) b = x; else a = x;
            LD sp = cv.speed(x, segid);
            LD xn = sp > 0 ? x - f / sp : (a + b) / 2;
            if (!(xn > a && xn < b)) xn = (a + b) / 2;
            if (std::fabs(xn - x) <= 1e-19L * (std::fabs(x) + 1e-30L) || b - a < 1e-19L) {
                x = xn;
                break;
            }
            x = xn;
        }
        return x;
    }
};

}  // namespace orc

using namespace orc;

// ===========================================================================
// C ABI (ctypes) — test infrastructure
// ===========================================================================
extern "C" {

void orc_gl_table(int order, double* x, double* w) {
    for (int j = 0; j < order; j++) {
        x[j] = GLX[order - 1][j];
        w[j] = GLW[order - 1][j];
    }
}

double orc_eps_device(int precision, double tol, int n_sub, int level) {
    return precision ? eps_level<double>(tol, n_sub, level) : (double)eps_level<float>(tol, n_sub, level);
}

int64_t orc_n_intervals(const orc_batch* B) {
    int64_t n = 0;
    for (int c = 0; c < B->n_curve; c++) n += n_intervals(*B, c);
    return n;
}

double orc_curve_scale(const orc_batch* B, int c) { return curve_scale(*B, c); }

// |C'(u)| of curve c on local interval segid, in precision `precision`,
// reference formulation (for derivative pins).
double orc_speed(const orc_batch* B, int c, int segid, double u, int precision) {
    if (precision) return load_curve<double>(*B, c, false).speed(u, segid);
    return (double)load_curve<float>(*B, c, false).speed((float)u, segid);
}

void orc_deriv(const orc_batch* B, int c, int segid, double u, double* out) {
    Curve<double> cv = load_curve<double>(*B, c, false);
    if (B->kind == 0) {
        out[0] = -cv.a * std::sin(u);
        out[1] = cv.b * std::cos(u);
        return;
    }
    cv.deriv(u, segid + cv.p, out);
}

void orc_truth_deriv(const orc_batch* B, int c, int segid, double u, double* out) {
    TruthCurve t = load_truth(*B, c, false);
    LD d[3];
    t.deriv((LD)u, segid, d);
    int D = B->kind == 0 ? 2 : B->dim[c];
    for (int k = 0; k < D; k++) out[k] = (double)d[k];
}

// ---- RefCpu ---------------------------------------------------------------
struct orc_cpu {
    int prec;
    void* eng;
};

orc_cpu* orc_cpu_create(const orc_batch* B, const orc_opts* o) {
    orc_cpu* h = new orc_cpu;
    h->prec = o->precision;
    if (o->precision) h->eng = new RefCpu<double>(*B, *o);
    else h->eng = new RefCpu<float>(*B, *o);
    return h;
}
void orc_cpu_destroy(orc_cpu* h) {
    if (!h) return;
    if (h->prec) delete (RefCpu<double>*)h->eng; else delete (RefCpu<float>*)h->eng;
    delete h;
}
// Returns the wall time in seconds (steady_clock) of the build.
double orc_cpu_build(orc_cpu* h, int n_threads) {
    auto t0 = std::chrono::steady_clock::now();
    if (h->prec) ((RefCpu<double>*)h->eng)->build(n_threads);
    else ((RefCpu<float>*)h->eng)->build(n_threads);
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}
// Lengths in curve units (evaluated lengths times the normalisation scale).
void orc_cpu_lengths(orc_cpu* h, int n_curve, double* out) {
    for (int c = 0; c < n_curve; c++)
        out[c] = h->prec ? ((RefCpu<double>*)h->eng)->length[c] * ((RefCpu<double>*)h->eng)->curves[c].scale
                         : ((RefCpu<float>*)h->eng)->length[c] * ((RefCpu<float>*)h->eng)->curves[c].scale;
}
int64_t orc_cpu_table_size(orc_cpu* h, int c) {
    return h->prec ? (int64_t)((RefCpu<double>*)h->eng)->s_tab[c].size()
                   : (int64_t)((RefCpu<float>*)h->eng)->s_tab[c].size();
}
void orc_cpu_table(orc_cpu* h, int c, double* u, double* s, int32_t* seg) {
    if (h->prec) {
        RefCpu<double>* e = (RefCpu<double>*)h->eng;
        for (size_t i = 0; i < e->s_tab[c].size(); i++) {
            u[i] = e->u_tab[c][i];
            s[i] = e->s_tab[c][i];
            seg[i] = e->seg_tab[c][i];
        }
    } else {
        RefCpu<float>* e = (RefCpu<float>*)h->eng;
        for (size_t i = 0; i < e->s_tab[c].size(); i++) {
            u[i] = e->u_tab[c][i];
            s[i] = e->s_tab[c][i];
            seg[i] = e->seg_tab[c][i];
        }
    }
}
// off: n_curve+1 sample offsets; pos: positions. Returns wall seconds.
double orc_cpu_reverse(orc_cpu* h, int n_curve, const int64_t* off, const double* pos, int is_relative,
                       double* out, int n_threads) {
    auto t0 = std::chrono::steady_clock::now();
    auto run = [&](int c0, int c1) {
        if (h->prec) ((RefCpu<double>*)h->eng)->reverse_range(off, pos, is_relative, out, c0, c1);
        else ((RefCpu<float>*)h->eng)->reverse_range(off, pos, is_relative, out, c0, c1);
    };
    if (n_threads <= 1) {
        run(0, n_curve);
    } else {
        // Each curve goes to the thread whose sample range holds its first sample.
        std::vector<std::thread> th;
        int64_t total = off[n_curve];
        for (int t = 0; t < n_threads; t++) {
            th.emplace_back([&, t]() {
                int64_t lo = total * t / n_threads, hi = total * (t + 1) / n_threads;
                int c0 = (int)(std::lower_bound(off, off + n_curve, lo) - off);
                int c1 = (int)(std::lower_bound(off, off + n_curve, hi) - off);
                run(c0, c1);
            });
        }
        for (auto& x : th) x.join();
    }
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// ---- RefGpuPlain ------------------------------------------------------------
struct orc_gpu {
    int prec;
    void* eng;
};
orc_gpu* orc_gpu_create(const orc_batch* B, const orc_opts* o) {
    orc_gpu* h = new orc_gpu;
    h->prec = o->precision;
    if (o->precision) h->eng = new RefGpuPlain<double>(*B, *o);
    else h->eng = new RefGpuPlain<float>(*B, *o);
    return h;
}
void orc_gpu_destroy(orc_gpu* h) {
    if (!h) return;
    if (h->prec) delete (RefGpuPlain<double>*)h->eng; else delete (RefGpuPlain<float>*)h->eng;
    delete h;
}
double orc_gpu_build(orc_gpu* h) {
    auto t0 = std::chrono::steady_clock::now();
    if (h->prec) ((RefGpuPlain<double>*)h->eng)->build();
    else ((RefGpuPlain<float>*)h->eng)->build();
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}
int64_t orc_gpu_n_tasks(orc_gpu* h) {
    return h->prec ? ((RefGpuPlain<double>*)h->eng)->n_tasks : ((RefGpuPlain<float>*)h->eng)->n_tasks;
}
// lengths[n_curve] and seg_cum[n_int] in curve units; sub_prefix[n_int*n_sub]
// in evaluated (normalised) units; root_depth[n_int*n_sub].
void orc_gpu_tables(orc_gpu* h, double* lengths, double* seg_cum, double* sub_prefix, int32_t* root_depth) {
    auto copy = [&](auto* e) {
        for (size_t i = 0; i < e->length.size(); i++) if (lengths) lengths[i] = e->length[i] * e->curves[i].scale;
        for (size_t c = 0; c < e->curves.size(); c++)
            for (int64_t i = e->ioff[c]; i < e->ioff[c + 1]; i++)
                if (seg_cum) seg_cum[i] = e->seg_cum[i] * e->curves[c].scale;
        for (size_t i = 0; i < e->sub_prefix.size(); i++) if (sub_prefix) sub_prefix[i] = e->sub_prefix[i];
        for (size_t i = 0; i < e->root_depth.size(); i++) if (root_depth) root_depth[i] = e->root_depth[i];
    };
    if (h->prec) copy((RefGpuPlain<double>*)h->eng); else copy((RefGpuPlain<float>*)h->eng);
}
void orc_gpu_reverse(orc_gpu* h, int n_curve, const int64_t* off, const double* pos, int is_relative,
                     double* out, int32_t* seg_out) {
    if (h->prec) ((RefGpuPlain<double>*)h->eng)->reverse_range(off, pos, is_relative, out, seg_out, 0, n_curve);
    else ((RefGpuPlain<float>*)h->eng)->reverse_range(off, pos, is_relative, out, seg_out, 0, n_curve);
}

// ---- Truth ------------------------------------------------------------------
orc::Truth* orc_truth_create(const orc_batch* B) {
    orc::Truth* t = new orc::Truth;
    t->build(*B);
    return t;
}
void orc_truth_destroy(orc::Truth* t) { delete t; }
void orc_truth_lengths(orc::Truth* t, double* lengths, double* int_cum) {
    for (size_t i = 0; i < t->length.size(); i++) if (lengths) lengths[i] = (double)t->length[i];
    for (size_t i = 0; i < t->int_cum.size(); i++) if (int_cum) int_cum[i] = (double)t->int_cum[i];
}
// Partial arc length from the curve start to parameter u on local interval segid.
double orc_truth_partial(orc::Truth* t, int c, int segid, double u) {
    const orc::TruthCurve& cv = t->curves[c];
    int64_t g = t->ioff[c] + segid;
    LD base = g > t->ioff[c] ? t->int_cum[g - 1] : 0;
    LD u0, u1;
    cv.interval(segid, u0, u1);
    const std::vector<TruthLeaf>& lv = t->leaves[g];
    if (lv.empty()) return (double)base;
    size_t k = 0;
    while (k + 1 < lv.size() && lv[k + 1].u0 <= (LD)u) k++;
    return (double)(base + lv[k].cum_before + cv.gl(lv[k].u0, (LD)u, segid));
}
// Sample parameters for targets given as positions (relative to the truth
// length if is_relative) — off: n_curve+1 sample offsets.
void orc_truth_reverse(orc::Truth* t, int n_curve, const int64_t* off, const double* pos, int is_relative,
                       double* out, int32_t* seg_out, int n_threads) {
    auto run = [&](int c0, int c1) {
        for (int c = c0; c < c1; c++)
            for (int64_t i = off[c]; i < off[c + 1]; i++) {
                LD target = is_relative ? (LD)pos[i] * t->length[c] : (LD)pos[i];
                out[i] = (double)t->invert(c, target, seg_out ? &seg_out[i] : nullptr);
            }
    };
    if (n_threads <= 1) { run(0, n_curve); return; }
    std::vector<std::thread> th;
    for (int k = 0; k < n_threads; k++)
        th.emplace_back([&, k]() { run((int)((int64_t)n_curve * k / n_threads), (int)((int64_t)n_curve * (k + 1) / n_threads)); });
    for (auto& x : th) x.join();
}

}  // extern "C"
