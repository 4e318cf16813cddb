// Header-only GPU engine for user curve plug-ins: the template
// gcurval::ArclengthCalculateGpu<Curve, CurveTraits, max_recursive_depth,
// n_subdivided> with the constructors and methods of the reference's
// (gcurval/arclength_calculate.h:10-18, gcurval/detail/cuda/arclength_calculate.h:141-237),
// compiled by hipcc in the user's translation unit.
//
// The plug-in's parser, derivative / point functors, length rule and
// per-thread scratch allocator are adapted into the gfx950 kernels the
// prebuilt curve families use (wave-local adaptive subdivision, FP64 scans,
// Hermite-start Newton sampler). Memory, copies and the stream come from the
// C ABI (libgcurval_amd.so): a process-wide context whose arena replaces RMM.
#pragma once

#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "base.h"
#include "../gcurval_c.h"
#include "detail/kernels.h"

namespace gcurval {

// ---------------------------------------------------------------------------
// Process-wide context, allocation and copies (gMalloc / gFree / gFree_s of
// gcurval/detail/cuda/util.h:41-71; copyToDevice / copyToHost replace the raw
// cudaMemcpy calls of the reference's plug-in subclasses).
// ---------------------------------------------------------------------------
inline void check(int rc, const char* what, gcv_ctx* c = nullptr) {
    if (rc != 0) {
        std::string msg = std::string("gcurval: ") + what + " failed (status " + std::to_string(rc) + ")";
        if (c) msg += std::string(": ") + gcv_last_error(c);
        throw std::runtime_error(msg);
    }
}

inline gcv_ctx* default_context() {
    static gcv_ctx* ctx = [] {
        gcv_ctx* c = nullptr;
        check(gcv_ctx_create(0, GCV_ARENA_POOL, &c), "gcv_ctx_create");
        return c;
    }();
    return ctx;
}

inline hipStream_t default_stream() {
    void* s = nullptr;
    check(gcv_ctx_stream(default_context(), &s), "gcv_ctx_stream", default_context());
    return (hipStream_t)s;
}

template <typename T>
inline void gMalloc(T** p, size_t bytes) {
    void* v = nullptr;
    check(gcv_alloc(default_context(), bytes ? bytes : 1, &v), "gcv_alloc", default_context());
    *p = (T*)v;
}
template <typename T>
inline void gFree(T* p) {
    if (p) gcv_free(default_context(), (void*)p);
}
template <typename T>
inline void gFree_s(T*& p) {
    if (p) {
        gcv_free(default_context(), (void*)p);
        p = nullptr;
    }
}
template <typename T>
inline void copyToDevice(T* d_dst, const T* h_src, size_t count) {
    check(gcv_copy_h2d(default_context(), d_dst, h_src, sizeof(T) * count), "gcv_copy_h2d", default_context());
}
template <typename T>
inline void copyToHost(T* h_dst, const T* d_src, size_t count) {
    check(gcv_copy_d2h(default_context(), h_dst, d_src, sizeof(T) * count), "gcv_copy_d2h", default_context());
}

namespace detail {

template <typename A>
struct HasScratch {
    static constexpr bool value = !std::is_same<A, NoSharedMemAllocator>::value;
};

// Plug-in scratch lives in dynamic LDS after the kernel's own use.
__device__ __forceinline__ uint8_t* dynamic_lds() {
    extern __shared__ __align__(16) uint8_t gcv_plugin_lds[];
    return gcv_plugin_lds;
}

// Whether a length rule integrates GetDerivative (so the sampler's Newton
// derivative is |dC/du| |du/dx|) or measures points (the derivative is then
// taken from the rule itself by a central difference).
template <class Rule>
struct IsDerivativeRule : std::true_type {};
template <template <typename> class R, typename Dt, typename Pack, typename G>
struct IsDerivativeRule<CalculateFunctionImpl<R, Dt, Pack, G>>
    : std::integral_constant<bool, std::is_same<R<Dt>, GaussLegendreQuadrature<Dt>>::value> {};

// Evaluator adapting a plug-in to the engine kernels. Kernel coordinates are
// interval-local x in [0, 1]; the plug-in works on its parameter u in the
// pack's [begin, end]. T is the arithmetic the kernel asks for; the plug-in
// itself computes in its DataType.
template <class Curve, class Traits>
struct PluginEval {
    using Data = typename Traits::DataType;
    using Pack = typename Traits::ParaPack;
    using Parser = typename Traits::ParaPackParser;
    using Alloc = typename Traits::SharedMemAllocator;
    using Rule = typename Traits::CalculateFunction;
    using Deriv = typename Traits::GetDerivative;
    static_assert(std::is_trivially_copyable<Pack>::value && sizeof(Pack) % 4 == 0,
                  "ParaPack must be trivially copyable with a size that is a multiple of 4 bytes");
    static constexpr bool kRule = true;

    struct Ctx {
        Pack pack;
    };

    Curve curve;
    const int32_t* int_curve;
    const int64_t* ioff;
    const double* u0;
    const double* h;
    unsigned scratch_off;  // byte offset of the plug-in scratch in dynamic LDS

    __device__ Ctx load(int i) const {
        const int c = int_curve[i];
        return Ctx{Parser{}(curve, c, (int)(i - ioff[c]), true)};
    }
    __device__ Ctx load(int i, double) const { return load(i); }
    __device__ double scale(int) const { return 1.0; }
    __device__ double to_u(int i, double x) const { return u0[i] + h[i] * x; }

    __device__ Data u_of(const Pack& p, Data x) const {
        if (x >= Data(1)) return p.end;
        return p.begin + (p.end - p.begin) * x;
    }
    template <typename T>
    __device__ T length(const Ctx& c, T x0, T x1) const {
        const Data ua = u_of(c.pack, Data(x0)), ub = u_of(c.pack, Data(x1));
        if constexpr (HasScratch<Alloc>::value) {
            auto view = Alloc{}(dynamic_lds() + scratch_off);
            return T(Rule{}(ua, ub, c.pack, view));
        } else {
            return T(Rule{}(ua, ub, c.pack));
        }
    }
    template <typename T>
    __device__ T speed(const Ctx& c, T x) const {
        if constexpr (!IsDerivativeRule<Rule>::value || std::is_same<Deriv, void>::value) {
            const T d = T(1.0 / 1024.0);
            const T a = x - d > T(0) ? x - d : T(0), b = x + d < T(1) ? x + d : T(1);
            return length(c, a, b) / (b - a);
        } else {
            const Data w = c.pack.end - c.pack.begin;
            const Data u = u_of(c.pack, Data(x));
            if constexpr (HasScratch<Alloc>::value) {
                auto view = Alloc{}(dynamic_lds() + scratch_off);
                return T(magnitude<Data>(Deriv{}(u, c.pack, view)) * abs(w));
            } else {
                return T(magnitude<Data>(Deriv{}(u, c.pack)) * abs(w));
            }
        }
    }
};

// Global knot-interval range of every curve from the plug-in's locator.
template <class Curve, class Locator>
__global__ void k_plugin_locate(Curve curve, int n_curve, int64_t* ioff) {
    const int c = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (c >= n_curve) return;
    const CurveIntervalPos r = Locator{}(curve, c);
    ioff[c] = r.begin;
    if (c == n_curve - 1) ioff[n_curve] = r.end;
}

// Interval -> curve map and each interval's parameter range (wave per curve).
template <class Curve, class Parser>
__global__ void k_plugin_intervals(Curve curve, int n_curve, const int64_t* ioff, int32_t* int_curve, double* u0,
                                   double* h) {
    const int lane = dev::lane_id();
    const int64_t w = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int64_t nw = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t c = w; c < n_curve; c += nw) {
        const int64_t i0 = ioff[c], i1 = ioff[c + 1];
        for (int64_t i = i0 + lane; i < i1; i += 64) {
            const auto p = Parser{}(curve, (int)c, (int)(i - i0), true);
            int_curve[i] = (int32_t)c;
            u0[i] = (double)p.begin;
            h[i] = (double)p.end - (double)p.begin;
        }
    }
}

// Curve points at sampled parameters (arclengthReverse returning points,
// kn_toPoints of gcurval/detail/cuda/kernel/reverse.inl.cuh:344-363).
template <class Curve, class Traits, class Pt>
__global__ void k_plugin_points(Curve curve, const int64_t* ioff, const int64_t* soff, int n_curve, int64_t n_total,
                                const int32_t* blk_curve, const double* u, const int32_t* seg, Pt* out) {
    using Parser = typename Traits::ParaPackParser;
    using Alloc = typename Traits::SharedMemAllocator;
    using GetPoint = typename Traits::GetPoint;
    using Data = typename Traits::DataType;
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n_total) return;
    dev::SampleArgs A{};
    A.n_curve = n_curve;
    A.n_total = n_total;
    A.soff = soff;
    const int c = dev::sample_curve(A, blk_curve, blockIdx.x, g);
    const auto p = Parser{}(curve, c, seg[g], true);
    if constexpr (HasScratch<Alloc>::value) {
        auto view = Alloc{}(dynamic_lds());
        out[g] = GetPoint{}(Data(u[g]), p, view);
    } else {
        out[g] = GetPoint{}(Data(u[g]), p);
    }
}

template <typename Dt>
__global__ void k_to_data(const double* in, int64_t n, Dt* out) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < n) out[g] = Dt(in[g]);
}

inline void launch_check(const char* what) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw std::runtime_error(std::string("gcurval: launch of ") + what + " failed: " +
                                                  hipGetErrorString(e));
}

}  // namespace detail

// ---------------------------------------------------------------------------
// The engine.
// ---------------------------------------------------------------------------
// BufferManager: accepted for the reference's five-parameter signature and
// ignored (BufferManagerBase in base.h).
template <typename Curve, typename CurveTraits,
          int max_recursive_depth = MaxRecursiveDepth<typename Curve::DataType>::value,
          int n_subdivided = N_SUBDIVIDED,
          typename BufferManager = BufferManagerBase<typename Curve::DataType, max_recursive_depth,
                                                     typename CurveTraits::ForwardSpace,
                                                     typename CurveTraits::BackwardSpace>>
class ArclengthCalculateGpu {
public:
    using Dt = typename CurveTraits::DataType;
    using PointType = typename CurveTraits::PointType;
    using Eval = detail::PluginEval<Curve, CurveTraits>;
    using Alloc = typename CurveTraits::SharedMemAllocator;
    static constexpr int MAX_RECURSIVE_DEPTH = max_recursive_depth;

    static_assert(n_subdivided >= 1 && n_subdivided <= 64 && (n_subdivided & (n_subdivided - 1)) == 0,
                  "n_subdivided must be a power of two in [1, 64]");
    static_assert(max_recursive_depth >= 1 && max_recursive_depth <= (sizeof(Dt) <= 4 ? 20 : 31),
                  "max_recursive_depth exceeds the task-key range (20 in FP32, 31 in FP64)");
    // FP32 task bounds are exact dyadic floats only while depth + log2(n_subdivided) <= 24
    static_assert(sizeof(Dt) > 4 || max_recursive_depth + (n_subdivided >= 64   ? 6
                                                            : n_subdivided >= 32 ? 5
                                                            : n_subdivided >= 16 ? 4
                                                            : n_subdivided >= 8  ? 3
                                                            : n_subdivided >= 4  ? 2
                                                            : n_subdivided >= 2  ? 1
                                                                                 : 0) <= 24,
                  "FP32: max_recursive_depth + log2(n_subdivided) must not exceed 24");

    // n_seg: total knot intervals over all curves (the plug-in's locator
    // ranges must tile [0, n_seg) in curve order).
    explicit ArclengthCalculateGpu(int n_curve, int n_seg) : n_seg_(n_seg) {
        if (n_curve < 0 || n_seg < 0) throw std::invalid_argument("gcurval: negative curve or interval count");
        d_curve.n_curve = n_curve;
        allocate();
    }
    explicit ArclengthCalculateGpu(int n_curve) : ArclengthCalculateGpu(n_curve, n_curve) {}

    virtual ~ArclengthCalculateGpu() {
        for (void* p : owned_) gcv_free(default_context(), p);
        freeArray_s(h_arclength_);
    }
    ArclengthCalculateGpu(const ArclengthCalculateGpu&) = delete;
    ArclengthCalculateGpu& operator=(const ArclengthCalculateGpu&) = delete;

    // Absolute error tolerance per knot interval (level 0 tests each of the
    // n_subdivided roots with tolerance / n_subdivided). Default 1e-6 (FP32),
    // 1e-14 (FP64).
    void setTolerance(double tol) { tol_ = tol; }

    // createArclengthTable (gcurval/detail/cuda/arclength_calculate.h:185-210):
    // enqueued on the context stream, no host synchronisation after the first call.
    void createArclengthTable() {
        hipStream_t s = default_stream();
        const int n_curve = d_curve.n_curve;
        if (!located_) locate(s);
        hipLaunchKernelGGL(dev::k_fill_u64, dim3((dev::kStatWords + 255) / 256), dim3(256), 0, s, stats_,
                           dev::kStatWords, 0ull);
        detail::launch_check("k_fill_u64");
        if (n_seg_ > 0) subdivide(s);
        if (n_curve > 0) {
            const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>((n_curve + 3) / 4, (int64_t)n_cu_ * 16));
            hipLaunchKernelGGL(dev::k_curve_scan, dim3((unsigned)blocks), dim3(256), 0, s, n_curve, ioff_, int_cum_,
                               curve_len_);
            detail::launch_check("k_curve_scan");
        }
        built_ = true;
    }

    // Per-curve lengths, host array the caller delete[]s (:212-222, 372-406).
    Dt* getAllCurveArclengths() {
        require_built();
        std::vector<double> L(d_curve.n_curve);
        if (!L.empty()) copyToHost(L.data(), curve_len_, L.size());
        Dt* out = new Dt[L.size()];
        for (size_t i = 0; i < L.size(); i++) out[i] = Dt(L[i]);
        return out;
    }

    Dt getArclength(int cid) {
        require_built();
        if (cid < 0 || cid >= d_curve.n_curve) throw std::out_of_range("gcurval: curve id out of range");
        if (!h_arclength_) h_arclength_ = getAllCurveArclengths();
        return h_arclength_[cid];
    }

    // arclengthReverse (:224-231, 408-505). n_ptrs_for_curves: host, samples
    // per curve; sample_position: host positions concatenated over curves
    // (fractions of the curve length when is_relative, else arc lengths), or
    // nullptr for the uniform positions j / (n - 1). The result is a device
    // array (free with gFree) or a host array (delete[]) of parameters u or of
    // curve points.
    ArclengthReverseResult<Dt, PointType> arclengthReverse(int* n_ptrs_for_curves, Dt* sample_position,
                                                           bool is_relative,
                                                           GcurvalEnum where = gcurvalDevice,
                                                           GcurvalEnum what = gcurvalReverseReturnUs) {
        require_built();
        const int n_curve = d_curve.n_curve;
        std::vector<int64_t> off(n_curve + 1, 0);
        for (int c = 0; c < n_curve; c++) {
            if (n_ptrs_for_curves[c] < 0) throw std::invalid_argument("gcurval: negative sample count");
            off[c + 1] = off[c] + n_ptrs_for_curves[c];
        }
        const int64_t n_total = off[n_curve];
        ArclengthReverseResult<Dt, PointType> res;
        res.n_points = (int)n_total;
        res.store_where = where;
        res.store_what = what;
        if (n_total == 0) return res;
        hipStream_t s = default_stream();

        int64_t* d_off = nullptr;
        double* d_pos = nullptr;
        double* d_u = nullptr;
        int32_t* d_seg = nullptr;
        int32_t* d_blk = nullptr;
        const int64_t n_blk = (n_total + dev::kSampleBlock - 1) / dev::kSampleBlock;
        gMalloc(&d_off, sizeof(int64_t) * (n_curve + 1));
        gMalloc(&d_u, sizeof(double) * n_total);
        gMalloc(&d_seg, sizeof(int32_t) * n_total);
        gMalloc(&d_blk, sizeof(int32_t) * n_blk);
        copyToDevice(d_off, off.data(), off.size());
        if (sample_position) {
            std::vector<double> pos(sample_position, sample_position + n_total);
            gMalloc(&d_pos, sizeof(double) * n_total);
            copyToDevice(d_pos, pos.data(), pos.size());
        }
        hipLaunchKernelGGL(dev::k_block_curve, dim3((unsigned)((n_curve + 255) / 256)), dim3(256), 0, s, n_curve,
                           (const int64_t*)d_off, n_blk, d_blk);
        detail::launch_check("k_block_curve");

        dev::SampleArgs A{};
        A.n_curve = n_curve;
        A.n_total = n_total;
        A.soff = d_off;
        A.pos = d_pos;
        A.uniform_closed = 1;
        A.is_relative = is_relative ? 1 : 0;
        A.ioff = ioff_;
        A.int_cum = int_cum_;
        A.curve_len = curve_len_;
        A.root_prefix = root_prefix_;
        A.root_k = nullptr;
        A.deep = deep_;
        A.log2_nsub = kLog2Sub;
        A.eps0 = eps0();
        A.ls_plain = kLog2Sub;
        A.max_depth = max_recursive_depth;
        A.newton_iters = sizeof(Dt) <= 4 ? 2 : 3;
        A.fp64_step = 0;
        A.out_u = d_u;
        A.out_seg = d_seg;
        A.class_of_curve = nullptr;
        Eval ev = eval(0);
        const size_t lds = (size_t)Alloc::size(dev::kSampleBlock);
        hipLaunchKernelGGL((dev::k_sample<Dt, 3, Eval, Eval>), dim3((unsigned)n_blk), dim3(dev::kSampleBlock), lds, s,
                           ev, ev, A, (const int32_t*)d_blk);
        detail::launch_check("k_sample");

        if (what == gcurvalReverseReturnPoints) {
            if constexpr (std::is_same<typename CurveTraits::GetPoint, void>::value) {
                throw std::invalid_argument("gcurval: points requested from a plug-in without GetPoint");
            } else {
            PointType* d_pts = nullptr;
            gMalloc(&d_pts, sizeof(PointType) * n_total);
            hipLaunchKernelGGL((detail::k_plugin_points<Curve, CurveTraits, PointType>), dim3((unsigned)n_blk),
                               dim3(dev::kSampleBlock), lds, s, d_curve, (const int64_t*)ioff_,
                               (const int64_t*)d_off, n_curve, n_total, (const int32_t*)d_blk, (const double*)d_u,
                               (const int32_t*)d_seg, d_pts);
            detail::launch_check("k_plugin_points");
            if (where == gcurvalHost) {
                res.points = new PointType[n_total];
                copyToHost(res.points, d_pts, n_total);
                gFree(d_pts);
            } else {
                res.points = d_pts;
            }
            }
        } else {
            Dt* d_us = nullptr;
            gMalloc(&d_us, sizeof(Dt) * n_total);
            hipLaunchKernelGGL(detail::k_to_data<Dt>, dim3((unsigned)((n_total + 255) / 256)), dim3(256), 0, s,
                               (const double*)d_u, n_total, d_us);
            detail::launch_check("k_to_data");
            if (where == gcurvalHost) {
                res.us = new Dt[n_total];
                copyToHost(res.us, d_us, n_total);
                gFree(d_us);
            } else {
                res.us = d_us;
            }
        }
        check(gcv_sync(default_context()), "gcv_sync", default_context());
        gFree(d_off);
        gFree(d_pos);
        gFree(d_u);
        gFree(d_seg);
        gFree(d_blk);
        return res;
    }

    // clearArclength: drops the host copy of the lengths (the device table is
    // kept until the engine is destroyed or rebuilt).
    void clearArclength() { freeArray_s(h_arclength_); }

    // Device views for callers that keep results on the GPU: FP64 cumulative
    // interval lengths within each curve, per-curve lengths, interval offsets.
    const double* deviceIntervalCumulative() const { return int_cum_; }
    const double* deviceCurveLengths() const { return curve_len_; }
    const int64_t* deviceIntervalOffsets() const { return ioff_; }

protected:
    Curve d_curve;

private:
    static constexpr int kLog2Sub = n_subdivided == 1 ? 0 : 31 - __builtin_clz((unsigned)n_subdivided);

    int n_seg_;
    int n_cu_ = 256;
    double tol_ = default_max_error<Dt>();
    bool located_ = false, built_ = false;
    std::vector<void*> owned_;
    int64_t* ioff_ = nullptr;
    int32_t* int_curve_ = nullptr;
    double* u0_ = nullptr;
    double* h_ = nullptr;
    double* int_cum_ = nullptr;
    double* curve_len_ = nullptr;
    double* root_prefix_ = nullptr;
    uint64_t* deep_ = nullptr;
    unsigned long long* stats_ = nullptr;
    Dt* h_arclength_ = nullptr;

    template <typename T>
    void alloc(T** p, size_t count) {
        gMalloc(p, sizeof(T) * (count ? count : 1));
        owned_.push_back((void*)*p);
    }
    void allocate() {
        const int n_curve = d_curve.n_curve;
        alloc(&ioff_, (size_t)n_curve + 1);
        alloc(&int_curve_, n_seg_);
        alloc(&u0_, n_seg_);
        alloc(&h_, n_seg_);
        alloc(&int_cum_, n_seg_);
        alloc(&curve_len_, n_curve);
        alloc(&root_prefix_, (size_t)n_seg_ << kLog2Sub);
        alloc(&deep_, n_seg_);
        alloc(&stats_, dev::kStatWords);
        int n_cu = 0;
        char name[64];
        if (gcv_device_info(default_context(), &n_cu, name, sizeof name) == 0 && n_cu > 0) n_cu_ = n_cu;
    }
    void require_built() const {
        if (!built_) throw std::logic_error("gcurval: createArclengthTable must run first");
    }
    Dt eps0() const { return Dt(tol_ / (double)n_subdivided); }
    Eval eval(unsigned scratch_off) const {
        return Eval{d_curve, int_curve_, ioff_, u0_, h_, scratch_off};
    }

    // Interval ranges from the locator (checked once on the host: they must
    // tile [0, n_seg) in curve order), then the interval -> curve map.
    void locate(hipStream_t s) {
        const int n_curve = d_curve.n_curve;
        if (n_curve > 0) {
            hipLaunchKernelGGL((detail::k_plugin_locate<Curve, typename CurveTraits::CurveLocator>),
                               dim3((unsigned)((n_curve + 255) / 256)), dim3(256), 0, s, d_curve, n_curve, ioff_);
            detail::launch_check("k_plugin_locate");
            std::vector<int64_t> off(n_curve + 1);
            copyToHost(off.data(), ioff_, off.size());
            if (off[0] != 0 || off[n_curve] != n_seg_)
                throw std::invalid_argument("gcurval: curve locator ranges do not cover [0, n_seg)");
            for (int c = 0; c < n_curve; c++)
                if (off[c + 1] < off[c]) throw std::invalid_argument("gcurval: curve locator ranges are not ordered");
            const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>((n_curve + 3) / 4, (int64_t)n_cu_ * 8));
            hipLaunchKernelGGL((detail::k_plugin_intervals<Curve, typename CurveTraits::ParaPackParser>),
                               dim3((unsigned)blocks), dim3(256), 0, s, d_curve, n_curve, (const int64_t*)ioff_,
                               int_curve_, u0_, h_);
            detail::launch_check("k_plugin_intervals");
        } else {
            const int64_t zero = 0;
            copyToDevice(ioff_, &zero, 1);
        }
        located_ = true;
    }

    void subdivide(hipStream_t s) {
        dev::SubdivArgs<Dt> A{};
        A.n_slots = n_seg_;
        A.int_list = nullptr;
        A.slot_range = nullptr;
        A.log2_nsub = kLog2Sub;
        A.eps0 = eps0();
        A.max_depth = max_recursive_depth;
        A.root_prefix = root_prefix_;
        A.int_len = int_cum_;
        A.curve_len = nullptr;
        A.deep = deep_;
        A.int_cost = nullptr;
        A.stack_cap = 64 * (max_recursive_depth + 2);
        A.stats = stats_;
        // 4 waves per block unless the stacks plus the plug-in scratch exceed 64 KiB
        const int per_wave = dev::subdiv_lds_per_wave<Dt>(A.stack_cap);
        int wpb = 4;
        while (wpb > 1 && (size_t)wpb * per_wave + (size_t)Alloc::size(64 * wpb) > 64 * 1024) wpb >>= 1;
        const size_t stacks = (size_t)wpb * per_wave;
        const size_t lds = stacks + (size_t)Alloc::size(64 * wpb);
        if (lds > 64 * 1024) throw std::invalid_argument("gcurval: plug-in scratch does not fit in LDS");
        const int ipw = 64 >> kLog2Sub;
        const int64_t n_tiles = ((int64_t)n_seg_ + ipw - 1) / ipw;
        int per_cu = 0;
        check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, dev::k_subdivide<Dt, 3, Eval, kLog2Sub == 6>,
                                                           64 * wpb, lds) ==
                      hipSuccess
                  ? 0
                  : 3,
              "occupancy query");
        if (per_cu < 1) per_cu = 1;
        const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>((int64_t)n_cu_ * per_cu, (n_tiles + wpb - 1) / wpb));
        hipLaunchKernelGGL((dev::k_subdivide<Dt, 3, Eval, kLog2Sub == 6>), dim3((unsigned)blocks), dim3(64 * wpb), lds, s,
                           eval((unsigned)stacks), A);
        detail::launch_check("k_subdivide");
    }
};

}  // namespace gcurval
