// Internal launcher interface between the C ABI (capi.hip) and the kernel
// instantiations (inst_<precision>_gl<order>.hip). Each instantiation unit
// compiles the kernels of one (arithmetic type, Gauss-Legendre order) pair.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/gcurval/detail/kernels.h"
#include "../../include/gcurval/detail/seeded.h"
#include "../../include/gcurval/detail/mlp.h"
#include "../../include/gcurval/detail/points.h"

namespace gcurval {
namespace rt {

// Launch context: stream, device size, and profiling hooks of the owner.
struct LaunchCfg {
    hipStream_t stream;
    int n_cu;
    void* owner;           // gcv_ctx*
    int sub_wpb = 0;       // subdivision waves per block (0: as many as LDS allows, up to 4)
    int sub_grid_pct = 100; // persistent subdivision grid, percent of the resident-block limit
};

// Profiling hooks implemented by the context (events around a launch on
// `stream`, the context's own stream when null).
void prof_begin(void* owner, const char* name, hipStream_t stream = nullptr);
void prof_end(void* owner, hipStream_t stream = nullptr);

// Device arrays of a B-spline table needed by the evaluators.
struct BSplineTableDev {
    const double* coef64;  // FP64 builds
    const float* coef32;   // FP32 builds
    int cstride;           // values per interval block
    const double* int_scale;
    const double* u0;
    const double* h;
};

// Class of a B-spline interval: 10 * degree + dim.
inline int bspline_class(int p, int D) { return 10 * p + D; }

template <typename Dt, int NGL>
struct Launch {
    // Subdivision of ellipses (one interval per curve).
    static hipError_t subdivide_ellipse(const LaunchCfg& c, const double* a, const double* b, int normalize,
                                        dev::SubdivArgs<Dt> A);
    // Subdivision of one B-spline class (degree p, dimension D).
    static hipError_t subdivide_bspline(const LaunchCfg& c, int p, int D, const BSplineTableDev& T,
                                        dev::SubdivArgs<Dt> A);
    // Seeded subdivision of the bucketed intervals of one class: the bins
    // k = 0..5, bin 6, then the re-run of tiles that outgrew the small stack.
    static hipError_t seeded_ellipse(const LaunchCfg& c, const double* a, const double* b, int normalize,
                                     const dev::SeededLaunch<Dt>& S);
    static hipError_t seeded_bspline(const LaunchCfg& c, int p, int D, const BSplineTableDev& T,
                                     const dev::SeededLaunch<Dt>& S);
    // blk_curve: curve of the first sample of every kSampleBlock samples (k_block_curve).
    static hipError_t sample_ellipse(const LaunchCfg& c, const double* a, const double* b, int normalize,
                                     const dev::SampleArgs& A, const int32_t* blk_curve);
    // Sampling of B-splines; class per curve from A.class_of_curve, or
    // uniform_class when that is null.
    static hipError_t sample_bspline(const LaunchCfg& c, int uniform_class, const BSplineTableDev& T,
                                     const dev::SampleArgs& A, const int32_t* blk_curve);
    // Training labels of the learned Newton start (FP32 GL-3 builds only):
    // per interval, log2 of the worst quintic-start error over its roots.
    static hipError_t start_labels(const LaunchCfg& c, int uniform_class, const BSplineTableDev& T,
                                   const int32_t* int_curve, const dev::SampleArgs& A, float* out);
    // LDS per block and waves per block of the subdivision kernel.
    static void subdivide_shape(int stack_cap, int want_wpb, int* waves_per_block, size_t* lds_bytes);

};

}  // namespace rt
}  // namespace gcurval
