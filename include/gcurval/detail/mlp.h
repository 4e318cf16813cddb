// Learned seed of the table build and of the sampler's Newton start: a
// 16 -> 64 -> 64 -> 4 multilayer perceptron (ReLU) that predicts, per knot
// interval, log2 of the initial root count (output 0, regression), the
// probability that every root of that seed passes the level-0 error test
// (output 1, logit), the scaled log2 of the worst error of the quintic
// Hermite start of the inverse map over the interval's roots (output 2,
// regression) and the probability that this start is close enough for the
// final FP64 Newton step to follow it directly (output 3, logit).
//
// Every GEMM runs on BF16 matrix cores with FP32 accumulation
// (v_mfma_f32_16x16x16bf16_1k, wave64): a wave owns a tile of 64 rows
// (4 M-tiles of 16), activations move between layers through LDS, and the
// weight gradients accumulate over the wave's rows in MFMA accumulators
// before one atomic flush per wave.
//
// MFMA 16x16x16 operand layout (lane l, j = 0..3):
//   A[m][k]: m = l % 16, k = 4 (l / 16) + j
//   B[k][n]: n = l % 16, k = 4 (l / 16) + j
//   D[m][n]: n = l % 16, m = 4 (l / 16) + j
#pragma once

#include "device.h"
#include "curves.h"
#include "seeded.h"
#include "features.h"
#include "kernels.h"

namespace gcurval {
namespace dev {

constexpr int kHid = 64;
constexpr int kOut = 4;
constexpr int kParams = kFeat * kHid + kHid + kHid * kHid + kHid + kHid * kOut + kOut;  // 5508

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Parameter block (FP32 master copy + BF16 operands), offsets into one array.
struct MlpParams {
    float* p;          // [kParams] W1[16][64], b1[64], W2[64][64], b2[64], W3[64][4], b3[4]
    float* feat_mean;  // [16]
    float* feat_istd;  // [16]
    uint16_t* w1b;     // [16][64]
    uint16_t* w2b;     // [64][64]
    uint16_t* w3b;     // [64][16]  (columns >= kOut are zero)
};
__host__ __device__ constexpr int off_W1() { return 0; }
__host__ __device__ constexpr int off_b1() { return kFeat * kHid; }
__host__ __device__ constexpr int off_W2() { return off_b1() + kHid; }
__host__ __device__ constexpr int off_b2() { return off_W2() + kHid * kHid; }
__host__ __device__ constexpr int off_W3() { return off_b2() + kHid; }
__host__ __device__ constexpr int off_b3() { return off_W3() + kHid * kOut; }

__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// K = 32 products (v_mfma_f32_16x16x32_bf16): each lane supplies 8 values of
// A's row / B's column. Operands built as the concatenation of two K = 16
// fragments (cat8) contract over the same 32 k indices in A and B, so the
// product equals the two K = 16 products summed.
typedef short bf16x8 __attribute__((ext_vector_type(8)));
__device__ __forceinline__ bf16x8 cat8(bf16x4 a, bf16x4 b) { return __builtin_shufflevector(a, b, 0, 1, 2, 3, 4, 5, 6, 7); }
__device__ __forceinline__ f32x4 mfma32(bf16x8 a, bf16x8 b, f32x4 c) {
    typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, a), __builtin_bit_cast(bf16x8_t, b), c,
                                                  0, 0, 0);
}

// A operand from an LDS tile T[16][ld] (bf16), columns k0..k0+15.
__device__ __forceinline__ bf16x4 load_a(const uint16_t* T, int ld, int k0) {
    const int l = lane_id();
    const uint16_t* r = T + (l & 15) * ld + k0 + 4 * (l >> 4);
    bf16x4 v;
    v[0] = (short)r[0];
    v[1] = (short)r[1];
    v[2] = (short)r[2];
    v[3] = (short)r[3];
    return v;
}
// A operand of the transpose of an LDS tile T[16 rows][ld]: A[m][k] = T[k][m0 + m].
__device__ __forceinline__ bf16x4 load_at(const uint16_t* T, int ld, int m0) {
    const int l = lane_id();
    const int m = m0 + (l & 15), k = 4 * (l >> 4);
    bf16x4 v;
    v[0] = (short)T[(k + 0) * ld + m];
    v[1] = (short)T[(k + 1) * ld + m];
    v[2] = (short)T[(k + 2) * ld + m];
    v[3] = (short)T[(k + 3) * ld + m];
    return v;
}
// B operand from a row-major global matrix W[K][ldw] (bf16): rows k0.., cols n0..
__device__ __forceinline__ bf16x4 load_b(const uint16_t* W, int ldw, int k0, int n0) {
    const int l = lane_id();
    const int n = n0 + (l & 15), k = k0 + 4 * (l >> 4);
    bf16x4 v;
    v[0] = (short)W[(k + 0) * ldw + n];
    v[1] = (short)W[(k + 1) * ldw + n];
    v[2] = (short)W[(k + 2) * ldw + n];
    v[3] = (short)W[(k + 3) * ldw + n];
    return v;
}
// B operand of the transpose of W[N][ldw]: B[k][n] = W[n0 + n][k0 + k].
__device__ __forceinline__ bf16x4 load_bt(const uint16_t* W, int ldw, int k0, int n0) {
    const int l = lane_id();
    const int n = n0 + (l & 15), k = k0 + 4 * (l >> 4);
    const uint16_t* r = W + n * ldw + k;
    bf16x4 v;
    v[0] = (short)r[0];
    v[1] = (short)r[1];
    v[2] = (short)r[2];
    v[3] = (short)r[3];
    return v;
}
// B operand from an LDS tile T[16 rows][ld], rows k (batch), columns n0..
__device__ __forceinline__ bf16x4 load_b_lds(const uint16_t* T, int ld, int n0) {
    const int l = lane_id();
    const int n = n0 + (l & 15), k = 4 * (l >> 4);
    bf16x4 v;
    v[0] = (short)T[(k + 0) * ld + n];
    v[1] = (short)T[(k + 1) * ld + n];
    v[2] = (short)T[(k + 2) * ld + n];
    v[3] = (short)T[(k + 3) * ld + n];
    return v;
}

// Store a D tile (+bias, ReLU if asked) as bf16 into LDS T[16][ld] at column n0;
// optionally keep the FP32 values.
__device__ __forceinline__ void store_d(f32x4 d, const float* bias, bool relu, uint16_t* T, int ld, int n0,
                                        float* keep, int keep_ld) {
    const int l = lane_id();
    const int n = n0 + (l & 15);
    const float bb = bias ? bias[n] : 0.0f;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int m = 4 * (l >> 4) + j;
        float v = d[j] + bb;
        if (relu) v = v > 0.0f ? v : 0.0f;
        T[m * ld + n] = f2bf(v);
        if (keep) keep[m * keep_ld + n] = v;
    }
}

// Inference runs the network transposed, H^T = W^T X^T, with 16 rows as the
// N dimension of each product: the accumulator of one layer (lane l holds
// hidden units 4 (l / 16) + j of row l % 16) is then exactly the B operand of
// the next, so activations never leave registers, and the input operand is 8
// consecutive bytes of a row-major feature row. The weight A operands
// W^T[m][k] = W[k][m] are loaded once per wave (48 VGPRs).
struct MlpFrags {
    bf16x4 w1[4];
    bf16x8 w2[8], w3[2];  // w2[c * 4 + mt]: k blocks 2c, 2c + 1 of output tile mt
    f32x4 b1[4], b2[4];
    float b3[kOut];
};
__device__ __forceinline__ MlpFrags load_frags(const MlpParams& P) {
    MlpFrags F;
    const int h0 = 4 * (lane_id() >> 4);
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
        F.w1[mt] = load_b(P.w1b, kHid, 0, mt * 16);
#pragma unroll
        for (int c = 0; c < 2; c++)
            F.w2[c * 4 + mt] = cat8(load_b(P.w2b, kHid, 2 * c * 16, mt * 16), load_b(P.w2b, kHid, (2 * c + 1) * 16, mt * 16));
#pragma unroll
        for (int j = 0; j < 4; j++) {
            F.b1[mt][j] = P.p[off_b1() + mt * 16 + h0 + j];
            F.b2[mt][j] = P.p[off_b2() + mt * 16 + h0 + j];
        }
    }
#pragma unroll
    for (int c = 0; c < 2; c++) F.w3[c] = cat8(load_b(P.w3b, 16, 2 * c * 16, 0), load_b(P.w3b, 16, (2 * c + 1) * 16, 0));
#pragma unroll
    for (int j = 0; j < kOut; j++) F.b3[j] = P.p[off_b3() + j];
    return F;
}

// ReLU as a signed-integer maximum of the bit pattern (negative floats and -0
// become +0; no canonicalising of the MFMA results), then round to BF16
// (nearest even) two values per v_cvt_pk_bf16_f32.
__device__ __forceinline__ float relu_bits(float x) {
    return __builtin_bit_cast(float, __builtin_elementwise_max(__builtin_bit_cast(int, x), 0));
}
__device__ __forceinline__ uint32_t relu_pk_bf16(float a, float b) {
    typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
    typedef float f32x2_t __attribute__((ext_vector_type(2)));
    const f32x2_t v = {relu_bits(a), relu_bits(b)};
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2_t));
}
__device__ __forceinline__ bf16x4 relu_bf16(f32x4 d) {
    typedef uint32_t u32x2_t __attribute__((ext_vector_type(2)));
    const u32x2_t p = {relu_pk_bf16(d[0], d[1]), relu_pk_bf16(d[2], d[3])};
    return __builtin_bit_cast(bf16x4, p);
}

// Forward pass of 16 rows: x is the B operand X^T (lane l: row l % 16,
// features 4 (l / 16) .. +3). Lanes 0..15 receive outputs 0..3 of row l.
// Each layer's accumulator starts from its bias.
__device__ __forceinline__ f32x4 mlp_forward16(const MlpFrags& W, bf16x4 x) {
    bf16x4 a1[4], a2[4];
#pragma unroll
    for (int mt = 0; mt < 4; mt++) a1[mt] = relu_bf16(mfma16(W.w1[mt], x, W.b1[mt]));
    f32x4 h2[4];
#pragma unroll
    for (int mt = 0; mt < 4; mt++) h2[mt] = W.b2[mt];
#pragma unroll
    for (int c = 0; c < 2; c++) {
        const bf16x8 b = cat8(a1[2 * c], a1[2 * c + 1]);
#pragma unroll
        for (int mt = 0; mt < 4; mt++) h2[mt] = mfma32(W.w2[c * 4 + mt], b, h2[mt]);
    }
#pragma unroll
    for (int mt = 0; mt < 4; mt++) a2[mt] = relu_bf16(h2[mt]);
    f32x4 o = {W.b3[0], W.b3[1], W.b3[2], W.b3[3]};
#pragma unroll
    for (int c = 0; c < 2; c++) o = mfma32(W.w3[c], cat8(a2[2 * c], a2[2 * c + 1]), o);
    return o;
}

// The network's operand fragments in LDS for the inference kernels, in the
// order the forward pass reads them: lane l's fragment of each operand is
// contiguous (conflict-free 8- and 16-byte reads; the bias fragments repeat
// over lane groups and are read as broadcasts). Reading them per pass keeps
// the weights out of the registers the interval work needs.
struct MlpLds {
    bf16x4 w1[4][64];
    bf16x8 w2[8][64];
    bf16x8 w3[2][64];
    f32x4 b1[4][4];  // [mt][l / 16]
    f32x4 b2[4][4];
    float b3[4];
};

// Stage the fragments: every thread of the block calls this, each copying
// a share of the elements (element j of lane l's fragment of output tile mt,
// as load_frags assembles it: A[m][k] = W[k][m], m = 16 mt + l % 16,
// k = 4 (l / 16) + j within each 16-wide k block).
__device__ __forceinline__ void stage_mlp(const MlpParams& P, MlpLds* L) {
    const int tid = (int)threadIdx.x, nt = (int)blockDim.x;
    uint16_t* w1 = (uint16_t*)L->w1;
    for (int e = tid; e < 4 * 64 * 4; e += nt) {  // [mt][l][j]
        const int j = e & 3, l = (e >> 2) & 63, mt = e >> 8;
        w1[e] = P.w1b[(4 * (l >> 4) + j) * kHid + mt * 16 + (l & 15)];
    }
    uint16_t* w2 = (uint16_t*)L->w2;
    for (int e = tid; e < 8 * 64 * 8; e += nt) {  // [c * 4 + mt][l][j], j < 4: k block 2c, else 2c + 1
        const int j = e & 7, l = (e >> 3) & 63, q = e >> 9, c = q >> 2, mt = q & 3;
        const int k = (2 * c + (j >> 2)) * 16 + 4 * (l >> 4) + (j & 3);
        w2[e] = P.w2b[k * kHid + mt * 16 + (l & 15)];
    }
    uint16_t* w3 = (uint16_t*)L->w3;
    for (int e = tid; e < 2 * 64 * 8; e += nt) {  // [c][l][j]
        const int j = e & 7, l = (e >> 3) & 63, c = e >> 9;
        const int k = (2 * c + (j >> 2)) * 16 + 4 * (l >> 4) + (j & 3);
        w3[e] = P.w3b[k * 16 + (l & 15)];
    }
    float* b1 = (float*)L->b1;
    float* b2 = (float*)L->b2;
    for (int e = tid; e < 4 * 4 * 4; e += nt) {  // [mt][g][j]: unit 16 mt + 4 g + j (g = l / 16)
        b1[e] = P.p[off_b1() + e];
        b2[e] = P.p[off_b2() + e];
    }
    if (tid < kOut) L->b3[tid] = P.p[off_b3() + tid];
    __syncthreads();
}

// mlp_forward16 with the operands read from LDS as each product needs them
// (scheduling barriers between the layers keep the reads of one layer from
// being hoisted above the products of the one before, which would hold all
// the fragments in registers at once).
__device__ __forceinline__ f32x4 mlp_forward16_lds(const MlpLds* L, bf16x4 x) {
    const int l = lane_id();
    bf16x4 a1[4], a2[4];
#pragma unroll
    for (int mt = 0; mt < 4; mt++) a1[mt] = relu_bf16(mfma16(L->w1[mt][l], x, L->b1[mt][l >> 4]));
    __builtin_amdgcn_sched_barrier(0);
    f32x4 h2[4];
#pragma unroll
    for (int mt = 0; mt < 4; mt++) h2[mt] = L->b2[mt][l >> 4];
#pragma unroll
    for (int c = 0; c < 2; c++) {
        const bf16x8 b = cat8(a1[2 * c], a1[2 * c + 1]);
#pragma unroll
        for (int mt = 0; mt < 4; mt++) h2[mt] = mfma32(L->w2[c * 4 + mt][l], b, h2[mt]);
        __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int mt = 0; mt < 4; mt++) a2[mt] = relu_bf16(h2[mt]);
    f32x4 o = {L->b3[0], L->b3[1], L->b3[2], L->b3[3]};
#pragma unroll
    for (int c = 0; c < 2; c++) o = mfma32(L->w3[c][l], cat8(a2[2 * c], a2[2 * c + 1]), o);
    return o;
}

// B operand X^T from a row-major BF16 row block (rows r0.., 16 features).
__device__ __forceinline__ bf16x4 load_xt(const uint16_t* rows, int64_t r0, int64_t n) {
    const int l = lane_id();
    const int64_t r = r0 + (l & 15);
    bf16x4 v = {0, 0, 0, 0};
    if (r < n) v = *(const bf16x4*)(rows + r * kFeat + 4 * (l >> 4));
    return v;
}

// Output 3 > 0: the quintic start of the interval's roots may go straight to
// the FP64 Newton step (the sampler still checks each sample's step).
__device__ __forceinline__ int8_t start_from_outputs(float o3) { return o3 > 0.f ? (int8_t)1 : (int8_t)0; }

__device__ __forceinline__ int8_t k_from_outputs(float o0, float o1, int hedge) {
    int k = (int)rintf(o0);
    if (hedge && o1 < 0.f) k += 1;
    return (int8_t)clamp_k(k);
}

struct FeatArgs {
    int64_t n;                 // intervals
    int kind;                  // 0 ellipse, 1 B-spline
    const double* a;           // ellipse
    const double* b;
    int normalize;
    // B-spline
    const int32_t* int_curve;
    const int32_t* degree;
    const int32_t* dim;
    const int64_t* ioff;
    const int64_t* ctrl_off;
    const double* ctrl;
    int ctrl_stride;
    const double* coef64;      // normalised power-basis coefficients
    const float* coef32;       // FP32 table (FP32 builds), preferred when present
    int cstride;               // values per interval block
    int ls_plain;              // log2 n_subdivided (safety floor of the plain root count)
    const double* h;
    float tol0;                // level-0 tolerance (normalised units)
};

// Features of interval i and its squared min / max speed ratio (safety floor).
__device__ __forceinline__ void interval_features(const FeatArgs& F, int64_t i, float* f, float* r2) {
    if (F.kind == 0) {
        const double aa = F.a[i], bb = F.b[i];
        const double s = F.normalize ? fmax(fmax(fabs(aa), fabs(bb)), 1e-300) : 1.0;
        features_ellipse(aa / s, bb / s, F.tol0, f);
        *r2 = ellipse_speed2_ratio(aa, bb);
        return;
    }
    const int c = F.int_curve[i];
    const int p = F.degree[c], D = F.dim[c];
    const int segid = (int)(i - F.ioff[c]);
    const int nint = (int)(F.ioff[c + 1] - F.ioff[c]);
    const double* ctrl = F.ctrl + F.ctrl_off[c] * F.ctrl_stride;
    float cf[kCoefStride];
    load_coef_f32(F.coef32, F.coef64, F.cstride, i, p * D, cf);
    *r2 = 1.0f;
    // control-polygon features are ratios, so raw control points serve
#define GCV_FEAT_CASE(PP, DD)                                                                        \
    case 10 * PP + DD:                                                                               \
        features_bspline<PP, DD>(cf, F.h[i], nint, ctrl, F.ctrl_stride, segid + PP, F.tol0, f, r2); \
        break;
    switch (10 * p + D) {
        GCV_FEAT_CASE(1, 2)
        GCV_FEAT_CASE(1, 3)
        GCV_FEAT_CASE(2, 2)
        GCV_FEAT_CASE(2, 3)
        GCV_FEAT_CASE(3, 2)
        GCV_FEAT_CASE(3, 3)
        GCV_FEAT_CASE(4, 2)
        GCV_FEAT_CASE(4, 3)
        GCV_FEAT_CASE(5, 2)
        GCV_FEAT_CASE(5, 3)
        default:
            for (int j = 0; j < kFeat; j++) f[j] = 0.f;
    }
#undef GCV_FEAT_CASE
}

// Raw features of every interval: out[i * 16 + j].
static __global__ __launch_bounds__(256) void k_features(FeatArgs F, float* out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= F.n) return;
    float f[kFeat], r2;
    interval_features(F, i, f, &r2);
#pragma unroll
    for (int j = 0; j < kFeat; j++) out[i * kFeat + j] = f[j];
}

// ---------------------------------------------------------------------------
// Inference: features -> standardise -> MLP -> k = clamp(rint(out0)) (+1 when
// the pass head predicts a level-0 failure and hedge is set), with the
// features computed where the interval's data already is, so no feature rows
// go through memory:
//   k_mlp_predict        ellipse tables: features from the axes
//   k_bspline_coef_seed  B-spline tables: the coefficient pass of one class
//                        (the same tables as k_bspline_coef), features from
//                        the coefficients in registers
// Each wave takes 64 intervals at a time: lane j standardises interval j's
// features into an LDS row, and four 16-row forward passes read them back.
// ---------------------------------------------------------------------------
constexpr int kRowLd = kFeat + 4;  // LDS row stride (elements): 40 bytes, conflict-free 8-byte reads
constexpr int kWaveLds = 64 * kRowLd * 2;
constexpr int kMlpLdsBytes = (int)sizeof(MlpLds);

// Where the predictions go: kk[i] (k of interval i), optional raw outputs,
// the safety floor kmin[i].
struct SeedOut {
    int8_t* kk;
    float* raw;     // optional [n][4]
    int8_t* kmin;
    int hedge;
    int ls_plain;   // log2 n_subdivided (safety floor of the plain root count)
    float tol0;     // level-0 tolerance (normalised units)
    int8_t* start_ok = nullptr;  // optional [n]: the Newton-start decision (output 3)
};

// Lane's standardised BF16 row into X[lane] (zeros for a lane without an
// interval: its outputs are never stored).
__device__ __forceinline__ void stage_row(uint16_t* X, const float* f, const MlpParams& P, bool valid) {
    const int lane = lane_id();
    typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
    typedef float f32x2_t __attribute__((ext_vector_type(2)));
    uint32_t w[kFeat / 2];
#pragma unroll
    for (int j = 0; j < kFeat / 2; j++) {
        // round to nearest even, two values per conversion (finite features:
        // the same bits as f2bf)
        const f32x2_t v = {(f[2 * j] - P.feat_mean[2 * j]) * P.feat_istd[2 * j],
                           (f[2 * j + 1] - P.feat_mean[2 * j + 1]) * P.feat_istd[2 * j + 1]};
        w[j] = valid ? __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2_t)) : 0u;
    }
    uint2* row = (uint2*)(X + lane * kRowLd);
#pragma unroll
    for (int j = 0; j < kFeat / 4; j++) row[j] = make_uint2(w[2 * j], w[2 * j + 1]);
}

// Forward passes of the wave's 64 staged rows; row r belongs to lane r,
// whose interval is `i` and which is valid when r < n_valid.
__device__ __forceinline__ void predict_rows(const MlpLds* Wl, const uint16_t* X, int64_t i, int n_valid,
                                             const SeedOut& O) {
    const int lane = lane_id();
    wave_lds_sync();
    // an offset the compiler cannot see through keeps the fragment reads in
    // each pass (hoisted out of the caller's loop they would hold 48 or more
    // registers through the interval work)
    int opaque = 0;
    asm volatile("" : "+s"(opaque));
    Wl = (const MlpLds*)((const char*)Wl + opaque);
#pragma unroll 1
    for (int nt = 0; nt < 4; nt++) {
        const uint2 q = *(const uint2*)(X + (nt * 16 + (lane & 15)) * kRowLd + 4 * (lane >> 4));
        bf16x4 x;
        x[0] = (short)(q.x & 0xFFFFu);
        x[1] = (short)(q.x >> 16);
        x[2] = (short)(q.y & 0xFFFFu);
        x[3] = (short)(q.y >> 16);
        const f32x4 o = mlp_forward16_lds(Wl, x);
        const int r = nt * 16 + (lane & 15);
        const int64_t ir = (int64_t)__shfl((long long)i, r, 64);
        if (lane < 16 && r < n_valid) {
            O.kk[ir] = k_from_outputs(o[0], o[1], O.hedge);
            if (O.start_ok) O.start_ok[ir] = start_from_outputs(o[3]);
            if (O.raw) *(f32x4*)(O.raw + ir * kOut) = o;
        }
    }
    wave_lds_sync();
}

static __global__ __launch_bounds__(256) void k_mlp_predict(FeatArgs F, MlpParams P, SeedOut O) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = lane_id();
    const int w = threadIdx.x >> 6;
    MlpLds* Wl = (MlpLds*)smem;
    uint16_t* X = (uint16_t*)(smem + sizeof(MlpLds) + (size_t)w * kWaveLds);  // [64][kRowLd]
    const int64_t gw = (int64_t)blockIdx.x * (blockDim.x >> 6) + w;
    const int64_t nw = (int64_t)gridDim.x * (blockDim.x >> 6);
    stage_mlp(P, Wl);
    for (int64_t t0 = gw * 64; t0 < F.n; t0 += nw * 64) {
        const int64_t i = t0 + lane;
        const bool valid = i < F.n;
        float f[kFeat], r2 = 1.0f;
        if (valid) interval_features(F, i, f, &r2);
        else
            for (int j = 0; j < kFeat; j++) f[j] = 0.f;
        if (valid && O.kmin) O.kmin[i] = floor_from_ratio(r2, O.ls_plain);
        stage_row(X, f, P, valid);
        predict_rows(Wl, X, i, (int)(F.n - t0 < 64 ? F.n - t0 : 64), O);
    }
}

// Coefficient pass of one B-spline class (degree P, dimension D) fused with
// the learned seed: writes what k_bspline_coef writes, and the interval's k
// and safety floor. Grid-stride over 64-slot wave groups (the network's
// weights are loaded once per wave).
template <int P, int D>
__global__ __launch_bounds__(256) void k_bspline_coef_seed(BSplineDev B, const int32_t* list, int64_t n_slots,
                                                           const int64_t* ioff, const int32_t* int_curve,
                                                           const double* int_scale, double* coef64, float* coef32,
                                                           int cstride, double* u0, double* h, MlpParams Pm,
                                                           SeedOut O) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = lane_id();
    const int w = threadIdx.x >> 6;
    MlpLds* Wl = (MlpLds*)smem;
    uint16_t* X = (uint16_t*)(smem + sizeof(MlpLds) + (size_t)w * kWaveLds);
    const int64_t gw = (int64_t)blockIdx.x * (blockDim.x >> 6) + w;
    const int64_t nw = (int64_t)gridDim.x * (blockDim.x >> 6);
    stage_mlp(Pm, Wl);
    for (int64_t t0 = gw * 64; t0 < n_slots; t0 += nw * 64) {
        const int64_t slot = t0 + lane;
        const bool valid = slot < n_slots;
        int64_t i = 0;
        float f[kFeat], r2 = 1.0f;
        if (valid) {
            i = list ? (int64_t)list[slot] : slot;
            const int c = int_curve[i];
            const int64_t i0 = ioff[c];
            const int nint = (int)(ioff[c + 1] - i0);
            const int s = (int)(i - i0) + P;
            const double* T = B.knots + B.knot_off[c];
            const double* Pc = B.ctrl + B.ctrl_off[c] * B.ctrl_stride;
            float cf[P * D];
            double hh;
            coef_block<P, D>(s, T, Pc, B.ctrl_stride, rcp_f64(int_scale[i]), coef64 ? coef64 + i * cstride : nullptr,
                             coef32 ? coef32 + i * cstride : nullptr, &u0[i], &h[i], cf, &hh);
            // control-polygon features are ratios, so raw control points serve
            features_bspline<P, D>(cf, hh, nint, Pc, B.ctrl_stride, s, O.tol0, f, &r2);
            O.kmin[i] = floor_from_ratio(r2, O.ls_plain);
        } else {
#pragma unroll
            for (int j = 0; j < kFeat; j++) f[j] = 0.f;
        }
        stage_row(X, f, Pm, valid);
        predict_rows(Wl, X, i, (int)(n_slots - t0 < 64 ? n_slots - t0 : 64), O);
    }
}

// ---------------------------------------------------------------------------
// Training: forward + backward over a batch of standardised features X (bf16
// [n][16]) with targets y ([n][4]: log2 seed, pass flag, scaled log2 start
// error, direct-start flag). Loss per row:
//   (out0 - y0)^2 + w_pass BCE(sigmoid(out1), y1)
//     + (out2 - y2)^2 + w_pass BCE(sigmoid(out3), y3),  averaged over n;
// a NaN target leaves its term (and gradient) out.
// Gradients accumulate into grad[kParams] (FP32, zeroed beforehand); the
// summed loss into loss[0].
// ---------------------------------------------------------------------------
constexpr int kTrainLds = (16 * kFeat + 3 * 16 * kHid + 16 * 16) * 2 + (2 * 16 * kHid) * 4;

static __global__ __launch_bounds__(64) void k_mlp_train(const uint16_t* Xg, const float* yg, int64_t n, MlpParams P,
                                                         float w_pass, float* grad, float* loss) {
    extern __shared__ __align__(16) unsigned char smem[];
    uint16_t* X = (uint16_t*)smem;          // [16][16]
    uint16_t* H1 = X + 16 * kFeat;          // [16][64]
    uint16_t* H2 = H1 + 16 * kHid;          // [16][64]
    uint16_t* dH = H2 + 16 * kHid;          // [16][64] dH2 then dH1 (bf16)
    uint16_t* dO = dH + 16 * kHid;          // [16][16] dOut padded (bf16)
    float* H1f = (float*)(dO + 16 * 16);    // [16][64] FP32 activations (ReLU masks)
    float* H2f = H1f + 16 * kHid;           // [16][64]
    const int l = lane_id();
    const float* b1 = P.p + off_b1();
    const float* b2 = P.p + off_b2();
    const float* b3 = P.p + off_b3();

    // weight-gradient accumulators over this wave's tiles
    f32x4 gW1[4], gW2[16], gW3[4];
#pragma unroll
    for (int t = 0; t < 4; t++) gW1[t] = gW3[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int t = 0; t < 16; t++) gW2[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    float gb1[4] = {0.f, 0.f, 0.f, 0.f}, gb2[4] = {0.f, 0.f, 0.f, 0.f};  // per column col + 16 nt
    float gb3 = 0.f, lsum = 0.f;

    const int64_t ntile = (n + 15) / 16;
    for (int64_t tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
        const int64_t r0 = tile * 16;
        // stage X tile (rows past n are zero and carry zero gradient)
        for (int e = l; e < 16 * kFeat; e += 64) {
            const int64_t r = r0 + e / kFeat;
            X[e] = r < n ? Xg[r0 * kFeat + e] : (uint16_t)0;
        }
        wave_lds_sync();
        // forward, keeping FP32 activations
        const bf16x4 ax = load_a(X, kFeat, 0);
#pragma unroll
        for (int nt = 0; nt < 4; nt++) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            acc = mfma16(ax, load_b(P.w1b, kHid, 0, nt * 16), acc);
            store_d(acc, b1, true, H1, kHid, nt * 16, H1f, kHid);
        }
        wave_lds_sync();
#pragma unroll
        for (int nt = 0; nt < 4; nt++) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kt = 0; kt < 4; kt++)
                acc = mfma16(load_a(H1, kHid, kt * 16), load_b(P.w2b, kHid, kt * 16, nt * 16), acc);
            store_d(acc, b2, true, H2, kHid, nt * 16, H2f, kHid);
        }
        wave_lds_sync();
        f32x4 o = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kt = 0; kt < 4; kt++) o = mfma16(load_a(H2, kHid, kt * 16), load_b(P.w3b, 16, kt * 16, 0), o);
        // output gradient, lane holds rows 4(l/16)+j of column l%16
        const int col = l & 15;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int m = 4 * (l >> 4) + j;
            const int64_t r = r0 + m;
            float g = 0.f;
            const float y = r < n && col < kOut ? yg[r * kOut + col] : 0.f;
            if (r < n && col < kOut && y == y) {
                const float out = o[j] + b3[col];
                if ((col & 1) == 0) {
                    const float d = out - y;
                    lsum += d * d;
                    g = 2.f * d;
                } else {
                    const float sgm = 1.f / (1.f + expf(-out));
                    lsum += w_pass * (fmaxf(out, 0.f) - out * y + log1pf(expf(-fabsf(out))));
                    g = w_pass * (sgm - y);
                }
                g *= 1.0f / (float)n;
            }
            dO[m * 16 + col] = f2bf(g);
            if (col < kOut) gb3 += g;
        }
        wave_lds_sync();
        // dW3 += H2^T dO  (M = 64 hidden in 4 tiles, K = 16 rows, N = 16 -> 4 used)
        const bf16x4 bo = load_b_lds(dO, 16, 0);
#pragma unroll
        for (int mt = 0; mt < 4; mt++) gW3[mt] = mfma16(load_at(H2, kHid, mt * 16), bo, gW3[mt]);
        // dH2 = (dO W3^T) * [H2 > 0]  (K = 16 with only 4 nonzero)
        const bf16x4 ad = load_a(dO, 16, 0);
#pragma unroll
        for (int nt = 0; nt < 4; nt++) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            acc = mfma16(ad, load_bt(P.w3b, 16, 0, nt * 16), acc);
            const int nn = nt * 16 + col;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int m = 4 * (l >> 4) + j;
                const float v = H2f[m * kHid + nn] > 0.f ? acc[j] : 0.f;
                dH[m * kHid + nn] = f2bf(v);
                gb2[nt] += v;
            }
        }
        wave_lds_sync();
        // dW2 += H1^T dH2  (M = 64 in 4 tiles, K = 16 rows, N = 64 in 4 tiles)
#pragma unroll
        for (int mt = 0; mt < 4; mt++) {
            const bf16x4 a = load_at(H1, kHid, mt * 16);
#pragma unroll
            for (int nt = 0; nt < 4; nt++) gW2[mt * 4 + nt] = mfma16(a, load_b_lds(dH, kHid, nt * 16), gW2[mt * 4 + nt]);
        }
        // dH1 = (dH2 W2^T) * [H1 > 0]
        f32x4 d1[4];
#pragma unroll
        for (int nt = 0; nt < 4; nt++) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kt = 0; kt < 4; kt++)
                acc = mfma16(load_a(dH, kHid, kt * 16), load_bt(P.w2b, kHid, kt * 16, nt * 16), acc);
            d1[nt] = acc;
        }
        wave_lds_sync();
#pragma unroll
        for (int nt = 0; nt < 4; nt++) {
            const int nn = nt * 16 + col;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int m = 4 * (l >> 4) + j;
                const float v = H1f[m * kHid + nn] > 0.f ? d1[nt][j] : 0.f;
                dH[m * kHid + nn] = f2bf(v);
                gb1[nt] += v;
            }
        }
        wave_lds_sync();
        // dW1 += X^T dH1  (M = 16 features, K = 16 rows, N = 64 in 4 tiles)
        const bf16x4 axt = load_at(X, kFeat, 0);
#pragma unroll
        for (int nt = 0; nt < 4; nt++) gW1[nt] = mfma16(axt, load_b_lds(dH, kHid, nt * 16), gW1[nt]);
        wave_lds_sync();
    }
    // flush: D tiles are [m = 4(l/16)+j][n = l%16]
    const int col = l & 15;
#pragma unroll
    for (int nt = 0; nt < 4; nt++)
#pragma unroll
        for (int j = 0; j < 4; j++)
            atomicAdd(&grad[off_W1() + (4 * (l >> 4) + j) * kHid + nt * 16 + col], gW1[nt][j]);
#pragma unroll
    for (int mt = 0; mt < 4; mt++)
#pragma unroll
        for (int nt = 0; nt < 4; nt++)
#pragma unroll
            for (int j = 0; j < 4; j++)
                atomicAdd(&grad[off_W2() + (mt * 16 + 4 * (l >> 4) + j) * kHid + nt * 16 + col], gW2[mt * 4 + nt][j]);
    if (col < kOut) {
#pragma unroll
        for (int mt = 0; mt < 4; mt++)
#pragma unroll
            for (int j = 0; j < 4; j++)
                atomicAdd(&grad[off_W3() + (mt * 16 + 4 * (l >> 4) + j) * kOut + col], gW3[mt][j]);
    }
    // bias gradients: lanes l, l^16, l^32, l^48 hold the same columns; reduce over them
    for (int o = 16; o < 64; o <<= 1) {
#pragma unroll
        for (int nt = 0; nt < 4; nt++) {
            gb1[nt] += __shfl_xor(gb1[nt], o, 64);
            gb2[nt] += __shfl_xor(gb2[nt], o, 64);
        }
        gb3 += __shfl_xor(gb3, o, 64);
    }
    if (l < 16) {
#pragma unroll
        for (int nt = 0; nt < 4; nt++) {
            atomicAdd(&grad[off_b1() + nt * 16 + l], gb1[nt]);
            atomicAdd(&grad[off_b2() + nt * 16 + l], gb2[nt]);
        }
        if (l < kOut) atomicAdd(&grad[off_b3() + l], gb3);
    }
    for (int o = 1; o < 64; o <<= 1) lsum += __shfl_xor(lsum, o, 64);
    if (l == 0) atomicAdd(loss, lsum);
}

// BF16 operand copies of the weights (W3 padded to 16 columns).
__device__ __forceinline__ void sync_bf16_one(const MlpParams& P, int i) {
    if (i < off_b1()) {
        P.w1b[i] = f2bf(P.p[i]);
    } else if (i >= off_W2() && i < off_b2()) {
        P.w2b[i - off_W2()] = f2bf(P.p[i]);
    } else if (i >= off_W3() && i < off_b3()) {
        const int r = (i - off_W3()) / kOut, cc = (i - off_W3()) % kOut;
        P.w3b[r * 16 + cc] = f2bf(P.p[i]);
    }
}

static __global__ void k_mlp_sync_bf16(MlpParams P) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kHid * 16 && (i % 16) >= kOut) P.w3b[i] = 0;  // padding columns of W3
    if (i < kParams) sync_bf16_one(P, i);
}

// Adam step on the FP32 master weights, then their BF16 copies.
static __global__ void k_mlp_adam(MlpParams P, float* m, float* v, const float* grad, float lr, float beta1,
                                  float beta2, float eps, float bc1, float bc2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kParams) return;
    const float g = grad[i];
    const float mi = beta1 * m[i] + (1.f - beta1) * g;
    const float vi = beta2 * v[i] + (1.f - beta2) * g * g;
    m[i] = mi;
    v[i] = vi;
    P.p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
    sync_bf16_one(P, i);
}

// Forward only over n standardised bf16 rows (tests and evaluation): out[n][4].
static __global__ __launch_bounds__(64) void k_mlp_forward(const uint16_t* Xg, int64_t n, MlpParams P, float* outg) {
    const int l = lane_id();
    const int64_t ntile = (n + 15) / 16;
    const MlpFrags Wf = load_frags(P);
    for (int64_t tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
        const int64_t r0 = tile * 16;
        const f32x4 o = mlp_forward16(Wf, load_xt(Xg, r0, n));
        if (l < 16 && r0 + l < n) *(f32x4*)(outg + (r0 + l) * kOut) = o;
    }
}

}  // namespace dev
}  // namespace gcurval
