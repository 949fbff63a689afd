// (c) Element-grouped atomic MLPs on gfx950 matrix cores.
//
// Replaces _compute_energy_per_atom (potentials/nnp/energy.py:23-39):
// scaler (descriptors/scaler.py:206-246) -> [Dense -> act] x hidden ->
// Dense(1) -> act_out (models/nn/model.py:24-58, y = x @ W + b), plus the
// reverse-mode gradient dE_i/dG_i that the force adjoint consumes.
//
// Atoms of one element are contiguous (internal order), so each layer is a
// plain GEMM over a 16-atom tile:
//   PT_MLP_F64    v_mfma_f64_16x16x4_f64   (FP64 in, FP64 accumulate)
//   PT_MLP_BF16   v_mfma_f32_16x16x32_bf16 (BF16 in, FP32 accumulate)
//   PT_MLP_BF16X3 the same with split operands x = hi + lo (hi*hi + hi*lo +
//                 lo*hi), giving about 16 mantissa bits per product.
// One wavefront owns a tile; activations and activation derivatives of
// each hidden layer go through LDS, weights come from a packed device
// buffer (both W and W^T layouts, so that every MFMA operand fragment is a
// contiguous run of K elements) built once by pt_mlp_pack.
//
// Operand layouts (16x16 output tile, lane l, accumulator element i = 0..3):
//   A[row = l % 16][k = kc + l / 16 (f64) | kc + 8 (l / 16) + 0..7 (bf16)]
//   B[k (same as A)][col = l % 16]
//   C[row][col = l % 16] with row = l / 16 + 4 i   (v_mfma_f64_16x16x4_f64)
//                         or  row = 4 (l / 16) + i (v_mfma_f32_16x16x32_bf16)
#include <algorithm>
#include <cmath>

#include "common.h"

typedef double pt_d4 __attribute__((ext_vector_type(4)));
typedef float pt_f4 __attribute__((ext_vector_type(4)));
typedef __bf16 pt_bf8 __attribute__((ext_vector_type(8)));

namespace pt {
namespace {

constexpr int kPad = 32;     // K/N padding of every layer
constexpr int kMaxW = 128;   // widest padded layer
constexpr int kMaxHidden = 3;

inline int padw(int w) { return (w + kPad - 1) / kPad * kPad; }

struct MlpShape {
    int n_hidden;
    int in_p;
    int w[kMaxHidden];    // true widths
    int wp[kMaxHidden];   // padded widths
    int act[kMaxHidden];
    int act_out;
    int kout;             // padded input width of the output layer
    int wmax;
    // packed buffer offsets (bytes)
    size_t W64[kMaxHidden], b64[kMaxHidden], wout64, bout64;
    size_t WTh[kMaxHidden], WTl[kMaxHidden], Wh[kMaxHidden], Wl[kMaxHidden], b32[kMaxHidden];
    size_t wout32, bout32;
    size_t total;
    // offsets into the unpacked FP64 parameter vector (W then b per layer)
    size_t pW[kMaxHidden + 1], pb[kMaxHidden + 1];
    int in_dim;
};

int make_shape(const pt_mlp_desc* d, MlpShape& s) {
    if (!d || d->n_layers < 1 || d->n_layers > kMaxHidden + 1) {
        set_error("mlp: n_layers must be in [1, %d]", kMaxHidden + 1);
        return PT_ERR_UNSUPPORTED;
    }
    if (d->width[d->n_layers - 1] != 1) {
        set_error("mlp: the output layer must have width 1");
        return PT_ERR_ARG;
    }
    s = MlpShape{};
    s.n_hidden = d->n_layers - 1;
    s.in_dim = d->in_dim;
    s.in_p = padw(d->in_dim);
    if (d->in_dim < 1 || s.in_p > kMaxW) {
        set_error("mlp: input width %d unsupported (max %d)", d->in_dim, kMaxW);
        return PT_ERR_UNSUPPORTED;
    }
    s.wmax = s.in_p;
    for (int l = 0; l < s.n_hidden; ++l) {
        s.w[l] = d->width[l];
        s.wp[l] = padw(d->width[l]);
        s.act[l] = d->act[l];
        if (s.w[l] < 1 || s.wp[l] > kMaxW) {
            set_error("mlp: hidden width %d unsupported (max %d)", s.w[l], kMaxW);
            return PT_ERR_UNSUPPORTED;
        }
        s.wmax = std::max(s.wmax, s.wp[l]);
    }
    s.act_out = d->act[d->n_layers - 1];
    s.kout = s.n_hidden > 0 ? s.wp[s.n_hidden - 1] : s.in_p;
    size_t o = 0, p = 0;
    auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
    int fan_in = d->in_dim;
    for (int l = 0; l <= s.n_hidden; ++l) {
        int fo = l < s.n_hidden ? s.w[l] : 1;
        s.pW[l] = p; p += (size_t)fan_in * fo;
        s.pb[l] = p; p += (size_t)fo;
        fan_in = fo;
    }
    int kin = s.in_p;
    for (int l = 0; l < s.n_hidden; ++l) {
        const size_t mn = (size_t)kin * s.wp[l];
        s.W64[l] = o; o = al(o + 8 * mn);
        s.b64[l] = o; o = al(o + 8 * (size_t)s.wp[l]);
        s.WTh[l] = o; o = al(o + 2 * mn);
        s.WTl[l] = o; o = al(o + 2 * mn);
        s.Wh[l] = o; o = al(o + 2 * mn);
        s.Wl[l] = o; o = al(o + 2 * mn);
        s.b32[l] = o; o = al(o + 4 * (size_t)s.wp[l]);
        kin = s.wp[l];
    }
    s.wout64 = o; o = al(o + 8 * (size_t)s.kout);
    s.bout64 = o; o = al(o + 8);
    s.wout32 = o; o = al(o + 4 * (size_t)s.kout);
    s.bout32 = o; o = al(o + 4);
    s.total = o;
    return PT_OK;
}

// ---- activations (models/nn/activation.py:7-60) ---------------------------
template <typename T>
__device__ __forceinline__ void activate(int a, T z, T& y, T& dy) {
    switch (a) {
    case PT_ACT_IDENTITY: y = z; dy = T(1); return;
    case PT_ACT_TANH: { T t = tanh(z); y = t; dy = T(1) - t * t; return; }
    case PT_ACT_LOGISTIC: { T s = T(1) / (T(1) + exp(-z)); y = s; dy = s * (T(1) - s); return; }
    case PT_ACT_SOFTPLUS: {
        // log(1 + e^z) computed stably; derivative = logistic(z)
        T az = z > T(0) ? z : -z;
        y = (z > T(0) ? z : T(0)) + log1p(exp(-az));
        dy = T(1) / (T(1) + exp(-z));
        return;
    }
    case PT_ACT_RELU: y = z > T(0) ? z : T(0); dy = z > T(0) ? T(1) : T(0); return;
    case PT_ACT_GAUSSIAN: { T g = exp(T(-0.5) * z * z); y = g; dy = -z * g; return; }
    case PT_ACT_COS: y = cos(z); dy = -sin(z); return;
    case PT_ACT_EXP: { T e = exp(-z); y = e; dy = -e; return; }
    case PT_ACT_HARMONIC: y = z * z; dy = T(2) * z; return;
    }
    y = z; dy = T(1);
}

// Activations whose derivative is a function of the output y: the inference
// kernel then keeps only y and turns it into dy/dz in place once the next
// layer has read it.
__host__ __device__ constexpr bool act_from_output(int a) {
    return a == PT_ACT_IDENTITY || a == PT_ACT_TANH || a == PT_ACT_LOGISTIC || a == PT_ACT_SOFTPLUS ||
           a == PT_ACT_RELU || a == PT_ACT_EXP;
}

template <typename T>
__device__ __forceinline__ T dact_from_output(int a, T y) {
    switch (a) {
    case PT_ACT_TANH: return T(1) - y * y;
    case PT_ACT_LOGISTIC: return y * (T(1) - y);
    case PT_ACT_SOFTPLUS: return -expm1(-y);          // logistic(z) = 1 - e^{-y}
    case PT_ACT_RELU: return y > T(0) ? T(1) : T(0);   // y = max(z, 0)
    case PT_ACT_EXP: return -y;                        // y = e^{-z}
    }
    return T(1);
}

// y, dy/dz and d2y/dz2 (the training kernel differentiates the input gradient)
template <typename T>
__device__ __forceinline__ void activate2(int a, T z, T& y, T& dy, T& ddy) {
    switch (a) {
    case PT_ACT_IDENTITY: y = z; dy = T(1); ddy = T(0); return;
    case PT_ACT_TANH: { T t = tanh(z); y = t; dy = T(1) - t * t; ddy = T(-2) * t * dy; return; }
    case PT_ACT_LOGISTIC: {
        T s = T(1) / (T(1) + exp(-z)); y = s; dy = s * (T(1) - s); ddy = dy * (T(1) - T(2) * s); return;
    }
    case PT_ACT_SOFTPLUS: {
        T az = z > T(0) ? z : -z;
        y = (z > T(0) ? z : T(0)) + log1p(exp(-az));
        T s = T(1) / (T(1) + exp(-z));
        dy = s; ddy = s * (T(1) - s);
        return;
    }
    case PT_ACT_RELU: y = z > T(0) ? z : T(0); dy = z > T(0) ? T(1) : T(0); ddy = T(0); return;
    case PT_ACT_GAUSSIAN: { T g = exp(T(-0.5) * z * z); y = g; dy = -z * g; ddy = (z * z - T(1)) * g; return; }
    case PT_ACT_COS: y = cos(z); dy = -sin(z); ddy = -y; return;
    case PT_ACT_EXP: { T e = exp(-z); y = e; dy = -e; ddy = e; return; }
    case PT_ACT_HARMONIC: y = z * z; dy = T(2) * z; ddy = T(2); return;
    }
    y = z; dy = T(1); ddy = T(0);
}

__device__ __forceinline__ unsigned short bf16_bits(float x) {
    // round to nearest even
    unsigned u = __float_as_uint(x);
    unsigned r = u + 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(r >> 16);
}
__device__ __forceinline__ float bf16_to_f(unsigned short b) { return __uint_as_float(((unsigned)b) << 16); }

__device__ __forceinline__ void split_bf16(float x, unsigned short& hi, unsigned short& lo) {
    hi = bf16_bits(x);
    lo = bf16_bits(x - bf16_to_f(hi));
}

// ---- packing: FP64 params -> padded W/W^T (f64, bf16 hi/lo), biases ------
__global__ void k_pack(MlpShape s, const double* __restrict__ params, char* __restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    int kin = s.in_p, fan_in = s.in_dim;
    for (int l = 0; l < s.n_hidden; ++l) {
        const int n = s.wp[l];
        const int mn = kin * n;
        if (t < mn) {
            const int k = t / n, c = t % n;  // W[k][c]
            const double v = (k < fan_in && c < s.w[l]) ? params[s.pW[l] + (size_t)k * s.w[l] + c] : 0.0;
            ((double*)(out + s.W64[l]))[t] = v;
            unsigned short hi, lo;
            split_bf16((float)v, hi, lo);
            ((unsigned short*)(out + s.Wh[l]))[t] = hi;
            ((unsigned short*)(out + s.Wl[l]))[t] = lo;
            ((unsigned short*)(out + s.WTh[l]))[c * kin + k] = hi;
            ((unsigned short*)(out + s.WTl[l]))[c * kin + k] = lo;
        }
        if (t < n) {
            const double b = t < s.w[l] ? params[s.pb[l] + t] : 0.0;
            ((double*)(out + s.b64[l]))[t] = b;
            ((float*)(out + s.b32[l]))[t] = (float)b;
        }
        fan_in = s.w[l];
        kin = n;
    }
    const int L = s.n_hidden;
    if (t < s.kout) {
        const double v = t < fan_in ? params[s.pW[L] + t] : 0.0;
        ((double*)(out + s.wout64))[t] = v;
        ((float*)(out + s.wout32))[t] = (float)v;
    }
    if (t == 0) {
        ((double*)(out + s.bout64))[0] = params[s.pb[L]];
        ((float*)(out + s.bout32))[0] = (float)params[s.pb[L]];
    }
}

// ---- the MLP kernel -----------------------------------------------------------
template <int MODE>
struct MT { typedef float T; };
template <>
struct MT<PT_MLP_F64> { typedef double T; };

struct MlpArgs {
    MlpShape s;
    int n_atoms;
    const void* G;
    int ldg;
    int g_f64;
    const double* scale;  // [2][in_dim]: sub, mul
    double scale_add;
    const char* packed;
    double* e_atom;
    void* dEdG;
    int stride;           // LDS row stride (elements)
    int nbuf;             // LDS buffers per wave
    int inplace;          // every hidden activation has act_from_output
};

// C = A(16 x K, LDS row-major, stride) * B, accumulating one 16-column tile.
// Row of accumulator element i in lane group grp = l / 16.
template <int MODE>
__device__ __forceinline__ int crow(int grp, int i) {
    return MODE == PT_MLP_F64 ? grp + 4 * i : 4 * grp + i;
}

template <int MODE>
__device__ __forceinline__ void gemm_tile(const typename MT<MODE>::T* A, int stride, int K,
                                          const char* packed, size_t off_hi, size_t off_lo,
                                          int ldb, bool bt, int n0, typename MT<MODE>::T acc[4]) {
    const int l = wave_lane();
    const int row = l & 15, grp = l >> 4;
    if constexpr (MODE == PT_MLP_F64) {
        pt_d4 c = {acc[0], acc[1], acc[2], acc[3]};
        const double* W = (const double*)(packed + off_hi);
        for (int k0 = 0; k0 < K; k0 += 4) {
            const int k = k0 + grp;
            const double a = A[row * stride + k];
            // bt == false: B[k][n] = W[k][n] (W is K x ldb);  bt == true: B[k][n] = W[n][k]
            const double b = bt ? W[(size_t)(n0 + row) * ldb + k] : W[(size_t)k * ldb + n0 + row];
            c = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
        }
        acc[0] = c[0]; acc[1] = c[1]; acc[2] = c[2]; acc[3] = c[3];
    } else {
        pt_f4 c = {acc[0], acc[1], acc[2], acc[3]};
        const unsigned short* Bh = (const unsigned short*)(packed + off_hi);
        const unsigned short* Bl = (const unsigned short*)(packed + off_lo);
        for (int k0 = 0; k0 < K; k0 += 32) {
            const int kb = k0 + 8 * grp;
            pt_bf8 ah, al, bh, bl;
            const float* ap = A + row * stride + kb;
            // B fragment: 8 consecutive k of column n0 + row, from the k-contiguous layout
            const size_t bo = (size_t)(n0 + row) * ldb + kb;
            const uint4 vh = *reinterpret_cast<const uint4*>(Bh + bo);
            const uint4 vl = *reinterpret_cast<const uint4*>(Bl + bo);
            const unsigned short* ph = reinterpret_cast<const unsigned short*>(&vh);
            const unsigned short* pl = reinterpret_cast<const unsigned short*>(&vl);
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                unsigned short hi, lo;
                split_bf16(ap[q], hi, lo);
                ah[q] = __builtin_bit_cast(__bf16, hi);
                al[q] = __builtin_bit_cast(__bf16, lo);
                bh[q] = __builtin_bit_cast(__bf16, ph[q]);
                bl[q] = __builtin_bit_cast(__bf16, pl[q]);
            }
            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, c, 0, 0, 0);
            if (MODE == PT_MLP_BF16X3) {
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, c, 0, 0, 0);
            }
        }
        acc[0] = c[0]; acc[1] = c[1]; acc[2] = c[2]; acc[3] = c[3];
    }
    (void)bt;
}

// C(16 x 16 nt) += A(16 x K, LDS row-major) * B over the column tiles
// n0 + 16 t, t < nt <= 2, of one layer.  The A fragment of a k step is read
// and split once for all tiles, and the weight fragments of all tiles are
// loaded before their products: the wave waits on weight loads once per k
// step rather than once per tile and step.  Per tile the products run in the
// same k order as gemm_tile.  Two tiles per pass: four held the weight
// fragments of too many tiles at once for four resident waves per SIMD.
template <int MODE>
__device__ __forceinline__ void gemm_cols(const typename MT<MODE>::T* A, int stride, int K, const char* packed,
                                          size_t off_hi, size_t off_lo, int ldb, bool bt, int n0, int nt,
                                          typename MT<MODE>::T (&acc)[2][4]) {
    const int l = wave_lane();
    const int row = l & 15, grp = l >> 4;
    if constexpr (MODE == PT_MLP_F64) {
        pt_d4 c[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) c[t] = pt_d4{acc[t][0], acc[t][1], acc[t][2], acc[t][3]};
        const double* W = (const double*)(packed + off_hi);
#pragma unroll 2
        for (int k0 = 0; k0 < K; k0 += 4) {
            const int k = k0 + grp;
            const double av = A[row * stride + k];
            double b[2];
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const int n = n0 + 16 * t + row;
                b[t] = t < nt ? (bt ? W[(size_t)n * ldb + k] : W[(size_t)k * ldb + n]) : 0.0;
            }
#pragma unroll
            for (int t = 0; t < 2; ++t)
                if (t < nt) c[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, b[t], c[t], 0, 0, 0);
        }
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            acc[t][0] = c[t][0]; acc[t][1] = c[t][1]; acc[t][2] = c[t][2]; acc[t][3] = c[t][3];
        }
    } else {
        pt_f4 c[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) c[t] = pt_f4{acc[t][0], acc[t][1], acc[t][2], acc[t][3]};
        const unsigned short* Bh = (const unsigned short*)(packed + off_hi);
        const unsigned short* Bl = (const unsigned short*)(packed + off_lo);
        for (int k0 = 0; k0 < K; k0 += 32) {
            const int kb = k0 + 8 * grp;
            // B fragments: 8 consecutive k of column n0 + 16 t + row (k-contiguous layout)
            uint4 vh[2], vl[2];
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                vh[t] = make_uint4(0u, 0u, 0u, 0u);
                vl[t] = vh[t];
                if (t < nt) {
                    const size_t bo = (size_t)(n0 + 16 * t + row) * ldb + kb;
                    vh[t] = *reinterpret_cast<const uint4*>(Bh + bo);
                    if (MODE == PT_MLP_BF16X3) vl[t] = *reinterpret_cast<const uint4*>(Bl + bo);
                }
            }
            pt_bf8 ah, al;
            const float* ap = A + row * stride + kb;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                unsigned short hi, lo;
                split_bf16(ap[q], hi, lo);
                ah[q] = __builtin_bit_cast(__bf16, hi);
                al[q] = __builtin_bit_cast(__bf16, lo);
            }
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                if (t < nt) {
                    const pt_bf8 bh = __builtin_bit_cast(pt_bf8, vh[t]);
                    const pt_bf8 bl = __builtin_bit_cast(pt_bf8, vl[t]);
                    c[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, c[t], 0, 0, 0);
                    if (MODE == PT_MLP_BF16X3) {
                        c[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, c[t], 0, 0, 0);
                        c[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, c[t], 0, 0, 0);
                    }
                }
            }
        }
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            acc[t][0] = c[t][0]; acc[t][1] = c[t][1]; acc[t][2] = c[t][2]; acc[t][3] = c[t][3];
        }
    }
    (void)bt;
}

// Lanes over the cells of a 16-row tile, for elementwise passes: a chunk of 64
// columns gives lane l column k0 + l of all 16 rows; a last chunk of 32 columns
// (widths are multiples of 32) gives lanes 32-63 the odd rows, 8 rows each.
struct TileCols {
    int k;       // this lane's column
    int r0;      // first row
    int rstep;   // row step
    int nr;      // rows per lane (16 or 8)
};

__device__ __forceinline__ TileCols tile_cols(int k0, int width) {
    const int l = wave_lane();
    if (width - k0 >= kWave) return TileCols{k0 + l, 0, 1, 16};
    return TileCols{k0 + (l & 31), l >> 5, 2, 8};
}

template <int MODE>
__global__ __launch_bounds__(256) void k_mlp(MlpArgs a) {
    typedef typename MT<MODE>::T T;
    extern __shared__ __align__(16) char smem[];
    const MlpShape& s = a.s;
    const int l = wave_lane();
    const int wv = uniform(threadIdx.x >> 6);
    const int wpb = blockDim.x >> 6;
    const int stride = a.stride;
    T* base = reinterpret_cast<T*>(smem) + (size_t)wv * ((size_t)a.nbuf * 16 * stride + 16);
    // 16-row buffers: the input in slot 0, hidden output h in slot ys(h)
    // (slot 1, then slot 0 once the input is dead, then 2, 3, ...).  With
    // a.inplace the derivative of layer h replaces its output in place after
    // layer h + 1 has read it (max(2, H) slots); otherwise it has a slot of
    // its own after those.
    const int H = s.n_hidden;
    const int ny = H > 2 ? H : 2;
    auto ys = [](int h) { return h == 0 ? 1 : (h == 1 ? 0 : h); };
    T* D[kMaxHidden];                              // activation derivatives
    for (int h = 0; h < kMaxHidden; ++h)
        D[h] = base + (size_t)(a.inplace ? ys(h) : ny + h) * 16 * stride;
    T* gout = base + (size_t)a.nbuf * 16 * stride;  // 16 output derivatives
    const T* wout = (MODE == PT_MLP_F64) ? (const T*)(a.packed + s.wout64) : (const T*)(a.packed + s.wout32);
    const T bout = (MODE == PT_MLP_F64) ? (T)((const double*)(a.packed + s.bout64))[0]
                                        : (T)((const float*)(a.packed + s.bout32))[0];

    const int row = l & 15, grp = l >> 4;
    const int ntiles = (a.n_atoms + 15) / 16;
    for (int tile = blockIdx.x * wpb + wv; tile < ntiles; tile += gridDim.x * wpb) {
        const int a0 = tile * 16;
        // 1. scaled input -> slot 0 (zero padded): every row's load of a
        //    column is issued before the first store
        for (int k0 = 0; k0 < s.in_p; k0 += kWave) {
            const TileCols tc = tile_cols(k0, s.in_p);
            const int k = tc.k;
            const bool col_ok = k < s.in_dim;
            const double sub = col_ok ? a.scale[k] : 0.0;
            const double mul = col_ok ? a.scale[s.in_dim + k] : 0.0;
            for (int j0 = 0; j0 < tc.nr; j0 += 8) {
                double g[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const int atom = a0 + tc.r0 + (j0 + j) * tc.rstep;
                    g[j] = 0.0;
                    if (col_ok && atom < a.n_atoms)
                        g[j] = a.g_f64 ? ((const double*)a.G)[(size_t)atom * a.ldg + k]
                                       : (double)((const float*)a.G)[(size_t)atom * a.ldg + k];
                }
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const int r = tc.r0 + (j0 + j) * tc.rstep;
                    base[r * stride + k] = (col_ok && a0 + r < a.n_atoms) ? (T)((g[j] - sub) * mul + a.scale_add)
                                                                         : T(0);
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();

        // 2. hidden layers, 32 columns (two MFMA tiles) at a time
        int kin = s.in_p;
        T* xin = base;
        for (int h = 0; h < H; ++h) {
            const int wp = s.wp[h];
            const int act = s.act[h];
            T* yout = base + (size_t)ys(h) * 16 * stride;
            for (int n0 = 0; n0 < wp; n0 += 32) {
                const int nt = min(2, (wp - n0) >> 4);
                T acc[2][4] = {};
                if (MODE == PT_MLP_F64)
                    gemm_cols<MODE>(xin, stride, kin, a.packed, s.W64[h], 0, wp, false, n0, nt, acc);
                else
                    gemm_cols<MODE>(xin, stride, kin, a.packed, s.WTh[h], s.WTl[h], kin, true, n0, nt, acc);
#pragma unroll
                for (int t = 0; t < 2; ++t) {
                    if (t < nt) {
                        const int col = n0 + 16 * t + row;
                        const T bias = (MODE == PT_MLP_F64) ? (T)((const double*)(a.packed + s.b64[h]))[col]
                                                            : (T)((const float*)(a.packed + s.b32[h]))[col];
#pragma unroll
                        for (int i = 0; i < 4; ++i) {
                            const int r = crow<MODE>(grp, i);
                            T y, dy;
                            activate<T>(act, acc[t][i] + bias, y, dy);
                            if (col >= s.w[h]) { y = T(0); dy = T(0); }
                            yout[r * stride + col] = y;
                            if (!a.inplace) D[h][r * stride + col] = dy;
                        }
                    }
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            __builtin_amdgcn_wave_barrier();
            if (a.inplace && h > 0) {
                // layer h has read the output of layer h - 1: turn it into its derivative
                const int wq = s.wp[h - 1], wt = s.w[h - 1], aq = s.act[h - 1];
                for (int k0 = 0; k0 < wq; k0 += kWave) {
                    const TileCols tc = tile_cols(k0, wq);
                    for (int j0 = 0; j0 < tc.nr; j0 += 8) {
                        T v[8];
#pragma unroll
                        for (int j = 0; j < 8; ++j) v[j] = D[h - 1][(tc.r0 + (j0 + j) * tc.rstep) * stride + tc.k];
#pragma unroll
                        for (int j = 0; j < 8; ++j)
                            D[h - 1][(tc.r0 + (j0 + j) * tc.rstep) * stride + tc.k] =
                                tc.k < wt ? dact_from_output<T>(aq, v[j]) : T(0);
                    }
                }
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                __builtin_amdgcn_wave_barrier();
            }
            xin = yout;
            kin = wp;
        }
        T* xlast = xin;

        // 3. output layer (width 1): 4 lanes per atom, lane part takes k = part + 4 j
        {
            const int r = l >> 2, part = l & 3;
            T sum = T(0);
            for (int j0 = 0; j0 < kin; j0 += 32) {
                T w[8], x[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const int k = j0 + part + 4 * j;
                    w[j] = k < kin ? wout[k] : T(0);
                    x[j] = k < kin ? xlast[r * stride + k] : T(0);
                }
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    if (j0 + part + 4 * j < kin) sum += x[j] * w[j];
            }
            sum += __shfl_xor(sum, 1, kWave);
            sum += __shfl_xor(sum, 2, kWave);
            T y, dy;
            activate<T>(s.act_out, sum + bout, y, dy);
            if (part == 0) {
                gout[r] = dy;
                if (a0 + r < a.n_atoms) a.e_atom[a0 + r] = (double)y;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (!a.dEdG) continue;

        // 4. backward: dZ of the last hidden layer in place in D[h]
        if (s.n_hidden > 0) {
            const int h = s.n_hidden - 1;
            const int wq = s.wp[h], wt = s.w[h], aq = s.act[h];
            for (int k0 = 0; k0 < wq; k0 += kWave) {
                const TileCols tc = tile_cols(k0, wq);
                const T wk = wout[tc.k];
                for (int j0 = 0; j0 < tc.nr; j0 += 8) {
                    T v[8], go[8];
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        const int r = tc.r0 + (j0 + j) * tc.rstep;
                        v[j] = D[h][r * stride + tc.k];
                        go[j] = gout[r];
                    }
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        // in place: the slot still holds the last hidden output
                        const T dy = !a.inplace ? v[j] : (tc.k < wt ? dact_from_output<T>(aq, v[j]) : T(0));
                        D[h][(tc.r0 + (j0 + j) * tc.rstep) * stride + tc.k] = dy * go[j] * wk;
                    }
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            __builtin_amdgcn_wave_barrier();
            for (int hh = h; hh >= 1; --hh) {
                // dZ_{hh-1} = (dZ_hh . W_hh^T) * D[hh-1]
                const int wprev = s.wp[hh - 1];
                for (int n0 = 0; n0 < wprev; n0 += 32) {
                    const int nt = min(2, (wprev - n0) >> 4);
                    T acc[2][4] = {};
                    if (MODE == PT_MLP_F64)
                        gemm_cols<MODE>(D[hh], stride, s.wp[hh], a.packed, s.W64[hh], 0, s.wp[hh], true, n0, nt, acc);
                    else
                        gemm_cols<MODE>(D[hh], stride, s.wp[hh], a.packed, s.Wh[hh], s.Wl[hh], s.wp[hh], true, n0,
                                        nt, acc);
#pragma unroll
                    for (int t = 0; t < 2; ++t)
                        if (t < nt) {
#pragma unroll
                            for (int i = 0; i < 4; ++i)
                                D[hh - 1][crow<MODE>(grp, i) * stride + n0 + 16 * t + row] *= acc[t][i];
                        }
                }
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                __builtin_amdgcn_wave_barrier();
            }
        }
        // 5. dE/dx' = dZ_0 . W_0^T (or wout for a linear model), then chain the scaler
        for (int n0 = 0; n0 < s.in_p; n0 += 32) {
            const int nt = min(2, (s.in_p - n0) >> 4);
            T acc[2][4] = {};
            if (s.n_hidden > 0) {
                if (MODE == PT_MLP_F64)
                    gemm_cols<MODE>(D[0], stride, s.wp[0], a.packed, s.W64[0], 0, s.wp[0], true, n0, nt, acc);
                else
                    gemm_cols<MODE>(D[0], stride, s.wp[0], a.packed, s.Wh[0], s.Wl[0], s.wp[0], true, n0, nt, acc);
            } else {
#pragma unroll
                for (int t = 0; t < 2; ++t)
                    if (t < nt) {
#pragma unroll
                        for (int i = 0; i < 4; ++i) acc[t][i] = gout[crow<MODE>(grp, i)] * wout[n0 + 16 * t + row];
                    }
            }
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const int k = n0 + 16 * t + row;
                if (t < nt && k < s.in_dim) {
                    const double mk = a.scale[s.in_dim + k];
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const int atom = a0 + crow<MODE>(grp, i);
                        if (atom < a.n_atoms) {
                            const double gk = (double)acc[t][i] * mk;
                            if (a.g_f64) ((double*)a.dEdG)[(size_t)atom * a.ldg + k] = gk;
                            else ((float*)a.dEdG)[(size_t)atom * a.ldg + k] = (float)gk;
                        }
                    }
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
}

// ---- training: forward-over-reverse through the MLP -------------------------
//
// Per atom the loss gradient w.r.t. the weights is
//   gy * dy/dtheta + tseed * d(ydot)/dtheta,
// y = MLP(x'), ydot = grad_x' MLP . xdot' (xdot' = T * mul, the descriptor
// tangent in scaled space).  Rows 0..15 of a 32-row tile carry the 16 atoms'
// primal values, rows 16..31 their tangents, so every layer (and its reverse)
// is one GEMM over 32 rows on the matrix cores; the weight gradients are
// GEMMs over those 32 rows (K = atoms x {primal, tangent}).

// C(16 x 16) = A^T B over the 32 rows of two LDS tiles: A[r][i0 + ..], B[r][o0 + ..].
template <int MODE>
__device__ __forceinline__ void wgrad_tile(const typename MT<MODE>::T* A, const typename MT<MODE>::T* B,
                                           int stride, int i0, int o0, typename MT<MODE>::T out[4]) {
    const int l = wave_lane();
    const int row = l & 15, grp = l >> 4;
    if constexpr (MODE == PT_MLP_F64) {
        pt_d4 c = {0.0, 0.0, 0.0, 0.0};
        for (int k0 = 0; k0 < 32; k0 += 4) {
            const int k = k0 + grp;
            c = __builtin_amdgcn_mfma_f64_16x16x4f64(A[k * stride + i0 + row], B[k * stride + o0 + row], c, 0, 0, 0);
        }
        out[0] = c[0]; out[1] = c[1]; out[2] = c[2]; out[3] = c[3];
    } else {
        pt_f4 c = {0.f, 0.f, 0.f, 0.f};
        pt_bf8 ah, al, bh, bl;
        const int kb = 8 * grp;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            unsigned short h1, l1, h2, l2;
            split_bf16(A[(kb + q) * stride + i0 + row], h1, l1);
            split_bf16(B[(kb + q) * stride + o0 + row], h2, l2);
            ah[q] = __builtin_bit_cast(__bf16, h1);
            al[q] = __builtin_bit_cast(__bf16, l1);
            bh[q] = __builtin_bit_cast(__bf16, h2);
            bl[q] = __builtin_bit_cast(__bf16, l2);
        }
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, c, 0, 0, 0);
        if (MODE == PT_MLP_BF16X3) {
            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, c, 0, 0, 0);
        }
        out[0] = c[0]; out[1] = c[1]; out[2] = c[2]; out[3] = c[3];
    }
}

struct TrainArgs {
    MlpShape s;
    int n_atoms;
    const void* G;
    const void* T;
    int ldg;
    int g_f64;
    const double* scale;
    double scale_add;
    const char* packed;
    const double* gy;     // per-atom seed of y
    double tseed;         // seed of ydot (same for every atom)
    double* grad;         // parameter gradient, FP64, layout of the unpacked params
    int stride;
    int nparam;
};

__device__ __forceinline__ void wave_fence() {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

template <int MODE>
__global__ __launch_bounds__(256) void k_mlp_train(TrainArgs a) {
    typedef typename MT<MODE>::T T;
    extern __shared__ __align__(16) char smem[];
    const MlpShape& s = a.s;
    const int H = s.n_hidden;
    const int l = wave_lane();
    const int wv = threadIdx.x >> 6;
    const int wpb = blockDim.x >> 6;
    const int stride = a.stride;
    const size_t tile = (size_t)32 * stride;
    const size_t per_wave = (size_t)(2 * H + 3) * tile + 64 + a.nparam;
    T* base = reinterpret_cast<T*>(smem) + (size_t)wv * per_wave;
    T* Abuf[kMaxHidden + 1];
    T* Zbuf[kMaxHidden];
    for (int h = 0; h <= H; ++h) Abuf[h] = base + h * tile;
    for (int h = 0; h < H; ++h) Zbuf[h] = base + (H + 1 + h) * tile;
    T* P[2] = {base + (2 * H + 1) * tile, base + (2 * H + 2) * tile};
    T* outv = base + (2 * H + 3) * tile;  // [0,16): z_L, [16,32): zdot_L, [32,48): zbar_L, [48,64): zdotbar_L
    T* gacc = outv + 64;
    for (int k = l; k < a.nparam; k += kWave) gacc[k] = T(0);
    const int row = l & 15, grp = l >> 4;
    const T* wout = (MODE == PT_MLP_F64) ? (const T*)(a.packed + s.wout64) : (const T*)(a.packed + s.wout32);
    const T bout = (MODE == PT_MLP_F64) ? (T)((const double*)(a.packed + s.bout64))[0]
                                        : (T)((const float*)(a.packed + s.bout32))[0];
    const int ntiles = (a.n_atoms + 15) / 16;
    for (int tl = blockIdx.x * wpb + wv; tl < ntiles; tl += gridDim.x * wpb) {
        const int a0 = tl * 16;
        // scaled input and its tangent
        for (int e = l; e < 16 * s.in_p; e += kWave) {
            const int r = e / s.in_p, k = e % s.in_p;
            const int atom = a0 + r;
            T v = T(0), t = T(0);
            if (atom < a.n_atoms && k < s.in_dim) {
                const size_t o = (size_t)atom * a.ldg + k;
                const double g = a.g_f64 ? ((const double*)a.G)[o] : (double)((const float*)a.G)[o];
                const double gt = a.g_f64 ? ((const double*)a.T)[o] : (double)((const float*)a.T)[o];
                v = (T)((g - a.scale[k]) * a.scale[s.in_dim + k] + a.scale_add);
                t = (T)(gt * a.scale[s.in_dim + k]);
            }
            Abuf[0][r * stride + k] = v;
            Abuf[0][(16 + r) * stride + k] = t;
        }
        wave_fence();
        // forward: primal rows 0..15, tangent rows 16..31
        int kin = s.in_p;
        for (int h = 0; h < H; ++h) {
            const int wp = s.wp[h];
            for (int n0 = 0; n0 < wp; n0 += 16) {
                T az[4] = {T(0), T(0), T(0), T(0)}, at[4] = {T(0), T(0), T(0), T(0)};
                if (MODE == PT_MLP_F64) {
                    gemm_tile<MODE>(Abuf[h], stride, kin, a.packed, s.W64[h], 0, wp, false, n0, az);
                    gemm_tile<MODE>(Abuf[h] + 16 * stride, stride, kin, a.packed, s.W64[h], 0, wp, false, n0, at);
                } else {
                    gemm_tile<MODE>(Abuf[h], stride, kin, a.packed, s.WTh[h], s.WTl[h], kin, true, n0, az);
                    gemm_tile<MODE>(Abuf[h] + 16 * stride, stride, kin, a.packed, s.WTh[h], s.WTl[h], kin, true, n0, at);
                }
                const int col = n0 + row;
                const T bias = (MODE == PT_MLP_F64) ? (T)((const double*)(a.packed + s.b64[h]))[col]
                                                    : (T)((const float*)(a.packed + s.b32[h]))[col];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int r = crow<MODE>(grp, i);
                    const T z = az[i] + bias, zd = at[i];
                    T y, dy;
                    activate<T>(s.act[h], z, y, dy);
                    const bool pad = col >= s.w[h];
                    Zbuf[h][r * stride + col] = pad ? T(0) : z;
                    Zbuf[h][(16 + r) * stride + col] = pad ? T(0) : zd;
                    Abuf[h + 1][r * stride + col] = pad ? T(0) : y;
                    Abuf[h + 1][(16 + r) * stride + col] = pad ? T(0) : dy * zd;
                }
            }
            wave_fence();
            kin = wp;
        }
        // output layer and its reverse seeds: 4 lanes per atom
        {
            const int r = l >> 2, part = l & 3;
            T z = T(0), zd = T(0);
            for (int k = part; k < kin; k += 4) {
                z += Abuf[H][r * stride + k] * wout[k];
                zd += Abuf[H][(16 + r) * stride + k] * wout[k];
            }
            z += __shfl_xor(z, 1, kWave); z += __shfl_xor(z, 2, kWave);
            zd += __shfl_xor(zd, 1, kWave); zd += __shfl_xor(zd, 2, kWave);
            z += bout;
            T y, dy, ddy;
            activate2<T>(s.act_out, z, y, dy, ddy);
            const int atom = a0 + r;
            const bool ok = atom < a.n_atoms;
            const T gy = ok ? (T)a.gy[atom] : T(0);
            const T gt = ok ? (T)a.tseed : T(0);
            if (part == 0) {
                outv[r] = z;
                outv[16 + r] = zd;
                outv[32 + r] = gy * dy + gt * ddy * zd;  // zbar_L
                outv[48 + r] = gt * dy;                  // zdotbar_L
            }
        }
        wave_fence();
        // gradients of the output layer
        const int L = H;
        for (int k = l; k < kin; k += kWave) {
            T gw = T(0);
            for (int r = 0; r < 16; ++r)
                gw += Abuf[H][r * stride + k] * outv[32 + r] + Abuf[H][(16 + r) * stride + k] * outv[48 + r];
            const int fan_in = H > 0 ? s.w[H - 1] : s.in_dim;
            if (k < fan_in) gacc[s.pW[L] + k] += gw;
        }
        if (l == 0) {
            T gb = T(0);
            for (int r = 0; r < 16; ++r) gb += outv[32 + r];
            gacc[s.pb[L]] += gb;
        }
        // abar, adotbar of the last hidden output -> P[0]
        for (int e = l; e < 16 * kin; e += kWave) {
            const int r = e / kin, k = e % kin;
            P[0][r * stride + k] = outv[32 + r] * wout[k];
            P[0][(16 + r) * stride + k] = outv[48 + r] * wout[k];
        }
        wave_fence();
        int cur = 0;
        for (int h = H - 1; h >= 0; --h) {
            const int wp = s.wp[h];
            const int kin_h = h > 0 ? s.wp[h - 1] : s.in_p;
            // zbar = abar f'(z) + adotbar f''(z) zdot ; zdotbar = adotbar f'(z)   (in place)
            for (int e = l; e < 16 * wp; e += kWave) {
                const int r = e / wp, k = e % wp;
                const T z = Zbuf[h][r * stride + k], zd = Zbuf[h][(16 + r) * stride + k];
                T y, dy, ddy;
                activate2<T>(s.act[h], z, y, dy, ddy);
                const T ab = P[cur][r * stride + k], adb = P[cur][(16 + r) * stride + k];
                const bool pad = k >= s.w[h];
                P[cur][r * stride + k] = pad ? T(0) : ab * dy + adb * ddy * zd;
                P[cur][(16 + r) * stride + k] = pad ? T(0) : adb * dy;
            }
            wave_fence();
            // dW_h += A_h^T [zbar; zdotbar]  (K = 32 rows)
            const int fan_in = h > 0 ? s.w[h - 1] : s.in_dim;
            for (int i0 = 0; i0 < kin_h; i0 += 16) {
                for (int o0 = 0; o0 < wp; o0 += 16) {
                    T c[4];
                    wgrad_tile<MODE>(Abuf[h], P[cur], stride, i0, o0, c);
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const int in = i0 + crow<MODE>(grp, i), out = o0 + row;
                        if (in < fan_in && out < s.w[h]) gacc[s.pW[h] + (size_t)in * s.w[h] + out] += c[i];
                    }
                }
            }
            for (int k = l; k < s.w[h]; k += kWave) {
                T gb = T(0);
                for (int r = 0; r < 16; ++r) gb += P[cur][r * stride + k];
                gacc[s.pb[h] + k] += gb;
            }
            if (h > 0) {
                // [abar; adotbar] of layer h's input = [zbar; zdotbar] W_h^T  -> P[cur ^ 1]
                for (int n0 = 0; n0 < kin_h; n0 += 16) {
                    T c0[4] = {T(0), T(0), T(0), T(0)}, c1[4] = {T(0), T(0), T(0), T(0)};
                    if (MODE == PT_MLP_F64) {
                        gemm_tile<MODE>(P[cur], stride, wp, a.packed, s.W64[h], 0, wp, true, n0, c0);
                        gemm_tile<MODE>(P[cur] + 16 * stride, stride, wp, a.packed, s.W64[h], 0, wp, true, n0, c1);
                    } else {
                        gemm_tile<MODE>(P[cur], stride, wp, a.packed, s.Wh[h], s.Wl[h], wp, true, n0, c0);
                        gemm_tile<MODE>(P[cur] + 16 * stride, stride, wp, a.packed, s.Wh[h], s.Wl[h], wp, true, n0, c1);
                    }
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const int r = crow<MODE>(grp, i);
                        P[cur ^ 1][r * stride + n0 + row] = c0[i];
                        P[cur ^ 1][(16 + r) * stride + n0 + row] = c1[i];
                    }
                }
                cur ^= 1;
            }
            wave_fence();
        }
    }
    wave_fence();
    for (int k = l; k < a.nparam; k += kWave) {
        const double v = (double)gacc[k];
        if (v != 0.0) atomicAdd(&a.grad[k], v);
    }
}

template <int MODE>
int launch_train(TrainArgs a, hipStream_t st) {
    typedef typename MT<MODE>::T T;
    const int H = a.s.n_hidden;
    a.stride = a.s.wmax + (sizeof(T) == 8 ? 1 : 4);
    const size_t per_wave = sizeof(T) * ((size_t)(2 * H + 3) * 32 * a.stride + 64 + a.nparam);
    const size_t budget = lds_limit();
    if (per_wave > budget) {
        set_error("mlp training: %zu bytes of LDS per wavefront exceed the device limit", per_wave);
        return PT_ERR_UNSUPPORTED;
    }
    const int wpb = (int)std::min<size_t>(4, budget / per_wave);
    const int ntiles = (a.n_atoms + 15) / 16;
    const int blocks = std::max(1, std::min((ntiles + wpb - 1) / wpb, 1024));
    hipLaunchKernelGGL(k_mlp_train<MODE>, dim3(blocks), dim3(64 * wpb), wpb * per_wave, st, a);
    return launch_check("k_mlp_train");
}

// ---- per-row Jacobian of the input-gradient projection (FP64) -----------------
//
// Row (atom r, direction d) of J is  sign * d(ydot_d)/dtheta,  with
// ydot_d = grad_x' MLP . (T_d * mul) the MLP output's derivative along the
// scaled tangent of direction d.  The forward-over-reverse sweep is the one of
// k_mlp_train (primal rows 0..15 once per tile, tangent rows 16..31 once per
// direction), but nothing is summed over atoms: a dense layer's block of a
// row is the outer product  a[r] (x) zbar[r] + adot[r] (x) zdotbar[r]  and a
// bias block is zbar[r], written straight to HBM with consecutive lanes on
// consecutive parameters of one row.
struct JacArgs {
    MlpShape s;
    int n_atoms;
    const double* G;
    const double* T;      // T[d * t_dir + atom * ldg + k]
    long long t_dir;
    int n_dir;
    int ldg;
    const double* scale;
    double scale_add;
    const char* packed;
    double sign;
    double* J;            // row (atom * n_dir + d), unpacked parameter layout
    long long ldj;
    int stride;
};

__global__ __launch_bounds__(256) void k_mlp_jacobian(JacArgs a) {
    typedef double T;
    constexpr int MODE = PT_MLP_F64;
    extern __shared__ __align__(16) char smem[];
    const MlpShape& s = a.s;
    const int H = s.n_hidden;
    const int l = wave_lane();
    const int wv = threadIdx.x >> 6;
    const int wpb = blockDim.x >> 6;
    const int stride = a.stride;
    const size_t tile = (size_t)32 * stride;
    const size_t per_wave = (size_t)(2 * H + 3) * tile + 64;
    T* base = reinterpret_cast<T*>(smem) + (size_t)wv * per_wave;
    T* Abuf[kMaxHidden + 1];
    T* Zbuf[kMaxHidden];
    for (int h = 0; h <= H; ++h) Abuf[h] = base + h * tile;
    for (int h = 0; h < H; ++h) Zbuf[h] = base + (H + 1 + h) * tile;
    T* P[2] = {base + (2 * H + 1) * tile, base + (2 * H + 2) * tile};
    T* outv = base + (2 * H + 3) * tile;  // [0,16): z_L, [32,48): zbar_L, [48,64): zdotbar_L
    const int row = l & 15, grp = l >> 4;
    const T* wout = (const T*)(a.packed + s.wout64);
    const T bout = ((const double*)(a.packed + s.bout64))[0];
    const int L = H;
    const int fan_L = H > 0 ? s.w[H - 1] : s.in_dim;
    const int ntiles = (a.n_atoms + 15) / 16;
    // one work item = (16-atom tile, direction): the primal sweep is redone per direction
    for (int it = blockIdx.x * wpb + wv; it < ntiles * a.n_dir; it += gridDim.x * wpb) {
        const int a0 = (it / a.n_dir) * 16;
        const int d = it % a.n_dir;
        const int nrows = min(16, a.n_atoms - a0);
        for (int e = l; e < 16 * s.in_p; e += kWave) {
            const int r = e / s.in_p, k = e % s.in_p;
            T v = T(0);
            if (r < nrows && k < s.in_dim)
                v = (a.G[(size_t)(a0 + r) * a.ldg + k] - a.scale[k]) * a.scale[s.in_dim + k] + a.scale_add;
            Abuf[0][r * stride + k] = v;
        }
        wave_fence();
        // primal forward, rows 0..15 (kept for every direction)
        int kin = s.in_p;
        for (int h = 0; h < H; ++h) {
            const int wp = s.wp[h];
            for (int n0 = 0; n0 < wp; n0 += 16) {
                T az[4] = {T(0), T(0), T(0), T(0)};
                gemm_tile<MODE>(Abuf[h], stride, kin, a.packed, s.W64[h], 0, wp, false, n0, az);
                const int col = n0 + row;
                const T bias = ((const double*)(a.packed + s.b64[h]))[col];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int r = crow<MODE>(grp, i);
                    T y, dy;
                    const T z = az[i] + bias;
                    activate<T>(s.act[h], z, y, dy);
                    const bool pad = col >= s.w[h];
                    Zbuf[h][r * stride + col] = pad ? T(0) : z;
                    Abuf[h + 1][r * stride + col] = pad ? T(0) : y;
                }
            }
            wave_fence();
            kin = wp;
        }
        {
            const int r = l >> 2, part = l & 3;
            T z = T(0);
            for (int k = part; k < kin; k += 4) z += Abuf[H][r * stride + k] * wout[k];
            z += __shfl_xor(z, 1, kWave);
            z += __shfl_xor(z, 2, kWave);
            if (part == 0) outv[r] = z + bout;
        }
        wave_fence();
        {
            // tangent input, rows 16..31
            for (int e = l; e < 16 * s.in_p; e += kWave) {
                const int r = e / s.in_p, k = e % s.in_p;
                T t = T(0);
                if (r < nrows && k < s.in_dim)
                    t = a.T[(size_t)d * a.t_dir + (size_t)(a0 + r) * a.ldg + k] * a.scale[s.in_dim + k];
                Abuf[0][(16 + r) * stride + k] = t;
            }
            wave_fence();
            kin = s.in_p;
            for (int h = 0; h < H; ++h) {
                const int wp = s.wp[h];
                for (int n0 = 0; n0 < wp; n0 += 16) {
                    T at[4] = {T(0), T(0), T(0), T(0)};
                    gemm_tile<MODE>(Abuf[h] + 16 * stride, stride, kin, a.packed, s.W64[h], 0, wp, false, n0, at);
                    const int col = n0 + row;
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const int r = crow<MODE>(grp, i);
                        T y, dy;
                        activate<T>(s.act[h], Zbuf[h][r * stride + col], y, dy);
                        const bool pad = col >= s.w[h];
                        Zbuf[h][(16 + r) * stride + col] = pad ? T(0) : at[i];
                        Abuf[h + 1][(16 + r) * stride + col] = pad ? T(0) : dy * at[i];
                    }
                }
                wave_fence();
                kin = wp;
            }
            // output layer: seeds of the projection's reverse sweep
            {
                const int r = l >> 2, part = l & 3;
                T zd = T(0);
                for (int k = part; k < kin; k += 4) zd += Abuf[H][(16 + r) * stride + k] * wout[k];
                zd += __shfl_xor(zd, 1, kWave);
                zd += __shfl_xor(zd, 2, kWave);
                T y, dy, ddy;
                activate2<T>(s.act_out, outv[r], y, dy, ddy);
                const T gt = r < nrows ? (T)a.sign : T(0);
                if (part == 0) {
                    outv[32 + r] = gt * ddy * zd;  // zbar_L
                    outv[48 + r] = gt * dy;        // zdotbar_L
                }
            }
            wave_fence();
            // rows of J: output layer
            for (int e = l; e < 16 * fan_L; e += kWave) {
                const int r = e / fan_L, k = e % fan_L;
                if (r < nrows)
                    a.J[(size_t)((a0 + r) * a.n_dir + d) * a.ldj + s.pW[L] + k] =
                        Abuf[H][r * stride + k] * outv[32 + r] + Abuf[H][(16 + r) * stride + k] * outv[48 + r];
            }
            if (l < nrows) a.J[(size_t)((a0 + l) * a.n_dir + d) * a.ldj + s.pb[L]] = outv[32 + l];
            for (int e = l; e < 16 * kin; e += kWave) {
                const int r = e / kin, k = e % kin;
                P[0][r * stride + k] = outv[32 + r] * wout[k];
                P[0][(16 + r) * stride + k] = outv[48 + r] * wout[k];
            }
            wave_fence();
            int cur = 0;
            for (int h = H - 1; h >= 0; --h) {
                const int wp = s.wp[h];
                const int kin_h = h > 0 ? s.wp[h - 1] : s.in_p;
                for (int e = l; e < 16 * wp; e += kWave) {
                    const int r = e / wp, k = e % wp;
                    const T z = Zbuf[h][r * stride + k], zd = Zbuf[h][(16 + r) * stride + k];
                    T y, dy, ddy;
                    activate2<T>(s.act[h], z, y, dy, ddy);
                    const T ab = P[cur][r * stride + k], adb = P[cur][(16 + r) * stride + k];
                    const bool pad = k >= s.w[h];
                    P[cur][r * stride + k] = pad ? T(0) : ab * dy + adb * ddy * zd;
                    P[cur][(16 + r) * stride + k] = pad ? T(0) : adb * dy;
                }
                wave_fence();
                // rows of J: W_h[in][out] and b_h[out]
                const int fan_in = h > 0 ? s.w[h - 1] : s.in_dim;
                const int wo = s.w[h];
                const int nW = fan_in * wo;
                for (int e = l; e < 16 * nW; e += kWave) {
                    const int r = e / nW, q = e % nW;
                    const int in = q / wo, out = q % wo;
                    if (r < nrows)
                        a.J[(size_t)((a0 + r) * a.n_dir + d) * a.ldj + s.pW[h] + q] =
                            Abuf[h][r * stride + in] * P[cur][r * stride + out] +
                            Abuf[h][(16 + r) * stride + in] * P[cur][(16 + r) * stride + out];
                }
                for (int e = l; e < 16 * wo; e += kWave) {
                    const int r = e / wo, out = e % wo;
                    if (r < nrows) a.J[(size_t)((a0 + r) * a.n_dir + d) * a.ldj + s.pb[h] + out] = P[cur][r * stride + out];
                }
                if (h > 0) {
                    for (int n0 = 0; n0 < kin_h; n0 += 16) {
                        T c0[4] = {T(0), T(0), T(0), T(0)}, c1[4] = {T(0), T(0), T(0), T(0)};
                        gemm_tile<MODE>(P[cur], stride, wp, a.packed, s.W64[h], 0, wp, true, n0, c0);
                        gemm_tile<MODE>(P[cur] + 16 * stride, stride, wp, a.packed, s.W64[h], 0, wp, true, n0, c1);
#pragma unroll
                        for (int i = 0; i < 4; ++i) {
                            const int r = crow<MODE>(grp, i);
                            P[cur ^ 1][r * stride + n0 + row] = c0[i];
                            P[cur ^ 1][(16 + r) * stride + n0 + row] = c1[i];
                        }
                    }
                    cur ^= 1;
                }
                wave_fence();
            }
        }
    }
}

// ---- loss seeds and Adam --------------------------------------------------------

__global__ void k_train_seeds(int n, const double* __restrict__ f, const double* __restrict__ fref,
                              const double* __restrict__ energy, double eref, double we, double wf,
                              double* __restrict__ delta, double* __restrict__ gy, double* __restrict__ loss) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const double de = energy[0] - eref;
    double s = 0.0;
    if (i < n) {
        for (int d = 0; d < 3; ++d) {
            const double r = f[3 * i + d] - fref[3 * i + d];
            delta[3 * i + d] = wf * 2.0 * r / (3.0 * n);
            s += r * r;
        }
        gy[i] = we * 2.0 * de / ((double)n * n);
    }
    s = wave_sum(s);  // all lanes of the (full) block take part
    if (wave_lane() == 0 && s != 0.0) atomicAdd(&loss[1], wf * s / (3.0 * n));
    if (i == 0) atomicAdd(&loss[0], we * de * de / ((double)n * n));
}

// Adam (Kingma & Ba 2015) with optional decoupled weight decay (AdamW).
__global__ void k_adam(int n, double* __restrict__ p, const double* __restrict__ g, double* __restrict__ m,
                       double* __restrict__ v, double lr, double b1, double b2, double eps, double wd,
                       double c1, double c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double gi = g[i];
    const double mi = b1 * m[i] + (1.0 - b1) * gi;
    const double vi = b2 * v[i] + (1.0 - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const double mh = mi / c1, vh = vi / c2;
    p[i] = p[i] - lr * (mh / (sqrt(vh) + eps) + wd * p[i]);
}

template <int MODE>
int launch_mlp(MlpArgs a, hipStream_t st) {
    typedef typename MT<MODE>::T T;
    a.stride = a.s.wmax + (sizeof(T) == 8 ? 1 : 4);
    a.inplace = 1;
    for (int h = 0; h < a.s.n_hidden; ++h) a.inplace = a.inplace && act_from_output(a.s.act[h]);
    const int ny = a.s.n_hidden > 2 ? a.s.n_hidden : 2;
    a.nbuf = ny + (a.inplace ? 0 : a.s.n_hidden);
    // nbuf activation buffers of 16 rows plus 16 output derivatives
    const size_t per_wave = sizeof(T) * ((size_t)a.nbuf * 16 * a.stride + 16);
    const size_t budget = lds_limit();
    if (per_wave > budget) {
        set_error("mlp: layer widths need %zu bytes of LDS per wavefront", per_wave);
        return PT_ERR_UNSUPPORTED;
    }
    const int wpb = (int)std::min<size_t>(4, budget / per_wave);
    const int ntiles = (a.n_atoms + 15) / 16;
    const int blocks = std::max(1, std::min((ntiles + wpb - 1) / wpb, 8192));
    hipLaunchKernelGGL(k_mlp<MODE>, dim3(blocks), dim3(64 * wpb), wpb * per_wave, st, a);
    return launch_check("k_mlp");
}

}  // namespace
}  // namespace pt

using namespace pt;

extern "C" size_t pt_mlp_packed_bytes(const pt_mlp_desc* desc) {
    MlpShape s;
    if (make_shape(desc, s) != PT_OK) return 0;
    return s.total;
}

extern "C" int pt_mlp_pack(const pt_mlp_desc* desc, const double* params, void* packed,
                           pt_stream_t stream) {
    PT_REQUIRE(desc && params && packed, "pt_mlp_pack: null argument");
    MlpShape s;
    if (int rc = make_shape(desc, s)) return rc;
    int work = s.kout;
    int kin = s.in_p;
    for (int l = 0; l < s.n_hidden; ++l) { work = std::max(work, kin * s.wp[l]); kin = s.wp[l]; }
    hipLaunchKernelGGL(k_pack, dim3((work + 255) / 256), dim3(256), 0, (hipStream_t)stream, s, params,
                       (char*)packed);
    return launch_check("pt_mlp_pack");
}

extern "C" int pt_mlp_energy_grad(int mode, int n_atoms, const void* G, int ldg, int g_precision,
                                  const double* scale, const pt_mlp_desc* desc, const void* packed,
                                  double* e_atom, void* dEdG, pt_stream_t stream) {
    PT_REQUIRE(n_atoms >= 0, "pt_mlp_energy_grad: n_atoms < 0");
    if (n_atoms == 0) return PT_OK;
    PT_REQUIRE(G && scale && desc && packed && e_atom, "pt_mlp_energy_grad: null argument");
    PT_REQUIRE(g_precision == PT_F64 || g_precision == PT_F32, "pt_mlp_energy_grad: bad precision");
    MlpArgs a{};
    if (int rc = make_shape(desc, a.s)) return rc;
    PT_REQUIRE(ldg >= desc->in_dim, "pt_mlp_energy_grad: ldg < in_dim");
    a.n_atoms = n_atoms;
    a.G = G;
    a.ldg = ldg;
    a.g_f64 = g_precision == PT_F64;
    a.scale = scale;
    a.scale_add = desc->scale_add;
    a.packed = (const char*)packed;
    a.e_atom = e_atom;
    a.dEdG = dEdG;
    hipStream_t st = (hipStream_t)stream;
    switch (mode) {
    case PT_MLP_F64: return launch_mlp<PT_MLP_F64>(a, st);
    case PT_MLP_BF16: return launch_mlp<PT_MLP_BF16>(a, st);
    case PT_MLP_BF16X3: return launch_mlp<PT_MLP_BF16X3>(a, st);
    }
    set_error("pt_mlp_energy_grad: unknown mode %d", mode);
    return PT_ERR_ARG;
}

extern "C" size_t pt_mlp_param_count(const pt_mlp_desc* desc) {
    MlpShape s;
    if (make_shape(desc, s) != PT_OK) return 0;
    return s.pb[s.n_hidden] + 1;
}

extern "C" int pt_mlp_train_grad(int mode, int n_atoms, const void* G, const void* T, int ldg,
                                 int g_precision, const double* scale, const pt_mlp_desc* desc,
                                 const void* packed, const double* gy, double tseed, double* grad,
                                 pt_stream_t stream) {
    PT_REQUIRE(n_atoms >= 0, "pt_mlp_train_grad: n_atoms < 0");
    if (n_atoms == 0) return PT_OK;
    PT_REQUIRE(G && T && scale && desc && packed && gy && grad, "pt_mlp_train_grad: null argument");
    PT_REQUIRE(g_precision == PT_F64 || g_precision == PT_F32, "pt_mlp_train_grad: bad precision");
    TrainArgs a{};
    if (int rc = make_shape(desc, a.s)) return rc;
    PT_REQUIRE(ldg >= desc->in_dim, "pt_mlp_train_grad: ldg < in_dim");
    a.n_atoms = n_atoms;
    a.G = G;
    a.T = T;
    a.ldg = ldg;
    a.g_f64 = g_precision == PT_F64;
    a.scale = scale;
    a.scale_add = desc->scale_add;
    a.packed = (const char*)packed;
    a.gy = gy;
    a.tseed = tseed;
    a.grad = grad;
    a.nparam = (int)(a.s.pb[a.s.n_hidden] + 1);
    hipStream_t st = (hipStream_t)stream;
    switch (mode) {
    case PT_MLP_F64: return launch_train<PT_MLP_F64>(a, st);
    case PT_MLP_BF16: return launch_train<PT_MLP_BF16>(a, st);
    case PT_MLP_BF16X3: return launch_train<PT_MLP_BF16X3>(a, st);
    }
    set_error("pt_mlp_train_grad: unknown mode %d", mode);
    return PT_ERR_ARG;
}

extern "C" int pt_mlp_jacobian(int n_atoms, int n_dir, const double* G, const double* T, long long t_dir, int ldg,
                               const double* scale, const pt_mlp_desc* desc, const void* packed, double sign,
                               double* J, long long ldj, pt_stream_t stream) {
    PT_REQUIRE(n_atoms >= 0 && n_dir >= 1, "pt_mlp_jacobian: n_atoms < 0 or n_dir < 1");
    if (n_atoms == 0) return PT_OK;
    PT_REQUIRE(G && T && scale && desc && packed && J, "pt_mlp_jacobian: null argument");
    JacArgs a{};
    if (int rc = make_shape(desc, a.s)) return rc;
    PT_REQUIRE(ldg >= desc->in_dim, "pt_mlp_jacobian: ldg < in_dim");
    PT_REQUIRE(t_dir >= (long long)(n_atoms - 1) * ldg + desc->in_dim || n_dir == 1,
               "pt_mlp_jacobian: tangent directions overlap");
    const long long np = (long long)(a.s.pb[a.s.n_hidden] + 1);
    PT_REQUIRE(ldj >= np, "pt_mlp_jacobian: ldj below the parameter count");
    PT_REQUIRE((long long)n_atoms * n_dir < (1LL << 31), "pt_mlp_jacobian: too many rows");
    a.n_atoms = n_atoms;
    a.n_dir = n_dir;
    a.G = G;
    a.T = T;
    a.t_dir = t_dir;
    a.ldg = ldg;
    a.scale = scale;
    a.scale_add = desc->scale_add;
    a.packed = (const char*)packed;
    a.sign = sign;
    a.J = J;
    a.ldj = ldj;
    a.stride = a.s.wmax + 1;
    const size_t per_wave = sizeof(double) * ((size_t)(2 * a.s.n_hidden + 3) * 32 * a.stride + 64);
    const size_t budget = lds_limit();
    if (per_wave > budget) {
        set_error("pt_mlp_jacobian: %zu bytes of LDS per wavefront exceed the device limit", per_wave);
        return PT_ERR_UNSUPPORTED;
    }
    const int wpb = (int)std::min<size_t>(4, budget / per_wave);
    const long long items = (long long)((n_atoms + 15) / 16) * n_dir;
    const int blocks = (int)std::max<long long>(1, std::min<long long>((items + wpb - 1) / wpb, 8192));
    hipLaunchKernelGGL(k_mlp_jacobian, dim3(blocks), dim3(64 * wpb), wpb * per_wave, (hipStream_t)stream, a);
    return launch_check("k_mlp_jacobian");
}

extern "C" int pt_train_seeds(int n, const double* force, const double* force_ref, const double* energy,
                              double energy_ref, double w_energy, double w_force, double* delta, double* gy,
                              double* loss, pt_stream_t stream) {
    PT_REQUIRE(n > 0 && force && force_ref && energy && delta && gy && loss, "pt_train_seeds: bad arguments");
    hipLaunchKernelGGL(k_train_seeds, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, n, force,
                       force_ref, energy, energy_ref, w_energy, w_force, delta, gy, loss);
    return launch_check("pt_train_seeds");
}

extern "C" int pt_adam_step(int n, double* params, const double* grad, double* m, double* v, double lr,
                            double beta1, double beta2, double eps, double weight_decay, int step,
                            pt_stream_t stream) {
    PT_REQUIRE(n >= 0 && params && grad && m && v && step >= 1, "pt_adam_step: bad arguments");
    if (n == 0) return PT_OK;
    const double c1 = 1.0 - pow(beta1, step), c2 = 1.0 - pow(beta2, step);
    hipLaunchKernelGGL(k_adam, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, n, params, grad, m, v,
                       lr, beta1, beta2, eps, weight_decay, c1, c2);
    return launch_check("pt_adam_step");
}
