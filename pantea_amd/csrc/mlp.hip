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

template <int MODE>
__global__ __launch_bounds__(256) void k_mlp(MlpArgs a) {
    typedef typename MT<MODE>::T T;
    extern __shared__ __align__(16) char smem[];
    const MlpShape& s = a.s;
    const int l = wave_lane();
    const int wv = threadIdx.x >> 6;
    const int wpb = blockDim.x >> 6;
    const int stride = a.stride;
    T* base = reinterpret_cast<T*>(smem) + (size_t)wv * ((size_t)a.nbuf * 16 * stride + 16);
    T* X[2] = {base, base + 16 * stride};         // ping-pong activations
    T* D[kMaxHidden];                              // activation derivatives
    for (int h = 0; h < kMaxHidden; ++h) D[h] = base + (2 + h) * 16 * stride;
    T* gout = base + (2 + s.n_hidden) * 16 * stride;  // 16 output derivatives

    const int row = l & 15, grp = l >> 4;
    const int ntiles = (a.n_atoms + 15) / 16;
    for (int tile = blockIdx.x * wpb + wv; tile < ntiles; tile += gridDim.x * wpb) {
        const int a0 = tile * 16;
        // 1. scaled input -> X[0] (zero padded)
        for (int e = l; e < 16 * s.in_p; e += kWave) {
            const int r = e / s.in_p, k = e % s.in_p;
            const int atom = a0 + r;
            T v = T(0);
            if (atom < a.n_atoms && k < s.in_dim) {
                const double g = a.g_f64 ? ((const double*)a.G)[(size_t)atom * a.ldg + k]
                                         : (double)((const float*)a.G)[(size_t)atom * a.ldg + k];
                v = (T)((g - a.scale[k]) * a.scale[s.in_dim + k] + a.scale_add);
            }
            X[0][r * stride + k] = v;
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();

        // 2. hidden layers
        int cur = 0, kin = s.in_p;
        for (int h = 0; h < s.n_hidden; ++h) {
            const int wp = s.wp[h];
            for (int n0 = 0; n0 < wp; n0 += 16) {
                T acc[4] = {T(0), T(0), T(0), T(0)};
                if (MODE == PT_MLP_F64)
                    gemm_tile<MODE>(X[cur], stride, kin, a.packed, s.W64[h], 0, wp, false, n0, acc);
                else
                    gemm_tile<MODE>(X[cur], stride, kin, a.packed, s.WTh[h], s.WTl[h], kin, true, n0, acc);
                const int col = n0 + row;
                const T bias = (MODE == PT_MLP_F64) ? (T)((const double*)(a.packed + s.b64[h]))[col]
                                                    : (T)((const float*)(a.packed + s.b32[h]))[col];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int r = crow<MODE>(grp, i);
                    T y, dy;
                    activate<T>(s.act[h], acc[i] + bias, y, dy);
                    if (col >= s.w[h]) { y = T(0); dy = T(0); }
                    X[cur ^ 1][r * stride + col] = y;
                    D[h][r * stride + col] = dy;
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            __builtin_amdgcn_wave_barrier();
            cur ^= 1;
            kin = wp;
        }

        // 3. output layer (width 1): 4 lanes per atom
        {
            const T* wout = (MODE == PT_MLP_F64) ? (const T*)(a.packed + s.wout64)
                                                 : (const T*)(a.packed + s.wout32);
            const T bout = (MODE == PT_MLP_F64) ? (T)((const double*)(a.packed + s.bout64))[0]
                                                : (T)((const float*)(a.packed + s.bout32))[0];
            const int r = l >> 2, part = l & 3;
            T sum = T(0);
            for (int k = part; k < kin; k += 4) sum += X[cur][r * stride + k] * wout[k];
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
        const T* wout = (MODE == PT_MLP_F64) ? (const T*)(a.packed + s.wout64)
                                             : (const T*)(a.packed + s.wout32);
        if (s.n_hidden > 0) {
            const int h = s.n_hidden - 1;
            for (int e = l; e < 16 * s.wp[h]; e += kWave) {
                const int r = e / s.wp[h], k = e % s.wp[h];
                D[h][r * stride + k] *= gout[r] * wout[k];
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            __builtin_amdgcn_wave_barrier();
            for (int hh = h; hh >= 1; --hh) {
                // dZ_{hh-1} = (dZ_hh . W_hh^T) * D[hh-1]
                const int wprev = s.wp[hh - 1];
                for (int n0 = 0; n0 < wprev; n0 += 16) {
                    T acc[4] = {T(0), T(0), T(0), T(0)};
                    if (MODE == PT_MLP_F64)
                        gemm_tile<MODE>(D[hh], stride, s.wp[hh], a.packed, s.W64[hh], 0, s.wp[hh], true, n0, acc);
                    else
                        gemm_tile<MODE>(D[hh], stride, s.wp[hh], a.packed, s.Wh[hh], s.Wl[hh], s.wp[hh], true, n0, acc);
#pragma unroll
                    for (int i = 0; i < 4; ++i) D[hh - 1][crow<MODE>(grp, i) * stride + n0 + row] *= acc[i];
                }
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                __builtin_amdgcn_wave_barrier();
            }
        }
        // 5. dE/dx' = dZ_0 . W_0^T (or wout for a linear model), then chain the scaler
        for (int n0 = 0; n0 < s.in_p; n0 += 16) {
            T acc[4] = {T(0), T(0), T(0), T(0)};
            if (s.n_hidden > 0) {
                if (MODE == PT_MLP_F64)
                    gemm_tile<MODE>(D[0], stride, s.wp[0], a.packed, s.W64[0], 0, s.wp[0], true, n0, acc);
                else
                    gemm_tile<MODE>(D[0], stride, s.wp[0], a.packed, s.Wh[0], s.Wl[0], s.wp[0], true, n0, acc);
            } else {
#pragma unroll
                for (int i = 0; i < 4; ++i) acc[i] = gout[crow<MODE>(grp, i)] * wout[n0 + row];
            }
            const int k = n0 + row;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int atom = a0 + crow<MODE>(grp, i);
                if (atom < a.n_atoms && k < s.in_dim) {
                    const double gk = (double)acc[i] * a.scale[s.in_dim + k];
                    if (a.g_f64) ((double*)a.dEdG)[(size_t)atom * a.ldg + k] = gk;
                    else ((float*)a.dEdG)[(size_t)atom * a.ldg + k] = (float)gk;
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
}

template <int MODE>
int launch_mlp(MlpArgs a, hipStream_t st) {
    typedef typename MT<MODE>::T T;
    a.stride = a.s.wmax + (sizeof(T) == 8 ? 1 : 4);
    a.nbuf = 2 + a.s.n_hidden;
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
