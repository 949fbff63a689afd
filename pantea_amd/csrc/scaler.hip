// Descriptor scaler statistics and the extrapolation counter on the device
// (descriptors/scaler.py:84-142 fit / partial_fit / check_warnings,
// :249-295 _fit, _partial_fit, _calculate_number_of_warnings).
//
// Statistics are per column of the descriptor rows G (n_rows x n_cols,
// leading dimension ldg, FP64 or FP32): mean, population standard deviation
// (ddof 0, as jnp.std), minimum and maximum.  Two passes (sums, then squared
// deviations from the mean), each as per-workgroup partials over fixed row
// ranges reduced in block order: the result does not depend on scheduling.
// The merge with earlier statistics follows _partial_fit term by term.
//
// Layout: params (device, 4 x n_cols FP64) = mean | sigma | minval | maxval,
// the field order of ScalerParams.
#include <cmath>

#include "common.h"

namespace pt {
namespace {

constexpr int kColTile = kWave;        // a wave covers 64 columns (lane = column)
constexpr int kStatWaves = 4;          // waves per workgroup, interleaved over rows
constexpr int kRowsPerBlock = 2048;    // rows per workgroup of a partial pass

template <typename R>
__device__ __forceinline__ double load_g(const void* G, size_t i) {
    return (double)((const R*)G)[i];
}

// min / max that propagate NaN (jnp.min / jnp.max do)
__device__ __forceinline__ double nan_min(double a, double b) { return (a != a || b != b) ? NAN : (a < b ? a : b); }
__device__ __forceinline__ double nan_max(double a, double b) { return (a != a || b != b) ? NAN : (a > b ? a : b); }

// PASS 0: column sums, minima and maxima; PASS 1: sums of (x - mean)^2.
// part: [block][3][n_cols] (PASS 0) or [block][n_cols] (PASS 1).
template <typename R, int PASS>
__global__ __launch_bounds__(256) void k_stat_partial(int n_rows, int n_cols, const void* __restrict__ G,
                                                      int ldg, const double* __restrict__ mean,
                                                      double* __restrict__ part) {
    const int lane = wave_lane();
    const int w = threadIdx.x >> 6;
    const int col = blockIdx.y * kColTile + lane;
    const int r0 = blockIdx.x * kRowsPerBlock;
    const int r1 = min(n_rows, r0 + kRowsPerBlock);
    double s = 0.0, lo = INFINITY, hi = -INFINITY;
    if (col < n_cols) {
        const double mu = PASS == 1 ? mean[col] : 0.0;
        for (int r = r0 + w; r < r1; r += kStatWaves) {
            const double x = load_g<R>(G, (size_t)r * ldg + col);
            if (PASS == 0) {
                s += x;
                lo = nan_min(lo, x);
                hi = nan_max(hi, x);
            } else {
                const double d = x - mu;
                s += d * d;
            }
        }
    }
    __shared__ double sh[3][kStatWaves][kColTile];
    sh[0][w][lane] = s;
    sh[1][w][lane] = lo;
    sh[2][w][lane] = hi;
    __syncthreads();
    if (w == 0 && col < n_cols) {
        double ss = 0.0, l = INFINITY, h = -INFINITY;
        for (int q = 0; q < kStatWaves; ++q) {
            ss += sh[0][q][lane];
            l = nan_min(l, sh[1][q][lane]);
            h = nan_max(h, sh[2][q][lane]);
        }
        if (PASS == 0) {
            double* p = part + (size_t)blockIdx.x * 3 * n_cols;
            p[col] = ss;
            p[n_cols + col] = l;
            p[2 * n_cols + col] = h;
        } else {
            part[(size_t)blockIdx.x * n_cols + col] = ss;
        }
    }
}

// Block partials -> mean, min, max (tmp: 3 x n_cols).
__global__ __launch_bounds__(64) void k_stat_moments(int nb, int n_rows, int n_cols, const double* __restrict__ part,
                                                     double* __restrict__ tmp) {
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= n_cols) return;
    double s = 0.0, l = INFINITY, h = -INFINITY;
    for (int b = 0; b < nb; ++b) {
        const double* p = part + (size_t)b * 3 * n_cols;
        s += p[col];
        l = nan_min(l, p[n_cols + col]);
        h = nan_max(h, p[2 * n_cols + col]);
    }
    tmp[col] = s / (double)n_rows;
    tmp[n_cols + col] = l;
    tmp[2 * n_cols + col] = h;
}

// Squared-deviation partials -> sigma, then the fit (nsamples == 0) or the
// merge of _partial_fit (scaler.py:262-283) into params.
__global__ __launch_bounds__(64) void k_stat_merge(int nb, int n_rows, int n_cols, const double* __restrict__ part,
                                                   const double* __restrict__ tmp, double nsamples,
                                                   double* __restrict__ params) {
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= n_cols) return;
    double ss = 0.0;
    for (int b = 0; b < nb; ++b) ss += part[(size_t)b * n_cols + col];
    const double mean = tmp[col], mn = tmp[n_cols + col], mx = tmp[2 * n_cols + col];
    const double sigma = sqrt(ss / (double)n_rows);
    double* p_mean = params;
    double* p_sigma = params + n_cols;
    double* p_min = params + 2 * n_cols;
    double* p_max = params + 3 * n_cols;
    if (!(nsamples > 0.0)) {
        p_mean[col] = mean;
        p_sigma[col] = sigma;
        p_min[col] = mn;
        p_max[col] = mx;
        return;
    }
    const double m = nsamples, n = (double)n_rows;
    const double fm = m / (m + n), fn = n / (m + n);
    const double pm = p_mean[col], ps = p_sigma[col];
    const double diff = pm - mean;
    p_mean[col] = fm * pm + fn * mean;
    p_sigma[col] = sqrt((fm * ps) * ps + (fn * sigma) * sigma + (fm * diff) * (fn * diff));
    p_min[col] = nan_min(p_min[col], mn);
    p_max[col] = nan_max(p_max[col], mx);
}

__device__ __forceinline__ double wave_max_nonneg(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}

// One wave per row (grid-stride), lanes over columns: a value outside
// [min - margin (max - min), max + margin (max - min)] (NaN included) is out of
// range.  counters[0] += rows with an out-of-range value, counters[1] +=
// out-of-range values, counters[2] = max(counters[2], bits of the largest
// excess in units of the column's range (non-negative doubles order like
// their bit patterns)).
template <typename R>
__global__ __launch_bounds__(256) void k_extrapolation(int n_rows, int n_cols, const void* __restrict__ G, int ldg,
                                                       const double* __restrict__ params, double margin,
                                                       unsigned long long* __restrict__ counters) {
    const int lane = wave_lane();
    const int wpb = blockDim.x >> 6;
    const double* p_min = params + 2 * n_cols;
    const double* p_max = params + 3 * n_cols;
    double rows_out = 0.0, vals_out = 0.0, worst = 0.0;
    for (int r = blockIdx.x * wpb + (threadIdx.x >> 6); r < n_rows; r += gridDim.x * wpb) {
        bool any = false;
        for (int c = lane; c < n_cols; c += kWave) {
            const double x = load_g<R>(G, (size_t)r * ldg + c);
            const double mn = p_min[c], mx = p_max[c];
            const double wd = mx - mn;
            const double lo = mn - margin * wd, hi = mx + margin * wd;
            if (!(x >= lo && x <= hi)) {
                any = true;
                vals_out += 1.0;
                const double ex = x > hi ? (x - mx) : (x < lo ? (mn - x) : INFINITY);
                worst = fmax(worst, wd > 0.0 ? ex / wd : ex);
            }
        }
        if (__ballot(any) != 0ull) rows_out += 1.0;
    }
    const double v = wave_sum(vals_out);   // exact: integer counts far below 2^53
    worst = wave_max_nonneg(worst);
    if (lane == 0) {
        if (rows_out > 0.0) atomicAdd(&counters[0], (unsigned long long)rows_out);
        if (v > 0.0) atomicAdd(&counters[1], (unsigned long long)v);
        if (worst > 0.0) atomicMax(&counters[2], (unsigned long long)__builtin_bit_cast(unsigned long long, worst));
    }
}

size_t stat_blocks(int n_rows) { return (size_t)((n_rows + kRowsPerBlock - 1) / kRowsPerBlock); }

}  // namespace
}  // namespace pt

using namespace pt;

extern "C" size_t pt_scaler_workspace_bytes(int n_rows, int n_cols) {
    if (n_rows < 0 || n_cols < 0) return 0;
    const size_t nb = stat_blocks(n_rows > 0 ? n_rows : 1);
    return sizeof(double) * (3 * nb * (size_t)n_cols + 3 * (size_t)n_cols);
}

extern "C" int pt_scaler_partial_fit(int n_rows, int n_cols, const void* G, int ldg, int g_precision,
                                     double nsamples, double* params, void* ws, size_t ws_bytes,
                                     pt_stream_t stream_) {
    hipStream_t stream = (hipStream_t)stream_;
    PT_REQUIRE(n_rows >= 1 && n_cols >= 1 && ldg >= n_cols && G && params && ws,
               "pt_scaler_partial_fit: bad arguments");
    PT_REQUIRE(g_precision == PT_F64 || g_precision == PT_F32, "pt_scaler_partial_fit: unknown precision");
    PT_REQUIRE(nsamples >= 0.0, "pt_scaler_partial_fit: nsamples must be >= 0");
    PT_REQUIRE(ws_bytes >= pt_scaler_workspace_bytes(n_rows, n_cols), "pt_scaler_partial_fit: workspace too small");
    const int nb = (int)stat_blocks(n_rows);
    double* part = (double*)ws;
    double* tmp = part + 3 * (size_t)nb * n_cols;
    const dim3 grid(nb, (n_cols + kColTile - 1) / kColTile);
    const dim3 fin((n_cols + 63) / 64);
    if (g_precision == PT_F64) {
        hipLaunchKernelGGL((k_stat_partial<double, 0>), grid, dim3(256), 0, stream, n_rows, n_cols, G, ldg,
                           (const double*)nullptr, part);
    } else {
        hipLaunchKernelGGL((k_stat_partial<float, 0>), grid, dim3(256), 0, stream, n_rows, n_cols, G, ldg,
                           (const double*)nullptr, part);
    }
    hipLaunchKernelGGL(k_stat_moments, fin, dim3(64), 0, stream, nb, n_rows, n_cols, (const double*)part, tmp);
    if (g_precision == PT_F64) {
        hipLaunchKernelGGL((k_stat_partial<double, 1>), grid, dim3(256), 0, stream, n_rows, n_cols, G, ldg,
                           (const double*)tmp, part);
    } else {
        hipLaunchKernelGGL((k_stat_partial<float, 1>), grid, dim3(256), 0, stream, n_rows, n_cols, G, ldg,
                           (const double*)tmp, part);
    }
    hipLaunchKernelGGL(k_stat_merge, fin, dim3(64), 0, stream, nb, n_rows, n_cols, (const double*)part,
                       (const double*)tmp, nsamples, params);
    return launch_check("pt_scaler_partial_fit");
}

extern "C" int pt_scaler_extrapolation(int n_rows, int n_cols, const void* G, int ldg, int g_precision,
                                       const double* params, double margin, unsigned long long* counters,
                                       pt_stream_t stream_) {
    hipStream_t stream = (hipStream_t)stream_;
    PT_REQUIRE(n_rows >= 0 && n_cols >= 1 && ldg >= n_cols && params && counters && (n_rows == 0 || G),
               "pt_scaler_extrapolation: bad arguments");
    PT_REQUIRE(g_precision == PT_F64 || g_precision == PT_F32, "pt_scaler_extrapolation: unknown precision");
    PT_REQUIRE(margin >= 0.0 && std::isfinite(margin), "pt_scaler_extrapolation: margin must be finite and >= 0");
    if (n_rows == 0) return PT_OK;
    const int blocks = (int)std::min<long long>(((long long)n_rows + 3) / 4, 4096);
    if (g_precision == PT_F64) {
        hipLaunchKernelGGL(k_extrapolation<double>, dim3(blocks), dim3(256), 0, stream, n_rows, n_cols, G, ldg,
                           params, margin, counters);
    } else {
        hipLaunchKernelGGL(k_extrapolation<float>, dim3(blocks), dim3(256), 0, stream, n_rows, n_cols, G, ldg,
                           params, margin, counters);
    }
    return launch_check("pt_scaler_extrapolation");
}
