// Shared device/host helpers for the gfx950 NNP kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#include "pantea_hip.h"

namespace pt {

constexpr int kWave = 64;

// ---------------------------------------------------------------------------
// Error reporting (thread-local message behind pt_last_error()).
// ---------------------------------------------------------------------------
void set_error(const char* fmt, ...);

#define PT_CHECK_HIP(expr)                                                  \
    do {                                                                    \
        hipError_t _e = (expr);                                             \
        if (_e != hipSuccess) {                                             \
            ::pt::set_error("%s failed: %s (%s:%d)", #expr,                 \
                            hipGetErrorString(_e), __FILE__, __LINE__);     \
            return PT_ERR_HIP;                                              \
        }                                                                   \
    } while (0)

#define PT_REQUIRE(cond, ...)                                               \
    do {                                                                    \
        if (!(cond)) {                                                      \
            ::pt::set_error(__VA_ARGS__);                                   \
            return PT_ERR_ARG;                                              \
        }                                                                   \
    } while (0)

// Largest dynamic LDS allocation of one workgroup on the current device.
size_t lds_limit();

inline int launch_check(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("launch of %s failed: %s", what, hipGetErrorString(e));
        return PT_ERR_HIP;
    }
    return PT_OK;
}

// ---------------------------------------------------------------------------
// Minimum image and distance, bit-compatible with the FP64 definition
//   dx = a - b; if (dx > L/2) dx -= L; if (dx < -L/2) dx += L;
//   r  = (dx == 0 in all components) ? 0 : sqrt((dx0^2 + dx1^2) + dx2^2)
// (atoms/box.py:112-117, atoms/distance.py:63-78).  Contraction into FMA is
// switched off so the neighbour test r <= rc decides exactly as the FP64
// definition does.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double mic1(double dx, double L, double halfL) {
#pragma clang fp contract(off)
    dx = (dx > halfL) ? dx - L : dx;
    dx = (dx < -halfL) ? dx + L : dx;
    return dx;
}

struct BoxArg {
    double L[3];
    double half[3];
    int periodic;
};

__device__ __forceinline__ double exact_dist(double& dx, double& dy, double& dz,
                                             const BoxArg& box) {
#pragma clang fp contract(off)
    if (box.periodic) {
        dx = mic1(dx, box.L[0], box.half[0]);
        dy = mic1(dy, box.L[1], box.half[1]);
        dz = mic1(dz, box.L[2], box.half[2]);
    }
    if (dx == 0.0 && dy == 0.0 && dz == 0.0) return 0.0;
    double s = dx * dx;
    s = s + dy * dy;
    s = s + dz * dz;
    return __builtin_sqrt(s);
}

// ---------------------------------------------------------------------------
// Wave-level reductions.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

__device__ __forceinline__ int wave_lane() { return threadIdx.x & (kWave - 1); }

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Exclusive prefix count of a ballot mask below this lane.
__device__ __forceinline__ int lanes_below(unsigned long long mask) {
    unsigned lo = (unsigned)mask, hi = (unsigned)(mask >> 32);
    unsigned c = __builtin_amdgcn_mbcnt_lo(lo, 0u);
    return (int)__builtin_amdgcn_mbcnt_hi(hi, c);
}

}  // namespace pt
