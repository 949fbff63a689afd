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
// LDS of one compute unit on the current device (all resident workgroups).
size_t lds_per_cu();

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
// Sum over the 64 lanes with DPP moves (no LDS traffic): pairs (xor 1),
// quads (xor 2), half rows and rows (mirrors), then rows 0+1 into row 1 and
// 2+3 into row 3 (row_bcast:15), then all into row 3 (row_bcast:31).  The
// total ends in lane 63 and is returned to every lane.  All 64 lanes must be
// active.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_f32(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, ROW_MASK,
                                                                 0xF, false));
}

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_f64(double v) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(u & 0xffffffffu), CTRL, ROW_MASK, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), CTRL, ROW_MASK, 0xF, false);
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

enum : int {
    kDppXor1 = 0xB1,      // quad_perm [1, 0, 3, 2]
    kDppXor2 = 0x4E,      // quad_perm [2, 3, 0, 1]
    kDppHalfMirror = 0x141,
    kDppMirror = 0x140,
    kDppBcast15 = 0x142,
    kDppBcast31 = 0x143,
};

__device__ __forceinline__ float wave_sum(float v) {
    v += dpp_f32<kDppXor1, 0xF>(v);
    v += dpp_f32<kDppXor2, 0xF>(v);
    v += dpp_f32<kDppHalfMirror, 0xF>(v);
    v += dpp_f32<kDppMirror, 0xF>(v);
    v += dpp_f32<kDppBcast15, 0xA>(v);
    v += dpp_f32<kDppBcast31, 0xC>(v);
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 63));
}

__device__ __forceinline__ double wave_sum(double v) {
    v += dpp_f64<kDppXor1, 0xF>(v);
    v += dpp_f64<kDppXor2, 0xF>(v);
    v += dpp_f64<kDppHalfMirror, 0xF>(v);
    v += dpp_f64<kDppMirror, 0xF>(v);
    v += dpp_f64<kDppBcast15, 0xA>(v);
    v += dpp_f64<kDppBcast31, 0xC>(v);
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u & 0xffffffffu), 63);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), 63);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

__device__ __forceinline__ int wave_lane() { return threadIdx.x & (kWave - 1); }

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Workgroups are dispatched to the 8 XCDs of the chip round-robin in
// blockIdx order, and each XCD has its own L2.  This maps the launch so that
// each XCD takes a contiguous eighth of the work items instead of every
// eighth one, which keeps spatially close items (shared neighbours) in one
// L2.  A bijection on [0, gridDim.x); a tail of fewer than 8 is unchanged.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_block() {
    const int b = blockIdx.x, nb = gridDim.x;
    const int full = nb / kXcds * kXcds;
    return b < full ? (b % kXcds) * (full / kXcds) + b / kXcds : b;
}

// Exclusive prefix count of a ballot mask below this lane.
__device__ __forceinline__ int lanes_below(unsigned long long mask) {
    unsigned lo = (unsigned)mask, hi = (unsigned)(mask >> 32);
    unsigned c = __builtin_amdgcn_mbcnt_lo(lo, 0u);
    return (int)__builtin_amdgcn_mbcnt_hi(hi, c);
}

}  // namespace pt
