// (d) FP64 reductions, the velocity-Verlet step, row permutations and the
// scatter-free force gather.
//
// Reductions are deterministic: fixed-size blocks reduce in a fixed tree
// order into per-block partials, and one block sums the partials in order.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <mutex>

#include "common.h"

namespace pt {

static thread_local char g_err[512];

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

size_t lds_limit() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 64 * 1024;
    static thread_local int cached_dev = -1;
    static thread_local size_t cached = 0;
    if (dev != cached_dev) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
            v = 64 * 1024;
        cached = (size_t)v;
        cached_dev = dev;
    }
    return cached;
}

size_t lds_per_cu() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 64 * 1024;
    static thread_local int cached_dev = -1;
    static thread_local size_t cached = 0;
    if (dev != cached_dev) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) != hipSuccess ||
            v <= 0)
            v = (int)lds_limit();
        cached = (size_t)v;
        cached_dev = dev;
    }
    return cached;
}

namespace {

constexpr int kRedThreads = 256;
constexpr int kRedItems = 4;
constexpr int kRedTile = kRedThreads * kRedItems;

__device__ __forceinline__ double block_sum(double v) {
    __shared__ double part[kRedThreads / kWave];
    v = wave_sum(v);
    const int w = threadIdx.x >> 6;
    if (wave_lane() == 0) part[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int q = 0; q < kRedThreads / kWave; ++q) s += part[q];
    return s;  // valid in thread 0
}

// mode 0: x[i]; mode 1: 0.5 * m[i] * |v_i|^2 with v row-major n x 3
template <int MODE>
__global__ __launch_bounds__(kRedThreads) void k_reduce_partial(int n, const double* x,
                                                                const double* m, double* partial) {
    double s = 0.0;
    const int base = blockIdx.x * kRedTile;
    for (int it = 0; it < kRedItems; ++it) {
        const int i = base + it * kRedThreads + threadIdx.x;
        if (i < n) {
            if (MODE == 0) s += x[i];
            else {
                const double vx = x[3 * i], vy = x[3 * i + 1], vz = x[3 * i + 2];
                s += 0.5 * m[i] * ((vx * vx + vy * vy) + vz * vz);
            }
        }
    }
    s = block_sum(s);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ __launch_bounds__(kRedThreads) void k_reduce_final(int nb, const double* partial,
                                                              double* out) {
    double s = 0.0;
    for (int k = threadIdx.x; k < nb; k += kRedThreads) s += partial[k];
    s = block_sum(s);
    if (threadIdx.x == 0) out[0] = s;
}

__device__ __forceinline__ double wrap1(double x, double L) {
    double r = fmod(x, L);
    if (r != 0.0 && ((r < 0.0) != (L < 0.0))) r = r + L;
    return r;
}

__global__ void k_verlet_pos(int n, double* pos4, const double* vel, const double* force,
                             const double* inv_mass, double dt, BoxArg box) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double im = inv_mass ? inv_mass[i] : 1.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        // positions + velocities * dt + 0.5 * a * dt * dt (molecular_dynamics.py:20)
        const double acc = inv_mass ? force[3 * i + d] * im : force[3 * i + d];
        double x = pos4[4 * i + d] + vel[3 * i + d] * dt + 0.5 * acc * dt * dt;
        if (box.periodic) x = wrap1(x, box.L[d]);
        pos4[4 * i + d] = x;
    }
}

__global__ void k_wrap(int n, double* pos, BoxArg box) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * n) return;
    pos[t] = wrap1(pos[t], box.L[t % 3]);
}

__global__ void k_verlet_vel(int n, double* vel, const double* f_old, const double* f_new,
                             const double* inv_mass, double dt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double im = inv_mass ? inv_mass[i] : 1.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        // velocities + 0.5 * (forces + new_forces) * dt (molecular_dynamics.py:30)
        const double ao = inv_mass ? f_old[3 * i + d] * im : f_old[3 * i + d];
        const double an = inv_mass ? f_new[3 * i + d] * im : f_new[3 * i + d];
        vel[3 * i + d] = vel[3 * i + d] + 0.5 * (ao + an) * dt;
    }
}

__global__ void k_berendsen(int n, double* vel, const double* ke, double dt, double tau, double t0,
                            double kb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double T = 2.0 * ke[0] / (3.0 * n * kb);
    const double lam = 1.0 / sqrt(1.0 + (dt / tau) * (T / t0 - 1.0));
    vel[3 * i] *= lam; vel[3 * i + 1] *= lam; vel[3 * i + 2] *= lam;
}

// The two largest squared displacements (of two different atoms): one pair
// per block, then one block over the pairs; out[0], out[1] = the largest and
// the second largest |MIC(pos4_i - ref_i)|.
__device__ __forceinline__ void top2_merge(double& a1, double& a2, double b1, double b2) {
    const double lo = fmin(a1, b1);
    a1 = fmax(a1, b1);
    a2 = fmax(lo, fmax(a2, b2));
}

__device__ void top2_block(double& m1, double& m2) {
    __shared__ double r1[256], r2[256];
    r1[threadIdx.x] = m1;
    r2[threadIdx.x] = m2;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            double a1 = r1[threadIdx.x], a2 = r2[threadIdx.x];
            top2_merge(a1, a2, r1[threadIdx.x + s], r2[threadIdx.x + s]);
            r1[threadIdx.x] = a1;
            r2[threadIdx.x] = a2;
        }
        __syncthreads();
    }
    m1 = r1[0];
    m2 = r2[0];
}

__global__ __launch_bounds__(256) void k_max_disp_partial(int n, const double* pos4, const double* ref,
                                                          BoxArg box, double* partial) {
    double m1 = 0.0, m2 = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        double dx = pos4[4 * i] - ref[4 * i], dy = pos4[4 * i + 1] - ref[4 * i + 1],
               dz = pos4[4 * i + 2] - ref[4 * i + 2];
        if (box.periodic) {
            dx = mic1(dx, box.L[0], box.half[0]);
            dy = mic1(dy, box.L[1], box.half[1]);
            dz = mic1(dz, box.L[2], box.half[2]);
        }
        top2_merge(m1, m2, dx * dx + dy * dy + dz * dz, 0.0);
    }
    top2_block(m1, m2);
    if (threadIdx.x == 0) { partial[2 * blockIdx.x] = m1; partial[2 * blockIdx.x + 1] = m2; }
}

__global__ __launch_bounds__(256) void k_max_disp_final(int nb, const double* partial, double* out) {
    double m1 = 0.0, m2 = 0.0;
    for (int k = threadIdx.x; k < nb; k += 256) top2_merge(m1, m2, partial[2 * k], partial[2 * k + 1]);
    top2_block(m1, m2);
    if (threadIdx.x == 0) { out[0] = sqrt(m1); out[1] = sqrt(m2); }
}

template <typename T>
__global__ void k_gather(int n, int width, const int* idx, const T* src, T* dst) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)n * width) return;
    const int i = (int)(t / width), k = (int)(t - (long long)i * width);
    dst[t] = src[(long long)idx[i] * width + k];
}

template <typename T>
__global__ void k_scatter(int n, int width, const int* idx, const T* src, T* dst) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)n * width) return;
    const int i = (int)(t / width), k = (int)(t - (long long)i * width);
    dst[(long long)idx[i] * width + k] = src[t];
}

// F[i] = sum_{e in row i} (g[e] - g[rev[e]]), one wavefront per atom.
template <typename E>
__global__ __launch_bounds__(256) void k_force_gather(int n, const long long* __restrict__ row_off,
                                                      const int* __restrict__ rev,
                                                      const E* __restrict__ g,
                                                      double* __restrict__ f) {
    const int lane = wave_lane();
    const int i = uniform(xcd_block() * 4 + (threadIdx.x >> 6));
    if (i >= n) return;
    const long long b = row_off[i], e = row_off[i + 1];
    double fx = 0, fy = 0, fz = 0;
    for (long long k = b + lane; k < e; k += kWave) {
        const long long r = rev[k];
        fx += (double)g[3 * k];
        fy += (double)g[3 * k + 1];
        fz += (double)g[3 * k + 2];
        if (r >= 0) {  // always true for a symmetric list (checked by the tests)
            fx -= (double)g[3 * r];
            fy -= (double)g[3 * r + 1];
            fz -= (double)g[3 * r + 2];
        }
    }
    fx = wave_sum(fx); fy = wave_sum(fy); fz = wave_sum(fz);
    if (lane == 0) { f[3 * i] = fx; f[3 * i + 1] = fy; f[3 * i + 2] = fz; }
}

// Transposed edge buffer: gT[e(i->j)] = dE_j/dr_i (written by centre j at the
// reverse slot), so F_i = f_self_i - sum_{e in row i} gT[e] reads row i
// contiguously, without the reverse index.
template <typename E>
__global__ __launch_bounds__(256) void k_force_gather_t(int n, const long long* __restrict__ row_off,
                                                        const double* __restrict__ f_self,
                                                        const E* __restrict__ gT, double* __restrict__ f) {
    const int lane = wave_lane();
    const int i = uniform(xcd_block() * 4 + (threadIdx.x >> 6));
    if (i >= n) return;
    const long long b = row_off[i], e = row_off[i + 1];
    double fx = 0, fy = 0, fz = 0;
    for (long long k = b + lane; k < e; k += kWave) {
        fx += (double)gT[3 * k];
        fy += (double)gT[3 * k + 1];
        fz += (double)gT[3 * k + 2];
    }
    fx = wave_sum(fx); fy = wave_sum(fy); fz = wave_sum(fz);
    if (lane == 0) {
        f[3 * i] = f_self[3 * i] - fx;
        f[3 * i + 1] = f_self[3 * i + 1] - fy;
        f[3 * i + 2] = f_self[3 * i + 2] - fz;
    }
}

inline int nblocks(long long work, int threads) { return (int)((work + threads - 1) / threads); }

// ---- gated state updates of an MD step that never waits on the host ----------
// *flag = 1 when the list must be rebuilt: the two largest displacements since
// the last build sum to more than the skin (a NaN displacement rebuilds too).
__global__ void k_skin_flag(const double* __restrict__ d, double skin, int* __restrict__ flag) {
    if (blockIdx.x == 0 && threadIdx.x == 0) flag[0] = (skin > 0.0 && d[0] + d[1] <= skin) ? 0 : 1;
}

__device__ __forceinline__ bool gate_closed(const int* gate) { return gate != nullptr && *gate == 0; }

template <typename T>
__global__ void k_copy_gated(const int* __restrict__ gate, long long n, const T* __restrict__ src,
                             T* __restrict__ dst) {
    if (gate_closed(gate)) return;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
        dst[t] = src[t];
}

template <typename T>
__global__ void k_gather_gated(const int* __restrict__ gate, int n, int width, const int* __restrict__ idx,
                               const T* __restrict__ src, T* __restrict__ dst) {
    if (gate_closed(gate)) return;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)n * width) return;
    const int i = (int)(t / width), k = (int)(t - (long long)i * width);
    dst[t] = src[(long long)idx[i] * width + k];
}

struct PermuteArgs {
    int n_arrays;
    char* buf[PT_PERMUTE_MAX];
    long long tmp_off[PT_PERMUTE_MAX];   // bytes into tmp
    int words[PT_PERMUTE_MAX];           // words per row
    int wbytes[PT_PERMUTE_MAX];          // 4 or 8
};

// thread per atom: row i of every array <- row perm[i], into tmp (rows written contiguously)
__global__ void k_permute_gather(const int* __restrict__ gate, int n, const int* __restrict__ perm, PermuteArgs a,
                                 char* __restrict__ tmp) {
    if (gate_closed(gate)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long long j = perm[i];
    for (int k = 0; k < a.n_arrays; ++k) {
        const int w = a.words[k];
        if (a.wbytes[k] == 8) {
            const double* src = reinterpret_cast<const double*>(a.buf[k]) + j * w;
            double* dst = reinterpret_cast<double*>(tmp + a.tmp_off[k]) + (long long)i * w;
            for (int q = 0; q < w; ++q) dst[q] = src[q];
        } else {
            const int* src = reinterpret_cast<const int*>(a.buf[k]) + j * w;
            int* dst = reinterpret_cast<int*>(tmp + a.tmp_off[k]) + (long long)i * w;
            for (int q = 0; q < w; ++q) dst[q] = src[q];
        }
    }
}

// every array back from tmp, one 4-byte word per thread over the concatenation
__global__ void k_permute_copy(const int* __restrict__ gate, int n, PermuteArgs a, const char* __restrict__ tmp,
                               long long total_words) {
    if (gate_closed(gate)) return;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total_words;
         t += (long long)gridDim.x * blockDim.x) {
        // array k holds n * words[k] * wbytes[k] / 4 four-byte words starting at tmp_off[k] / 4
        int k = 0;
        while (k + 1 < a.n_arrays && t >= a.tmp_off[k + 1] / 4) ++k;
        const long long u = t - a.tmp_off[k] / 4;
        reinterpret_cast<int*>(a.buf[k])[u] = reinterpret_cast<const int*>(tmp)[t];
    }
}

}  // namespace
}  // namespace pt

using namespace pt;

extern "C" size_t pt_permute_state_tmp_bytes(int n, int n_arrays, const pt_permute_array* arrays) {
    if (n < 0 || n_arrays < 0 || n_arrays > PT_PERMUTE_MAX || (n_arrays > 0 && !arrays)) return 0;
    size_t b = 0;
    for (int k = 0; k < n_arrays; ++k) b += (size_t)n * arrays[k].width * arrays[k].word_bytes;
    return b;
}

extern "C" int pt_permute_state_gated(int n, const int* perm, int n_arrays, const pt_permute_array* arrays,
                                      void* tmp, size_t tmp_bytes, const int* gate, pt_stream_t stream) {
    PT_REQUIRE(n >= 0 && n_arrays >= 0 && n_arrays <= PT_PERMUTE_MAX, "pt_permute_state_gated: bad sizes");
    if (n == 0 || n_arrays == 0) return PT_OK;
    PT_REQUIRE(perm && arrays && tmp, "pt_permute_state_gated: null argument");
    PermuteArgs a{};
    a.n_arrays = n_arrays;
    long long off = 0;
    for (int k = 0; k < n_arrays; ++k) {
        const pt_permute_array& r = arrays[k];
        PT_REQUIRE(r.buf && r.width > 0 && (r.word_bytes == 4 || r.word_bytes == 8),
                   "pt_permute_state_gated: bad array %d", k);
        PT_REQUIRE((char*)r.buf + (size_t)n * r.width * r.word_bytes <= (char*)tmp ||
                       (char*)tmp + tmp_bytes <= (char*)r.buf,
                   "pt_permute_state_gated: array %d overlaps tmp", k);
        PT_REQUIRE(off % r.word_bytes == 0, "pt_permute_state_gated: array %d would be misaligned in tmp "
                   "(put the 8-byte arrays first)", k);
        a.buf[k] = (char*)r.buf;
        a.tmp_off[k] = off;
        a.words[k] = r.width;
        a.wbytes[k] = r.word_bytes;
        off += (long long)n * r.width * r.word_bytes;
    }
    PT_REQUIRE((size_t)off <= tmp_bytes, "pt_permute_state_gated: tmp too small");
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_permute_gather, dim3(nblocks(n, 256)), dim3(256), 0, st, gate, n, perm, a, (char*)tmp);
    if (int rc = launch_check("pt_permute_state_gated")) return rc;
    const long long words = off / 4;
    const int blocks = (int)std::min<long long>(nblocks(words, 256), 8192);
    hipLaunchKernelGGL(k_permute_copy, dim3(blocks), dim3(256), 0, st, gate, n, a, (const char*)tmp, words);
    return launch_check("pt_permute_state_gated");
}

extern "C" const char* pt_last_error(void) { return pt::g_err; }
extern "C" int pt_version(void) { return 1; }

extern "C" size_t pt_reduce_workspace_bytes(int n) {
    const int nb = (n + kRedTile - 1) / kRedTile;
    return sizeof(double) * (size_t)(nb > 0 ? nb : 1);
}

extern "C" int pt_reduce_sum_f64(int n, const double* x, double* out, void* ws, size_t ws_bytes,
                                 pt_stream_t stream_) {
    hipStream_t stream = (hipStream_t)stream_;
    PT_REQUIRE(n >= 0 && out && ws, "pt_reduce_sum_f64: bad arguments");
    PT_REQUIRE(ws_bytes >= pt_reduce_workspace_bytes(n), "pt_reduce_sum_f64: workspace too small");
    const int nb = std::max(1, (n + kRedTile - 1) / kRedTile);
    hipLaunchKernelGGL(k_reduce_partial<0>, dim3(nb), dim3(kRedThreads), 0, stream, n, x,
                       (const double*)nullptr, (double*)ws);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kRedThreads), 0, stream, nb, (const double*)ws,
                       out);
    return launch_check("pt_reduce_sum_f64");
}

// ---- FP64 sum without allocation ----------------------------------------------
// The partials of pt_reduce_sum_f64_noalloc live in `scratch`, a file-scope
// device array: every GPU the library runs on holds its own copy (512 KiB)
// until the process ends.  Its address on a device comes from
// hipGetSymbolAddress, taken once per device with that device current; host
// code uses only that address (the array's name is not a device pointer on
// the host).  k_reduce_partial writes every partial that k_reduce_final reads,
// so nothing needs clearing between calls; calls that overlap on one GPU
// (several streams or threads) share the array and race.
namespace pt {
constexpr int kNoAllocMaxN = 1 << 26;                       // 67,108,864 elements
constexpr int kNoAllocPartials = kNoAllocMaxN / kRedTile;   // 65,536 partial sums
}  // namespace pt

static __device__ __align__(256) double scratch[pt::kNoAllocPartials];

namespace pt {
namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up on its first use there).
int scratch_on_current_device(double** out) {
    int dev = -1;
    PT_CHECK_HIP(hipGetDevice(&dev));
    PT_REQUIRE(dev >= 0 && dev < kMaxDevices, "device index %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        PT_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        PT_REQUIRE(p != nullptr, "hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_ptr[dev] = (double*)p;
    }
    *out = g_scratch_ptr[dev];
    return PT_OK;
}

}  // namespace
}  // namespace pt

extern "C" int pt_reduce_sum_f64_noalloc(int n, const double* x, double* out, pt_stream_t stream_) {
    hipStream_t stream = (hipStream_t)stream_;
    PT_REQUIRE(n >= 0 && out && (n == 0 || x), "pt_reduce_sum_f64_noalloc: bad arguments");
    PT_REQUIRE(n <= pt::kNoAllocMaxN,
               "pt_reduce_sum_f64_noalloc: n = %d exceeds the maximum of %d elements "
               "(pt_reduce_sum_f64 takes a workspace for larger sums)", n, pt::kNoAllocMaxN);
    double* part = nullptr;
    if (int rc = pt::scratch_on_current_device(&part)) return rc;
    const int nb = std::max(1, (n + kRedTile - 1) / kRedTile);
    hipLaunchKernelGGL(k_reduce_partial<0>, dim3(nb), dim3(kRedThreads), 0, stream, n, x,
                       (const double*)nullptr, part);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kRedThreads), 0, stream, nb, (const double*)part,
                       out);
    return launch_check("pt_reduce_sum_f64_noalloc");
}

extern "C" int pt_kinetic_energy(int n, const double* vel, const double* mass, double* out,
                                 void* ws, size_t ws_bytes, pt_stream_t stream_) {
    hipStream_t stream = (hipStream_t)stream_;
    PT_REQUIRE(n >= 0 && vel && mass && out && ws, "pt_kinetic_energy: bad arguments");
    PT_REQUIRE(ws_bytes >= pt_reduce_workspace_bytes(n), "pt_kinetic_energy: workspace too small");
    const int nb = std::max(1, (n + kRedTile - 1) / kRedTile);
    hipLaunchKernelGGL(k_reduce_partial<1>, dim3(nb), dim3(kRedThreads), 0, stream, n, vel, mass,
                       (double*)ws);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kRedThreads), 0, stream, nb, (const double*)ws,
                       out);
    return launch_check("pt_kinetic_energy");
}

extern "C" int pt_verlet_positions(int n, double* pos4, const double* vel, const double* force,
                                   const double* inv_mass, double dt, const double* box,
                                   pt_stream_t stream) {
    PT_REQUIRE(n >= 0 && pos4 && vel && force, "pt_verlet_positions: bad arguments");
    if (n == 0) return PT_OK;
    BoxArg b{};
    b.periodic = box != nullptr;
    for (int d = 0; d < 3; ++d) { b.L[d] = box ? box[d] : 0; b.half[d] = 0.5 * b.L[d]; }
    hipLaunchKernelGGL(k_verlet_pos, dim3(nblocks(n, 256)), dim3(256), 0, (hipStream_t)stream, n,
                       pos4, vel, force, inv_mass, dt, b);
    return launch_check("pt_verlet_positions");
}

extern "C" int pt_wrap_positions(int n, double* pos, const double* box, pt_stream_t stream) {
    PT_REQUIRE(n >= 0 && pos && box, "pt_wrap_positions: bad arguments");
    if (n == 0) return PT_OK;
    BoxArg b{};
    b.periodic = 1;
    for (int d = 0; d < 3; ++d) { b.L[d] = box[d]; b.half[d] = 0.5 * box[d]; }
    hipLaunchKernelGGL(k_wrap, dim3(nblocks(3LL * n, 256)), dim3(256), 0, (hipStream_t)stream, n, pos, b);
    return launch_check("pt_wrap_positions");
}

extern "C" int pt_verlet_velocities(int n, double* vel, const double* f_old, const double* f_new,
                                    const double* inv_mass, double dt, pt_stream_t stream) {
    PT_REQUIRE(n >= 0 && vel && f_old && f_new, "pt_verlet_velocities: bad arguments");
    if (n == 0) return PT_OK;
    hipLaunchKernelGGL(k_verlet_vel, dim3(nblocks(n, 256)), dim3(256), 0, (hipStream_t)stream, n,
                       vel, f_old, f_new, inv_mass, dt);
    return launch_check("pt_verlet_velocities");
}

extern "C" int pt_berendsen_rescale(int n, double* vel, const double* ke, double dt, double tau,
                                    double t0, double kb, pt_stream_t stream) {
    PT_REQUIRE(n >= 0 && vel && ke && tau != 0.0 && t0 != 0.0, "pt_berendsen_rescale: bad arguments");
    if (n == 0) return PT_OK;
    hipLaunchKernelGGL(k_berendsen, dim3(nblocks(n, 256)), dim3(256), 0, (hipStream_t)stream, n,
                       vel, ke, dt, tau, t0, kb);
    return launch_check("pt_berendsen_rescale");
}

extern "C" int pt_max_displacement(int n, const double* pos4, const double* ref, const double* box,
                                   double* out, pt_stream_t stream) {
    PT_REQUIRE(n >= 0 && pos4 && ref && out, "pt_max_displacement: bad arguments");
    BoxArg b{};
    b.periodic = box != nullptr;
    for (int d = 0; d < 3; ++d) { b.L[d] = box ? box[d] : 0; b.half[d] = 0.5 * b.L[d]; }
    // out[2 .. 2 + 2 nb) holds the block pairs (out must hold 2 + 2 * 256 doubles)
    const int nb = std::max(1, std::min(256, (n + 255) / 256));
    hipLaunchKernelGGL(k_max_disp_partial, dim3(nb), dim3(256), 0, (hipStream_t)stream, n, pos4, ref, b,
                       out + 2);
    hipLaunchKernelGGL(k_max_disp_final, dim3(1), dim3(256), 0, (hipStream_t)stream, nb,
                       (const double*)(out + 2), out);
    return launch_check("pt_max_displacement");
}

extern "C" int pt_skin_check(int n, const double* pos4, const double* ref, const double* box, double skin,
                             double* scratch, int* flag, pt_stream_t stream) {
    PT_REQUIRE(n >= 0 && pos4 && ref && scratch && flag, "pt_skin_check: bad arguments");
    if (int rc = pt_max_displacement(n, pos4, ref, box, scratch, stream)) return rc;
    hipLaunchKernelGGL(k_skin_flag, dim3(1), dim3(64), 0, (hipStream_t)stream, (const double*)scratch, skin, flag);
    return launch_check("pt_skin_check");
}

extern "C" int pt_copy_gated(long long n_words, int word_bytes, const void* src, void* dst, const int* gate,
                             pt_stream_t stream) {
    PT_REQUIRE(n_words >= 0 && (word_bytes == 4 || word_bytes == 8) && (n_words == 0 || (src && dst)),
               "pt_copy_gated: bad arguments");
    if (n_words == 0) return PT_OK;
    const int blocks = (int)std::min<long long>(nblocks(n_words, 256), 4096);
    if (word_bytes == 8)
        hipLaunchKernelGGL(k_copy_gated<double>, dim3(blocks), dim3(256), 0, (hipStream_t)stream, gate, n_words,
                           (const double*)src, (double*)dst);
    else
        hipLaunchKernelGGL(k_copy_gated<int>, dim3(blocks), dim3(256), 0, (hipStream_t)stream, gate, n_words,
                           (const int*)src, (int*)dst);
    return launch_check("pt_copy_gated");
}

extern "C" int pt_permute_rows_gated(int n, int width, int word_bytes, const int* perm, void* buf, void* tmp,
                                     const int* gate, pt_stream_t stream) {
    PT_REQUIRE(n >= 0 && width > 0 && (word_bytes == 4 || word_bytes == 8) && (n == 0 || (perm && buf && tmp)),
               "pt_permute_rows_gated: bad arguments");
    PT_REQUIRE(buf != tmp, "pt_permute_rows_gated: buf and tmp must differ");
    if (n == 0) return PT_OK;
    const long long words = (long long)n * width;
    if (word_bytes == 8)
        hipLaunchKernelGGL(k_gather_gated<double>, dim3(nblocks(words, 256)), dim3(256), 0, (hipStream_t)stream,
                           gate, n, width, perm, (const double*)buf, (double*)tmp);
    else
        hipLaunchKernelGGL(k_gather_gated<int>, dim3(nblocks(words, 256)), dim3(256), 0, (hipStream_t)stream,
                           gate, n, width, perm, (const int*)buf, (int*)tmp);
    if (int rc = launch_check("pt_permute_rows_gated")) return rc;
    return pt_copy_gated(words, word_bytes, tmp, buf, gate, stream);
}

extern "C" int pt_gather_rows_f64(int n, int width, const int* idx, const double* src, double* dst,
                                  pt_stream_t stream) {
    PT_REQUIRE(n >= 0 && width > 0 && idx && src && dst, "pt_gather_rows_f64: bad arguments");
    if (n == 0) return PT_OK;
    hipLaunchKernelGGL(k_gather<double>, dim3(nblocks((long long)n * width, 256)), dim3(256), 0,
                       (hipStream_t)stream, n, width, idx, src, dst);
    return launch_check("pt_gather_rows_f64");
}

extern "C" int pt_scatter_rows_f64(int n, int width, const int* idx, const double* src, double* dst,
                                   pt_stream_t stream) {
    PT_REQUIRE(n >= 0 && width > 0 && idx && src && dst, "pt_scatter_rows_f64: bad arguments");
    if (n == 0) return PT_OK;
    hipLaunchKernelGGL(k_scatter<double>, dim3(nblocks((long long)n * width, 256)), dim3(256), 0,
                       (hipStream_t)stream, n, width, idx, src, dst);
    return launch_check("pt_scatter_rows_f64");
}

extern "C" int pt_gather_rows_i32(int n, int width, const int* idx, const int* src, int* dst,
                                  pt_stream_t stream) {
    PT_REQUIRE(n >= 0 && width > 0 && idx && src && dst, "pt_gather_rows_i32: bad arguments");
    if (n == 0) return PT_OK;
    hipLaunchKernelGGL(k_gather<int>, dim3(nblocks((long long)n * width, 256)), dim3(256), 0,
                       (hipStream_t)stream, n, width, idx, src, dst);
    return launch_check("pt_gather_rows_i32");
}

extern "C" int pt_force_gather_transposed(int n, const long long* row_off, const double* f_self,
                                          const void* edge_grad_t, int edge_precision, double* f_full,
                                          pt_stream_t stream) {
    PT_REQUIRE(n >= 0 && row_off && f_self && edge_grad_t && f_full, "pt_force_gather_transposed: bad arguments");
    if (n == 0) return PT_OK;
    const int blocks = (n + 3) / 4;
    if (edge_precision == PT_F64)
        hipLaunchKernelGGL(k_force_gather_t<double>, dim3(blocks), dim3(256), 0, (hipStream_t)stream, n,
                           row_off, f_self, (const double*)edge_grad_t, f_full);
    else if (edge_precision == PT_F32)
        hipLaunchKernelGGL(k_force_gather_t<float>, dim3(blocks), dim3(256), 0, (hipStream_t)stream, n,
                           row_off, f_self, (const float*)edge_grad_t, f_full);
    else {
        set_error("pt_force_gather_transposed: unknown precision %d", edge_precision);
        return PT_ERR_ARG;
    }
    return launch_check("pt_force_gather_transposed");
}

extern "C" int pt_force_gather(int n, const long long* row_off, const int* rev, const void* edge_grad,
                               int edge_precision, double* f_full, pt_stream_t stream) {
    PT_REQUIRE(n >= 0 && row_off && rev && edge_grad && f_full, "pt_force_gather: bad arguments");
    if (n == 0) return PT_OK;
    const int blocks = (n + 3) / 4;
    if (edge_precision == PT_F64)
        hipLaunchKernelGGL(k_force_gather<double>, dim3(blocks), dim3(256), 0, (hipStream_t)stream, n,
                           row_off, rev, (const double*)edge_grad, f_full);
    else if (edge_precision == PT_F32)
        hipLaunchKernelGGL(k_force_gather<float>, dim3(blocks), dim3(256), 0, (hipStream_t)stream, n,
                           row_off, rev, (const float*)edge_grad, f_full);
    else {
        set_error("pt_force_gather: unknown precision %d", edge_precision);
        return PT_ERR_ARG;
    }
    return launch_check("pt_force_gather");
}
