// (a) Neighbour list on gfx950: cell binning, stable LSD radix sort of
// (type, cell) keys, and a CSR list built by one wavefront per centre.
//
// Replaces the dense masks of atoms/neighbor.py:47-64,102-107: row i holds
// every j with r_ij <= rc and r_ij > 0, r_ij being the single-shift
// minimum-image distance of atoms/distance.py:63-78 / atoms/box.py:112-117.
// Candidates come from the cells within one cell of i's cell (cells are at
// least rc wide, stencil offsets deduplicated when a dimension has fewer
// than three cells); the distance itself is always the explicit MIC formula,
// so the row is the reference mask exactly, whatever the box/cutoff ratio.
//
// Rows are emitted in ascending internal index: the stencil cells are
// visited in ascending cell id, type by type, and internal order is sorted
// by (type, cell) with ties kept in input order (stable sort).
//
// The cutoff test r <= rc is done as s <= T on s = dx^2 + dy^2 + dz^2
// (same operation order, no contraction), T being the largest double whose
// correctly rounded square root is <= rc (found on the host): sqrt is
// monotone, so this is the reference's test bit for bit without a sqrt per
// candidate.  One candidate pass stages each row (up to kStageW entries) and
// counts it; after the row-offset scan one kernel copies the staged rows into
// the CSR arrays and pairs every edge with its reverse.  A list whose longest
// row exceeds kStageW takes a second candidate pass instead.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "common.h"

namespace pt {
namespace {

constexpr int kMaxCells = 1 << 18;
constexpr int kSortThreads = 256;
constexpr int kSortItems = 8;
constexpr int kTile = kSortThreads * kSortItems;
constexpr int kMaxTypes = 8;
constexpr int kMaxRanges = 27 * kMaxTypes;
constexpr int kScanBlock = 1024;  // items per block of the row-offset scan
constexpr int kStageW = 256;      // staged entries per row
constexpr int kBucket = 16;       // entries per bucket of a staged row
constexpr int kNBucket = kStageW / kBucket;
constexpr int kCellThreads = 512;  // threads of a candidate-pass workgroup (one per cell)

struct Grid {
    int nc[3];
    int ncell;
    int nst[3];      // number of distinct stencil offsets per dimension
    int off[3][3];   // the offsets
    int periodic;
    double lo[3];
    double inv_w[3];
};

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ int cell_coord(double x, double lo, double inv_w, int nc, int periodic) {
    double u = (x - lo) * inv_w;
    int c = (int)floor(u);
    if (periodic) {
        c %= nc;
        if (c < 0) c += nc;
    } else {
        c = c < 0 ? 0 : (c >= nc ? nc - 1 : c);
    }
    return c;
}

__global__ void k_bbox(int n, const double* pos, int ps, double* out) {
    // single block; out[0..2] = min, out[3..5] = max
    __shared__ double smin[3][256], smax[3][256];
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = threadIdx.x; i < n; i += blockDim.x)
        for (int d = 0; d < 3; ++d) {
            double v = pos[(size_t)ps * i + d];
            mn[d] = fmin(mn[d], v);
            mx[d] = fmax(mx[d], v);
        }
    for (int d = 0; d < 3; ++d) { smin[d][threadIdx.x] = mn[d]; smax[d][threadIdx.x] = mx[d]; }
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int d = 0; d < 3; ++d) {
                smin[d][threadIdx.x] = fmin(smin[d][threadIdx.x], smin[d][threadIdx.x + s]);
                smax[d][threadIdx.x] = fmax(smax[d][threadIdx.x], smax[d][threadIdx.x + s]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int d = 0; d < 3; ++d) { out[d] = smin[d][0]; out[3 + d] = smax[d][0]; }
}

// A build gated by a device flag (the skin check of an MD step) does nothing
// when the flag is 0: every kernel of the build returns at once, so the list,
// the order and the positions stay those of the last build.
__device__ __forceinline__ bool gated_off(const int* gate) { return gate != nullptr && *gate == 0; }

__global__ void k_bin(const int* __restrict__ gate, int n, const double* pos, int ps, const int* types, Grid g,
                      int* keys, int* vals, unsigned long long* bmax) {
    if (gated_off(gate)) return;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (i == 0) *bmax = 0ull;   // max |coordinate|, raised by k_gather_pos4
    int c0 = cell_coord(pos[(size_t)ps * i + 0], g.lo[0], g.inv_w[0], g.nc[0], g.periodic);
    int c1 = cell_coord(pos[(size_t)ps * i + 1], g.lo[1], g.inv_w[1], g.nc[1], g.periodic);
    int c2 = cell_coord(pos[(size_t)ps * i + 2], g.lo[2], g.inv_w[2], g.nc[2], g.periodic);
    int cell = (c2 * g.nc[1] + c1) * g.nc[0] + c0;
    keys[i] = (types[i] - 1) * g.ncell + cell;
    vals[i] = i;
}

// ---- stable LSD radix sort, 8-bit digits -----------------------------------

__global__ __launch_bounds__(kSortThreads) void k_radix_hist(const int* __restrict__ gate, int n, const int* keys,
                                                             int shift, int nblocks, int* ghist) {
    if (gated_off(gate)) return;
    __shared__ int hist[256];
    hist[threadIdx.x] = 0;
    __syncthreads();
    int base = blockIdx.x * kTile;
    for (int it = 0; it < kSortItems; ++it) {
        int idx = base + it * kSortThreads + threadIdx.x;
        if (idx < n) atomicAdd(&hist[(keys[idx] >> shift) & 255], 1);
    }
    __syncthreads();
    ghist[threadIdx.x * nblocks + blockIdx.x] = hist[threadIdx.x];
}

// Exclusive scan of m ints (up to a few 1e5) in place by one block: each
// thread scans a contiguous chunk of ceil(m / 1024), the chunk totals are
// scanned across the block (wave shuffles, then the 16 wave totals).  Up to
// kScanLds entries go through LDS, loaded and stored coalesced (the chunks'
// strided reads from HBM kept one load in flight per thread: 12.9 us for the
// 12,288 radix counters of 98,304 atoms).
constexpr int kScanLds = 16384;

template <bool LDS>
__device__ __forceinline__ void scan_chunks(int m, int* data, int* wtot) {
    const int lane = threadIdx.x & (kWave - 1);
    const int w = threadIdx.x >> 6;
    const int per = (m + 1023) / 1024;
    const int b = threadIdx.x * per;
    const int e = min(b + per, m);
    int sum = 0;
    for (int k = b; k < e; ++k) sum += data[k];
    int incl = sum;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const int o = __shfl_up(incl, off, kWave);
        if (lane >= off) incl += o;
    }
    if (lane == kWave - 1) wtot[w] = incl;
    __syncthreads();
    if (w == 0) {
        int v = lane < 1024 / kWave ? wtot[lane] : 0;
        int vi = v;
#pragma unroll
        for (int off = 1; off < 1024 / kWave; off <<= 1) {
            const int o = __shfl_up(vi, off, kWave);
            if (lane >= off) vi += o;
        }
        if (lane < 1024 / kWave) wtot[lane] = vi - v;   // exclusive over waves
    }
    __syncthreads();
    int run = wtot[w] + incl - sum;
    for (int k = b; k < e; ++k) {
        const int v = data[k];
        data[k] = run;
        run += v;
    }
    (void)LDS;
}

__global__ __launch_bounds__(1024) void k_scan_small(const int* __restrict__ gate, int m, int* data) {
    if (gated_off(gate)) return;
    __shared__ int wtot[1024 / kWave];
    __shared__ int buf[kScanLds];
    if (m <= kScanLds) {
        for (int k = threadIdx.x; k < m; k += 1024) buf[k] = data[k];
        __syncthreads();
        scan_chunks<true>(m, buf, wtot);
        __syncthreads();
        for (int k = threadIdx.x; k < m; k += 1024) data[k] = buf[k];
    } else {
        scan_chunks<false>(m, data, wtot);
    }
}

__global__ __launch_bounds__(kSortThreads) void k_radix_scatter(const int* __restrict__ gate, int n,
                                                                const int* keys_in, const int* vals_in,
                                                                int shift, int nblocks, const int* gscan,
                                                                int* keys_out, int* vals_out) {
    if (gated_off(gate)) return;
    __shared__ int run[256];
    __shared__ int wcount[kSortThreads / kWave][256];
    const int lane = threadIdx.x & (kWave - 1);
    const int w = threadIdx.x >> 6;
    run[threadIdx.x] = gscan[threadIdx.x * nblocks + blockIdx.x];
    for (int q = 0; q < kSortThreads / kWave; ++q) wcount[q][threadIdx.x] = 0;
    __syncthreads();
    int base = blockIdx.x * kTile;
    for (int it = 0; it < kSortItems; ++it) {
        int idx = base + it * kSortThreads + threadIdx.x;
        bool valid = idx < n;
        int key = valid ? keys_in[idx] : 0;
        int val = valid ? vals_in[idx] : 0;
        int d = (key >> shift) & 255;
        // lanes of this wave holding the same digit
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            unsigned long long bal = __ballot((d >> b) & 1);
            peers &= ((d >> b) & 1) ? bal : ~bal;
        }
        int rank = lanes_below(peers);
        if (valid && rank == 0) wcount[w][d] = __popcll(peers);
        __syncthreads();
        if (valid) {
            int dst = run[d] + rank;
            for (int q = 0; q < w; ++q) dst += wcount[q][d];
            keys_out[dst] = key;
            vals_out[dst] = val;
        }
        __syncthreads();
        int tot = 0;
        for (int q = 0; q < kSortThreads / kWave; ++q) { tot += wcount[q][threadIdx.x]; wcount[q][threadIdx.x] = 0; }
        run[threadIdx.x] += tot;
        __syncthreads();
    }
}

__global__ void k_cell_start(const int* __restrict__ gate, int nkeys, int n, const int* skeys, int* cell_start) {
    if (gated_off(gate)) return;
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > nkeys) return;
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (skeys[mid] < k) lo = mid + 1; else hi = mid;
    }
    cell_start[k] = lo;
}

// Internal-order positions: FP64 (x, y, z, type) and an FP32 copy for the
// candidate screen; bmax gets the bit pattern of max |coordinate| (the bit
// patterns of non-negative doubles order like the values).
__global__ __launch_bounds__(256) void k_gather_pos4(const int* __restrict__ gate, int n, const int* perm,
                                                     const double* pos, int ps, const int* types, double* pos4,
                                                     float4* pos4f, unsigned long long* bmax, int* perm_out) {
    if (gated_off(gate)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) perm_out[i] = perm[i];
    double m = 0.0;
    if (i < n) {
        const int v = perm[i];
        double4 p;
        p.x = pos[(size_t)ps * v]; p.y = pos[(size_t)ps * v + 1]; p.z = pos[(size_t)ps * v + 2];
        p.w = (double)types[v];
        reinterpret_cast<double4*>(pos4)[i] = p;
        pos4f[i] = make_float4((float)p.x, (float)p.y, (float)p.z, 0.0f);
        // a NaN coordinate makes the bound NaN: every candidate then takes the exact test
        m = fmax(fmax(fabs(p.x), fabs(p.y)), fabs(p.z));
        if (p.x != p.x || p.y != p.y || p.z != p.z) m = __builtin_nan("");
    }
    unsigned long long b = __builtin_bit_cast(unsigned long long, m) & ~(1ull << 63);
    for (int off = 1; off < kWave; off <<= 1) {
        const unsigned long long o = __shfl_xor(b, off, kWave);
        b = o > b ? o : b;
    }
    // one atomic per block, and only when it raises the bound
    __shared__ unsigned long long wmax[256 / kWave];
    if ((threadIdx.x & (kWave - 1)) == 0) wmax[threadIdx.x >> 6] = b;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < (int)(blockDim.x >> 6); ++q) b = wmax[q] > b ? wmax[q] : b;
        if (b > __hip_atomic_load(bmax, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(bmax, b);
    }
}

// ---- one wavefront per centre: count, fill (CSR) or stage ------------------

enum { PASS_FILL = 1, PASS_STAGE = 2 };

// squared single-shift minimum-image distance, operation order of
// atoms/distance.py:63-78 (no contraction)
__device__ __forceinline__ double mic_dist2(double dx, double dy, double dz, const BoxArg& box) {
#pragma clang fp contract(off)
    if (box.periodic) {
        dx = mic1(dx, box.L[0], box.half[0]);
        dy = mic1(dy, box.L[1], box.half[1]);
        dz = mic1(dz, box.L[2], box.half[2]);
    }
    double s = dx * dx;
    s = s + dy * dy;
    s = s + dz * dz;
    return s;
}

// The neighbouring float of f towards +inf (up) or -inf; f is not NaN.
__device__ __forceinline__ float f32_step(float f, bool up) {
    if (f == 0.0f) return up ? 0x1p-149f : -0x1p-149f;
    int b = __builtin_bit_cast(int, f);
    b += ((f > 0.0f) == up) ? 1 : -1;
    return __builtin_bit_cast(float, b);
}

// Conservative FP32 bounds of a double (below / above).
__device__ __forceinline__ float f32_below(double x) {
    const float f = (float)x;
    return (double)f > x ? f32_step(f, false) : f;
}

__device__ __forceinline__ float f32_above(double x) {
    const float f = (float)x;
    return (double)f < x ? f32_step(f, true) : f;
}

// FP32 screen of a candidate: s32 is x_i - x_j, imaged and squared in FP32
// from FP32 copies of the positions.  With u = 2^-24, B = max |coordinate|
// and L the largest box length, each |d| component is within
// delta = 16 u (B + L) of its FP64 value (rounding of both positions, the
// difference, the box length and the shift, and an image chosen differently
// within that distance of L/2, where |d| and |d - L| agree), so
// |s32 - s| <= delta (2 sqrt(3 s) + 3 delta) + 4 u s.  With M twice that
// bound at s = T, s32 in (M, T - M) is a hit and s32 >= T + M a miss for
// certain; anything else (and NaN) is settled by the FP64 test.
struct Screen {
    float hit_lo, hit_hi, miss;
};

__device__ __forceinline__ Screen make_screen(double thr, double bmax, const BoxArg& box) {
    const double u = 0x1p-24;
    const double L = box.periodic ? fmax(fmax(box.L[0], box.L[1]), box.L[2]) : 0.0;
    const double dl = 16.0 * u * (bmax + L);
    const double M = 2.0 * (dl * (2.0 * sqrt(3.0 * thr) + 3.0 * dl) + 4.0 * u * thr);
    Screen sc;
    sc.hit_lo = f32_above(M);
    sc.hit_hi = f32_below(thr - M);
    sc.miss = f32_above(thr + M);
    if (!(M == M)) { sc.hit_lo = INFINITY; sc.hit_hi = -INFINITY; sc.miss = INFINITY; }
    return sc;
}

constexpr int kCand = 1536;   // candidates of one cell staged in LDS, large configuration (a multiple of 64)

// One workgroup per cell: the candidate ranges of the cell's stencil (the
// same for all its centres) are built once, the candidates (FP32 position and
// index) staged in LDS when they fit, and the workgroup's wavefronts take
// the cell's centres in turn, 64 candidates per step.  A cell with more than
// kCand candidates reads them from global memory instead (index by binary
// search over the ranges).  Rows come out in candidate order: type-major,
// stencil cells ascending, i.e. ascending internal index.
// CAND candidates are staged in LDS per cell (a multiple of 64) and THREADS
// threads (THREADS / 64 wavefronts) share a cell: small cells (short cutoffs)
// take a small configuration so that more cells are resident per CU.
template <int PASS, int CAND, int THREADS>
__global__ __launch_bounds__(THREADS) void k_nbr_cell(const int* __restrict__ gate,
                                                  const long long* __restrict__ total_and_max, long long cap,
                                                  const double* __restrict__ pos4,
                                                  const float4* __restrict__ pos4f,
                                                  const unsigned long long* __restrict__ bmax,
                                                  const int* __restrict__ cell_start, Grid g,
                                                  BoxArg box, int n_types, double thr,
                                                  int* __restrict__ counts,
                                                  const long long* __restrict__ row_off,
                                                  int* __restrict__ out_idx,
                                                  int* __restrict__ heads) {
    // out_idx: the CSR neighbour array (FILL) or the n x kStageW stage (STAGE);
    // heads (STAGE): the first entry of each bucket of kBucket staged entries.
    // FILL with total_and_max set (a launch the host did not decide on): only
    // when some row is longer than the stage and the list fits the arrays.
    constexpr bool FILL = PASS == PASS_FILL;
    if (gated_off(gate)) return;
    if (FILL && total_and_max && (total_and_max[1] <= kStageW || total_and_max[0] > cap)) return;
    __shared__ int4 s_cand[CAND];             // (x, y, z) as FP32 bits, internal index
    __shared__ int s_cells[32];
    __shared__ int s_runf[32], s_runl[32];    // first / last cell of each run
    __shared__ int s_cimg[32], s_rimg[32];    // periodic image of each stencil cell / run (code below)
    __shared__ int s_nrun;
    __shared__ int s_rstart[kMaxRanges];
    __shared__ int s_rpre[kMaxRanges + 1];
    __shared__ int s_cpre[kMaxTypes + 1];     // centres of this cell, by type
    __shared__ int s_cbeg[kMaxTypes];
    __shared__ int s_nr;
    const int lane = threadIdx.x & (kWave - 1);
    const int w = threadIdx.x >> 6;
    const int cell = xcd_block();

    if (threadIdx.x == 0) {
        int acc = 0;
        s_cpre[0] = 0;
        for (int t = 0; t < n_types; ++t) {
            const int k = t * g.ncell + cell;
            s_cbeg[t] = cell_start[k];
            acc += cell_start[k + 1] - cell_start[k];
            s_cpre[t + 1] = acc;
        }
    }
    __syncthreads();
    const int ncent = s_cpre[n_types];
    if (ncent == 0) return;

    if (w == 0) {
        const int c0 = cell % g.nc[0];
        const int c1 = (cell / g.nc[0]) % g.nc[1];
        const int c2 = cell / (g.nc[0] * g.nc[1]);
        // stencil cells, sorted ascending
        const int ns = g.nst[0] * g.nst[1] * g.nst[2];
        int myc = -1;
        int img = 13;   // (sx + 1) + 3 (sy + 1) + 9 (sz + 1): the cell's positions are seen shifted by s L
        if (lane < ns) {
            int a = lane % g.nst[0], b = (lane / g.nst[0]) % g.nst[1], c = lane / (g.nst[0] * g.nst[1]);
            int x = c0 + g.off[0][a], y = c1 + g.off[1][b], z = c2 + g.off[2][c];
            bool ok = true;
            if (g.periodic) {
                const int sx = x < 0 ? -1 : (x >= g.nc[0] ? 1 : 0);
                const int sy = y < 0 ? -1 : (y >= g.nc[1] ? 1 : 0);
                const int sz = z < 0 ? -1 : (z >= g.nc[2] ? 1 : 0);
                img = (sx + 1) + 3 * (sy + 1) + 9 * (sz + 1);
                x = (x + g.nc[0]) % g.nc[0]; y = (y + g.nc[1]) % g.nc[1]; z = (z + g.nc[2]) % g.nc[2];
            } else {
                ok = x >= 0 && x < g.nc[0] && y >= 0 && y < g.nc[1] && z >= 0 && z < g.nc[2];
            }
            myc = ok ? (z * g.nc[1] + y) * g.nc[0] + x : -1;
        }
        int rank = 0;
        for (int m = 0; m < ns; ++m) {
            int o = __shfl(myc, m, kWave);
            rank += (o >= 0 && o < myc) ? 1 : 0;
        }
        const int nvalid = __popcll(__ballot(myc >= 0));
        if (myc >= 0) { s_cells[rank] = myc; s_cimg[rank] = img; }
        wave_sync();
        // runs of consecutive cell ids (x-adjacent cells) are contiguous per type; a
        // run also keeps to one periodic image (the last cell of a row and the first
        // of the next have consecutive ids but are seen in different images)
        const bool in = lane < nvalid;
        const int cv = in ? s_cells[lane] : 0;
        const int ci = in ? s_cimg[lane] : 0;
        const bool first = in && (lane == 0 || s_cells[lane - 1] != cv - 1 || s_cimg[lane - 1] != ci);
        const bool last = in && (lane + 1 == nvalid || s_cells[lane + 1] != cv + 1 || s_cimg[lane + 1] != ci);
        const unsigned long long fm = __ballot(first);
        const int nrun = __popcll(fm);
        const int run = __popcll(fm & ((2ull << lane) - 1ull)) - 1;   // lane < 32
        if (first) { s_runf[run] = cv; s_rimg[run] = ci; }
        if (last) s_runl[run] = cv;
        wave_sync();
        // ranges (type-major, runs ascending) and their inclusive prefix
        const int nr = n_types * nrun;
        int carry = 0;
        for (int base = 0; base < nr; base += kWave) {
            const int r = base + lane;
            int st = 0, len = 0;
            if (r < nr) {
                const int t = r / nrun, q = r - (r / nrun) * nrun;
                const int k0 = t * g.ncell + s_runf[q], k1 = t * g.ncell + s_runl[q] + 1;
                st = cell_start[k0];
                len = cell_start[k1] - st;
            }
            int incl = len;
#pragma unroll
            for (int off = 1; off < kWave; off <<= 1) {
                const int o = __shfl_up(incl, off, kWave);
                if (lane >= off) incl += o;
            }
            if (r < nr) { s_rstart[r] = st; s_rpre[r + 1] = carry + incl; }
            carry += __shfl(incl, kWave - 1, kWave);
        }
        if (lane == 0) { s_rpre[0] = 0; s_nr = nr; s_nrun = nrun; }
    }
    __syncthreads();
    const int nr = s_nr;
    const int total = s_rpre[nr];
    const bool staged = total <= CAND;
    // Every run keeps to one periodic image (runs are split where it changes).  With
    // at least 4 cells per dimension the 27 stencil cells are distinct and at least
    // rc wide, so a pair within the cutoff is always seen in the image of the
    // stencil cell that holds it: the staged candidates are stored at that image and
    // the distance test needs no per-pair image.  A pair seen in another image is
    // farther apart than its minimum image, i.e. beyond the cutoff either way.
    const bool direct = staged && (!box.periodic || (g.nc[0] >= 4 && g.nc[1] >= 4 && g.nc[2] >= 4));
    const float3 bf = make_float3((float)box.L[0], (float)box.L[1], (float)box.L[2]);
    if (staged) {
        const int nrun = s_nrun;
        for (int t = threadIdx.x; t < total; t += THREADS) {
            int lo = 0, hi = nr;  // largest r with s_rpre[r] <= t
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (s_rpre[mid] <= t) lo = mid; else hi = mid;
            }
            const int j = s_rstart[lo] + (t - s_rpre[lo]);
            float4 p = pos4f[j];
            if (direct && box.periodic) {
                const int im = s_rimg[lo % nrun];
                const int sx = im % 3 - 1, sy = (im / 3) % 3 - 1, sz = im / 9 - 1;
                p.x += (float)sx * bf.x;
                p.y += (float)sy * bf.y;
                p.z += (float)sz * bf.z;
            }
            s_cand[t] = make_int4(__builtin_bit_cast(int, p.x), __builtin_bit_cast(int, p.y),
                                  __builtin_bit_cast(int, p.z), j);
        }
    }
    __syncthreads();

    const Screen sc = make_screen(thr, __builtin_bit_cast(double, *bmax), box);
    // FP32 image: d - L rint(d / L), the single-shift image whenever |d| < 1.5 L
    // (a candidate with a larger component takes the FP64 test)
    const float3 ib = make_float3(1.0f / bf.x, 1.0f / bf.y, 1.0f / bf.z);
    const float3 lim = make_float3(1.45f * bf.x, 1.45f * bf.y, 1.45f * bf.z);
    for (int k = w; k < ncent; k += THREADS / kWave) {
        int t0 = 0;
        while (t0 + 1 < n_types && s_cpre[t0 + 1] <= k) ++t0;
        const int i = uniform(s_cbeg[t0] + (k - s_cpre[t0]));
        const double4 pi = reinterpret_cast<const double4*>(pos4)[i];
        const float4 pif = pos4f[i];
        const long long out = FILL ? row_off[i] : 0;
        int count = 0;
        for (int base = 0; base < total; base += kWave) {
            const int t = base + lane;
            int j = i;
            float4 pj = pif;
            if (staged) {
                // t < CAND: total <= CAND and CAND is a multiple of the wave size
                const int4 c = s_cand[t];
                pj = make_float4(__builtin_bit_cast(float, c.x), __builtin_bit_cast(float, c.y),
                                 __builtin_bit_cast(float, c.z), 0.0f);
                j = c.w;
            } else if (t < total) {
                int lo = 0, hi = nr;
                while (hi - lo > 1) {
                    const int mid = (lo + hi) >> 1;
                    if (s_rpre[mid] <= t) lo = mid; else hi = mid;
                }
                j = s_rstart[lo] + (t - s_rpre[lo]);
                pj = pos4f[j];
            }
            const bool ok = (t < total) & (j != i);   // r_ii = 0 is never a neighbour
            float dx = pif.x - pj.x, dy = pif.y - pj.y, dz = pif.z - pj.z;
            bool img = true;
            if (box.periodic && !direct) {
                img = (fabsf(dx) < lim.x) & (fabsf(dy) < lim.y) & (fabsf(dz) < lim.z);
                dx = fmaf(-bf.x, rintf(dx * ib.x), dx);
                dy = fmaf(-bf.y, rintf(dy * ib.y), dy);
                dz = fmaf(-bf.z, rintf(dz * ib.z), dz);
            }
            const float s32 = dx * dx + dy * dy + dz * dz;
            bool hit = ok & img & (s32 > sc.hit_lo) & (s32 < sc.hit_hi);
            if (ok & !hit & (!img | !(s32 >= sc.miss))) {
                const double4 pd = reinterpret_cast<const double4*>(pos4)[j];
                const double s = mic_dist2(pi.x - pd.x, pi.y - pd.y, pi.z - pd.z, box);
                hit = (s <= thr) && (s > 0.0);   // r <= rc and r > 0
            }
            const unsigned long long bal = __ballot(hit);
            if (hit) {
                const int slot = count + lanes_below(bal);
                if (FILL) out_idx[out + slot] = j;
                if (PASS == PASS_STAGE && slot < kStageW) {
                    out_idx[(size_t)i * kStageW + slot] = j;
                    if (slot % kBucket == 0) heads[(size_t)i * kNBucket + slot / kBucket] = j;
                }
            }
            count += __popcll(bal);
        }
        if (!FILL && lane == 0) counts[i] = count;
        if (PASS == PASS_STAGE && count <= kStageW) {
            // INT_MAX past the row end, to the end of its last bucket, and in
            // the heads of unused buckets: the reverse search then needs no
            // row length (INT_MAX exceeds every atom index)
            const int nb = (count + kBucket - 1) / kBucket;
            if (count + lane < nb * kBucket) out_idx[(size_t)i * kStageW + count + lane] = INT_MAX;
            if (lane >= nb && lane < kNBucket) heads[(size_t)i * kNBucket + lane] = INT_MAX;
        }
    }
}

// ---- row offsets: exclusive scan of counts (int64) + max row ---------------

__global__ __launch_bounds__(256) void k_scan_reduce(const int* __restrict__ gate, int n, const int* counts,
                                                     long long* psum, int* pmax) {
    if (gated_off(gate)) return;
    __shared__ long long ss[256];
    __shared__ int sm[256];
    long long s = 0;
    int m = 0;
    int base = blockIdx.x * kScanBlock;
    for (int k = threadIdx.x; k < kScanBlock; k += 256) {
        int idx = base + k;
        if (idx < n) { int c = counts[idx]; s += c; m = max(m, c); }
    }
    ss[threadIdx.x] = s; sm[threadIdx.x] = m;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            ss[threadIdx.x] += ss[threadIdx.x + o];
            sm[threadIdx.x] = max(sm[threadIdx.x], sm[threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) { psum[blockIdx.x] = ss[0]; pmax[blockIdx.x] = sm[0]; }
}

// totals (when set): the running totals of an asynchronous build: [0] pairs of the last
// build, [1] largest row and [2] largest pair count since the caller cleared them,
// [3] builds done
__global__ __launch_bounds__(1024) void k_scan_top(const int* __restrict__ gate, int nb, long long* psum,
                                                   const int* pmax, long long* total_and_max,
                                                   long long* __restrict__ totals) {
    if (gated_off(gate)) return;
    __shared__ long long buf[1024];
    __shared__ long long carry;
    __shared__ int mx[1024];
    if (threadIdx.x == 0) carry = 0;
    int m = 0;
    __syncthreads();
    for (int base = 0; base < nb; base += 1024) {
        int idx = base + threadIdx.x;
        long long v = idx < nb ? psum[idx] : 0;
        if (idx < nb) m = max(m, pmax[idx]);
        buf[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            long long t = (int)threadIdx.x >= off ? buf[threadIdx.x - off] : 0;
            __syncthreads();
            buf[threadIdx.x] += t;
            __syncthreads();
        }
        if (idx < nb) psum[idx] = carry + buf[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry += buf[1023];
        __syncthreads();
    }
    mx[threadIdx.x] = m;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) mx[threadIdx.x] = max(mx[threadIdx.x], mx[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        total_and_max[0] = carry;
        total_and_max[1] = mx[0];
        if (totals) {
            totals[0] = carry;
            totals[1] = mx[0] > totals[1] ? mx[0] : totals[1];
            totals[2] = carry > totals[2] ? carry : totals[2];
            totals[3] += 1;
        }
    }
}

__global__ __launch_bounds__(256) void k_scan_down(const int* __restrict__ gate, int n, const int* counts,
                                                   const long long* psum, const long long* total_and_max,
                                                   long long* row_off) {
    if (gated_off(gate)) return;
    __shared__ long long buf[256];
    constexpr int kPer = kScanBlock / 256;
    int base = blockIdx.x * kScanBlock + threadIdx.x * kPer;
    long long loc[kPer];
    long long s = 0;
    for (int k = 0; k < kPer; ++k) {
        int idx = base + k;
        loc[k] = s;
        s += idx < n ? counts[idx] : 0;
    }
    buf[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        long long t = (int)threadIdx.x >= off ? buf[threadIdx.x - off] : 0;
        __syncthreads();
        buf[threadIdx.x] += t;
        __syncthreads();
    }
    long long pre = psum[blockIdx.x] + buf[threadIdx.x] - s;
    for (int k = 0; k < kPer; ++k) {
        int idx = base + k;
        if (idx < n) row_off[idx] = pre + loc[k];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) row_off[n] = total_and_max[0];
}

// ---- reverse edge index: rev[e(i->j)] = e(j->i) ------------------------------

__global__ __launch_bounds__(256) void k_rev(const int* __restrict__ gate,
                                             const long long* __restrict__ total_and_max, long long cap, int n,
                                             const long long* __restrict__ row_off,
                                             const int* __restrict__ nbr, int* __restrict__ rev) {
    // total_and_max set: a launch the host did not decide on (see k_nbr_cell)
    if (gated_off(gate)) return;
    if (total_and_max && (total_and_max[1] <= kStageW || total_and_max[0] > cap)) return;
    const int lane = threadIdx.x & (kWave - 1);
    const int i = uniform(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (i >= n) return;
    long long b = row_off[i], e = row_off[i + 1];
    for (long long k = b + lane; k < e; k += kWave) {
        int j = nbr[k];
        long long lo = row_off[j], hi = row_off[j + 1];
        while (lo < hi) {
            long long mid = (lo + hi) >> 1;
            if (nbr[mid] < i) lo = mid + 1; else hi = mid;
        }
        bool found = lo < row_off[j + 1] && nbr[lo] == i;
        rev[k] = found ? (int)lo : -1;  // -1 only if the list were asymmetric
    }
}

// Staged rows -> CSR, and the reverse index of every edge.  Wave i copies
// its row; for each j > i (a suffix of the ascending row) it finds i in the
// staged row of j and writes both rev entries, so each unordered pair is
// searched once.  The search is two 64-byte reads, each shared by the four
// lanes of a quad: the bucket heads of row j (first entry of every kBucket
// entries) select the bucket, then the bucket itself.  The relation is
// symmetric bit for bit (x_i - x_j and the single-shift image are exactly
// antisymmetric), so every entry is written.
__device__ __forceinline__ int count_lt(const int4& v, int base, int lim, int key) {
    return (base + 0 < lim && v.x < key) + (base + 1 < lim && v.y < key) + (base + 2 < lim && v.z < key) +
           (base + 3 < lim && v.w < key);
}

__device__ __forceinline__ int count_le(const int4& v, int base, int lim, int key) {
    return (base + 0 < lim && v.x <= key) + (base + 1 < lim && v.y <= key) + (base + 2 < lim && v.z <= key) +
           (base + 3 < lim && v.w <= key);
}

__device__ __forceinline__ bool has_eq(const int4& v, int base, int lim, int key) {
    return (base + 0 < lim && v.x == key) || (base + 1 < lim && v.y == key) || (base + 2 < lim && v.z == key) ||
           (base + 3 < lim && v.w == key);
}

__global__ __launch_bounds__(256) void k_compact_rev(const int* __restrict__ gate, int n,
                                                     const long long* __restrict__ row_off,
                                                     const int* __restrict__ counts,
                                                     const int* __restrict__ stage,
                                                     const int* __restrict__ heads, int* __restrict__ nbr,
                                                     int* __restrict__ rev,
                                                     const long long* __restrict__ total_and_max,
                                                     long long cap) {
    static_assert(kBucket == 16 && kStageW % kBucket == 0, "four int4 per bucket");
    if (gated_off(gate)) return;
    // launched before the host has read the totals: nothing to do unless the
    // list fits the arrays and every row was staged whole
    if (total_and_max[0] > cap || total_and_max[1] > kStageW) return;
    const int lane = threadIdx.x & (kWave - 1);
    const int i = uniform(xcd_block() * 4 + (threadIdx.x >> 6));
    if (i >= n) return;
    const long long oi = row_off[i];
    const int ci = counts[i];
    const int* si = stage + (size_t)i * kStageW;
    // the whole staged row in registers, slot s holding entry s * 64 + lane:
    // its loads are issued together rather than one round trip per slot
    constexpr int kSlots = kStageW / kWave;
    int js[kSlots];
#pragma unroll
    for (int s = 0; s < kSlots; ++s) {
        const int t = s * kWave + lane;
        js[s] = (s * kWave < ci && t < ci) ? si[t] : 0;
    }
    // count first, store after: the count waits once for every load, and the
    // stores then issue back to back (a store placed between them would be
    // waited for by the next slot's wait)
    int below = 0;   // entries j < i
#pragma unroll
    for (int s = 0; s < kSlots; ++s) {
        const int t = s * kWave + lane;
        below += __popcll(__ballot(t < ci && js[s] < i));
    }
#pragma unroll
    for (int s = 0; s < kSlots; ++s) {
        const int t = s * kWave + lane;
        if (t < ci) nbr[oi + t] = js[s];
    }
    if (rev == nullptr) return;
    // the entries j > i (positions below .. ci-1), sixteen per step: the four
    // lanes of a quad share an entry and each reads one int4 quarter of the
    // 64-byte heads and bucket, so one load instruction touches 16 rows of
    // the vector cache instead of 64.  Staged rows carry INT_MAX past their
    // end and in unused heads, so every compare is against the full 16 heads
    // and the full 16 entries of the bucket.
    const int q = lane & 3;
    const int na = ci - below;
    for (int base = 0; base < na; base += kWave / 4) {
        const int e = base + (lane >> 2);
        const bool act = e < na;
        const int t = below + e;
        int j = 0;   // staged entry t, from the slot registers
#pragma unroll
        for (int s = 0; s < kSlots; ++s) {
            const int v = __shfl(js[s], t & (kWave - 1));
            if ((t >> 6) == s) j = v;
        }
        long long oj = 0;
        int4 h = make_int4(INT_MAX, INT_MAX, INT_MAX, INT_MAX);
        if (act) {
            oj = row_off[j];
            h = reinterpret_cast<const int4*>(heads + (size_t)j * kNBucket)[q];
        }
        int c = count_le(h, 0, kNBucket, i);
        c += __shfl_xor(c, 1);
        c += __shfl_xor(c, 2);
        const int b = c - 1;   // bucket whose head is the last <= i
        int lt = 0, eq = 0;
        if (act && b >= 0) {
            const int o = b * kBucket;
            const int4 v = reinterpret_cast<const int4*>(stage + (size_t)j * kStageW + o)[q];
            lt = count_lt(v, 0, kStageW, i);
            eq = has_eq(v, 0, kStageW, i);
        }
        lt += __shfl_xor(lt, 1);
        lt += __shfl_xor(lt, 2);
        eq |= __shfl_xor(eq, 1);
        eq |= __shfl_xor(eq, 2);
        if (act && q == 0) {
            if (eq) {
                const int p = b * kBucket + lt;
                rev[oi + t] = (int)(oj + p);
                rev[oj + p] = (int)(oi + t);
            } else {
                rev[oi + t] = -1;   // only if the list were asymmetric
            }
        }
    }
}

__global__ __launch_bounds__(256) void k_mark_rows(int n_sel, const int* __restrict__ sel,
                                                   const long long* __restrict__ row_off,
                                                   const int* __restrict__ nbr, unsigned char* __restrict__ flags) {
    const int lane = threadIdx.x & (kWave - 1);
    const int q = uniform(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (q >= n_sel) return;
    const int i = sel[q];
    if (lane == 0) flags[i] = 1;
    for (long long k = row_off[i] + lane; k < row_off[i + 1]; k += kWave) flags[nbr[k]] = 1;
}

struct WsLayout {
    size_t keys_a, keys_b, vals_a, vals_b, ghist, cell_start, counts, psum, pmax, misc, stage, bucket, pos4f, total;
};

inline size_t align_up(size_t x) { return (x + 255) & ~size_t(255); }

WsLayout ws_layout(int n, int n_types) {
    WsLayout L{};
    size_t nb = (size_t)(n + kTile - 1) / kTile;
    size_t nsb = (size_t)(n + kScanBlock - 1) / kScanBlock;
    size_t o = 0;
    L.keys_a = o; o = align_up(o + sizeof(int) * (size_t)n);
    L.keys_b = o; o = align_up(o + sizeof(int) * (size_t)n);
    L.vals_a = o; o = align_up(o + sizeof(int) * (size_t)n);
    L.vals_b = o; o = align_up(o + sizeof(int) * (size_t)n);
    L.ghist = o; o = align_up(o + sizeof(int) * 256 * (nb > 0 ? nb : 1));
    L.cell_start = o; o = align_up(o + sizeof(int) * ((size_t)kMaxCells * n_types + 1));
    L.counts = o; o = align_up(o + sizeof(int) * (size_t)n);
    L.psum = o; o = align_up(o + sizeof(long long) * (nsb + 1));
    L.pmax = o; o = align_up(o + sizeof(int) * (nsb + 1));
    L.misc = o; o = align_up(o + 64 * sizeof(double));
    L.stage = o; o = align_up(o + sizeof(int) * (size_t)n * kStageW);
    L.bucket = o; o = align_up(o + sizeof(int) * (size_t)n * kNBucket);
    L.pos4f = o; o = align_up(o + sizeof(float4) * (size_t)n);
    L.total = o;
    return L;
}

// Cells per dimension: as many as fit with width > rc (strictly, with a
// relative margin so that rounding in the cell index never breaks the
// one-cell-stencil guarantee).
int cells_for(double extent, double rc) {
    if (!(extent > 0.0)) return 1;
    int nc = (int)std::floor(extent / rc);
    while (nc > 1 && extent / nc <= rc * (1.0 + 1e-9)) --nc;
    return nc < 1 ? 1 : nc;
}

// Largest double T with sqrt(T) <= rc (IEEE sqrt is correctly rounded and
// monotone, so sqrt(s) <= rc exactly when s <= T).
double dist2_threshold(double rc) {
    double t = rc * rc;
    while (t > 0.0 && std::sqrt(t) > rc) t = std::nextafter(t, 0.0);
    for (;;) {
        const double u = std::nextafter(t, INFINITY);
        if (std::sqrt(u) <= rc) t = u; else break;
    }
    return t;
}

}  // namespace
}  // namespace pt

using namespace pt;

extern "C" size_t pt_nbr_workspace_bytes(int n, int n_types) {
    if (n < 0 || n_types < 1) return 0;
    return ws_layout(n, n_types).total;
}

namespace pt {
namespace {

struct BuildOut {
    int* perm;
    double* pos4;
    long long* row_off;
    int* nbr;
    int* rev;
    long long cap;
};

// The whole build.  host_totals != nullptr: synchronous (the host reads the
// pair count and the longest row and decides on the second pass); otherwise
// every launch is enqueued unconditionally, behind `gate`, and the totals go
// to dev_totals.
int build_impl(int n, const double* pos, int pos_stride, const int* types, int n_types, const double* box,
               double rc, const BuildOut& o, void* ws, size_t ws_bytes, const int* gate,
               long long* host_totals, long long* dev_totals, int row_limit, hipStream_t stream) {
    WsLayout L = ws_layout(n, n_types);
    PT_REQUIRE(ws_bytes >= L.total, "pt_nbr_build: workspace too small (%zu < %zu)", ws_bytes, L.total);
    char* w = (char*)ws;
    int* keys_a = (int*)(w + L.keys_a);
    int* keys_b = (int*)(w + L.keys_b);
    int* vals_a = (int*)(w + L.vals_a);
    int* vals_b = (int*)(w + L.vals_b);
    int* ghist = (int*)(w + L.ghist);
    int* cell_start = (int*)(w + L.cell_start);
    int* counts = (int*)(w + L.counts);
    long long* psum = (long long*)(w + L.psum);
    int* pmax = (int*)(w + L.pmax);
    double* misc = (double*)(w + L.misc);
    int* stage = (int*)(w + L.stage);
    int* heads = (int*)(w + L.bucket);
    float4* pos4f = (float4*)(w + L.pos4f);
    unsigned long long* bmax = (unsigned long long*)(misc + 8);   // misc[0..5]: bbox, then total / max row
    const double thr = dist2_threshold(rc);
    const bool sync = host_totals != nullptr;

    Grid g{};
    BoxArg ba{};
    g.periodic = box != nullptr;
    ba.periodic = g.periodic;
    double extent[3];
    if (g.periodic) {
        for (int d = 0; d < 3; ++d) {
            PT_REQUIRE(box[d] > 0.0, "pt_nbr_build: box lengths must be > 0");
            extent[d] = box[d];
            g.lo[d] = 0.0;
            ba.L[d] = box[d];
            ba.half[d] = 0.5 * box[d];
        }
    } else {
        // the grid of a non-periodic structure follows its bounding box, read back here
        PT_REQUIRE(sync, "pt_nbr_build_async: periodic boxes only");
        hipLaunchKernelGGL(k_bbox, dim3(1), dim3(256), 0, stream, n, pos, pos_stride, misc);
        if (int rcode = launch_check("k_bbox")) return rcode;
        double bb[6];
        PT_CHECK_HIP(hipMemcpyAsync(bb, misc, sizeof(bb), hipMemcpyDeviceToHost, stream));
        PT_CHECK_HIP(hipStreamSynchronize(stream));
        for (int d = 0; d < 3; ++d) {
            PT_REQUIRE(std::isfinite(bb[d]) && std::isfinite(bb[3 + d]),
                       "pt_nbr_build: non-finite positions");
            extent[d] = bb[3 + d] - bb[d];
            g.lo[d] = bb[d];
        }
    }
    for (int d = 0; d < 3; ++d) g.nc[d] = cells_for(extent[d], rc);
    while ((long long)g.nc[0] * g.nc[1] * g.nc[2] > kMaxCells) {
        int dm = 0;
        for (int d = 1; d < 3; ++d) if (g.nc[d] > g.nc[dm]) dm = d;
        g.nc[dm] = std::max(1, g.nc[dm] / 2);
    }
    g.ncell = g.nc[0] * g.nc[1] * g.nc[2];
    for (int d = 0; d < 3; ++d) {
        double width = extent[d] > 0.0 ? extent[d] / g.nc[d] : 1.0;
        g.inv_w[d] = 1.0 / width;
        if (g.periodic && g.nc[d] == 1) { g.nst[d] = 1; g.off[d][0] = 0; }
        else if (g.periodic && g.nc[d] == 2) { g.nst[d] = 2; g.off[d][0] = 0; g.off[d][1] = 1; }
        else { g.nst[d] = 3; g.off[d][0] = -1; g.off[d][1] = 0; g.off[d][2] = 1; }
    }
    const int nkeys = n_types * g.ncell;

    // bin + sort
    const int tb = 256;
    hipLaunchKernelGGL(k_bin, dim3((n + tb - 1) / tb), dim3(tb), 0, stream, gate, n, pos, pos_stride, types,
                       g, keys_a, vals_a, bmax);
    if (int rcode = launch_check("k_bin")) return rcode;
    int bits = 1;
    while ((1 << bits) < nkeys) ++bits;
    const int nblocks = (n + kTile - 1) / kTile;
    int* kin = keys_a; int* kout = keys_b; int* vin = vals_a; int* vout = vals_b;
    for (int shift = 0; shift < bits; shift += 8) {
        hipLaunchKernelGGL(k_radix_hist, dim3(nblocks), dim3(kSortThreads), 0, stream, gate, n, kin,
                           shift, nblocks, ghist);
        hipLaunchKernelGGL(k_scan_small, dim3(1), dim3(1024), 0, stream, gate, 256 * nblocks, ghist);
        hipLaunchKernelGGL(k_radix_scatter, dim3(nblocks), dim3(kSortThreads), 0, stream, gate, n, kin,
                           vin, shift, nblocks, ghist, kout, vout);
        if (int rcode = launch_check("radix sort")) return rcode;
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    // kin/vin now hold the sorted keys / permutation (copied to perm by k_gather_pos4)
    hipLaunchKernelGGL(k_cell_start, dim3((nkeys + 1 + tb - 1) / tb), dim3(tb), 0, stream, gate, nkeys, n,
                       kin, cell_start);
    hipLaunchKernelGGL(k_gather_pos4, dim3((n + tb - 1) / tb), dim3(tb), 0, stream, gate, n, vin, pos,
                       pos_stride, types, o.pos4, pos4f, bmax, o.perm);
    if (int rcode = launch_check("cell start / gather")) return rcode;

    // count + stage, scan, then compact (or a second pass for rows > kStageW)
    const int nwb = (n + 3) / 4;
    // candidates per cell on average: atoms per cell times the stencil's cells; a cell
    // above the chosen configuration's LDS stage reads its candidates from HBM instead
    const double avg_cand = (double)n / g.ncell * g.nst[0] * g.nst[1] * g.nst[2];
    const int ccfg = avg_cand * 1.6 <= 512 ? 0 : (avg_cand * 1.6 <= 1024 ? 1 : 2);
    auto launch_cell = [&](bool fill, const long long* tm, const long long* roff, int* out, int* hd) {
        auto go = [&](auto kern, int threads) {
            hipLaunchKernelGGL(kern, dim3(g.ncell), dim3(threads), 0, stream, gate, tm, o.cap, o.pos4,
                               (const float4*)pos4f, (const unsigned long long*)bmax, cell_start, g, ba, n_types,
                               thr, counts, roff, out, hd);
        };
        if (ccfg == 0) {
            if (fill) go(k_nbr_cell<PASS_FILL, 512, 128>, 128); else go(k_nbr_cell<PASS_STAGE, 512, 128>, 128);
        } else if (ccfg == 1) {
            if (fill) go(k_nbr_cell<PASS_FILL, 1024, 256>, 256); else go(k_nbr_cell<PASS_STAGE, 1024, 256>, 256);
        } else {
            if (fill) go(k_nbr_cell<PASS_FILL, kCand, kCellThreads>, kCellThreads);
            else go(k_nbr_cell<PASS_STAGE, kCand, kCellThreads>, kCellThreads);
        }
    };
    launch_cell(false, nullptr, nullptr, stage, heads);
    if (int rcode = launch_check("k_nbr_cell<stage>")) return rcode;
    const int nsb = (n + kScanBlock - 1) / kScanBlock;
    long long* tot_max = (long long*)misc;
    hipLaunchKernelGGL(k_scan_reduce, dim3(nsb), dim3(256), 0, stream, gate, n, counts, psum, pmax);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, stream, gate, nsb, psum, pmax, tot_max, dev_totals);
    hipLaunchKernelGGL(k_scan_down, dim3(nsb), dim3(256), 0, stream, gate, n, counts, psum, tot_max, o.row_off);
    if (int rcode = launch_check("row offsets")) return rcode;
    // the compaction checks the totals itself, so it is queued before the host
    // reads them (the GPU does not idle through the round trip)
    if (o.nbr) {
        hipLaunchKernelGGL(k_compact_rev, dim3(nwb), dim3(256), 0, stream, gate, n, (const long long*)o.row_off,
                           (const int*)counts, (const int*)stage, (const int*)heads, o.nbr, o.rev,
                           (const long long*)tot_max, o.cap);
        if (int rcode = launch_check("k_compact_rev")) return rcode;
    }
    if (!sync) {
        // rows longer than the stage: the second pass decides on the device whether it runs; it is
        // not enqueued when the caller reads no row that long (such a list is reported in totals[1])
        if (o.nbr && row_limit > kStageW) {
            launch_cell(true, (const long long*)tot_max, (const long long*)o.row_off, o.nbr, nullptr);
            if (o.rev)
                hipLaunchKernelGGL(k_rev, dim3(nwb), dim3(256), 0, stream, gate, (const long long*)tot_max, o.cap,
                                   n, (const long long*)o.row_off, (const int*)o.nbr, o.rev);
        }
        return launch_check("gated build tail");
    }
    // synchronous: the totals are read through the caller's stack (a pageable
    // copy, ordered on the stream) -- no pinned buffer, no event
    long long tm[2] = {0, 0};
    PT_CHECK_HIP(hipMemcpyAsync(tm, tot_max, sizeof(tm), hipMemcpyDeviceToHost, stream));
    PT_CHECK_HIP(hipStreamSynchronize(stream));
    host_totals[0] = tm[0];
    host_totals[1] = tm[1];
    if (tm[0] > o.cap) {
        set_error("pt_nbr_build: capacity %lld < %lld pairs", o.cap, tm[0]);
        return PT_ERR_CAPACITY;
    }
    PT_REQUIRE(o.nbr != nullptr || tm[0] == 0, "pt_nbr_build: nbr is null");
    if (tm[1] <= kStageW) return PT_OK;
    launch_cell(true, nullptr, (const long long*)o.row_off, o.nbr, nullptr);
    if (int rcode = launch_check("k_nbr_cell<fill>")) return rcode;
    if (o.rev) {
        hipLaunchKernelGGL(k_rev, dim3(nwb), dim3(256), 0, stream, gate, (const long long*)nullptr, o.cap, n,
                           (const long long*)o.row_off, (const int*)o.nbr, o.rev);
        if (int rcode = launch_check("k_rev")) return rcode;
    }
    return PT_OK;
}

}  // namespace
}  // namespace pt

extern "C" int pt_nbr_build(int n, const double* pos, int pos_stride, const int* types, int n_types,
                            const double* box, double rc, int* perm, double* pos4,
                            long long* row_off, int* nbr, int* rev, long long cap,
                            long long* n_pairs, int* max_row, void* ws, size_t ws_bytes,
                            pt_stream_t stream_) {
    PT_REQUIRE(n >= 1, "pt_nbr_build: n must be >= 1 (got %d)", n);
    PT_REQUIRE(n_types >= 1 && n_types <= kMaxTypes, "pt_nbr_build: n_types must be in [1, %d]",
               kMaxTypes);
    PT_REQUIRE(rc > 0.0 && std::isfinite(rc), "pt_nbr_build: rc must be finite and > 0");
    PT_REQUIRE(pos && types && perm && pos4 && row_off && n_pairs && max_row && ws,
               "pt_nbr_build: null pointer argument");
    PT_REQUIRE(pos_stride >= 3, "pt_nbr_build: pos_stride must be >= 3");
    PT_REQUIRE((const void*)pos != (const void*)pos4, "pt_nbr_build: pos and pos4 must not alias");
    BuildOut o{perm, pos4, row_off, nbr, rev, cap};
    long long tm[2] = {0, 0};
    const int rc_ = build_impl(n, pos, pos_stride, types, n_types, box, rc, o, ws, ws_bytes, nullptr, tm,
                               nullptr, INT_MAX, (hipStream_t)stream_);
    if (rc_ == PT_OK || rc_ == PT_ERR_CAPACITY) {
        *n_pairs = tm[0];
        *max_row = (int)tm[1];
    }
    return rc_;
}

extern "C" int pt_nbr_build_async(int n, const double* pos, int pos_stride, const int* types, int n_types,
                                  const double* box, double rc, int* perm, double* pos4, long long* row_off,
                                  int* nbr, int* rev, long long cap, const int* gate, long long* totals,
                                  int row_limit, void* ws, size_t ws_bytes, pt_stream_t stream_) {
    PT_REQUIRE(n >= 1, "pt_nbr_build_async: n must be >= 1 (got %d)", n);
    PT_REQUIRE(n_types >= 1 && n_types <= kMaxTypes, "pt_nbr_build_async: n_types must be in [1, %d]",
               kMaxTypes);
    PT_REQUIRE(rc > 0.0 && std::isfinite(rc), "pt_nbr_build_async: rc must be finite and > 0");
    PT_REQUIRE(box != nullptr, "pt_nbr_build_async: periodic boxes only");
    PT_REQUIRE(pos && types && perm && pos4 && row_off && nbr && totals && ws,
               "pt_nbr_build_async: null pointer argument");
    PT_REQUIRE(pos_stride >= 3, "pt_nbr_build_async: pos_stride must be >= 3");
    PT_REQUIRE((const void*)pos != (const void*)pos4, "pt_nbr_build_async: pos and pos4 must not alias");
    BuildOut o{perm, pos4, row_off, nbr, rev, cap};
    return build_impl(n, pos, pos_stride, types, n_types, box, rc, o, ws, ws_bytes, gate, nullptr, totals,
                      row_limit, (hipStream_t)stream_);
}

extern "C" int pt_nbr_mark_rows(int n_sel, const int* sel, const long long* row_off, const int* nbr,
                                unsigned char* flags, pt_stream_t stream) {
    PT_REQUIRE(n_sel >= 0 && (n_sel == 0 || (sel && row_off && nbr && flags)), "pt_nbr_mark_rows: bad arguments");
    if (n_sel == 0) return PT_OK;
    hipLaunchKernelGGL(k_mark_rows, dim3((n_sel + 3) / 4), dim3(256), 0, (hipStream_t)stream, n_sel, sel, row_off,
                       nbr, flags);
    return launch_check("k_mark_rows");
}
