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

__global__ void k_bin(int n, const double* pos, int ps, const int* types, Grid g, int* keys, int* vals) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int c0 = cell_coord(pos[(size_t)ps * i + 0], g.lo[0], g.inv_w[0], g.nc[0], g.periodic);
    int c1 = cell_coord(pos[(size_t)ps * i + 1], g.lo[1], g.inv_w[1], g.nc[1], g.periodic);
    int c2 = cell_coord(pos[(size_t)ps * i + 2], g.lo[2], g.inv_w[2], g.nc[2], g.periodic);
    int cell = (c2 * g.nc[1] + c1) * g.nc[0] + c0;
    keys[i] = (types[i] - 1) * g.ncell + cell;
    vals[i] = i;
}

// ---- stable LSD radix sort, 8-bit digits -----------------------------------

__global__ __launch_bounds__(kSortThreads) void k_radix_hist(int n, const int* keys, int shift,
                                                             int nblocks, int* ghist) {
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

// Exclusive scan of m ints in one block (m up to a few 1e5).
__global__ __launch_bounds__(1024) void k_scan_small(int m, int* data) {
    __shared__ int buf[1024];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < m; base += 1024) {
        int idx = base + threadIdx.x;
        int v = idx < m ? data[idx] : 0;
        buf[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            int t = (int)threadIdx.x >= off ? buf[threadIdx.x - off] : 0;
            __syncthreads();
            buf[threadIdx.x] += t;
            __syncthreads();
        }
        int incl = buf[threadIdx.x];
        if (idx < m) data[idx] = carry + incl - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry += buf[1023];
        __syncthreads();
    }
}

__global__ __launch_bounds__(kSortThreads) void k_radix_scatter(int n, const int* keys_in,
                                                                const int* vals_in, int shift,
                                                                int nblocks, const int* gscan,
                                                                int* keys_out, int* vals_out) {
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

__global__ void k_cell_start(int nkeys, int n, const int* skeys, int* cell_start) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > nkeys) return;
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (skeys[mid] < k) lo = mid + 1; else hi = mid;
    }
    cell_start[k] = lo;
}

__global__ void k_gather_pos4(int n, const int* perm, const double* pos, int ps, const int* types,
                              double* pos4) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int v = perm[i];
    double4 p;
    p.x = pos[(size_t)ps * v]; p.y = pos[(size_t)ps * v + 1]; p.z = pos[(size_t)ps * v + 2];
    p.w = (double)types[v];
    reinterpret_cast<double4*>(pos4)[i] = p;
}

// ---- one wavefront per centre: count (FILL=false) or fill (FILL=true) -----

template <bool FILL>
__global__ __launch_bounds__(256) void k_nbr_pass(int n, const double* __restrict__ pos4,
                                                  const int* __restrict__ skeys,
                                                  const int* __restrict__ cell_start, Grid g,
                                                  BoxArg box, int n_types, double rc,
                                                  int* __restrict__ counts,
                                                  const long long* __restrict__ row_off,
                                                  int* __restrict__ nbr) {
    __shared__ int s_cells[4][32];
    __shared__ int s_rstart[4][kMaxRanges];
    __shared__ int s_rpre[4][kMaxRanges + 1];
    const int lane = threadIdx.x & (kWave - 1);
    const int w = threadIdx.x >> 6;
    const int i = uniform(blockIdx.x * 4 + w);
    if (i >= n) return;

    const double4 pi = reinterpret_cast<const double4*>(pos4)[i];
    const int cell = uniform(skeys[i] % g.ncell);
    const int c0 = cell % g.nc[0];
    const int c1 = (cell / g.nc[0]) % g.nc[1];
    const int c2 = cell / (g.nc[0] * g.nc[1]);

    // stencil cells of this centre, sorted ascending
    const int ns = g.nst[0] * g.nst[1] * g.nst[2];
    int myc = -1;
    if (lane < ns) {
        int a = lane % g.nst[0], b = (lane / g.nst[0]) % g.nst[1], c = lane / (g.nst[0] * g.nst[1]);
        int x = c0 + g.off[0][a], y = c1 + g.off[1][b], z = c2 + g.off[2][c];
        bool ok = true;
        if (g.periodic) {
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
    if (myc >= 0) s_cells[w][rank] = myc;
    wave_sync();

    // candidate ranges: type-major, cells ascending
    const int nr = n_types * nvalid;
    for (int r = lane; r < nr; r += kWave) {
        int t = r / nvalid, q = r - t * nvalid;
        int k = t * g.ncell + s_cells[w][q];
        int st = cell_start[k];
        s_rstart[w][r] = st;
        s_rpre[w][r + 1] = cell_start[k + 1] - st;
    }
    if (lane == 0) s_rpre[w][0] = 0;
    wave_sync();
    // inclusive prefix over s_rpre[1..nr]
    int carry = 0;
    for (int base = 0; base < nr; base += kWave) {
        int r = base + lane;
        int v = r < nr ? s_rpre[w][r + 1] : 0;
        int incl = v;
#pragma unroll
        for (int off = 1; off < kWave; off <<= 1) {
            int t = __shfl_up(incl, off, kWave);
            if (lane >= off) incl += t;
        }
        if (r < nr) s_rpre[w][r + 1] = carry + incl;
        carry += __shfl(incl, kWave - 1, kWave);
    }
    wave_sync();
    const int total = carry;

    long long out = FILL ? row_off[i] : 0;
    int count = 0;
    for (int base = 0; base < total; base += kWave) {
        int t = base + lane;
        bool hit = false;
        int j = 0;
        if (t < total) {
            int lo = 0, hi = nr;  // largest r with s_rpre[r] <= t
            while (hi - lo > 1) {
                int mid = (lo + hi) >> 1;
                if (s_rpre[w][mid] <= t) lo = mid; else hi = mid;
            }
            j = s_rstart[w][lo] + (t - s_rpre[w][lo]);
            const double4 pj = reinterpret_cast<const double4*>(pos4)[j];
            double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
            double r = exact_dist(dx, dy, dz, box);
            hit = (r <= rc) && (r > 0.0);
        }
        unsigned long long bal = __ballot(hit);
        if (FILL && hit) nbr[out + count + lanes_below(bal)] = j;
        count += __popcll(bal);
    }
    if (!FILL && lane == 0) counts[i] = count;
}

// ---- row offsets: exclusive scan of counts (int64) + max row ---------------

__global__ __launch_bounds__(256) void k_scan_reduce(int n, const int* counts, long long* psum,
                                                     int* pmax) {
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

__global__ __launch_bounds__(1024) void k_scan_top(int nb, long long* psum, const int* pmax,
                                                   long long* total_and_max) {
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
    if (threadIdx.x == 0) { total_and_max[0] = carry; total_and_max[1] = mx[0]; }
}

__global__ __launch_bounds__(256) void k_scan_down(int n, const int* counts, const long long* psum,
                                                   const long long* total_and_max,
                                                   long long* row_off) {
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

__global__ __launch_bounds__(256) void k_rev(int n, const long long* __restrict__ row_off,
                                             const int* __restrict__ nbr, int* __restrict__ rev) {
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
    size_t keys_a, keys_b, vals_a, vals_b, ghist, cell_start, counts, psum, pmax, misc, total;
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

}  // namespace
}  // namespace pt

using namespace pt;

extern "C" size_t pt_nbr_workspace_bytes(int n, int n_types) {
    if (n < 0 || n_types < 1) return 0;
    return ws_layout(n, n_types).total;
}

extern "C" int pt_nbr_build(int n, const double* pos, int pos_stride, const int* types, int n_types,
                            const double* box, double rc, int* perm, double* pos4,
                            long long* row_off, int* nbr, int* rev, long long cap,
                            long long* n_pairs, int* max_row, void* ws, size_t ws_bytes,
                            pt_stream_t stream_) {
    hipStream_t stream = (hipStream_t)stream_;
    PT_REQUIRE(n >= 1, "pt_nbr_build: n must be >= 1 (got %d)", n);
    PT_REQUIRE(n_types >= 1 && n_types <= kMaxTypes, "pt_nbr_build: n_types must be in [1, %d]",
               kMaxTypes);
    PT_REQUIRE(rc > 0.0, "pt_nbr_build: rc must be > 0");
    PT_REQUIRE(pos && types && perm && pos4 && row_off && n_pairs && max_row && ws,
               "pt_nbr_build: null pointer argument");
    PT_REQUIRE(pos_stride >= 3, "pt_nbr_build: pos_stride must be >= 3");
    PT_REQUIRE((const void*)pos != (const void*)pos4, "pt_nbr_build: pos and pos4 must not alias");
    WsLayout L = ws_layout(n, n_types);
    PT_REQUIRE(ws_bytes >= L.total, "pt_nbr_build: workspace too small (%zu < %zu)", ws_bytes,
               L.total);
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
    hipLaunchKernelGGL(k_bin, dim3((n + tb - 1) / tb), dim3(tb), 0, stream, n, pos, pos_stride, types,
                       g, keys_a, vals_a);
    if (int rcode = launch_check("k_bin")) return rcode;
    int bits = 1;
    while ((1 << bits) < nkeys) ++bits;
    const int nblocks = (n + kTile - 1) / kTile;
    int* kin = keys_a; int* kout = keys_b; int* vin = vals_a; int* vout = vals_b;
    for (int shift = 0; shift < bits; shift += 8) {
        hipLaunchKernelGGL(k_radix_hist, dim3(nblocks), dim3(kSortThreads), 0, stream, n, kin,
                           shift, nblocks, ghist);
        hipLaunchKernelGGL(k_scan_small, dim3(1), dim3(1024), 0, stream, 256 * nblocks, ghist);
        hipLaunchKernelGGL(k_radix_scatter, dim3(nblocks), dim3(kSortThreads), 0, stream, n, kin,
                           vin, shift, nblocks, ghist, kout, vout);
        if (int rcode = launch_check("radix sort")) return rcode;
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    // kin/vin now hold the sorted keys / permutation
    PT_CHECK_HIP(hipMemcpyAsync(perm, vin, sizeof(int) * (size_t)n, hipMemcpyDeviceToDevice, stream));
    hipLaunchKernelGGL(k_cell_start, dim3((nkeys + 1 + tb - 1) / tb), dim3(tb), 0, stream, nkeys, n,
                       kin, cell_start);
    hipLaunchKernelGGL(k_gather_pos4, dim3((n + tb - 1) / tb), dim3(tb), 0, stream, n, vin, pos,
                       pos_stride, types, pos4);
    if (int rcode = launch_check("cell start / gather")) return rcode;

    // count, scan, fill
    const int nwb = (n + 3) / 4;
    hipLaunchKernelGGL(k_nbr_pass<false>, dim3(nwb), dim3(256), 0, stream, n, pos4, kin, cell_start,
                       g, ba, n_types, rc, counts, (const long long*)nullptr, (int*)nullptr);
    if (int rcode = launch_check("k_nbr_pass<count>")) return rcode;
    const int nsb = (n + kScanBlock - 1) / kScanBlock;
    long long* tot_max = (long long*)misc;
    hipLaunchKernelGGL(k_scan_reduce, dim3(nsb), dim3(256), 0, stream, n, counts, psum, pmax);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, stream, nsb, psum, pmax, tot_max);
    hipLaunchKernelGGL(k_scan_down, dim3(nsb), dim3(256), 0, stream, n, counts, psum, tot_max,
                       row_off);
    if (int rcode = launch_check("row offsets")) return rcode;
    long long host_tm[2];
    PT_CHECK_HIP(hipMemcpyAsync(host_tm, tot_max, sizeof(host_tm), hipMemcpyDeviceToHost, stream));
    PT_CHECK_HIP(hipStreamSynchronize(stream));
    *n_pairs = host_tm[0];
    *max_row = (int)host_tm[1];
    if (host_tm[0] > cap) {
        set_error("pt_nbr_build: capacity %lld < %lld pairs", cap, host_tm[0]);
        return PT_ERR_CAPACITY;
    }
    PT_REQUIRE(nbr != nullptr || host_tm[0] == 0, "pt_nbr_build: nbr is null");
    hipLaunchKernelGGL(k_nbr_pass<true>, dim3(nwb), dim3(256), 0, stream, n, pos4, kin, cell_start,
                       g, ba, n_types, rc, counts, row_off, nbr);
    if (int rcode = launch_check("k_nbr_pass<fill>")) return rcode;
    if (rev) {
        hipLaunchKernelGGL(k_rev, dim3(nwb), dim3(256), 0, stream, n, row_off, nbr, rev);
        if (int rcode = launch_check("k_rev")) return rcode;
    }
    return PT_OK;
}

extern "C" int pt_nbr_mark_rows(int n_sel, const int* sel, const long long* row_off, const int* nbr,
                                unsigned char* flags, pt_stream_t stream) {
    PT_REQUIRE(n_sel >= 0 && (n_sel == 0 || (sel && row_off && nbr && flags)), "pt_nbr_mark_rows: bad arguments");
    if (n_sel == 0) return PT_OK;
    hipLaunchKernelGGL(k_mark_rows, dim3((n_sel + 3) / 4), dim3(256), 0, (hipStream_t)stream, n_sel, sel, row_off,
                       nbr, flags);
    return launch_check("k_mark_rows");
}
