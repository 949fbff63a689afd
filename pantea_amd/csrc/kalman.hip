// (f) Global extended Kalman filter update of the MLP weights on the device.
//
// Replaces KalmanFilter.update_state_vector
// (potentials/nnp/kalman_filter.py:196-233).  With n_w weights (all elements,
// one state vector) and m observations of one structure (m = 1 for the
// energy, m = 3 N for the force components):
//
//   X = P H                           n_w x m    rocBLAS dgemm
//   A = H^T X + r I                   m x m      rocBLAS dgemm + k_add_diag
//   A = L L^T                                    chol_lower (below)
//   Y = A^-1 X^T         (= K^T)      m x n_w    k_transpose + rocSOLVER dpotrs
//   P = c (P - Y^T X^T) + q I                    rocBLAS dgemm + k_add_diag
//   W = W + Y^T xi                               rocBLAS dgemv
//
// r is the measurement noise (1/eta for the filter of type 0, lambda for
// type 1), c the fading-memory factor (1 or 1/lambda) and q the process
// noise; the schedules of eta, q and lambda are host scalars.  The
// reference forms inv(A) and multiplies; the Cholesky solve gives the same
// K (A is symmetric positive definite: H^T P H is semi-definite and r > 0)
// without forming the inverse.
//
// With more observations than weights (m > n_w: the force update of a large
// structure) the same update is computed through n_w x n_w factors instead
// of m x m ones.  With P = L L^T and G = L^T H, A = G^T G + r I, and by the
// push-through identity K = X A^-1 = L S^-1 G with S = G G^T + r I (n_w x n_w),
// so that P - K X^T = r L S^-1 L^T.  With S = Ls Ls^T and M = L Ls^-T:
//
//   L = chol(P)                                  chol_lower (n_w)
//   G = L^T H                        n_w x m     rocBLAS dtrmm
//   S = G G^T + r I,  Ls = chol(S)   n_w x n_w   rocBLAS dsyrk + chol_lower
//   [M; y^T] = [L; (G xi)^T] Ls^-T               dgemv + dtrsm (one solve: y = Ls^-1 G xi)
//   W = W + M y  (= L S^-1 G xi)                 dgemv
//   P = c r M M^T + q I                          dsyrk + k_symm_lower
//
// about 2 n_w^2 m + 8/3 n_w^3 flops against 4 n_w^2 m + 4 m^2 n_w + m^3 / 3,
// and factorisations of n_w rows instead of m.  All matrices are column-major FP64:
// P (n_w x n_w, ld n_w) and H (n_w x m, ld ldh) -- H column o is the
// gradient of observation o, i.e. a row-major (m x n_w) Jacobian.
#include <mutex>
#include <vector>

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "common.h"

namespace pt {
namespace {

__global__ void k_add_diag(int n, double* __restrict__ A, long long lda, double v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) A[(size_t)i * lda + i] += v;
}

// M ((n + 1) x n, ld ldm > n) = the lower triangle of A (zeros above) with g^T as its last row
__global__ void k_lower_copy(int n, const double* __restrict__ A, long long lda, const double* __restrict__ g,
                             double* __restrict__ M, long long ldm) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;   // row
    const int j = blockIdx.y;                             // column
    if (i < n) M[(size_t)j * ldm + i] = i >= j ? A[(size_t)j * lda + i] : 0.0;
    else if (i == n) M[(size_t)j * ldm + n] = g[j];
}

// A = its lower triangle mirrored into the upper one, plus v on the diagonal
__global__ void k_symm_lower(int n, double* __restrict__ A, long long lda, double v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;   // row
    const int j = blockIdx.y;                             // column
    if (i >= n) return;
    if (i > j) A[(size_t)i * lda + j] = A[(size_t)j * lda + i];
    else if (i == j) A[(size_t)j * lda + i] += v;
}

// the first of two factorisations' info codes that is not zero
__global__ void k_merge_info(int* __restrict__ info, const int* __restrict__ info2) {
    if (threadIdx.x == 0 && *info == 0) *info = *info2;
}

__global__ void k_zero_int(int* __restrict__ p) {
    if (threadIdx.x == 0) *p = 0;
}

// ---- Cholesky factorisation A = L L^T (lower, in place), right-looking in
// blocks of kCholNb columns.  Each diagonal block is factorised by one
// wavefront (lane i holds row i in registers), which also writes the block's inverse;
// the panel below is then L_21 = A_21 L_11^-T, a GEMM with that inverse, and
// the trailing matrix takes A_22 -= L_21 L_21^T (rocBLAS dsyrk).  rocSOLVER's
// dpotrf factorises its diagonal panels in a kernel of ~156 us each, most of
// the time of a factorisation of a few thousand rows.
constexpr int kCholNb = 64;

__device__ __forceinline__ double readlane_f64(double v, int lane) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u & 0xffffffffu), lane);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), lane);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// Lane l holds row l of the block in registers (r[c] = A(k0 + l, k0 + c)); the
// column entries a step needs from other rows are read with v_readlane, so the
// factorisation and the inverse run without LDS or barriers.
__global__ __launch_bounds__(kCholNb) void k_chol_diag(int n, double* __restrict__ A, long long lda, int k0,
                                                       double* __restrict__ Linv, int* __restrict__ info) {
    const int l = threadIdx.x;
    const int nb = min(kCholNb, n - k0);
    double r[kCholNb];
#pragma unroll
    for (int c = 0; c < kCholNb; ++c)
        r[c] = (l < nb && c < nb && l >= c) ? A[(size_t)(k0 + c) * lda + k0 + l] : 0.0;
    int bad = 0;   // 1-based column of the first pivot that is not positive (the same in every lane)
#pragma unroll
    for (int j = 0; j < kCholNb; ++j) {
        if (j < nb) {
            const double ajj = readlane_f64(r[j], j);
            if (!(ajj > 0.0) && bad == 0) bad = j + 1;
            const double d = sqrt(ajj);
            const double inv = 1.0 / d;
            if (l == j) r[j] = d;
            else if (l > j) r[j] *= inv;
            // A(l, c) -= L(l, j) L(c, j) for j < c <= l
#pragma unroll
            for (int c = j + 1; c < kCholNb; ++c) {
                const double lcj = readlane_f64(r[j], c);
                if (l >= c) r[c] -= r[j] * lcj;
            }
        }
    }
#pragma unroll
    for (int c = 0; c < kCholNb; ++c)
        if (l < nb && c < nb && l >= c) A[(size_t)(k0 + c) * lda + k0 + l] = r[c];
    // inverse: lane l forms column l of X = L^-1 by forward substitution,
    // x[i] = (delta(i, l) - sum_{k < i} L(i, k) x[k]) / L(i, i)
    double x[kCholNb];
#pragma unroll
    for (int i = 0; i < kCholNb; ++i) {
        double t = (i == l) ? 1.0 : 0.0;
#pragma unroll
        for (int k = 0; k < i; ++k) t -= readlane_f64(r[k], i) * x[k];
        x[i] = (i < nb) ? t / readlane_f64(r[i], i) : 0.0;
    }
#pragma unroll
    for (int i = 0; i < kCholNb; ++i) Linv[(size_t)l * kCholNb + i] = (l < nb) ? x[i] : 0.0;
    if (l == 0 && bad != 0 && *info == 0) *info = k0 + bad;
}

// B (cols x rows, ld ldb) = A^T, A rows x cols (ld lda), 32 x 32 tiles through LDS
__global__ void k_transpose(int rows, int cols, const double* __restrict__ A, long long lda,
                            double* __restrict__ B, long long ldb) {
    __shared__ double t[32][33];
    const int r0 = blockIdx.x * 32, c0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8 threads
    for (int j = ty; j < 32; j += 8) {
        const int r = r0 + tx, c = c0 + j;
        if (r < rows && c < cols) t[j][tx] = A[(size_t)c * lda + r];
    }
    __syncthreads();
    for (int j = ty; j < 32; j += 8) {
        const int c = c0 + tx, r = r0 + j;
        if (r < rows && c < cols) B[(size_t)r * ldb + c] = t[tx][j];
    }
}

// xi = (ref - pred) * scale, loss += sum xi^2 (one block of 256 per 256 entries)
__global__ void k_residual(int n, const double* __restrict__ pred, const double* __restrict__ ref, double ref_scalar,
                           double scale, double* __restrict__ xi, double* __restrict__ loss) {
    __shared__ double part[4];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v = 0.0;
    if (i < n) {
        v = ((ref ? ref[i] : ref_scalar) - pred[i]) * scale;
        xi[i] = v;
    }
    double s = wave_sum(v * v);
    if (wave_lane() == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(loss, part[0] + part[1] + part[2] + part[3]);
}

std::mutex g_mu;
std::vector<rocblas_handle> g_handles;

rocblas_handle handle_for_current_device(int* rc) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        set_error("pt_kalman_update: hipGetDevice failed");
        *rc = PT_ERR_HIP;
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_mu);
    if ((int)g_handles.size() <= dev) g_handles.resize(dev + 1, nullptr);
    if (!g_handles[dev]) {
        rocblas_handle h = nullptr;
        if (rocblas_create_handle(&h) != rocblas_status_success) {
            set_error("pt_kalman_update: rocblas_create_handle failed");
            *rc = PT_ERR_HIP;
            return nullptr;
        }
        rocblas_set_pointer_mode(h, rocblas_pointer_mode_host);
        g_handles[dev] = h;
    }
    return g_handles[dev];
}

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// the n_w x n_w form when the observations outnumber the weights
bool factored_form(int n_w, int m) { return m > n_w; }

// workspace of chol_lower for n rows: the block inverse and the panel
size_t chol_ws_bytes(int n) { return align256(8 * (size_t)kCholNb * kCholNb) + align256(8 * (size_t)n * kCholNb); }

#define PT_BLAS(call, what)                                                            \
    do {                                                                               \
        rocblas_status st_ = (call);                                                   \
        if (st_ != rocblas_status_success) {                                           \
            set_error("pt_kalman_update: %s failed (%s)", what, rocblas_status_to_string(st_)); \
            return PT_ERR_HIP;                                                         \
        }                                                                              \
    } while (0)

// A (n x n, lower triangle, ld lda) = L L^T in place; *info = 0 or the 1-based
// column of the first pivot that is not positive (as dpotrf reports it).
int chol_lower(rocblas_handle h, hipStream_t st, int n, double* A, int lda, char* ws, int* info) {
    double* Linv = (double*)ws;
    double* T = (double*)(ws + align256(8 * (size_t)kCholNb * kCholNb));
    const double one = 1.0, zero = 0.0, mone = -1.0;
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, st, info);
    if (int e = launch_check("k_zero_int")) return e;
    for (int k0 = 0; k0 < n; k0 += kCholNb) {
        const int nb = std::min(kCholNb, n - k0);
        hipLaunchKernelGGL(k_chol_diag, dim3(1), dim3(kCholNb), 0, st, n, A, (long long)lda, k0, Linv, info);
        if (int e = launch_check("k_chol_diag")) return e;
        const int mk = n - k0 - nb;
        if (mk <= 0) break;
        double* A21 = A + (size_t)k0 * lda + k0 + nb;
        double* A22 = A + (size_t)(k0 + nb) * lda + k0 + nb;
        // T = A_21 L_11^-T, A_22 -= T T^T, A_21 = T
        PT_BLAS(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_transpose, mk, nb, nb, &one, A21, lda,
                              Linv, kCholNb, &zero, T, mk),
                "dgemm panel");
        PT_BLAS(rocblas_dsyrk(h, rocblas_fill_lower, rocblas_operation_none, mk, nb, &mone, T, mk, &one, A22, lda),
                "dsyrk trailing");
        if (hipMemcpy2DAsync(A21, (size_t)lda * 8, T, (size_t)mk * 8, (size_t)mk * 8, nb, hipMemcpyDeviceToDevice,
                             st) != hipSuccess) {
            set_error("pt_kalman_update: panel copy failed");
            return PT_ERR_HIP;
        }
    }
    return PT_OK;
}

}  // namespace
}  // namespace pt

using namespace pt;

extern "C" size_t pt_kalman_workspace_bytes(int n_w, int m) {
    if (n_w < 1 || m < 1) return 0;
    const size_t nw = (size_t)n_w, mm = (size_t)m;
    if (factored_form(n_w, m))   // G, S, [M; y^T], g, info, Cholesky workspace
        return align256(8 * nw * mm) + align256(8 * nw * nw) + align256(8 * (nw + 1) * nw) + align256(8 * nw) +
               align256(sizeof(int)) + chol_ws_bytes(n_w);
    return align256(8 * nw * mm) + align256(8 * mm * mm) + align256(8 * mm * nw) + chol_ws_bytes(m);
}

extern "C" int pt_kalman_residual(int n, const double* pred, const double* ref, double ref_scalar, double scale,
                                  double* xi, double* loss, pt_stream_t stream) {
    PT_REQUIRE(n >= 1 && pred && xi && loss, "pt_kalman_residual: bad arguments");
    hipLaunchKernelGGL(k_residual, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, n, pred, ref,
                       ref_scalar, scale, xi, loss);
    return launch_check("pt_kalman_residual");
}

extern "C" int pt_kalman_update(int n_w, int m, double* P, const double* H, long long ldh, const double* xi,
                                double* W, double r, double c, double q, void* ws, size_t ws_bytes, int* info,
                                pt_stream_t stream) {
    PT_REQUIRE(n_w >= 1 && m >= 1, "pt_kalman_update: n_w and m must be positive");
    PT_REQUIRE(P && H && xi && W && ws && info, "pt_kalman_update: null argument");
    PT_REQUIRE(ldh >= n_w, "pt_kalman_update: ldh < n_w");
    PT_REQUIRE(r > 0.0, "pt_kalman_update: the measurement noise must be positive");
    PT_REQUIRE(ws_bytes >= pt_kalman_workspace_bytes(n_w, m), "pt_kalman_update: workspace too small");
    PT_REQUIRE((size_t)n_w * n_w < (1ull << 31) && (factored_form(n_w, m) || (size_t)m * m < (1ull << 31)) &&
                   (size_t)n_w * m < (1ull << 31) && ldh < (1LL << 31),
               "pt_kalman_update: matrices beyond 32-bit BLAS indexing");
    int rc = PT_OK;
    rocblas_handle h = handle_for_current_device(&rc);
    if (!h) return rc;
    hipStream_t st = (hipStream_t)stream;
    PT_BLAS(rocblas_set_stream(h, st), "rocblas_set_stream");
    char* base = (char*)ws;
    const double one = 1.0, zero = 0.0, mc = -c;
    if (factored_form(n_w, m)) {
        const size_t sz_g = align256(8 * (size_t)n_w * m), sz_s = align256(8 * (size_t)n_w * n_w);
        const size_t sz_m = align256(8 * ((size_t)n_w + 1) * n_w);
        const int ldm = n_w + 1;
        double* G = (double*)base;
        double* S = (double*)(base + sz_g);
        double* M = (double*)(base + sz_g + sz_s);
        double* g = (double*)(base + sz_g + sz_s + sz_m);
        int* info2 = (int*)(base + sz_g + sz_s + sz_m + align256(8 * (size_t)n_w));
        char* cws = base + sz_g + sz_s + sz_m + align256(8 * (size_t)n_w) + align256(sizeof(int));
        const dim3 grid2((n_w + 256) / 256, n_w);   // rows 0 .. n_w
        // L = chol(P), lower triangle in place
        if (int e = chol_lower(h, st, n_w, P, n_w, cws, info)) return e;
        // G = L^T H
        PT_BLAS(rocblas_dtrmm(h, rocblas_side_left, rocblas_fill_lower, rocblas_operation_transpose,
                              rocblas_diagonal_non_unit, n_w, m, &one, P, n_w, H, (rocblas_int)ldh, G, n_w),
                "dtrmm L^T H");
        // S = G G^T + r I = Ls Ls^T
        PT_BLAS(rocblas_dsyrk(h, rocblas_fill_lower, rocblas_operation_none, n_w, m, &one, G, n_w, &zero, S, n_w),
                "dsyrk G G^T");
        hipLaunchKernelGGL(k_add_diag, dim3((n_w + 255) / 256), dim3(256), 0, st, n_w, S, (long long)n_w, r);
        if (int e = launch_check("k_add_diag")) return e;
        if (int e = chol_lower(h, st, n_w, S, n_w, cws, info2)) return e;
        hipLaunchKernelGGL(k_merge_info, dim3(1), dim3(64), 0, st, info, (const int*)info2);
        if (int e = launch_check("k_merge_info")) return e;
        // [M; y^T] = [L; (G xi)^T] Ls^-T: M = L Ls^-T and y = Ls^-1 G xi in one solve
        PT_BLAS(rocblas_dgemv(h, rocblas_operation_none, n_w, m, &one, G, n_w, xi, 1, &zero, g, 1), "dgemv G xi");
        hipLaunchKernelGGL(k_lower_copy, grid2, dim3(256), 0, st, n_w, (const double*)P, (long long)n_w,
                           (const double*)g, M, (long long)ldm);
        if (int e = launch_check("k_lower_copy")) return e;
        PT_BLAS(rocblas_dtrsm(h, rocblas_side_right, rocblas_fill_lower, rocblas_operation_transpose,
                              rocblas_diagonal_non_unit, n_w + 1, n_w, &one, S, n_w, M, ldm),
                "dtrsm [L; g^T] Ls^-T");
        // W += M y (= L S^-1 G xi), then P = c r M M^T + q I
        PT_BLAS(rocblas_dgemv(h, rocblas_operation_none, n_w, n_w, &one, M, ldm, M + n_w, ldm, &one, W, 1),
                "dgemv M y");
        const double cr = c * r;
        PT_BLAS(rocblas_dsyrk(h, rocblas_fill_lower, rocblas_operation_none, n_w, n_w, &cr, M, ldm, &zero, P, n_w),
                "dsyrk M M^T");
        hipLaunchKernelGGL(k_symm_lower, grid2, dim3(256), 0, st, n_w, P, (long long)n_w, q);
        return launch_check("k_symm_lower");
    }
    double* X = (double*)base;
    double* A = (double*)(base + align256(8 * (size_t)n_w * m));
    double* Y = (double*)(base + align256(8 * (size_t)n_w * m) + align256(8 * (size_t)m * m));
    // X = P H
    PT_BLAS(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, n_w, m, n_w, &one, P, n_w, H,
                          (rocblas_int)ldh, &zero, X, n_w),
            "dgemm P H");
    // A = H^T X + r I
    PT_BLAS(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, m, m, n_w, &one, H,
                          (rocblas_int)ldh, X, n_w, &zero, A, m),
            "dgemm H^T X");
    hipLaunchKernelGGL(k_add_diag, dim3((m + 255) / 256), dim3(256), 0, st, m, A, (long long)m, r);
    if (int e = launch_check("k_add_diag")) return e;
    if (int e = chol_lower(h, st, m, A, m, base + align256(8 * (size_t)n_w * m) + align256(8 * (size_t)m * m) +
                                              align256(8 * (size_t)m * n_w), info))
        return e;
    // Y = X^T, then A Y = X^T
    hipLaunchKernelGGL(k_transpose, dim3((n_w + 31) / 32, (m + 31) / 32), dim3(256), 0, st, n_w, m, X,
                       (long long)n_w, Y, (long long)m);
    if (int e = launch_check("k_transpose")) return e;
    PT_BLAS(rocsolver_dpotrs(h, rocblas_fill_lower, m, n_w, A, m, Y, m), "dpotrs");
    // P = c (P - Y^T X^T) + q I
    PT_BLAS(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_transpose, n_w, n_w, m, &mc, Y, m, X,
                          n_w, &c, P, n_w),
            "dgemm K X^T");
    if (q != 0.0) {
        hipLaunchKernelGGL(k_add_diag, dim3((n_w + 255) / 256), dim3(256), 0, st, n_w, P, (long long)n_w, q);
        if (int e = launch_check("k_add_diag")) return e;
    }
    // W += Y^T xi
    PT_BLAS(rocblas_dgemv(h, rocblas_operation_transpose, m, n_w, &one, Y, m, xi, 1, &one, W, 1), "dgemv K xi");
    return PT_OK;
}
