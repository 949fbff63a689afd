// (f) Global extended Kalman filter update of the MLP weights on the device.
//
// Replaces KalmanFilter.update_state_vector
// (potentials/nnp/kalman_filter.py:196-233).  With n_w weights (all elements,
// one state vector) and m observations of one structure (m = 1 for the
// energy, m = 3 N for the force components):
//
//   X = P H                           n_w x m    rocBLAS dgemm
//   A = H^T X + r I                   m x m      rocBLAS dgemm + k_add_diag
//   A = L L^T                                    rocSOLVER dpotrf
//   Y = A^-1 X^T         (= K^T)      m x n_w    k_transpose + rocSOLVER dpotrs
//   P = c (P - Y^T X^T) + q I                    rocBLAS dgemm + k_add_diag
//   W = W + Y^T xi                               rocBLAS dgemv
//
// r is the measurement noise (1/eta for the filter of type 0, lambda for
// type 1), c the fading-memory factor (1 or 1/lambda) and q the process
// noise; the schedules of eta, q and lambda are host scalars.  The
// reference forms inv(A) and multiplies; the Cholesky solve gives the same
// K (A is symmetric positive definite: H^T P H is semi-definite and r > 0)
// without forming the inverse.  All matrices are column-major FP64:
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

}  // namespace
}  // namespace pt

using namespace pt;

extern "C" size_t pt_kalman_workspace_bytes(int n_w, int m) {
    if (n_w < 1 || m < 1) return 0;
    const size_t nw = (size_t)n_w, mm = (size_t)m;
    return align256(8 * nw * mm) + align256(8 * mm * mm) + align256(8 * mm * nw);
}

extern "C" int pt_kalman_residual(int n, const double* pred, const double* ref, double ref_scalar, double scale,
                                  double* xi, double* loss, pt_stream_t stream) {
    PT_REQUIRE(n >= 1 && pred && xi && loss, "pt_kalman_residual: bad arguments");
    hipLaunchKernelGGL(k_residual, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, n, pred, ref,
                       ref_scalar, scale, xi, loss);
    return launch_check("pt_kalman_residual");
}

#define PT_BLAS(call, what)                                                            \
    do {                                                                               \
        rocblas_status st_ = (call);                                                   \
        if (st_ != rocblas_status_success) {                                           \
            set_error("pt_kalman_update: %s failed (%s)", what, rocblas_status_to_string(st_)); \
            return PT_ERR_HIP;                                                         \
        }                                                                              \
    } while (0)

extern "C" int pt_kalman_update(int n_w, int m, double* P, const double* H, long long ldh, const double* xi,
                                double* W, double r, double c, double q, void* ws, size_t ws_bytes, int* info,
                                pt_stream_t stream) {
    PT_REQUIRE(n_w >= 1 && m >= 1, "pt_kalman_update: n_w and m must be positive");
    PT_REQUIRE(P && H && xi && W && ws && info, "pt_kalman_update: null argument");
    PT_REQUIRE(ldh >= n_w, "pt_kalman_update: ldh < n_w");
    PT_REQUIRE(r > 0.0, "pt_kalman_update: the measurement noise must be positive");
    PT_REQUIRE(ws_bytes >= pt_kalman_workspace_bytes(n_w, m), "pt_kalman_update: workspace too small");
    PT_REQUIRE((size_t)n_w * n_w < (1ull << 31) && (size_t)m * m < (1ull << 31) && (size_t)n_w * m < (1ull << 31) &&
                   ldh < (1LL << 31),
               "pt_kalman_update: matrices beyond 32-bit BLAS indexing");
    int rc = PT_OK;
    rocblas_handle h = handle_for_current_device(&rc);
    if (!h) return rc;
    hipStream_t st = (hipStream_t)stream;
    PT_BLAS(rocblas_set_stream(h, st), "rocblas_set_stream");
    char* base = (char*)ws;
    double* X = (double*)base;
    double* A = (double*)(base + align256(8 * (size_t)n_w * m));
    double* Y = (double*)(base + align256(8 * (size_t)n_w * m) + align256(8 * (size_t)m * m));
    const double one = 1.0, zero = 0.0, mc = -c;
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
    PT_BLAS(rocsolver_dpotrf(h, rocblas_fill_lower, m, A, m, info), "dpotrf");
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
