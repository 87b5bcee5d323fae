// scf.hip -- density purification for the device SCF loop.
//
// Trace-correcting purification (TC2: Niklasson, Phys. Rev. B 66, 155115
// (2002)) of a symmetric m x m matrix F in an orthonormal basis: X0 =
// (e_max - F)/(e_max - e_min) with Gershgorin bounds, then each step
// X <- X^2 or 2X - X^2, whichever brings tr X closer to the occupation.  The
// fixed point is the projector onto the nocc lowest eigenvectors of F.  One
// step is three launches (the product X X, one-block traces, elementwise
// branch), with the branch taken on the device so that no host round trip is
// needed between steps.  X X is a plain dense GEMM of a few hundred rows:
// rocBLAS dgemm, one handle per thread.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include "scf.h"

// bounds[0] = min_i (F_ii - r_i), bounds[1] = max_i (F_ii + r_i),
// r_i = sum_{j != i} |F_ij| (one block).
__global__ void __launch_bounds__(256) gersh_kernel(const double* __restrict__ F, int m, double* bounds)
{
    __shared__ double lo[256], hi[256];
    double mn = 1e300, mx = -1e300;
    for (int i = threadIdx.x; i < m; i += blockDim.x) {
        double r = 0.0;
        for (int j = 0; j < m; ++j)
            if (j != i) r += fabs(F[(size_t)i * m + j]);
        const double d = F[(size_t)i * m + i];
        mn = fmin(mn, d - r);
        mx = fmax(mx, d + r);
    }
    lo[threadIdx.x] = mn;
    hi[threadIdx.x] = mx;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            lo[threadIdx.x] = fmin(lo[threadIdx.x], lo[threadIdx.x + s]);
            hi[threadIdx.x] = fmax(hi[threadIdx.x], hi[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        bounds[0] = lo[0];
        bounds[1] = hi[0];
    }
}

__global__ void tc2_init_kernel(const double* __restrict__ F, int m, const double* __restrict__ bounds, double* X)
{
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= (long long)m * m) return;
    const int i = (int)(t / m), j = (int)(t % m);
    const double emin = bounds[0], emax = bounds[1];
    X[t] = ((i == j ? emax : 0.0) - F[t]) / (emax - emin);
}

// t[0] = tr X, t[1] = tr X2 (one block)
__global__ void __launch_bounds__(256) trace2_kernel(const double* __restrict__ X, const double* __restrict__ X2, int m,
                                                      double* t)
{
    __shared__ double a[256], b[256];
    double s1 = 0.0, s2 = 0.0;
    for (int i = threadIdx.x; i < m; i += blockDim.x) {
        s1 += X[(size_t)i * m + i];
        s2 += X2[(size_t)i * m + i];
    }
    a[threadIdx.x] = s1;
    b[threadIdx.x] = s2;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            a[threadIdx.x] += a[threadIdx.x + s];
            b[threadIdx.x] += b[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        t[0] = a[0];
        t[1] = b[0];
    }
}

// X <- X2 when |tr X2 - nocc| < |2 tr X - tr X2 - nocc|, else 2 X - X2
__global__ void tc2_select_kernel(double* X, const double* __restrict__ X2, long long n, double nocc,
                                  const double* __restrict__ t)
{
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double t1 = t[0], t2 = t[1];
    const bool up = fabs(t2 - nocc) < fabs(2.0 * t1 - t2 - nocc);
    const double x2 = X2[i];
    X[i] = up ? x2 : 2.0 * X[i] - x2;
}

// P = (X + X^T) / 2
__global__ void symmetrize_kernel(const double* __restrict__ X, double* P, int m)
{
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= (long long)m * m) return;
    const int i = (int)(t / m), j = (int)(t % m);
    P[t] = 0.5 * (X[(size_t)i * m + j] + X[(size_t)j * m + i]);
}

static inline unsigned nblk(long long n) { return (unsigned)((n + 255) / 256); }

int tc2_start(const double* F, int m, double* X, double* bounds, hipStream_t st)
{
    if (m <= 0) return -1;
    hipLaunchKernelGGL(gersh_kernel, dim3(1), dim3(256), 0, st, F, m, bounds);
    hipLaunchKernelGGL(tc2_init_kernel, dim3(nblk((long long)m * m)), dim3(256), 0, st, F, m, (const double*)bounds, X);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

static rocblas_handle blas_handle()
{
    static thread_local rocblas_handle h = nullptr;
    if (!h && rocblas_create_handle(&h) != rocblas_status_success) h = nullptr;
    return h;
}

int tc2_steps(double* X, double* X2, int m, int nocc, int steps, double* t, hipStream_t st)
{
    if (m <= 0 || steps < 0) return -1;
    const long long n = (long long)m * m;
    rocblas_handle h = blas_handle();
    if (!h || rocblas_set_stream(h, st) != rocblas_status_success) return -1;
    const double one = 1.0, zero = 0.0;
    for (int s = 0; s < steps; ++s) {
        // rocBLAS is column-major: it sees X^T and computes X^T X^T = (X X)^T,
        // which read row-major is X X
        if (rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, m, m, m, &one, X, m, X, m, &zero, X2,
                          m) != rocblas_status_success)
            return -1;
        hipLaunchKernelGGL(trace2_kernel, dim3(1), dim3(256), 0, st, (const double*)X, (const double*)X2, m, t);
        hipLaunchKernelGGL(tc2_select_kernel, dim3(nblk(n)), dim3(256), 0, st, X, (const double*)X2, n, (double)nocc,
                           (const double*)t);
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int tc2_finish(const double* X, double* P, int m, hipStream_t st)
{
    hipLaunchKernelGGL(symmetrize_kernel, dim3(nblk((long long)m * m)), dim3(256), 0, st, X, P, m);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
