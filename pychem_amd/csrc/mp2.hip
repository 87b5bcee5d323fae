// mp2.hip -- dense contractions of the MP2 AO->MO transform on the matrix
// cores, and the MP2 energy reduction.
//
// gemm_f64: batched C = op(A) op(B) (+ beta C) in FP64 on
// v_mfma_f64_16x16x4f64.  Block tile 64x64, four waves of 32x32 (2x2 MFMA
// blocks), K staged 16 at a time through LDS.  op(A)(m,k) = A[k*lda+m] when
// ta (A stored k-major) else A[m*lda+k]; op(B)(k,n) = B[n*ldb+k] when tb else
// B[k*ldb+n].
//
// MFMA 16x16x4 f64 operand layout (measured on gfx950): lane l supplies
// A[l%16][k0 + l/16] and B[k0 + l/16][l%16]; result register r of lane l is
// D[4*r + l/16][l%16], r = 0..3.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mp2.h"

typedef double double4_t __attribute__((ext_vector_type(4)));

template <bool TA, bool TB>
__global__ void __launch_bounds__(256) gemm_f64_kernel(const GemmF64Args g)
{
    __shared__ double As[16][65];
    __shared__ double Bs[16][65];
    const long long bz = blockIdx.z;
    const double* __restrict__ A = g.A + bz * g.sA;
    const double* __restrict__ B = g.B + bz * g.sB;
    double* __restrict__ C = g.C + bz * g.sC;
    const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    const int M = g.M, N = g.N, K = g.K;
    double4_t acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = (double4_t){0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < K; k0 += 16) {
#pragma unroll
        for (int e = tid; e < 1024; e += 256) {
            int kk, mm;
            if (TA) { kk = e >> 6; mm = e & 63; } else { mm = e >> 4; kk = e & 15; }
            const int gm = m0 + mm, gk = k0 + kk;
            double v = 0.0;
            if (gm < M && gk < K) v = TA ? A[(long long)gk * g.lda + gm] : A[(long long)gm * g.lda + gk];
            As[kk][mm] = v;
        }
#pragma unroll
        for (int e = tid; e < 1024; e += 256) {
            int kk, nn;
            if (TB) { nn = e >> 4; kk = e & 15; } else { kk = e >> 6; nn = e & 63; }
            const int gn = n0 + nn, gk = k0 + kk;
            double v = 0.0;
            if (gn < N && gk < K) v = TB ? B[(long long)gn * g.ldb + gk] : B[(long long)gk * g.ldb + gn];
            Bs[kk][nn] = v;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; kk += 4) {
            const int kr = kk + (lane >> 4);
            double a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) a[i] = As[kr][wm + i * 16 + (lane & 15)];
#pragma unroll
            for (int j = 0; j < 2; ++j) b[j] = Bs[kr][wn + j * 16 + (lane & 15)];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gm = m0 + wm + i * 16 + 4 * r + (lane >> 4);
                const int gn = n0 + wn + j * 16 + (lane & 15);
                if (gm < M && gn < N) {
                    double* dst = &C[(long long)gm * g.ldc + gn];
                    const double v = g.alpha * acc[i][j][r];
                    *dst = g.beta == 0.0 ? v : v + g.beta * *dst;
                }
            }
}

int gemm_f64_launch(const GemmF64Args& g, int batch, hipStream_t st)
{
    if (g.M <= 0 || g.N <= 0 || batch <= 0) return 0;
    if (g.K <= 0) return -1;
    dim3 grid((g.N + 63) / 64, (g.M + 63) / 64, batch), block(256);
    if (g.ta && g.tb) hipLaunchKernelGGL((gemm_f64_kernel<true, true>), grid, block, 0, st, g);
    else if (g.ta) hipLaunchKernelGGL((gemm_f64_kernel<true, false>), grid, block, 0, st, g);
    else if (g.tb) hipLaunchKernelGGL((gemm_f64_kernel<false, true>), grid, block, 0, st, g);
    else hipLaunchKernelGGL((gemm_f64_kernel<false, false>), grid, block, 0, st, g);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// MP2 energy over (ia|jb) stored [o1][v1][o2][v2]:
//   mode 0 (closed shell): sum v (2 v - (ib|ja)) / D          (o1 = o2, v1 = v2)
//   mode 1 (same spin):    1/4 sum [(ia|jb) - (ib|ja)]^2 / D  (Methods/mp2.py:77-88, i>=j, p>=q)
//   mode 2 (opposite spin): sum (ia|jb)^2 / D                 (Methods/mp2.py:90-94)
// D = e_i + e_j - e_a - e_b.  Block partial sums, one atomic per block.
__global__ void __launch_bounds__(256) mp2_energy_kernel(const double* __restrict__ ovov, const double* __restrict__ eo1,
                                                          const double* __restrict__ ev1, const double* __restrict__ eo2,
                                                          const double* __restrict__ ev2, int o1, int v1, int o2, int v2,
                                                          int mode, double* out)
{
    const long long total = (long long)o1 * v1 * o2 * v2;
    double s = 0.0;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total; t += (long long)gridDim.x * blockDim.x) {
        const int b = (int)(t % v2);
        long long r = t / v2;
        const int j = (int)(r % o2);
        r /= o2;
        const int a = (int)(r % v1);
        const int i = (int)(r / v1);
        const double v = ovov[t];
        const double den = eo1[i] + eo2[j] - ev1[a] - ev2[b];
        if (mode == 2) {
            s += v * v / den;
        } else {
            const double x = ovov[(((long long)i * v1 + b) * o2 + j) * v2 + a];   // (ib|ja)
            s += mode == 0 ? v * (2.0 * v - x) / den : 0.25 * (v - x) * (v - x) / den;
        }
    }
    __shared__ double red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) unsafeAtomicAdd(out, red[0]);
}

int mp2_energy_launch(const double* ovov, const double* eo1, const double* ev1, const double* eo2, const double* ev2,
                      int o1, int v1, int o2, int v2, int mode, double* out, hipStream_t st)
{
    const long long total = (long long)o1 * v1 * o2 * v2;
    long long blocks = (total + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(mp2_energy_kernel, dim3((unsigned)blocks), dim3(256), 0, st, ovov, eo1, ev1, eo2, ev2, o1, v1, o2,
                       v2, mode, out);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
