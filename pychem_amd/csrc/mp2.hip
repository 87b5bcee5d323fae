// mp2.hip -- dense contractions of the MP2 AO->MO transform on the matrix
// cores, and the MP2 energy reduction.
//
// gemm_f64: batched C = op(A) op(B) (+ beta C) in FP64 on
// v_mfma_f64_16x16x4f64.  Block tile 64 (M) x 128 (N), four waves of 32 x 64
// (2 x 4 MFMA blocks), K staged 16 at a time through LDS with the next K
// tile's global loads in flight during the products.  op(A)(m,k) = A[k*lda+m] when
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

#define XD_M 64
#define XD_N 128
#define XD_K 16

template <bool TA, bool TB>
__global__ void __launch_bounds__(256) gemm_f64_kernel(const GemmF64Args g)
{
    __shared__ double As[XD_K][XD_M + 1];
    __shared__ double Bs[XD_K][XD_N + 1];
    const long long bz = blockIdx.z;
    const double* __restrict__ A = g.A + bz * g.sA;
    const double* __restrict__ B = g.B + bz * g.sB;
    double* __restrict__ C = g.C + bz * g.sC;
    const int m0 = blockIdx.y * XD_M, n0 = blockIdx.x * XD_N;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 64;
    const int M = g.M, N = g.N, K = g.K;
    double ra[4], rb[8];
    // element e of the A tile (64 x 16) / B tile (16 x 128) handled by this thread
    auto a_mk = [&](int e, int& mm, int& kk) {
        if (TA) { kk = e >> 6; mm = e & 63; } else { mm = e >> 4; kk = e & 15; }
    };
    auto b_nk = [&](int e, int& nn, int& kk) {
        if (TB) { nn = e >> 4; kk = e & 15; } else { kk = e >> 7; nn = e & 127; }
    };
    auto load = [&](int k0) {
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            int mm, kk;
            a_mk(tid + 256 * t, mm, kk);
            const int gm = m0 + mm, gk = k0 + kk;
            ra[t] = (gm < M && gk < K) ? (TA ? A[(long long)gk * g.lda + gm] : A[(long long)gm * g.lda + gk]) : 0.0;
        }
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            int nn, kk;
            b_nk(tid + 256 * t, nn, kk);
            const int gn = n0 + nn, gk = k0 + kk;
            rb[t] = (gn < N && gk < K) ? (TB ? B[(long long)gn * g.ldb + gk] : B[(long long)gk * g.ldb + gn]) : 0.0;
        }
    };
    double4_t acc[2][4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = (double4_t){0.0, 0.0, 0.0, 0.0};
    load(0);
    for (int k0 = 0; k0 < K; k0 += XD_K) {
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            int mm, kk;
            a_mk(tid + 256 * t, mm, kk);
            As[kk][mm] = ra[t];
        }
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            int nn, kk;
            b_nk(tid + 256 * t, nn, kk);
            Bs[kk][nn] = rb[t];
        }
        __syncthreads();
        if (k0 + XD_K < K) load(k0 + XD_K);
#pragma unroll
        for (int kk = 0; kk < XD_K; kk += 4) {
            const int kr = kk + (lane >> 4);
            double a[2], b[4];
#pragma unroll
            for (int i = 0; i < 2; ++i) a[i] = As[kr][wm + i * 16 + (lane & 15)];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = Bs[kr][wn + j * 16 + (lane & 15)];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
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
    dim3 grid((g.N + XD_N - 1) / XD_N, (g.M + XD_M - 1) / XD_M, batch), block(256);
    if (g.ta && g.tb) hipLaunchKernelGGL((gemm_f64_kernel<true, true>), grid, block, 0, st, g);
    else if (g.ta) hipLaunchKernelGGL((gemm_f64_kernel<true, false>), grid, block, 0, st, g);
    else if (g.tb) hipLaunchKernelGGL((gemm_f64_kernel<false, true>), grid, block, 0, st, g);
    else hipLaunchKernelGGL((gemm_f64_kernel<false, false>), grid, block, 0, st, g);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// gemm_bf16x3: the same batched product with FP32-accumulated split-BF16
// matrix-core products.  Each FP64 operand element is rounded to FP32 and
// split into hi = bf16(x) and lo = bf16(x - hi); a*b is taken as
// hi_a hi_b + hi_a lo_b + lo_a hi_b (the lo_a lo_b term, ~2^-16 relative, is
// dropped) on v_mfma_f32_16x16x32_bf16, accumulated in FP32 and written as
// FP64.  Block tile 64 (M) x 128 (N), K staged 32 at a time through LDS with
// the next K tile's global loads in flight during the products; four waves
// of 32 x 64 (2 x 4 MFMA blocks).  The MFMA layout (measured on gfx950): lane
// l supplies A[l%16][8*(l/16) + i] and B[8*(l/16) + i][l%16], i = 0..7;
// result register r of lane l is D[4*(l/16) + r][l%16] (rows in the other
// order from the FP64 16x16x4 instruction's D[4*r + l/16][l%16]).
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float floatx4_t __attribute__((ext_vector_type(4)));

#define XB_M 64
#define XB_N 128
#define XB_K 32
#define XB_LD 40   // LDS row stride in bf16 (80 bytes: 16-byte aligned fragments)

__device__ __forceinline__ void split_bf16(double x, __bf16& hi, __bf16& lo)
{
    const float f = (float)x;
    hi = (__bf16)f;
    lo = (__bf16)(f - (float)hi);
}

template <bool TA, bool TB>
__global__ void __launch_bounds__(256) gemm_bf16x3_kernel(const GemmF64Args g)
{
    __shared__ __attribute__((aligned(16))) __bf16 Ah[XB_M * XB_LD], Al[XB_M * XB_LD];
    __shared__ __attribute__((aligned(16))) __bf16 Bh[XB_N * XB_LD], Bl[XB_N * XB_LD];
    const long long bz = blockIdx.z;
    const double* __restrict__ A = g.A + bz * g.sA;
    const double* __restrict__ B = g.B + bz * g.sB;
    double* __restrict__ C = g.C + bz * g.sC;
    const int m0 = blockIdx.y * XB_M, n0 = blockIdx.x * XB_N;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 64;
    const int M = g.M, N = g.N, K = g.K;
    double ra[8], rb[16];
    // element e of the A tile (64 x 32) / B tile (32 x 128) handled by this thread
    auto a_mk = [&](int e, int& mm, int& kk) {
        if (TA) { kk = e >> 6; mm = e & 63; } else { mm = e >> 5; kk = e & 31; }
    };
    auto b_nk = [&](int e, int& nn, int& kk) {
        if (TB) { nn = e >> 5; kk = e & 31; } else { kk = e >> 7; nn = e & 127; }
    };
    auto load = [&](int k0) {
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            int mm, kk;
            a_mk(tid + 256 * t, mm, kk);
            const int gm = m0 + mm, gk = k0 + kk;
            ra[t] = (gm < M && gk < K) ? (TA ? A[(long long)gk * g.lda + gm] : A[(long long)gm * g.lda + gk]) : 0.0;
        }
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            int nn, kk;
            b_nk(tid + 256 * t, nn, kk);
            const int gn = n0 + nn, gk = k0 + kk;
            rb[t] = (gn < N && gk < K) ? (TB ? B[(long long)gn * g.ldb + gk] : B[(long long)gk * g.ldb + gn]) : 0.0;
        }
    };
    floatx4_t acc[2][4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = (floatx4_t){0.f, 0.f, 0.f, 0.f};
    load(0);
    for (int k0 = 0; k0 < K; k0 += XB_K) {
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            int mm, kk;
            a_mk(tid + 256 * t, mm, kk);
            split_bf16(ra[t], Ah[mm * XB_LD + kk], Al[mm * XB_LD + kk]);
        }
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            int nn, kk;
            b_nk(tid + 256 * t, nn, kk);
            split_bf16(rb[t], Bh[nn * XB_LD + kk], Bl[nn * XB_LD + kk]);
        }
        __syncthreads();
        if (k0 + XB_K < K) load(k0 + XB_K);
        const int fr = lane & 15, fk = 8 * (lane >> 4);
        bf16x8_t ah[2], al[2], bh[4], bl[4];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            ah[i] = *(const bf16x8_t*)&Ah[(wm + i * 16 + fr) * XB_LD + fk];
            al[i] = *(const bf16x8_t*)&Al[(wm + i * 16 + fr) * XB_LD + fk];
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            bh[j] = *(const bf16x8_t*)&Bh[(wn + j * 16 + fr) * XB_LD + fk];
            bl[j] = *(const bf16x8_t*)&Bl[(wn + j * 16 + fr) * XB_LD + fk];
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
            }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gm = m0 + wm + i * 16 + 4 * (lane >> 4) + r;
                const int gn = n0 + wn + j * 16 + (lane & 15);
                if (gm < M && gn < N) {
                    double* dst = &C[(long long)gm * g.ldc + gn];
                    const double v = g.alpha * (double)acc[i][j][r];
                    *dst = g.beta == 0.0 ? v : v + g.beta * *dst;
                }
            }
}

int gemm_bf16x3_launch(const GemmF64Args& g, int batch, hipStream_t st)
{
    if (g.M <= 0 || g.N <= 0 || batch <= 0) return 0;
    if (g.K <= 0) return -1;
    dim3 grid((g.N + XB_N - 1) / XB_N, (g.M + XB_M - 1) / XB_M, batch), block(256);
    if (g.ta && g.tb) hipLaunchKernelGGL((gemm_bf16x3_kernel<true, true>), grid, block, 0, st, g);
    else if (g.ta) hipLaunchKernelGGL((gemm_bf16x3_kernel<true, false>), grid, block, 0, st, g);
    else if (g.tb) hipLaunchKernelGGL((gemm_bf16x3_kernel<false, true>), grid, block, 0, st, g);
    else hipLaunchKernelGGL((gemm_bf16x3_kernel<false, false>), grid, block, 0, st, g);
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
