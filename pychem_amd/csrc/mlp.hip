// mlp.hip -- training and inference kernels of the learned SCF guess (a
// small multilayer perceptron, FP32).
//
//   forward   Z = X W + b, A = act(Z)             (act: 0 identity, 1 tanh)
//   backward  dZ = dA * act'(Z); dX = dZ W^T; dW = X^T dZ; db = sum_rows dZ
//   loss      masked mean-squared error and its gradient
//   optimiser Adam
//
// Shapes: X [B x I], W [I x O] (row-major), b [O], Z/A/dA/dZ [B x O].  The
// products use 16x16 output tiles per 256-thread block with operand tiles of
// 16 staged through LDS; every global access is bounds-checked.
#include <hip/hip_runtime.h>
#include <math.h>

#include "mlp.h"

#define TS 16

// C[M x N] = op(A) op(B) with op(A)(m,k) = ta ? A[k*M+m] : A[m*K+k],
// op(B)(k,n) = tb ? B[n*K+k] : B[k*N+n]; optional epilogue bias + activation.
template <bool TA, bool TB>
__global__ void __launch_bounds__(256) mm_kernel(const float* __restrict__ A, const float* __restrict__ B,
                                                 const float* __restrict__ bias, float* __restrict__ Z,
                                                 float* __restrict__ Act, int M, int N, int K, int act)
{
    __shared__ float As[TS][TS + 1];
    __shared__ float Bs[TS][TS + 1];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int m = blockIdx.y * TS + ty, n = blockIdx.x * TS + tx;
    float acc = 0.0f;
    for (int k0 = 0; k0 < K; k0 += TS) {
        const int ka = k0 + tx, kb = k0 + ty;
        float a = 0.0f, b = 0.0f;
        if (m < M && ka < K) a = TA ? A[(long long)ka * M + m] : A[(long long)m * K + ka];
        if (n < N && kb < K) b = TB ? B[(long long)n * K + kb] : B[(long long)kb * N + n];
        As[ty][tx] = a;
        Bs[ty][tx] = b;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < TS; ++k) acc += As[ty][k] * Bs[k][tx];
        __syncthreads();
    }
    if (m < M && n < N) {
        if (bias) acc += bias[n];
        Z[(long long)m * N + n] = acc;
        if (Act) Act[(long long)m * N + n] = act == 1 ? tanhf(acc) : acc;
    }
}

static void mm(bool ta, bool tb, const float* A, const float* B, const float* bias, float* Z, float* Act, int M,
               int N, int K, int act, hipStream_t st)
{
    dim3 grid((N + TS - 1) / TS, (M + TS - 1) / TS), block(256);
    if (ta && tb) hipLaunchKernelGGL((mm_kernel<true, true>), grid, block, 0, st, A, B, bias, Z, Act, M, N, K, act);
    else if (ta) hipLaunchKernelGGL((mm_kernel<true, false>), grid, block, 0, st, A, B, bias, Z, Act, M, N, K, act);
    else if (tb) hipLaunchKernelGGL((mm_kernel<false, true>), grid, block, 0, st, A, B, bias, Z, Act, M, N, K, act);
    else hipLaunchKernelGGL((mm_kernel<false, false>), grid, block, 0, st, A, B, bias, Z, Act, M, N, K, act);
}

// dZ = dA * act'(Z) with tanh' = 1 - tanh(Z)^2 (A holds tanh(Z))
__global__ void act_bwd_kernel(const float* dA, const float* A, float* dZ, long long n, int act)
{
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    dZ[i] = act == 1 ? dA[i] * (1.0f - A[i] * A[i]) : dA[i];
}

// db[o] = sum_b dZ[b][o]
__global__ void colsum_kernel(const float* dZ, float* db, int B, int O)
{
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= O) return;
    float s = 0.0f;
    for (int b = 0; b < B; ++b) s += dZ[(long long)b * O + o];
    db[o] = s;
}

// loss = sum mask (Y - T)^2 / norm ; dY = 2 mask (Y - T) / norm
__global__ void mse_kernel(const float* Y, const float* T, const float* mask, float* dY, float* loss, long long n,
                           float inv_norm)
{
    __shared__ float red[256];
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    float l = 0.0f;
    if (i < n) {
        const float w = mask ? mask[i] : 1.0f;
        const float d = (Y[i] - T[i]) * w;
        dY[i] = 2.0f * d * inv_norm;
        l = d * d * inv_norm;
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, red[0]);
}

__global__ void adam_kernel(float* p, const float* g, float* m, float* v, long long n, float lr, float b1, float b2,
                            float eps, float c1, float c2)
{
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

int mlp_linear_fwd(const float* X, const float* W, const float* b, float* Z, float* A, int B, int I, int O, int act,
                   hipStream_t st)
{
    if (B <= 0 || I <= 0 || O <= 0) return -1;
    mm(false, false, X, W, b, Z, A, B, O, I, act, st);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int mlp_linear_bwd(const float* X, const float* W, const float* A, const float* dA, float* dZ, float* dX, float* dW,
                   float* db, int B, int I, int O, int act, hipStream_t st)
{
    if (B <= 0 || I <= 0 || O <= 0) return -1;
    const long long n = (long long)B * O;
    hipLaunchKernelGGL(act_bwd_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, dA, A, dZ, n, act);
    if (dX) mm(false, true, dZ, W, nullptr, dX, nullptr, B, I, O, 0, st);          // dX = dZ W^T
    mm(true, false, X, dZ, nullptr, dW, nullptr, I, O, B, 0, st);                    // dW = X^T dZ
    hipLaunchKernelGGL(colsum_kernel, dim3((O + 127) / 128), dim3(128), 0, st, dZ, db, B, O);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int mlp_mse(const float* Y, const float* T, const float* mask, float* dY, float* loss, long long n, float inv_norm,
            hipStream_t st)
{
    if (n <= 0) return -1;
    hipLaunchKernelGGL(mse_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, Y, T, mask, dY, loss, n, inv_norm);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int mlp_adam(float* p, const float* g, float* m, float* v, long long n, float lr, float b1, float b2, float eps,
             int step, hipStream_t st)
{
    if (n <= 0 || step <= 0) return -1;
    const float c1 = 1.0f / (1.0f - powf(b1, (float)step));
    const float c2 = 1.0f / (1.0f - powf(b2, (float)step));
    hipLaunchKernelGGL(adam_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, g, m, v, n, lr, b1, b2, eps,
                       c1, c2);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
