// mp2.h -- internal launchers of mp2.hip (matrix-core GEMM, MP2 energy).
#pragma once
#include <hip/hip_runtime.h>

struct GemmF64Args {
    const double* A;
    const double* B;
    double* C;
    long long lda, ldb, ldc;      // leading dimensions (elements)
    long long sA, sB, sC;         // batch strides (elements)
    int M, N, K;
    int ta, tb;                   // A stored k-major (K x M) / B stored n-major (N x K)
    double alpha, beta;
};

int gemm_f64_launch(const GemmF64Args& g, int batch, hipStream_t st);
// Same product with split-BF16 (hi*hi + hi*lo + lo*hi) MFMA, FP32 accumulation, FP64 in/out.
int gemm_bf16x3_launch(const GemmF64Args& g, int batch, hipStream_t st);
int mp2_energy_launch(const double* ovov, const double* eo1, const double* ev1, const double* eo2, const double* ev2,
                      int o1, int v1, int o2, int v2, int mode, double* out, hipStream_t st);
