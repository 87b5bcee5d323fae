// mlp.h -- internal launchers of mlp.hip (learned-guess MLP, FP32).
#pragma once
#include <hip/hip_runtime.h>

int mlp_linear_fwd(const float* X, const float* W, const float* b, float* Z, float* A, int B, int I, int O, int act,
                   hipStream_t st);
int mlp_linear_bwd(const float* X, const float* W, const float* A, const float* dA, float* dZ, float* dX, float* dW,
                   float* db, int B, int I, int O, int act, hipStream_t st);
int mlp_mse(const float* Y, const float* T, const float* mask, float* dY, float* loss, long long n, float inv_norm,
            hipStream_t st);
int mlp_adam(float* p, const float* g, float* m, float* v, long long n, float lr, float b1, float b2, float eps,
             int step, hipStream_t st);
