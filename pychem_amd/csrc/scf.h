// scf.h -- device density purification (TC2) for the SCF loop.
#pragma once
#include <hip/hip_runtime.h>

// X = (e_max - F)/(e_max - e_min) with Gershgorin bounds (bounds[2] scratch).
int tc2_start(const double* F, int m, double* X, double* bounds, hipStream_t st);
// `steps` TC2 steps on X (X2 and t[2] scratch; t ends as tr X, tr X^2 of the
// last step's input).
int tc2_steps(double* X, double* X2, int m, int nocc, int steps, double* t, hipStream_t st);
// P = (X + X^T)/2.
int tc2_finish(const double* X, double* P, int m, hipStream_t st);
