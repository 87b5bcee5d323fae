// eri_common.h -- device-side types, Boys function and per-class kernel
// templates shared by the generated class translation units and the host code.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#define ERI_BLOCK 128

struct PairHdr {
    double ab[3];   // X1 - X2 (primary minus secondary centre)
    int off;        // first primitive pair
    int n;          // primitive pairs kept
    int s1, s2;     // shells: primary, secondary
    int pcls;       // pair class 0..5: ss ps pp ds dp dd
    int mol;        // molecule of both shells
};

// One molecule of a (possibly batched) basis: its matrices are the n x n
// block at offset base of the packed matrix arrays; its atoms are
// [atom_lo, atom_hi).  Shell AO offsets are local to the molecule.
struct MolInfo {
    long long base;
    int n;
    int atom_lo, atom_hi;
    int pad;
};

template <typename real>
struct PrimPair {
    real Px, Py, Pz, PAx, PAy, PAz, p, inv2p, K, beta;
};

template <typename real>
struct EriArgs {
    const int2* quartets;
    const PairHdr* pairs;
    const PrimPair<double>* prims;   // FP64 for both precisions
    real* out;
    int nq;
    const int* order;   // thread t computes quartet order[t] (NULL: quartet t)
};

// Several classes in one launch: blocks [block0[j], block0[j + 1]) run
// class slot which[j] (eri_fused_k) on quartets cls[j].
#define ERI_MAXFUSED 8
template <typename real>
struct EriFusedArgs {
    EriArgs<real> cls[ERI_MAXFUSED];
    int which[ERI_MAXFUSED];
    int block0[ERI_MAXFUSED + 1];
    int n;
};

struct OneeArgs {
    const int* pair_list;
    const PairHdr* pairs;
    const PrimPair<double>* prims;
    const double* atoms;     // x, y, z, Z per atom
    const int* cart_off;
    const MolInfo* mols;
    double *S, *T, *V;       // packed per-molecule blocks
    int npair;
};
// the six pair classes in one launch: class p covers blocks [block0[p],
// block0[p + 1]) of 128 threads, G lanes per shell pair (G: power of two <= 64)
struct OneeFusedArgs {
    OneeArgs cls[6];
    int block0[7];
    int G;
};

template <typename real>
struct JkArgs {
    const int2* quartets;
    const PairHdr* pairs;
    const int* cart_off;
    const MolInfo* mols;
    const real* vals;
    const double* Dj;   // density for J (ncart x ncart)
    const double* Dk;   // density for K
    double* J;          // accumulates J' (unsymmetrised)
    double* K;          // accumulates K' (unsymmetrised, positive)
    int nq, n;
};

template <typename real>
struct ScatterArgs {
    const int2* quartets;
    const PairHdr* pairs;
    const int* cart_off;
    const real* vals;
    double* out;        // out[m][n][l - l0][s] for l in [l0, l0 + lc)
    int nq, n, l0, lc;
};

// --------------------------------------------------------------------------
// Boys function F_m(T), m = 0..L
//   T < 36: 9-term Taylor expansion of F_L about the nearest grid point
//           (tabulated F_n(k*0.1) from mpmath), then downward recursion
//           F_m = (2T F_{m+1} + e^-T)/(2m+1)
//   T >= 36: F_0 = sqrt(pi/T)/2, upward recursion F_{m+1} = ((2m+1)F_m - e^-T)/(2T)
// --------------------------------------------------------------------------
#include "boys_table.inc"

template <int L>
__device__ __forceinline__ void boys_eval(double T, double* F)
{
    if (T < BOYS_TMAX) {
        const int k = __double2int_rn(T * BOYS_INV_STEP);
        const double x = k * BOYS_STEP - T;
        const double* row = boys_table + k * BOYS_NCOL + L;
        double f = row[8];
#pragma unroll
        for (int j = 8; j >= 1; --j) f = row[j - 1] + f * (x * (1.0 / j));
        F[L] = f;
        if (L > 0) {
            const double e = exp(-T);
            const double t2 = 2.0 * T;
#pragma unroll
            for (int m = L - 1; m >= 0; --m) F[m] = (t2 * F[m + 1] + e) * (1.0 / (2 * m + 1));
        }
    } else {
        const double it = 1.0 / T;
        F[0] = 0.88622692545275801365 * sqrt(it);
        if (L > 0) {
            const double e = exp(-T);
            const double h = 0.5 * it;
#pragma unroll
            for (int m = 0; m < L; ++m) F[m + 1] = ((2 * m + 1) * F[m] - e) * h;
        }
    }
}

template <int L>
__device__ __forceinline__ void boys_eval(float T, float* F)
{
    if (T < (float)BOYS_TMAX) {
        const int k = __float2int_rn(T * (float)BOYS_INV_STEP);
        const float x = k * (float)BOYS_STEP - T;
        const double* row = boys_table + k * BOYS_NCOL + L;
        float f = (float)row[5];
#pragma unroll
        for (int j = 5; j >= 1; --j) f = (float)row[j - 1] + f * (x * (1.0f / j));
        F[L] = f;
        if (L > 0) {
            const float e = __expf(-T);
            const float t2 = 2.0f * T;
#pragma unroll
            for (int m = L - 1; m >= 0; --m) F[m] = (t2 * F[m + 1] + e) * (1.0f / (2 * m + 1));
        }
    } else {
        const float it = 1.0f / T;
        F[0] = 0.88622692545275801365f * sqrtf(it);
        if (L > 0) {
            const float e = __expf(-T);
            const float h = 0.5f * it;
#pragma unroll
            for (int m = 0; m < L; ++m) F[m + 1] = ((2 * m + 1) * F[m] - e) * h;
        }
    }
}

// Per primitive quartet: prefactor and recurrence coefficients for the
// unscaled Obara-Saika VRR.  W - P = -(q/(p+q)) (P-Q), W - Q = (p/(p+q)) (P-Q).
// The geometry (P-Q, T, prefactor) is always formed in FP64 from the FP64
// primitive-pair data; the FP32 kernels round only after that, so centre
// coordinates never lose digits to single precision.
#define ERI_PRIM_SETUP(LL)                                                          \
    const double pd_ = P.p, qd_ = Q.p;                                              \
    const double ipq_d = 1.0 / (pd_ + qd_);                                         \
    const double rho_d = pd_ * qd_ * ipq_d;                                         \
    const double PQx_d = P.Px - Q.Px, PQy_d = P.Py - Q.Py, PQz_d = P.Pz - Q.Pz;     \
    const double T_d = rho_d * (PQx_d * PQx_d + PQy_d * PQy_d + PQz_d * PQz_d);     \
    real F[(LL) + 1];                                                               \
    boys_eval<(LL)>((real)T_d, F);                                                  \
    const real pref = (real)(P.K * Q.K * 1.1283791670955126 * sqrt(rho_d));         \
    const real ipq = (real)ipq_d;                                                   \
    const real rho_p = (real)(qd_ * ipq_d), rho_q = (real)(pd_ * ipq_d);            \
    const real PQx = (real)PQx_d, PQy = (real)PQy_d, PQz = (real)PQz_d;             \
    const real WPx = -rho_p * PQx, WPy = -rho_p * PQy, WPz = -rho_p * PQz;          \
    const real WQx = rho_q * PQx, WQy = rho_q * PQy, WQz = rho_q * PQz;             \
    const real PAx = (real)P.PAx, PAy = (real)P.PAy, PAz = (real)P.PAz;             \
    const real QCx = (real)Q.PAx, QCy = (real)Q.PAy, QCz = (real)Q.PAz;             \
    const real oo2p = (real)P.inv2p, oo2q = (real)Q.inv2p, oo2pq = (real)0.5 * ipq; \
    (void)WPx; (void)WPy; (void)WPz; (void)WQx; (void)WQy; (void)WQz;               \
    (void)PAx; (void)PAy; (void)PAz; (void)QCx; (void)QCy; (void)QCz;               \
    (void)oo2p; (void)oo2q; (void)oo2pq; (void)rho_p; (void)rho_q;
