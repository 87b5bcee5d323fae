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
    const PrimPair<real>* prims;
    real* out;
    int nq;
};

struct OneeArgs {
    const int* pair_list;
    const PairHdr* pairs;
    const PrimPair<double>* prims;
    const double* atoms;     // x, y, z, Z per atom
    const int* cart_off;
    double *S, *T, *V;
    int npair, natom, ncart;
};

template <typename real>
struct JkArgs {
    const int2* quartets;
    const PairHdr* pairs;
    const int* cart_off;
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
    double* out;        // n^4 dense
    int nq, n;
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
#define ERI_PRIM_SETUP(LL)                                                          \
    const real p_ = P.p, q_ = Q.p;                                                  \
    const real ipq = (real)1 / (p_ + q_);                                           \
    const real rho = p_ * q_ * ipq;                                                 \
    const real PQx = P.Px - Q.Px, PQy = P.Py - Q.Py, PQz = P.Pz - Q.Pz;             \
    const real Tb = rho * (PQx * PQx + PQy * PQy + PQz * PQz);                      \
    real F[(LL) + 1];                                                               \
    boys_eval<(LL)>(Tb, F);                                                         \
    const real pref = P.K * Q.K * (real)1.1283791670955126 * sqrt(rho);             \
    const real rho_p = q_ * ipq, rho_q = p_ * ipq;                                  \
    const real WPx = -rho_p * PQx, WPy = -rho_p * PQy, WPz = -rho_p * PQz;          \
    const real WQx = rho_q * PQx, WQy = rho_q * PQy, WQz = rho_q * PQz;             \
    const real PAx = P.PAx, PAy = P.PAy, PAz = P.PAz;                               \
    const real QCx = Q.PAx, QCy = Q.PAy, QCz = Q.PAz;                               \
    const real oo2p = P.inv2p, oo2q = Q.inv2p, oo2pq = (real)0.5 * ipq;             \
    (void)WPx; (void)WPy; (void)WPz; (void)WQx; (void)WQy; (void)WQz;               \
    (void)PAx; (void)PAy; (void)PAz; (void)QCx; (void)QCy; (void)QCz;               \
    (void)oo2p; (void)oo2q; (void)oo2pq; (void)rho_p; (void)rho_q;

// Per primitive quartet: prefactor and recurrence coefficients for the
// unscaled Obara-Saika VRR.  W - P = -(q/(p+q)) (P-Q), W - Q = (p/(p+q)) (P-Q).
#define ERI_PRIM_SETUP(LL)                                                          \
    const real p_ = P.p, q_ = Q.p;                                                  \
    const real ipq = (real)1 / (p_ + q_);                                           \
    const real rho = p_ * q_ * ipq;                                                 \
    const real PQx = P.Px - Q.Px, PQy = P.Py - Q.Py, PQz = P.Pz - Q.Pz;             \
    const real Tb = rho * (PQx * PQx + PQy * PQy + PQz * PQz);                      \
    real F[(LL) + 1];                                                               \
    boys_eval<(LL)>(Tb, F);                                                         \
    const real pref = P.K * Q.K * (real)1.1283791670955126 * sqrt(rho);             \
    const real rho_p = q_ * ipq, rho_q = p_ * ipq;                                  \
    const real WPx = -rho_p * PQx, WPy = -rho_p * PQy, WPz = -rho_p * PQz;          \
    const real WQx = rho_q * PQx, WQy = rho_q * PQy, WQz = rho_q * PQz;             \
    const real PAx = P.PAx, PAy = P.PAy, PAz = P.PAz;                               \
    const real QCx = Q.PAx, QCy = Q.PAy, QCz = Q.PAz;                               \
    const real oo2p = P.inv2p, oo2q = Q.inv2p, oo2pq = (real)0.5 * ipq;             \
    (void)WPx; (void)WPy; (void)WPz; (void)WQx; (void)WQy; (void)WQz;               \
    (void)PAx; (void)PAy; (void)PAz; (void)QCx; (void)QCy; (void)QCz;               \
    (void)oo2p; (void)oo2q; (void)oo2pq; (void)rho_p; (void)rho_q;

// --------------------------------------------------------------------------
// J/K digestion of one stored quartet block (8-fold symmetry, shell-level
// degeneracy factor).  For each element v = (mn|ls) with symmetric D:
//   J'_mn += 4s Dj_ls v   J'_ls += 4s Dj_mn v
//   K'_ml += 2s Dk_ns v   K'_nl += 2s Dk_ms v   K'_ms += 2s Dk_nl v   K'_ns += 2s Dk_ml v
// and J = (J' + J'^T)/2, K = (K' + K'^T)/2 afterwards.
// --------------------------------------------------------------------------
__device__ __forceinline__ void atomic_addd(double* p, double v)
{
    unsafeAtomicAdd(p, v);
}

template <int LA, int LB, int LC, int LD, typename real>
__global__ void __launch_bounds__(128) jk_kernel(const JkArgs<real> a)
{
    constexpr int N1 = (LA + 1) * (LA + 2) / 2, N2 = (LB + 1) * (LB + 2) / 2;
    constexpr int N3 = (LC + 1) * (LC + 2) / 2, N4 = (LD + 1) * (LD + 2) / 2;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= a.nq) return;
    const int2 pr = a.quartets[q];
    const PairHdr bh = a.pairs[pr.x], kh = a.pairs[pr.y];
    const int o1 = a.cart_off[bh.s1], o2 = a.cart_off[bh.s2];
    const int o3 = a.cart_off[kh.s1], o4 = a.cart_off[kh.s2];
    double sc = 1.0;
    if (bh.s1 == bh.s2) sc *= 0.5;
    if (kh.s1 == kh.s2) sc *= 0.5;
    if (pr.x == pr.y) sc *= 0.5;
    const int n = a.n;
    const double* Dj = a.Dj;
    const double* Dk = a.Dk;
    const size_t nq = a.nq;
    double j34[N3 * N4], k23[N2 * N3], k24[N2 * N4];
#pragma unroll
    for (int i = 0; i < N3 * N4; ++i) j34[i] = 0.0;
#pragma unroll
    for (int i = 0; i < N2 * N3; ++i) k23[i] = 0.0;
#pragma unroll
    for (int i = 0; i < N2 * N4; ++i) k24[i] = 0.0;
#pragma unroll
    for (int ia = 0; ia < N1; ++ia) {
        double k13[N3], k14[N4];
#pragma unroll
        for (int i = 0; i < N3; ++i) k13[i] = 0.0;
#pragma unroll
        for (int i = 0; i < N4; ++i) k14[i] = 0.0;
#pragma unroll
        for (int ib = 0; ib < N2; ++ib) {
            const double d12 = Dj[(size_t)(o1 + ia) * n + o2 + ib];
            double j12 = 0.0;
#pragma unroll
            for (int ic = 0; ic < N3; ++ic) {
                const double d13 = Dk[(size_t)(o1 + ia) * n + o3 + ic];
                const double d23 = Dk[(size_t)(o2 + ib) * n + o3 + ic];
#pragma unroll
                for (int id = 0; id < N4; ++id) {
                    const double v = (double)a.vals[(size_t)(((ia * N2 + ib) * N3 + ic) * N4 + id) * nq + q];
                    j12 += Dj[(size_t)(o3 + ic) * n + o4 + id] * v;
                    j34[ic * N4 + id] += d12 * v;
                    k13[ic] += Dk[(size_t)(o2 + ib) * n + o4 + id] * v;
                    k23[ib * N3 + ic] += Dk[(size_t)(o1 + ia) * n + o4 + id] * v;
                    k14[id] += d23 * v;
                    k24[ib * N4 + id] += d13 * v;
                }
            }
            atomic_addd(&a.J[(size_t)(o1 + ia) * n + o2 + ib], 4.0 * sc * j12);
        }
#pragma unroll
        for (int ic = 0; ic < N3; ++ic) atomic_addd(&a.K[(size_t)(o1 + ia) * n + o3 + ic], 2.0 * sc * k13[ic]);
#pragma unroll
        for (int id = 0; id < N4; ++id) atomic_addd(&a.K[(size_t)(o1 + ia) * n + o4 + id], 2.0 * sc * k14[id]);
    }
#pragma unroll
    for (int ic = 0; ic < N3; ++ic)
#pragma unroll
        for (int id = 0; id < N4; ++id) atomic_addd(&a.J[(size_t)(o3 + ic) * n + o4 + id], 4.0 * sc * j34[ic * N4 + id]);
#pragma unroll
    for (int ib = 0; ib < N2; ++ib) {
#pragma unroll
        for (int ic = 0; ic < N3; ++ic) atomic_addd(&a.K[(size_t)(o2 + ib) * n + o3 + ic], 2.0 * sc * k23[ib * N3 + ic]);
#pragma unroll
        for (int id = 0; id < N4; ++id) atomic_addd(&a.K[(size_t)(o2 + ib) * n + o4 + id], 2.0 * sc * k24[ib * N4 + id]);
    }
}

// Dense scatter with 8-fold permutational symmetry (Methods/hartree_fock.py:314-325).
template <int LA, int LB, int LC, int LD, typename real>
__global__ void __launch_bounds__(128) scatter_kernel(const ScatterArgs<real> a)
{
    constexpr int N1 = (LA + 1) * (LA + 2) / 2, N2 = (LB + 1) * (LB + 2) / 2;
    constexpr int N3 = (LC + 1) * (LC + 2) / 2, N4 = (LD + 1) * (LD + 2) / 2;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= a.nq) return;
    const int2 pr = a.quartets[q];
    const PairHdr bh = a.pairs[pr.x], kh = a.pairs[pr.y];
    const size_t o1 = a.cart_off[bh.s1], o2 = a.cart_off[bh.s2];
    const size_t o3 = a.cart_off[kh.s1], o4 = a.cart_off[kh.s2];
    const size_t n = a.n, nq = a.nq;
    for (int ia = 0; ia < N1; ++ia)
        for (int ib = 0; ib < N2; ++ib)
            for (int ic = 0; ic < N3; ++ic)
                for (int id = 0; id < N4; ++id) {
                    const double v = (double)a.vals[(size_t)(((ia * N2 + ib) * N3 + ic) * N4 + id) * nq + q];
                    const size_t m = o1 + ia, nn = o2 + ib, l = o3 + ic, s = o4 + id;
                    a.out[((m * n + nn) * n + l) * n + s] = v;
                    a.out[((nn * n + m) * n + l) * n + s] = v;
                    a.out[((m * n + nn) * n + s) * n + l] = v;
                    a.out[((nn * n + m) * n + s) * n + l] = v;
                    a.out[((l * n + s) * n + m) * n + nn] = v;
                    a.out[((l * n + s) * n + nn) * n + m] = v;
                    a.out[((s * n + l) * n + m) * n + nn] = v;
                    a.out[((s * n + l) * n + nn) * n + m] = v;
                }
}

