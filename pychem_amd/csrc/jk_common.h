// jk_common.h -- J/K digestion and dense-scatter kernel templates over the
// stored ERI classes.
#pragma once
#include "eri_common.h"

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

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// One thread per stored quartet.  Lists are sorted by bra pair, so a wave
// usually shares one bra pair: then the bra-side J block is summed across the
// wave and written with one atomic per element instead of 64 colliding ones.
// Lanes past the end of the list stay active with zero weight so that the
// cross-lane sums see every lane of the wave.
template <int LA, int LB, int LC, int LD, typename real>
__global__ void __launch_bounds__(128) jk_kernel(const JkArgs<real> a)
{
    constexpr int N1 = (LA + 1) * (LA + 2) / 2, N2 = (LB + 1) * (LB + 2) / 2;
    constexpr int N3 = (LC + 1) * (LC + 2) / 2, N4 = (LD + 1) * (LD + 2) / 2;
    const int q0 = blockIdx.x * blockDim.x + threadIdx.x;
    if ((int)(blockIdx.x * blockDim.x + (threadIdx.x & ~63)) >= a.nq) return;   // whole wave past the end
    const bool valid = q0 < a.nq;
    const int q = valid ? q0 : a.nq - 1;
    const int2 pr = a.quartets[q];
    const PairHdr bh = a.pairs[pr.x], kh = a.pairs[pr.y];
    const int o1 = a.cart_off[bh.s1], o2 = a.cart_off[bh.s2];
    const int o3 = a.cart_off[kh.s1], o4 = a.cart_off[kh.s2];
    const MolInfo mi = a.mols[bh.mol];
    double sc = valid ? 1.0 : 0.0;
    if (bh.s1 == bh.s2) sc *= 0.5;
    if (kh.s1 == kh.s2) sc *= 0.5;
    if (pr.x == pr.y) sc *= 0.5;
    const int bra0 = __shfl(pr.x, 0, 64);
    const bool uniform = __all(pr.x == bra0);
    const int lane = threadIdx.x & 63;
    const int n = mi.n;
    const double* Dj = a.Dj + mi.base;
    const double* Dk = a.Dk + mi.base;
    double* const Jm = a.J + mi.base;
    double* const Km = a.K + mi.base;
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
            double* dst = &Jm[(size_t)(o1 + ia) * n + o2 + ib];
            if (uniform) {
                const double s = wave_sum(4.0 * sc * j12);
                if (lane == 0) atomic_addd(dst, s);
            } else if (valid) {
                atomic_addd(dst, 4.0 * sc * j12);
            }
        }
        if (valid) {
#pragma unroll
            for (int ic = 0; ic < N3; ++ic) atomic_addd(&Km[(size_t)(o1 + ia) * n + o3 + ic], 2.0 * sc * k13[ic]);
#pragma unroll
            for (int id = 0; id < N4; ++id) atomic_addd(&Km[(size_t)(o1 + ia) * n + o4 + id], 2.0 * sc * k14[id]);
        }
    }
    if (!valid) return;
#pragma unroll
    for (int ic = 0; ic < N3; ++ic)
#pragma unroll
        for (int id = 0; id < N4; ++id) atomic_addd(&Jm[(size_t)(o3 + ic) * n + o4 + id], 4.0 * sc * j34[ic * N4 + id]);
#pragma unroll
    for (int ib = 0; ib < N2; ++ib) {
#pragma unroll
        for (int ic = 0; ic < N3; ++ic) atomic_addd(&Km[(size_t)(o2 + ib) * n + o3 + ic], 2.0 * sc * k23[ib * N3 + ic]);
#pragma unroll
        for (int id = 0; id < N4; ++id) atomic_addd(&Km[(size_t)(o2 + ib) * n + o4 + id], 2.0 * sc * k24[ib * N4 + id]);
    }
}

// Run-based J/K: one wave per work item = a contiguous range of quartets that
// share one bra pair (a,b) (the class lists are sorted by bra pair).  Every K
// contribution of such a range lands in the rows of shells a and b, so those
// (N1+N2) rows are accumulated in LDS and flushed to global memory once per
// range; the bra J block is summed over the wave.  Only the ket J block is
// written with per-quartet global atomics.
template <typename real>
struct JkRunArgs {
    const int2* quartets;
    const int2* items;       // [start, end) per work item, one bra pair each
    const PairHdr* pairs;
    const int* cart_off;
    const MolInfo* mols;
    const real* vals;
    const double* Dj;
    const double* Dk;
    double* J;
    double* K;
    int nq, n, nitems, grid;   // grid: blocks launched (<= nitems; persistent when smaller)
};

template <int LA, int LB, int LC, int LD, typename real>
__global__ void __launch_bounds__(64) jk_run_kernel(const JkRunArgs<real> a)
{
    constexpr int N1 = (LA + 1) * (LA + 2) / 2, N2 = (LB + 1) * (LB + 2) / 2;
    constexpr int N3 = (LC + 1) * (LC + 2) / 2, N4 = (LD + 1) * (LD + 2) / 2;
    extern __shared__ double krows[];   // (N1 + N2) x n
    // persistent: the grid may be smaller than the number of work items
    for (int item = blockIdx.x; item < a.nitems; item += gridDim.x) {
    const int2 rng = a.items[item];
    const int lane = threadIdx.x;
    const int2 pr0 = a.quartets[rng.x];
    const PairHdr bh = a.pairs[pr0.x];
    const MolInfo mi = a.mols[bh.mol];
    const int n = mi.n;
    const int nrow = (N1 + N2) * n;
    for (int i = lane; i < nrow; i += 64) krows[i] = 0.0;
    __syncthreads();
    const int o1 = a.cart_off[bh.s1], o2 = a.cart_off[bh.s2];
    const double sbra = (bh.s1 == bh.s2) ? 0.5 : 1.0;
    const double* Dj = a.Dj + mi.base;
    const double* Dk = a.Dk + mi.base;
    double* const Jm = a.J + mi.base;
    double* const Km = a.K + mi.base;
    const size_t nq = a.nq;
    double j12[N1 * N2];
#pragma unroll
    for (int i = 0; i < N1 * N2; ++i) j12[i] = 0.0;
    for (int q = rng.x + lane; q < rng.y; q += 64) {
        const int2 pr = a.quartets[q];
        const PairHdr kh = a.pairs[pr.y];
        const int o3 = a.cart_off[kh.s1], o4 = a.cart_off[kh.s2];
        double sc = sbra;
        if (kh.s1 == kh.s2) sc *= 0.5;
        if (pr.x == pr.y) sc *= 0.5;
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
                double jab = 0.0;
#pragma unroll
                for (int ic = 0; ic < N3; ++ic) {
                    const double d13 = Dk[(size_t)(o1 + ia) * n + o3 + ic];
                    const double d23 = Dk[(size_t)(o2 + ib) * n + o3 + ic];
#pragma unroll
                    for (int id = 0; id < N4; ++id) {
                        const double v = (double)a.vals[(size_t)(((ia * N2 + ib) * N3 + ic) * N4 + id) * nq + q];
                        jab += Dj[(size_t)(o3 + ic) * n + o4 + id] * v;
                        j34[ic * N4 + id] += d12 * v;
                        k13[ic] += Dk[(size_t)(o2 + ib) * n + o4 + id] * v;
                        k23[ib * N3 + ic] += Dk[(size_t)(o1 + ia) * n + o4 + id] * v;
                        k14[id] += d23 * v;
                        k24[ib * N4 + id] += d13 * v;
                    }
                }
                j12[ia * N2 + ib] += sc * jab;
            }
#pragma unroll
            for (int ic = 0; ic < N3; ++ic) atomicAdd(&krows[ia * n + o3 + ic], 2.0 * sc * k13[ic]);
#pragma unroll
            for (int id = 0; id < N4; ++id) atomicAdd(&krows[ia * n + o4 + id], 2.0 * sc * k14[id]);
        }
#pragma unroll
        for (int ic = 0; ic < N3; ++ic)
#pragma unroll
            for (int id = 0; id < N4; ++id) atomic_addd(&Jm[(size_t)(o3 + ic) * n + o4 + id], 4.0 * sc * j34[ic * N4 + id]);
#pragma unroll
        for (int ib = 0; ib < N2; ++ib) {
#pragma unroll
            for (int ic = 0; ic < N3; ++ic) atomicAdd(&krows[(N1 + ib) * n + o3 + ic], 2.0 * sc * k23[ib * N3 + ic]);
#pragma unroll
            for (int id = 0; id < N4; ++id) atomicAdd(&krows[(N1 + ib) * n + o4 + id], 2.0 * sc * k24[ib * N4 + id]);
        }
    }
#pragma unroll
    for (int i = 0; i < N1 * N2; ++i) {
        const double s = wave_sum(4.0 * j12[i]);
        if (lane == 0) atomic_addd(&Jm[(size_t)(o1 + i / N2) * n + o2 + i % N2], s);
    }
    __syncthreads();
    for (int i = lane; i < nrow; i += 64) {
        const double v = krows[i];
        if (v != 0.0) {
            const int r = i / n, c = i - r * n;
            const int row = r < N1 ? o1 + r : o2 + (r - N1);
            atomic_addd(&Km[(size_t)row * n + c], v);
        }
    }
    __syncthreads();
    }
}

// Dense scatter with 8-fold permutational symmetry (Methods/hartree_fock.py:314-325)
// into a slab of third indices l in [l0, l0 + lc): out[m][n][l - l0][s].
template <int LA, int LB, int LC, int LD, typename real>
__global__ void __launch_bounds__(128) scatter_kernel(const ScatterArgs<real> a)
{
    constexpr int N1 = (LA + 1) * (LA + 2) / 2, N2 = (LB + 1) * (LB + 2) / 2;
    constexpr int N3 = (LC + 1) * (LC + 2) / 2, N4 = (LD + 1) * (LD + 2) / 2;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= a.nq) return;
    const int2 pr = a.quartets[q];
    const PairHdr bh = a.pairs[pr.x], kh = a.pairs[pr.y];
    const long long o1 = a.cart_off[bh.s1], o2 = a.cart_off[bh.s2];
    const long long o3 = a.cart_off[kh.s1], o4 = a.cart_off[kh.s2];
    const long long n = a.n, nq = a.nq, l0 = a.l0, lc = a.lc;
    auto put = [&](long long m, long long nn, long long l, long long s, double v) {
        const long long t = l - l0;
        if (t >= 0 && t < lc) a.out[((m * n + nn) * lc + t) * n + s] = v;
    };
    for (int ia = 0; ia < N1; ++ia)
        for (int ib = 0; ib < N2; ++ib)
            for (int ic = 0; ic < N3; ++ic)
                for (int id = 0; id < N4; ++id) {
                    const double v = (double)a.vals[(size_t)(((ia * N2 + ib) * N3 + ic) * N4 + id) * nq + q];
                    const long long m = o1 + ia, nn = o2 + ib, l = o3 + ic, s = o4 + id;
                    put(m, nn, l, s, v);
                    put(nn, m, l, s, v);
                    put(m, nn, s, l, v);
                    put(nn, m, s, l, v);
                    put(l, s, m, nn, v);
                    put(l, s, nn, m, v);
                    put(s, l, m, nn, v);
                    put(s, l, nn, m, v);
                }
}
