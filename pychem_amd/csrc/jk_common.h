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

// Tiled J/K digestion.  A work item is a contiguous range of quartets that
// share one bra pair (a,b) (the class lists are sorted by bra pair).  Every K
// contribution of such a range lands in rows of shells a and b, so those rows
// are accumulated in LDS and flushed to global memory once per item, and the
// bra rows of Dk that every K term reads are staged in LDS beside them.  The
// bra-side J block is summed over each wave.  The ket J block of every
// quartet is stored, not added: jkbuf[(tile * N3 * N4 + cd) * nq + q] holds
// 4 s sum_ab Dj_ab (ab|cd) for it, and a reduction over the bras that reach
// each ket (jket_reduce) adds those into J' with one atomic per ket element.
//
// The bra components are cut into tiles of TA x TB (first x second shell
// components); blockIdx.y selects the tile.  Each lane holds, per quartet,
// TA*TB*N3*N4 values and 2(TA+TB)(N4+1) + 2 TA TB accumulators and density
// elements, which is what bounds the register file: the generator picks the
// tile per class.  The ket loop runs c outer, d inner, so the ket J sum of
// element (c,d) is complete in a register when it is written.
struct JkItem {
    int start, end;   // quartet range [start, end)
    int bra;          // o1 | o2 << 15 | (s1 == s2) << 30, offsets local to the molecule
    int mol;
};

template <typename real>
struct JkTileList {
    const unsigned* kets;    // per quartet: o3 | o4 << 15 | (s3 == s4) << 30 | (P == Q) << 31
    const real* vals;
    double* jkbuf;           // [ntiles * N3 * N4][nq] ket J values per quartet
    int nq;
};

// Both stored lists of one class (FP64 values, and FP32 values below the
// Schwarz split) in one launch.  A work item is a chunk of one bra pair's
// kets: its FP64 quartets [start, end) then its FP32 quartets items32[item]
// (the bra's FP32 kets follow its FP64 kets in bound order), digested with
// the same staged rows.
struct JkTileArgs {
    const JkItem* items;
    const int2* items32;     // FP32 quartet range per item (NULL: FP64 list only)
    JkTileList<double> l64;
    JkTileList<float> l32;
    const MolInfo* mols;
    const double* Dj;
    const double* Dk;
    double* J;
    double* K;
    int nmax, nitems, grid;  // grid: blocks in x (<= nitems; persistent when smaller)
};

// The quartets [q0, q1) of one list for the current item and tile.  Quartet
// q0 + i goes to thread (i + shift) mod blockDim: with shift = the number of
// quartets of the FP64 range, the FP32 range continues where the FP64 range
// ended, so that in a short item no thread takes a quartet of each list.
template <int LA, int LB, int LC, int LD, int TA, int TB, typename real>
__device__ __forceinline__ void jk_tile_quartets(const JkTileList<real>& L, int q0, int q1, int shift,
                                                 const double* drows, double* krows, int n,
                                                 const double* __restrict__ Dj, double sbra,
                                                 const double (&d12)[TA][TB], double (&j12)[TA][TB], int a0, int b0,
                                                 int tile)
{
    constexpr int N2 = (LB + 1) * (LB + 2) / 2;
    constexpr int N3 = (LC + 1) * (LC + 2) / 2, N4 = (LD + 1) * (LD + 2) / 2;
    const int tid = threadIdx.x, nt = blockDim.x;
    const size_t nq = L.nq;
    const real* __restrict__ vals = L.vals;
    double* __restrict__ jkbuf = L.jkbuf;
    const int first = (tid - shift % nt + nt) % nt;
    for (int q = q0 + first; q < q1; q += nt) {
            const unsigned w = L.kets[q];
            const int o3 = w & 0x7fff, o4 = (w >> 15) & 0x7fff;
            double sc = sbra;
            if ((w >> 30) & 1u) sc *= 0.5;
            if (w >> 31) sc *= 0.5;
            double d14[TA][N4], d24[TB][N4], k14[TA][N4], k24[TB][N4];
#pragma unroll
            for (int d = 0; d < N4; ++d) {
#pragma unroll
                for (int x = 0; x < TA; ++x) {
                    d14[x][d] = drows[x * n + o4 + d];
                    k14[x][d] = 0.0;
                }
#pragma unroll
                for (int y = 0; y < TB; ++y) {
                    d24[y][d] = drows[(TA + y) * n + o4 + d];
                    k24[y][d] = 0.0;
                }
            }
#pragma unroll
            for (int c = 0; c < N3; ++c) {
                // all loads of this ket row first (values and Dj elements), so
                // that they are in flight together
                // values kept in the list's own type until use: an FP32 row
                // batch takes half the registers, so its loads still fit in
                // flight together
                real vv[N4][TA][TB];
                double dj[N4];
#pragma unroll
                for (int d = 0; d < N4; ++d) {
                    dj[d] = Dj[(size_t)(o3 + c) * n + o4 + d];
#pragma unroll
                    for (int x = 0; x < TA; ++x)
#pragma unroll
                        for (int y = 0; y < TB; ++y)
                            vv[d][x][y] = vals[(size_t)((((a0 + x) * N2 + b0 + y) * N3 + c) * N4 + d) * nq + q];
                }
                double d13[TA], d23[TB], k13[TA], k23[TB];
#pragma unroll
                for (int x = 0; x < TA; ++x) {
                    d13[x] = drows[x * n + o3 + c];
                    k13[x] = 0.0;
                }
#pragma unroll
                for (int y = 0; y < TB; ++y) {
                    d23[y] = drows[(TA + y) * n + o3 + c];
                    k23[y] = 0.0;
                }
#pragma unroll
                for (int d = 0; d < N4; ++d) {
                    double jcd = 0.0;
#pragma unroll
                    for (int x = 0; x < TA; ++x)
#pragma unroll
                        for (int y = 0; y < TB; ++y) {
                            const double v = sc * (double)vv[d][x][y];
                            j12[x][y] += dj[d] * v;
                            jcd += d12[x][y] * v;
                            k13[x] += d24[y][d] * v;
                            k14[x][d] += d23[y] * v;
                            k23[y] += d14[x][d] * v;
                            k24[y][d] += d13[x] * v;
                        }
                    jkbuf[(size_t)(tile * N3 * N4 + c * N4 + d) * nq + q] = 4.0 * jcd;
                }
#pragma unroll
                for (int x = 0; x < TA; ++x) atomicAdd(&krows[x * n + o3 + c], 2.0 * k13[x]);
#pragma unroll
                for (int y = 0; y < TB; ++y) atomicAdd(&krows[(TA + y) * n + o3 + c], 2.0 * k23[y]);
            }
#pragma unroll
            for (int d = 0; d < N4; ++d) {
#pragma unroll
                for (int x = 0; x < TA; ++x) atomicAdd(&krows[x * n + o4 + d], 2.0 * k14[x][d]);
#pragma unroll
                for (int y = 0; y < TB; ++y) atomicAdd(&krows[(TA + y) * n + o4 + d], 2.0 * k24[y][d]);
            }
            }
}

// Classes with at least three ket components (every ket but ss)
// digest one (quartet, ket row c) unit per thread instead of a quartet: N3
// times the independent units of an item, and only one row's values in
// registers.  Unit u of the range goes to row u / len, quartet q0 + u % len
// (consecutive lanes read consecutive quartets of one row), and to thread
// (u + shift) mod blockDim.  The K contributions of row c go to the LDS rows
// at once, the ket J values to their own jkbuf slots as in the quartet loop.
template <int LC, int LD>
struct JkRowSplit {
    static constexpr bool value = (LC + 1) * (LC + 2) / 2 * ((LD + 1) * (LD + 2) / 2) >= 3;
};

template <int LA, int LB, int LC, int LD, int TA, int TB, typename real>
__device__ __forceinline__ void jk_tile_rows(const JkTileList<real>& L, int q0, int q1, int shift,
                                             const double* drows, double* krows, int n,
                                             const double* __restrict__ Dj, double sbra,
                                             const double (&d12)[TA][TB], double (&j12)[TA][TB], int a0, int b0,
                                             int tile)
{
    constexpr int N2 = (LB + 1) * (LB + 2) / 2;
    constexpr int N3 = (LC + 1) * (LC + 2) / 2, N4 = (LD + 1) * (LD + 2) / 2;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int len = q1 - q0;
    if (len <= 0) return;
    const int total = N3 * len;
    const size_t nq = L.nq;
    const real* __restrict__ vals = L.vals;
    double* __restrict__ jkbuf = L.jkbuf;
    const int first = (tid - shift % nt + nt) % nt;
    for (int u = first; u < total; u += nt) {
        const int c = u / len;
        const int q = q0 + (u - c * len);
        const unsigned w = L.kets[q];
        const int o3 = w & 0x7fff, o4 = (w >> 15) & 0x7fff;
        double sc = sbra;
        if ((w >> 30) & 1u) sc *= 0.5;
        if (w >> 31) sc *= 0.5;
        real vv[N4][TA][TB];
        double dj[N4];
#pragma unroll
        for (int d = 0; d < N4; ++d) {
            dj[d] = Dj[(size_t)(o3 + c) * n + o4 + d];
#pragma unroll
            for (int x = 0; x < TA; ++x)
#pragma unroll
                for (int y = 0; y < TB; ++y)
                    vv[d][x][y] = vals[(size_t)((((a0 + x) * N2 + b0 + y) * N3 + c) * N4 + d) * nq + q];
        }
        double d13[TA], d23[TB], k13[TA], k23[TB];
#pragma unroll
        for (int x = 0; x < TA; ++x) {
            d13[x] = drows[x * n + o3 + c];
            k13[x] = 0.0;
        }
#pragma unroll
        for (int y = 0; y < TB; ++y) {
            d23[y] = drows[(TA + y) * n + o3 + c];
            k23[y] = 0.0;
        }
#pragma unroll
        for (int d = 0; d < N4; ++d) {
            double d14[TA], d24[TB], k14[TA], k24[TB];
#pragma unroll
            for (int x = 0; x < TA; ++x) {
                d14[x] = drows[x * n + o4 + d];
                k14[x] = 0.0;
            }
#pragma unroll
            for (int y = 0; y < TB; ++y) {
                d24[y] = drows[(TA + y) * n + o4 + d];
                k24[y] = 0.0;
            }
            double jcd = 0.0;
#pragma unroll
            for (int x = 0; x < TA; ++x)
#pragma unroll
                for (int y = 0; y < TB; ++y) {
                    const double v = sc * (double)vv[d][x][y];
                    j12[x][y] += dj[d] * v;
                    jcd += d12[x][y] * v;
                    k13[x] += d24[y] * v;
                    k14[x] += d23[y] * v;
                    k23[y] += d14[x] * v;
                    k24[y] += d13[x] * v;
                }
            jkbuf[(size_t)(tile * N3 * N4 + c * N4 + d) * nq + q] = 4.0 * jcd;
#pragma unroll
            for (int x = 0; x < TA; ++x) atomicAdd(&krows[x * n + o4 + d], 2.0 * k14[x]);
#pragma unroll
            for (int y = 0; y < TB; ++y) atomicAdd(&krows[(TA + y) * n + o4 + d], 2.0 * k24[y]);
        }
#pragma unroll
        for (int x = 0; x < TA; ++x) atomicAdd(&krows[x * n + o3 + c], 2.0 * k13[x]);
#pragma unroll
        for (int y = 0; y < TB; ++y) atomicAdd(&krows[(TA + y) * n + o3 + c], 2.0 * k23[y]);
    }
}

template <int LA, int LB, int LC, int LD, int TA, int TB>
__device__ __forceinline__ void jk_tile_item(const JkTileArgs& a, int item, double* lds, int tile)
{
    constexpr int N1 = (LA + 1) * (LA + 2) / 2, N2 = (LB + 1) * (LB + 2) / 2;
    static_assert(N1 % TA == 0 && N2 % TB == 0, "tile must divide the bra components");
    constexpr int NTB = N2 / TB, NR = TA + TB;
    const int a0 = (tile / NTB) * TA, b0 = (tile % NTB) * TB;
    const int tid = threadIdx.x, nt = blockDim.x;
    const JkItem it = a.items[item];
    const MolInfo mi = a.mols[it.mol];
    const int n = mi.n;
    double* drows = lds;            // NR rows of Dk: o1+a0.., then o2+b0..
    double* krows = lds + NR * n;   // K' rows, same order
    const int o1 = it.bra & 0x7fff, o2 = (it.bra >> 15) & 0x7fff;
    const double sbra = ((it.bra >> 30) & 1) ? 0.5 : 1.0;
    const double* __restrict__ Dj = a.Dj + mi.base;
    const double* __restrict__ Dk = a.Dk + mi.base;
    double* const Jm = a.J + mi.base;
    double* const Km = a.K + mi.base;
    // stage the Dk rows eight elements a thread at a time, all eight loads
    // issued before the first LDS write (the index is clamped into range so
    // that no branch separates them; out-of-range slots are not written)
    {
        constexpr int SU = 8;
        const int tot = NR * n;
        for (int i0 = tid; i0 < tot; i0 += SU * nt) {
            double v[SU];
#pragma unroll
            for (int u = 0; u < SU; ++u) {
                const int i = min(i0 + u * nt, tot - 1);
                const int r = i / n, c = i - r * n;
                const int row = r < TA ? o1 + a0 + r : o2 + b0 + (r - TA);
                v[u] = Dk[(size_t)row * n + c];
            }
#pragma unroll
            for (int u = 0; u < SU; ++u) {
                const int i = i0 + u * nt;
                if (i < tot) {
                    drows[i] = v[u];
                    krows[i] = 0.0;
                }
            }
        }
    }
    double d12[TA][TB], j12[TA][TB];
#pragma unroll
    for (int x = 0; x < TA; ++x)
#pragma unroll
        for (int y = 0; y < TB; ++y) {
            d12[x][y] = Dj[(size_t)(o1 + a0 + x) * n + o2 + b0 + y];
            j12[x][y] = 0.0;
        }
    __syncthreads();
    if (JkRowSplit<LC, LD>::value) {
        constexpr int N3 = (LC + 1) * (LC + 2) / 2;
        jk_tile_rows<LA, LB, LC, LD, TA, TB, double>(a.l64, it.start, it.end, 0, drows, krows, n, Dj, sbra, d12, j12,
                                                     a0, b0, tile);
        if (a.items32) {
            const int2 r32 = a.items32[item];
            jk_tile_rows<LA, LB, LC, LD, TA, TB, float>(a.l32, r32.x, r32.y, N3 * (it.end - it.start), drows, krows,
                                                        n, Dj, sbra, d12, j12, a0, b0, tile);
        }
    } else {
        jk_tile_quartets<LA, LB, LC, LD, TA, TB, double>(a.l64, it.start, it.end, 0, drows, krows, n, Dj, sbra, d12,
                                                         j12, a0, b0, tile);
        if (a.items32) {
            const int2 r32 = a.items32[item];
            jk_tile_quartets<LA, LB, LC, LD, TA, TB, float>(a.l32, r32.x, r32.y, it.end - it.start, drows, krows, n,
                                                            Dj, sbra, d12, j12, a0, b0, tile);
        }
    }
#pragma unroll
    for (int x = 0; x < TA; ++x)
#pragma unroll
        for (int y = 0; y < TB; ++y) {
            const double s = wave_sum(4.0 * j12[x][y]);
            if ((tid & 63) == 0 && s != 0.0) atomic_addd(&Jm[(size_t)(o1 + a0 + x) * n + o2 + b0 + y], s);
        }
    __syncthreads();
    for (int i = tid; i < NR * n; i += nt) {
        const double v = krows[i];
        if (v != 0.0) {
            const int r = i / n, c = i - r * n;
            const int row = r < TA ? o1 + a0 + r : o2 + b0 + (r - TA);
            atomic_addd(&Km[(size_t)row * n + c], v);
        }
    }
    __syncthreads();
}

template <int LA, int LB, int LC, int LD, int TA, int TB>
__global__ void __launch_bounds__(256) jk_tile_kernel(const JkTileArgs a)
{
    extern __shared__ double lds[];
    for (int item = blockIdx.x; item < a.nitems; item += gridDim.x)
        jk_tile_item<LA, LB, LC, LD, TA, TB>(a, item, lds, blockIdx.y);
}

// Dense scatter with 8-fold permutational symmetry (Methods/hartree_fock.py:314-325)
// into a slab of third indices l in [l0, l0 + lc): out[m][n][l - l0][s].  One
// thread per (component, quartet), component-major, so that consecutive
// threads read consecutive values: a single high-l quartet (an atom's
// (dd|dd)) spreads over its components instead of one thread writing them
// all.
template <int LA, int LB, int LC, int LD, typename real>
__global__ void __launch_bounds__(256) scatter_kernel(const ScatterArgs<real> a)
{
    constexpr int N2 = (LB + 1) * (LB + 2) / 2;
    constexpr int N3 = (LC + 1) * (LC + 2) / 2, N4 = (LD + 1) * (LD + 2) / 2;
    constexpr int N1 = (LA + 1) * (LA + 2) / 2, NC = N1 * N2 * N3 * N4;
    const long long nq = a.nq;
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= nq * NC) return;
    const int comp = (int)(t / nq);
    const long long q = t - comp * nq;
    const int id = comp % N4, ic = (comp / N4) % N3, ib = (comp / (N3 * N4)) % N2, ia = comp / (N2 * N3 * N4);
    const int2 pr = a.quartets[q];
    const PairHdr bh = a.pairs[pr.x], kh = a.pairs[pr.y];
    // Positions reached from two components (a pair of one shell with
    // itself, a quartet of one pair with itself) hold values equal up to
    // rounding; one canonical component writes them, so that the result
    // does not depend on the order of the threads.  The symmetric images
    // of the kept components still cover every position.
    if (bh.s1 == bh.s2 && ia < ib) return;
    if (kh.s1 == kh.s2 && ic < id) return;
    if (pr.x == pr.y && ia * N2 + ib < ic * N4 + id) return;
    const long long n = a.n, l0 = a.l0, lc = a.lc;
    const long long m = a.cart_off[bh.s1] + ia, nn = a.cart_off[bh.s2] + ib;
    const long long l = a.cart_off[kh.s1] + ic, s = a.cart_off[kh.s2] + id;
    const double v = (double)a.vals[(size_t)comp * nq + q];
    auto put = [&](long long i, long long j, long long k, long long r) {
        const long long u = k - l0;
        if (u >= 0 && u < lc) a.out[((i * n + j) * lc + u) * n + r] = v;
    };
    put(m, nn, l, s);
    put(nn, m, l, s);
    put(m, nn, s, l);
    put(nn, m, s, l);
    put(l, s, m, nn);
    put(l, s, nn, m);
    put(s, l, m, nn);
    put(s, l, nn, m);
}
