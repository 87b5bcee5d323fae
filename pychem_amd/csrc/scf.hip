// scf.hip -- density purification for the device SCF loop.
//
// Trace-correcting purification (TC2: Niklasson, Phys. Rev. B 66, 155115
// (2002)) of a symmetric m x m matrix F in an orthonormal basis: X0 =
// (e_max - F)/(e_max - e_min) with Gershgorin bounds, then each step
// X <- X^2 or 2X - X^2, whichever brings tr X closer to the occupation.  The
// fixed point is the projector onto the nocc lowest eigenvectors of F.
//
// One launch per step.  Step k reads the previous step's pair (X_{k-1},
// X_{k-1}^2) and traces, forms X_k = X_{k-1}^2 or 2 X_{k-1} - X_{k-1}^2 on the
// fly while loading operands, and writes the new pair (X_k, X_k^2) and
// traces; pairs alternate between two buffers.  The matrices are kept with
// a leading dimension ld = m rounded up to 32 and zero outside the m x m
// block (zero rows and columns stay zero under X^2 and 2X - X^2 and add
// nothing to the traces), so every tile is whole and every load is an
// aligned 32-byte run without bounds checks.  X stays exactly symmetric
// (the start symmetrises F, each step computes only the upper 32 x 32 output
// tiles and mirrors them), so the right operand's columns are read as rows
// of X too (tiles of 16 x 16 by default, 32 x 32 with PQ_TC2_TILE=32: the
// smaller tiles give four times the blocks, each with half the operand
// loads and a quarter of the products; 0.41 against 0.52 ms per
// purification at m = 194, 0.73 against 0.77 ms at m = 384, where the
// doubled total operand traffic, 16 ld^3 / edge bytes per step, starts to
// bind).  A block's eight waves each compute the whole tile over a
// share of k, straight from global memory with all of a wave's 16-wide k
// blocks (up to four) of loads in flight at once, and one LDS pass sums the
// partial tiles.
// Products on v_mfma_f64_16x16x4f64: lane l supplies A[l%16][l/16] and
// B[l/16][l%16], result register r of lane l is D[4r + l/16][l%16].  Sizes here are a few hundred rows, so a step is
// latency-bound: no barrier inside the k loop, one launch per step.
#include <hip/hip_runtime.h>
#include <stdlib.h>

#include <algorithm>

#include "scf.h"

typedef double tc2_d4 __attribute__((ext_vector_type(4)));


// Gershgorin interval of row i: lohi[i] = F_ii - r_i, lohi[m + i] = F_ii + r_i,
// r_i = sum_{j != i} |F_ij| (one wave per row).
__global__ void __launch_bounds__(256) gersh_rows_kernel(const double* __restrict__ F, int m, double* lohi)
{
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (i >= m) return;
    double r = 0.0;
    for (int j = lane; j < m; j += 64)
        if (j != i) r += fabs(F[(size_t)i * m + j]);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) r += __shfl_xor(r, o, 64);
    if (lane == 0) {
        const double d = F[(size_t)i * m + i];
        lohi[i] = d - r;
        lohi[m + i] = d + r;
    }
}

// bounds[0] = min_i lohi[i], bounds[1] = max_i lohi[m + i] (one block); also
// zeroes the nt trace slots and clears the stop word.
__global__ void __launch_bounds__(256) gersh_reduce_kernel(const double* __restrict__ lohi, int m, double* bounds,
                                                           double* t, int nt, int* stop)
{
    __shared__ double lo[256], hi[256];
    for (int i = threadIdx.x; i < nt; i += blockDim.x) t[i] = 0.0;
    if (threadIdx.x == 0) *stop = 0x7fffffff;
    double mn = 1e300, mx = -1e300;
    for (int i = threadIdx.x; i < m; i += blockDim.x) {
        mn = fmin(mn, lohi[i]);
        mx = fmax(mx, lohi[m + i]);
    }
    lo[threadIdx.x] = mn;
    hi[threadIdx.x] = mx;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            lo[threadIdx.x] = fmin(lo[threadIdx.x], lo[threadIdx.x + s]);
            hi[threadIdx.x] = fmax(hi[threadIdx.x], hi[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        bounds[0] = lo[0];
        bounds[1] = hi[0];
    }
}

// X0 = (e_max I - (F + F^T)/2) / (e_max - e_min) in the ld x ld layout, zero
// outside the m x m block
__global__ void tc2_init_kernel(const double* __restrict__ F, int m, int ld, const double* __restrict__ bounds,
                                double* X)
{
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= (long long)ld * ld) return;
    const int i = (int)(t / ld), j = (int)(t % ld);
    double x = 0.0;
    if (i < m && j < m) {
        const double emin = bounds[0], emax = bounds[1];
        const double f = 0.5 * (F[(size_t)i * m + j] + F[(size_t)j * m + i]);
        x = ((i == j ? emax : 0.0) - f) / (emax - emin);
    }
    X[t] = x;
}

struct Tc2Step {
    const double* P;      // X_{k-1}
    const double* Q;      // X_{k-1}^2 (unused when direct)
    const double* tin;    // tr X_{k-1}, tr X_{k-1}^2 (unused when direct)
    double* X;            // X_k
    double* X2;           // X_k^2
    double* tout;         // += tr X_k, tr X_k^2 (zero on entry)
    int* stop;            // index of the last step (0x7fffffff until known)
    double nocc, tol;
    int ld, nt;           // leading dimension (a multiple of 32), tiles per side
    int k;                // this step
    int direct;           // step 0: X_k = P
};

__device__ __forceinline__ double tc2_wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

#define TC2_W 8   // waves per block (each sums a share of k)

// Operands of one round: the C k blocks kb = kb0 + c * TC2_W of 16 k each
// (blocks past the end read the last block again; keep masks them to zero).
// Inside a block, MFMA step s pairs lane group q with k = 16 kb + 4 q + s, so
// a lane's four k values of one row are consecutive (one aligned 32-byte
// load; the product is a sum over k in any order, the same order for both
// operands).  Rows of the left operand: i0 + 16 u + li; columns of the right
// operand (read as rows of the symmetric X): j0 + 16 v + li.
template <bool DIRECT, int C, int NU>
struct Tc2Round {
    tc2_d4 pa[C][NU], pb[C][NU], qa[C][NU], qb[C][NU];
    double keep[C];
    __device__ __forceinline__ void load(const Tc2Step& a, int kb0, int nkb, int i0, int j0, int li, int lq)
    {
        const int ld = a.ld;
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const int kb = kb0 + c * TC2_W;
            const int k = 16 * min(kb, nkb - 1) + 4 * lq;
            keep[c] = (double)(kb < nkb);
#pragma unroll
            for (int u = 0; u < NU; ++u) {
                const size_t ia = (size_t)(i0 + 16 * u + li) * ld + k, ib = (size_t)(j0 + 16 * u + li) * ld + k;
                pa[c][u] = *(const tc2_d4*)(a.P + ia);
                pb[c][u] = *(const tc2_d4*)(a.P + ib);
                if (!DIRECT) {
                    qa[c][u] = *(const tc2_d4*)(a.Q + ia);
                    qb[c][u] = *(const tc2_d4*)(a.Q + ib);
                }
            }
        }
    }
    __device__ __forceinline__ void mul(double cp, double cq, tc2_d4 (&acc)[NU][NU]) const
    {
#pragma unroll
        for (int c = 0; c < C; ++c) {
            tc2_d4 fa[NU], fb[NU];
#pragma unroll
            for (int u = 0; u < NU; ++u) {
                fa[u] = DIRECT ? keep[c] * pa[c][u] : keep[c] * (cp * pa[c][u] + cq * qa[c][u]);
                fb[u] = DIRECT ? keep[c] * pb[c][u] : keep[c] * (cp * pb[c][u] + cq * qb[c][u]);
            }
#pragma unroll
            for (int s = 0; s < 4; ++s)
#pragma unroll
                for (int u = 0; u < NU; ++u)
#pragma unroll
                    for (int v = 0; v < NU; ++v)
                        acc[u][v] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[u][s], fb[v][s], acc[u][v], 0, 0, 0);
        }
    }
};

// Step k.  DIRECT (step 0): X_k = P.  Otherwise X_k = Q (X_{k-1}^2) when
// tr X_{k-1}^2 is the closer to nocc, else 2 P - Q.  The first round's
// operand loads (both matrices) are issued before the step's inputs (stop
// word, traces) are read, so deciding the branch costs no memory round trip
// of its own.  Once tr(X_{k-1} - X_{k-1}^2) < tol, step k is the last one
// (the stop word records k); steps after the recorded one store nothing, so
// the host may queue more steps than turn out to be needed.  Such a step
// still loads and multiplies (nothing waits on the stop word before the
// stores: an early exit there made the compiler wait for it before issuing
// the operand loads).
template <bool DIRECT, int C, int NU>
__device__ __forceinline__ void tc2_step_body(const Tc2Step& a, double (&red)[TC2_W][NU * NU][4][64])
{
    constexpr int TT = 16 * NU;   // output tile edge
    // upper tile (bi, bj), bi <= bj, from the linear block index
    int t = blockIdx.x, bi = 0;
    while (t >= a.nt - bi) {
        t -= a.nt - bi;
        ++bi;
    }
    const int bj = bi + t;
    const int i0 = bi * TT, j0 = bj * TT;
    const int ld = a.ld;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int li = lane & 15, lq = lane >> 4;
    // the block's own tile of X_k (and its mirror) is written at the end;
    // its elements are loaded now, with the operands
    constexpr int NX = (TT * TT + 64 * TC2_W - 1) / (64 * TC2_W);
    double xp[NX], xq[NX];
#pragma unroll
    for (int n = 0; n < NX; ++n) {
        const int e = min(tid + 64 * TC2_W * n, TT * TT - 1);   // past the tile: a repeat, not stored
        const size_t idx = (size_t)(i0 + e / TT) * ld + j0 + e % TT;
        xp[n] = a.P[idx];
        xq[n] = DIRECT ? 0.0 : a.Q[idx];
    }
    // the step's inputs (written by earlier launches: plain loads)
    int stop = 0x7fffffff;
    double t1 = 0.0, t2 = 0.0;
    if (!DIRECT) {
        stop = *a.stop;
        t1 = a.tin[0];
        t2 = a.tin[1];
    }
    const int nkb = ld / 16;
    Tc2Round<DIRECT, C, NU> rd;
    rd.load(a, wave, nkb, i0, j0, li, lq);
    double cp = 1.0, cq = 0.0;   // X_k = cp P + cq Q
    bool active = true;          // uniform over the grid
    if (!DIRECT) {
        active = !(stop < a.k);
        if (active && t1 - t2 < a.tol && blockIdx.x == 0 && threadIdx.x == 0) *a.stop = a.k;
        const bool up = fabs(t2 - a.nocc) < fabs(2.0 * t1 - t2 - a.nocc);
        cp = up ? 0.0 : 2.0;
        cq = up ? 1.0 : -1.0;
    }
    tc2_d4 acc[NU][NU];
#pragma unroll
    for (int u = 0; u < NU; ++u)
#pragma unroll
        for (int v = 0; v < NU; ++v) acc[u][v] = (tc2_d4){0.0, 0.0, 0.0, 0.0};
    rd.mul(cp, cq, acc);
    for (int kb0 = wave + TC2_W * C; kb0 < nkb; kb0 += TC2_W * C) {   // more rounds (ld > 128 C)
        rd.load(a, kb0, nkb, i0, j0, li, lq);
        rd.mul(cp, cq, acc);
    }
#pragma unroll
    for (int u = 0; u < NU; ++u)
#pragma unroll
        for (int v = 0; v < NU; ++v)
#pragma unroll
            for (int r = 0; r < 4; ++r) red[wave][NU * u + v][r][lane] = acc[u][v][r];
    __syncthreads();
    if (!active) return;
    // output slot e = (sub-tile g, register r, lane l), 1024 of them, summed
    // over the waves; sub-tile g: u = g / 2, v = g % 2
    double trx = 0.0, trx2 = 0.0;
    for (int e = tid; e < NU * NU * 256; e += 64 * TC2_W) {
        const int g = e >> 8, r = (e >> 6) & 3, l = e & 63, u = g / NU, v = g % NU;
        double x2 = 0.0;
#pragma unroll
        for (int w = 0; w < TC2_W; ++w) x2 += red[w][g][r][l];
        const int gi = i0 + 16 * u + 4 * r + (l >> 4), gj = j0 + 16 * v + (l & 15);
        a.X2[(size_t)gi * ld + gj] = x2;
        if (bi != bj) a.X2[(size_t)gj * ld + gi] = x2;
        else if (gi == gj) trx2 += x2;
    }
#pragma unroll
    for (int n = 0; n < NX; ++n) {
        const int e = tid + 64 * TC2_W * n;
        if (e >= TT * TT) break;
        const int gi = i0 + e / TT, gj = j0 + e % TT;
        const double x = DIRECT ? xp[n] : cp * xp[n] + cq * xq[n];
        a.X[(size_t)gi * ld + gj] = x;
        if (bi != bj) a.X[(size_t)gj * ld + gi] = x;
        else if (gi == gj) trx += x;
    }
    if (bi == bj) {
        trx = tc2_wave_sum(trx);
        trx2 = tc2_wave_sum(trx2);
        if (lane == 0) {
            atomicAdd(&a.tout[0], trx);
            atomicAdd(&a.tout[1], trx2);
        }
    }
}

// Candidate form (the default; the step is bound by its operand loads at
// every size measured, m = 194 and 384): step k stores both candidates for X_{k+1}, U_k = X_k^2 (in the X
// slot) and W_k = 2 X_k - X_k^2 (in the X^2 slot), with the traces of X_k
// and X_k^2.  Step k + 1 reads the traces first, chooses U or W and loads
// only that matrix: half the operand bytes of loading X and X^2 together,
// for one more memory round trip.  The block's own tile of X_k is read in
// the layout of the products, so that W = 2 X - X^2 is formed where X^2 is.
template <bool DIRECT, int C, int NU>
__device__ __forceinline__ void tc2_cand_body(const Tc2Step& a, double (&red)[TC2_W][NU * NU][4][64])
{
    int t = blockIdx.x, bi = 0;
    while (t >= a.nt - bi) {
        t -= a.nt - bi;
        ++bi;
    }
    const int bj = bi + t;
    const int i0 = bi * 16 * NU, j0 = bj * 16 * NU;
    const int ld = a.ld;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int li = lane & 15, lq = lane >> 4;
    const double* Xk = a.P;   // X_k: x0 at step 0, else the chosen candidate
    bool active = true;       // uniform over the grid
    if (!DIRECT) {
        const int stop = *a.stop;
        const double t1 = a.tin[0], t2 = a.tin[1];
        active = !(stop < a.k);
        if (active && t1 - t2 < a.tol && blockIdx.x == 0 && threadIdx.x == 0) *a.stop = a.k;
        const bool up = fabs(t2 - a.nocc) < fabs(2.0 * t1 - t2 - a.nocc);
        Xk = up ? a.P : a.Q;
    }
    constexpr int NOUT = NU * NU * 256;
    constexpr int NO = (NOUT + 64 * TC2_W - 1) / (64 * TC2_W);
    auto slot = [&](int e, int& gi, int& gj) {   // output slot e = (sub-tile g, register r, lane l)
        const int g = e >> 8, r = (e >> 6) & 3, l = e & 63, u = g / NU, v = g % NU;
        gi = i0 + 16 * u + 4 * r + (l >> 4);
        gj = j0 + 16 * v + (l & 15);
    };
    double xo[NO];
#pragma unroll
    for (int n = 0; n < NO; ++n) {
        int gi, gj;
        slot(min(tid + 64 * TC2_W * n, NOUT - 1), gi, gj);   // past the tile: a repeat, not stored
        xo[n] = Xk[(size_t)gi * ld + gj];
    }
    Tc2Step s = a;
    s.P = Xk;
    const int nkb = ld / 16;
    tc2_d4 acc[NU][NU];
#pragma unroll
    for (int u = 0; u < NU; ++u)
#pragma unroll
        for (int v = 0; v < NU; ++v) acc[u][v] = (tc2_d4){0.0, 0.0, 0.0, 0.0};
    Tc2Round<true, C, NU> rd;
    for (int kb0 = wave; kb0 < nkb; kb0 += TC2_W * C) {
        rd.load(s, kb0, nkb, i0, j0, li, lq);
        rd.mul(1.0, 0.0, acc);
    }
#pragma unroll
    for (int u = 0; u < NU; ++u)
#pragma unroll
        for (int v = 0; v < NU; ++v)
#pragma unroll
            for (int r = 0; r < 4; ++r) red[wave][NU * u + v][r][lane] = acc[u][v][r];
    __syncthreads();
    if (!active) return;
    double trx = 0.0, trx2 = 0.0;
#pragma unroll
    for (int n = 0; n < NO; ++n) {
        const int e = tid + 64 * TC2_W * n;
        if (e >= NOUT) break;
        const int g = e >> 8, r = (e >> 6) & 3, l = e & 63;
        double x2 = 0.0;
#pragma unroll
        for (int w = 0; w < TC2_W; ++w) x2 += red[w][g][r][l];
        int gi, gj;
        slot(e, gi, gj);
        const double x = xo[n], wv = 2.0 * x - x2;
        a.X[(size_t)gi * ld + gj] = x2;
        a.X2[(size_t)gi * ld + gj] = wv;
        if (bi != bj) {
            a.X[(size_t)gj * ld + gi] = x2;
            a.X2[(size_t)gj * ld + gi] = wv;
        } else if (gi == gj) {
            trx += x;
            trx2 += x2;
        }
    }
    if (bi == bj) {
        trx = tc2_wave_sum(trx);
        trx2 = tc2_wave_sum(trx2);
        if (lane == 0) {
            atomicAdd(&a.tout[0], trx);
            atomicAdd(&a.tout[1], trx2);
        }
    }
}

// C = k blocks of 16 per wave and round: ceil(ld / (16 TC2_W)), at most 4.
// EARLY: read the stop word first and return at once past the recorded last
// step (a queued step that is most likely not needed then costs only its
// launch; when it is needed it pays one more memory round trip).
template <int C, bool EARLY, int NU, bool CAND>
__global__ void __launch_bounds__(64 * TC2_W) tc2_step_kernel(const Tc2Step a)
{
    __shared__ double red[TC2_W][NU * NU][4][64];   // [wave][sub-tile][register][lane]
    if (a.direct) {
        if (CAND) tc2_cand_body<true, C, NU>(a, red);
        else tc2_step_body<true, C, NU>(a, red);
        return;
    }
    if (EARLY && *a.stop < a.k) return;
    if (CAND) tc2_cand_body<false, C, NU>(a, red);
    else tc2_step_body<false, C, NU>(a, red);
}

// P = sym(X_n) (m x m): n - 1 = the recorded last step, or the last queued
// step when none was recorded; X_n from step n - 1's pair (ld layout) and
// traces.
__global__ void tc2_finish_kernel(const double* __restrict__ pair0, const double* __restrict__ pair1,
                                  const double* __restrict__ tslots, const int* __restrict__ stop, int queued,
                                  double nocc, int m, int ld, int cand, double* P)
{
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= (long long)m * m) return;
    const int i = (int)(t / m), j = (int)(t % m);
    const int last = min(*stop, queued - 1);
    const double* Xp = (last & 1) ? pair1 : pair0;
    const double* X2p = Xp + (size_t)ld * ld;
    const double t1 = tslots[2 * last], t2 = tslots[2 * last + 1];
    const bool up = fabs(t2 - nocc) < fabs(2.0 * t1 - t2 - nocc);
    const size_t ij = (size_t)i * ld + j, ji = (size_t)j * ld + i;
    // candidate form: the pair holds U = X^2 and W = 2 X - X^2 themselves
    const double a = cand ? (up ? Xp[ij] : X2p[ij]) : (up ? X2p[ij] : 2.0 * Xp[ij] - X2p[ij]);
    const double b = cand ? (up ? Xp[ji] : X2p[ji]) : (up ? X2p[ji] : 2.0 * Xp[ji] - X2p[ji]);
    P[t] = 0.5 * (a + b);
}

static inline unsigned nblk(long long n) { return (unsigned)((n + 255) / 256); }
// k blocks of 16 per wave and round of the step kernel
static inline int tc2_rounds(int ld) { return std::min(4, std::max(1, (ld / 16 + TC2_W - 1) / TC2_W)); }

static inline double* tc2_X(const Tc2Work& w, int k) { return w.pair[k & 1]; }
static inline double* tc2_X2(const Tc2Work& w, int k) { return w.pair[k & 1] + (size_t)w.ld * w.ld; }

Tc2Work tc2_work(double* work, int m, int max_steps)
{
    Tc2Work w;
    w.m = m;
    w.ld = tc2_ld(m);
    w.max_steps = max_steps;
    const size_t ll = (size_t)w.ld * w.ld;
    w.pair[0] = work;
    w.pair[1] = work + 2 * ll;
    w.bounds = work + 4 * ll;
    w.t = w.bounds + 2;
    w.stop = (int*)(w.t + 2 * max_steps);
    w.x0 = nullptr;
    static const int cand_env = [] {
        const char* e = getenv("PQ_TC2_CAND");
        return e ? atoi(e) : -1;
    }();
    w.cand = cand_env != 0;
    return w;
}

int tc2_start(const double* F, Tc2Work& w, hipStream_t st)
{
    const int m = w.m, ld = w.ld;
    if (m <= 0 || w.max_steps < 1) return -1;
    // X0 goes into the X slot of pair 1, read by step 0 as its direct input;
    // the X^2 slot of pair 1 is free until step 1: row intervals go there
    w.x0 = w.pair[1];
    double* lohi = w.pair[1] + (size_t)ld * ld;
    hipLaunchKernelGGL(gersh_rows_kernel, dim3((m + 3) / 4), dim3(256), 0, st, F, m, lohi);
    hipLaunchKernelGGL(gersh_reduce_kernel, dim3(1), dim3(256), 0, st, (const double*)lohi, m, w.bounds, w.t,
                       2 * w.max_steps, w.stop);
    hipLaunchKernelGGL(tc2_init_kernel, dim3(nblk((long long)ld * ld)), dim3(256), 0, st, F, m, ld,
                       (const double*)w.bounds, w.x0);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Output tile edge of the step kernel: 16 (NU = 1) or 32 (NU = 2);
// PQ_TC2_TILE=32 selects the larger tiles.
static int tc2_nu()
{
    static const int nu = [] {
        const char* e = getenv("PQ_TC2_TILE");
        return (e && atoi(e) == 32) ? 2 : 1;
    }();
    return nu;
}

template <int NU, bool CAND>
static void tc2_launch(int c, bool early, dim3 grid, dim3 block, const Tc2Step& a, hipStream_t st)
{
    if (early) {
        if (c == 1) hipLaunchKernelGGL((tc2_step_kernel<1, true, NU, CAND>), grid, block, 0, st, a);
        else if (c == 2) hipLaunchKernelGGL((tc2_step_kernel<2, true, NU, CAND>), grid, block, 0, st, a);
        else if (c == 3) hipLaunchKernelGGL((tc2_step_kernel<3, true, NU, CAND>), grid, block, 0, st, a);
        else hipLaunchKernelGGL((tc2_step_kernel<4, true, NU, CAND>), grid, block, 0, st, a);
    } else {
        if (c == 1) hipLaunchKernelGGL((tc2_step_kernel<1, false, NU, CAND>), grid, block, 0, st, a);
        else if (c == 2) hipLaunchKernelGGL((tc2_step_kernel<2, false, NU, CAND>), grid, block, 0, st, a);
        else if (c == 3) hipLaunchKernelGGL((tc2_step_kernel<3, false, NU, CAND>), grid, block, 0, st, a);
        else hipLaunchKernelGGL((tc2_step_kernel<4, false, NU, CAND>), grid, block, 0, st, a);
    }
}

int tc2_steps(const Tc2Work& w, int nocc, double tol, int k0, int steps, hipStream_t st, bool early)
{
    if (k0 < 0 || steps < 0 || k0 + steps > w.max_steps) return -1;
    const int nu = tc2_nu();
    const int nt = w.ld / (16 * nu);   // ld is a multiple of 32
    for (int k = k0; k < k0 + steps; ++k) {
        Tc2Step a;
        a.direct = k == 0;
        a.k = k;
        a.P = a.direct ? w.x0 : tc2_X(w, k - 1);
        a.Q = a.direct ? nullptr : tc2_X2(w, k - 1);
        a.tin = a.direct ? nullptr : w.t + 2 * (k - 1);
        a.X = tc2_X(w, k);
        a.X2 = tc2_X2(w, k);
        a.tout = w.t + 2 * k;
        a.stop = w.stop;
        a.nocc = (double)nocc;
        a.tol = tol;
        a.ld = w.ld;
        a.nt = nt;
        const dim3 grid(nt * (nt + 1) / 2), block(64 * TC2_W);
        const int c = tc2_rounds(w.ld);
        if (w.cand) {
            if (nu == 1) tc2_launch<1, true>(c, early, grid, block, a, st);
            else tc2_launch<2, true>(c, early, grid, block, a, st);
        } else {
            if (nu == 1) tc2_launch<1, false>(c, early, grid, block, a, st);
            else tc2_launch<2, false>(c, early, grid, block, a, st);
        }
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int tc2_finish(const Tc2Work& w, int nocc, int queued, double* P, hipStream_t st)
{
    const int m = w.m;
    if (queued < 1 || queued > w.max_steps) return -1;
    hipLaunchKernelGGL(tc2_finish_kernel, dim3(nblk((long long)m * m)), dim3(256), 0, st,
                       (const double*)w.pair[0], (const double*)w.pair[1], (const double*)w.t, (const int*)w.stop,
                       queued, (double)nocc, m, w.ld, w.cand, P);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// --------------------------------------------------------------------------
// Scalars of one SCF cycle in two launches: dot products of up to
// SCF_MAXDOT (x, y) pairs and the max / max-abs of up to two arrays, all over
// n elements.  Pass 1: every block reduces a grid-strided share into
// partial[block][out]; pass 2 (one block) reduces the partials.
// --------------------------------------------------------------------------
__device__ __forceinline__ double scf_block_reduce(double v, double* red, bool is_max)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double w = __shfl_xor(v, o, 64);
        v = is_max ? fmax(v, w) : v + w;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double r = red[0];
    for (int k = 1; k < nw; ++k) r = is_max ? fmax(r, red[k]) : r + red[k];
    return r;
}

__global__ void __launch_bounds__(256) scf_dots_kernel(const ScfDotArgs a, double* partial)
{
    __shared__ double red[4];
    const int nout = a.npair + 2 * a.nr;
    for (int o = 0; o < nout; ++o) {
        double v;
        const bool is_max = o >= a.npair;
        if (!is_max) {
            v = 0.0;
            const double* __restrict__ x = a.x[o];
            const double* __restrict__ y = a.y[o];
            for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < a.n;
                 e += (long long)gridDim.x * blockDim.x)
                v += x[e] * y[e];
        } else {
            const int j = (o - a.npair) >> 1;
            const bool absval = (o - a.npair) & 1;
            const double* __restrict__ r = a.r[j];
            v = -1e300;
            for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < a.n;
                 e += (long long)gridDim.x * blockDim.x)
                v = fmax(v, absval ? fabs(r[e]) : r[e]);
        }
        v = scf_block_reduce(v, red, is_max);
        if (threadIdx.x == 0) partial[(size_t)blockIdx.x * nout + o] = v;
    }
}

__global__ void __launch_bounds__(256) scf_dots_final_kernel(const double* __restrict__ partial, int nblocks, int npair,
                                                             int nout, double* out)
{
    __shared__ double red[4];
    for (int o = 0; o < nout; ++o) {
        const bool is_max = o >= npair;
        double v = is_max ? -1e300 : 0.0;
        for (int b = threadIdx.x; b < nblocks; b += blockDim.x) {
            const double w = partial[(size_t)b * nout + o];
            v = is_max ? fmax(v, w) : v + w;
        }
        v = scf_block_reduce(v, red, is_max);
        if (threadIdx.x == 0) out[o] = v;
    }
}

int scf_dots(const ScfDotArgs& a, double* work, hipStream_t st)
{
    if (a.npair < 0 || a.npair > SCF_MAXDOT || a.nr < 0 || a.nr > 2 || a.n <= 0) return -1;
    const int nout = a.npair + 2 * a.nr;
    if (nout == 0) return 0;
    const int nb = scf_dots_blocks(a.n);
    double* partial = work;
    double* out = work + (size_t)nb * nout;
    hipLaunchKernelGGL(scf_dots_kernel, dim3(nb), dim3(256), 0, st, a, partial);
    hipLaunchKernelGGL(scf_dots_final_kernel, dim3(1), dim3(256), 0, st, (const double*)partial, nb, a.npair, nout,
                       out);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// out = sum_i c_i m_i over n elements (the DIIS extrapolation).
__global__ void __launch_bounds__(256) lincomb_kernel(const LinCombArgs a, double* out)
{
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < a.n; e += (long long)gridDim.x * blockDim.x) {
        double v = 0.0;
        for (int i = 0; i < a.k; ++i) v += a.c[i] * a.m[i][e];
        out[e] = v;
    }
}

int lincomb(const LinCombArgs& a, double* out, hipStream_t st)
{
    if (a.k < 1 || a.k > SCF_MAXCOMB || a.n <= 0) return -1;
    const long long nb = std::min<long long>(1024, (a.n + 255) / 256);
    hipLaunchKernelGGL(lincomb_kernel, dim3((unsigned)nb), dim3(256), 0, st, a, out);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
