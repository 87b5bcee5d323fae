// scf.h -- device density purification (TC2) for the SCF loop.
#pragma once
#include <hip/hip_runtime.h>

// Work layout of one purification (4 ld^2 + 2 max_steps + 3 doubles, ld = m
// rounded up to 32): two pairs of ld x ld matrices (zero outside the m x m
// block; (X^2, 2 X - X^2) or (X, X^2)) that alternate between steps, the Gershgorin bounds, per
// step k the traces (tr X_k, tr X_k^2), and the stop word (index of the last
// step).
inline int tc2_ld(int m) { return (m + 31) / 32 * 32; }
struct Tc2Work {
    double* pair[2];    // pair[p]: X at +0, X^2 at +ld*ld
    double* bounds;     // [2]
    double* t;          // [2 * max_steps]
    int* stop;          // one double's room, right after the traces
    double* x0;         // start matrix (the X slot of pair 1 until step 1 overwrites it)
    int m, ld, max_steps;
    // candidate form (default; PQ_TC2_CAND=0 selects the X, X^2 pairs):
    // pair[p] holds X^2 at +0 and 2 X - X^2 at +ld*ld, and a step loads only
    // the one its input traces choose
    int cand;
};
inline long long tc2_work_doubles(int m, int max_steps)
{
    const long long ld = tc2_ld(m);
    return 4 * ld * ld + 2LL * max_steps + 3;
}
Tc2Work tc2_work(double* work, int m, int max_steps);

// Gershgorin bounds, zeroed trace slots and X0 = (e_max - F)/(e_max - e_min)
// (F symmetrised).
int tc2_start(const double* F, Tc2Work& w, hipStream_t st);
// Steps k0 .. k0 + steps - 1 (one launch each): step k forms X_k from the
// pair and traces of step k - 1 (X0 for k = 0) and writes its pair (X_k^2
// and 2 X_k - X_k^2, or X_k and X_k^2) and the traces of X_k and X_k^2 into
// slot k.  The step after the one whose input has
// tr(X - X^2) < tol is the last; it records its index in the stop word and
// later queued steps do nothing (early: such steps return at once; otherwise
// they still load and multiply, and a needed step saves a memory round trip).
int tc2_steps(const Tc2Work& w, int nocc, double tol, int k0, int steps, hipStream_t st, bool early = false);
// P = X_n symmetrised, n - 1 the recorded last step (or queued - 1 when no
// step recorded one), X_n from step n - 1's pair and traces.
int tc2_finish(const Tc2Work& w, int nocc, int queued, double* P, hipStream_t st);

// Scalars of one SCF cycle: dot products sum_e x_i[e] y_i[e] of npair
// pairs, then max and max-abs of nr arrays (out: npair + 2 nr values, in
// that order), over n elements.  work: scf_dots_work(n, nout) doubles; the
// results land at work + scf_dots_blocks(n) * nout.
#define SCF_MAXDOT 24
struct ScfDotArgs {
    const double* x[SCF_MAXDOT];
    const double* y[SCF_MAXDOT];
    const double* r[2];
    int npair, nr;
    long long n;
};
inline int scf_dots_blocks(long long n)
{
    const long long b = (n + 2047) / 2048;
    return (int)(b < 1 ? 1 : (b > 256 ? 256 : b));
}
inline long long scf_dots_work(long long n, int nout) { return (long long)scf_dots_blocks(n) * nout + nout; }
int scf_dots(const ScfDotArgs& a, double* work, hipStream_t st);

// out = sum_i c_i m_i (k <= SCF_MAXCOMB matrices of n elements).
#define SCF_MAXCOMB 16
struct LinCombArgs {
    const double* m[SCF_MAXCOMB];
    double c[SCF_MAXCOMB];
    int k;
    long long n;
};
int lincomb(const LinCombArgs& a, double* out, hipStream_t st);

// Sets the calling thread's error message (pq_last_error) and returns -1.
int pq_set_error(const char* m);
