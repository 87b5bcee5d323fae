// scf_run.hip -- the SCF iteration as one library call.
//
// The loop of the reference's SCF driver (Methods/hartree_fock.py:79-140:
// Fock build, DIIS, new density, energy, convergence on |dE|) with the
// density from purification instead of a diagonalisation and the energy test
// on the variational energy of each cycle's input density, all launched from
// the host side of the library: per cycle one device-to-host copy of the
// cycle's scalars (energy terms, residual maxima, DIIS overlaps) and the
// purification's own step-count read, and no other host round trip.
//
// Per cycle and spin s (restricted: one spin, J from 2 D):
//   J, K_s      = pq_jk(D_total, D_s)                (Cartesian; spherical bases
//                                                     go through T on both sides)
//   F_s         = H + J + K_s                        (stored in the DIIS slot)
//   F'_s        = X^T F_s X                          (orthonormal basis)
//   r_s         = X^T (S D_s F_s - F_s D_s S) X      (orbital gradient)
//               = P_s F'_s - F'_s P_s                when D_s = X P_s X^T, i.e.
//                                                     every cycle after the first
//                                                     (X^T S X = 1)
//   E           = sum D_s (H + F_s) / 2 (summed over spins; restricted
//                 sum D (H + F))
//   test        |E - E_prev| <= conv and max|r| <= sqrt(conv) -> done; the
//                 result is this cycle's input density and its own Fock
//   F'_ext      = Pulay DIIS over the stored (r, F') pairs (the same
//                 admission, trimming and C1 system as
//                 Methods/hf_extensions/diis.py:14-155, with the caller's
//                 threshold, condition limit and size)
//   P_s         = purify(F'_ext, n_s);  D_s = X P_s X^T
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/pychem_hip.h"
#include "mp2.h"
#include "scf.h"

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------

// F = H + js * J + K (n2 elements).
__global__ void __launch_bounds__(256) scf_fock_kernel(const double* __restrict__ H, const double* __restrict__ J,
                                                       double js, const double* __restrict__ K, double* __restrict__ F,
                                                       long long n2)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n2) F[i] = H[i] + js * J[i] + K[i];
}

// C = A + B (n2 elements).
__global__ void __launch_bounds__(256) scf_add_kernel(const double* __restrict__ A, const double* __restrict__ B,
                                                      double* __restrict__ C, long long n2)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n2) C[i] = A[i] + B[i];
}

// r (n x n, row stride n) = A - A^T on the leading m x m block (A: m x m,
// row stride m), zero elsewhere.
__global__ void __launch_bounds__(256) scf_antisym_kernel(const double* __restrict__ A, int m, double* __restrict__ r,
                                                          int n)
{
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long long)n * n) return;
    const int i = (int)(e / n), j = (int)(e % n);
    r[e] = (i < m && j < m) ? A[(long long)i * m + j] - A[(long long)j * m + i] : 0.0;
}

// P (m x m) = d * identity.
__global__ void __launch_bounds__(256) scf_diag_kernel(double* __restrict__ P, int m, double d)
{
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long long)m * m) return;
    P[e] = (e / m == e % m) ? d : 0.0;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
namespace {

int bad(const std::string& m)
{
    return pq_set_error(m.c_str());
}

#define SCFCHK(x)                                                                          \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) return bad(std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)
#define SCFCALL(x)              \
    do {                        \
        if ((x) != 0) return -1; \
    } while (0)

inline unsigned blocks_for(long long n) { return (unsigned)((n + 255) / 256); }

// C (M x N, ldc) = op(A) op(B): ta = A stored K x M, tb = B stored N x K.
int gemm(int ta, int tb, int M, int N, int K, const double* A, long long lda, const double* B, long long ldb,
         double* C, long long ldc, hipStream_t st)
{
    GemmF64Args g{A, B, C, lda, ldb, ldc, 0, 0, 0, M, N, K, ta, tb, 1.0, 0.0};
    if (gemm_f64_launch(g, 1, st)) return bad("pq_scf_run: GEMM launch failed");
    return 0;
}

// Work layout: chunks of whole 256-byte lines.
struct Carve {
    double* base;
    long long used = 0;
    double* take(long long n)
    {
        double* p = base ? base + used : nullptr;
        used += (n + 31) / 32 * 32;
        return p;
    }
};

#define SCF_MAXV 16         // DIIS slots per spin (vectors kept + the new one)
#define SCF_MAX_STEPS 200   // purification step limit

struct Bufs {
    // Cartesian J/K and transforms
    double *J, *Ka, *Kb, *J2, *Dt, *Dca, *Dcb, *Dct, *Tmp;
    // per spin
    double *F[2][SCF_MAXV], *Fo[2][SCF_MAXV], *R[2][SCF_MAXV];
    double *P[2], *Fx[2];
    // temporaries
    double *T1, *A, *SD, *B;
    double *pur, *dots;
};

long long lay_out(Bufs& b, double* base, int n, int nx, int nc, int nsp, int nslot)
{
    Carve c{base};
    const long long n2 = (long long)n * n, x2 = (long long)nx * nx, c2 = (long long)nc * nc;
    const long long big = std::max(n2, c2);
    b.J = c.take(c2);
    b.Ka = c.take(c2);
    b.Kb = nsp > 1 ? c.take(c2) : nullptr;
    b.J2 = nsp > 1 ? c.take(c2) : nullptr;
    b.Dt = nsp > 1 ? c.take(n2) : nullptr;
    const bool sph = nc != n;
    b.Dca = sph ? c.take(c2) : nullptr;
    b.Dcb = (sph && nsp > 1) ? c.take(c2) : nullptr;
    b.Dct = (sph && nsp > 1) ? c.take(c2) : nullptr;
    b.Tmp = sph ? c.take((long long)n * nc) : nullptr;
    for (int s = 0; s < 2; ++s)
        for (int v = 0; v < SCF_MAXV; ++v) {
            const bool on = s < nsp && v < nslot;
            b.F[s][v] = on ? c.take(n2) : nullptr;
            b.Fo[s][v] = on ? c.take(x2) : nullptr;
            b.R[s][v] = on ? c.take(n2) : nullptr;
        }
    for (int s = 0; s < 2; ++s) {
        b.P[s] = s < nsp ? c.take(x2) : nullptr;
        b.Fx[s] = s < nsp ? c.take(x2) : nullptr;
    }
    b.T1 = c.take((long long)n * nx);
    b.A = c.take(big);
    b.SD = c.take(big);
    b.B = c.take(big);
    b.pur = c.take(tc2_work_doubles(nx, SCF_MAX_STEPS));
    b.dots = c.take(scf_dots_work(n2, SCF_MAXDOT + 4));
    return c.used;
}

// Eigenvalues of a symmetric k x k matrix (cyclic Jacobi).
void sym_eig(std::vector<double> a, int k, std::vector<double>& ev, std::vector<double>* vec)
{
    std::vector<double> v(vec ? (size_t)k * k : 0, 0.0);
    if (vec)
        for (int i = 0; i < k; ++i) v[(size_t)i * k + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int i = 0; i < k; ++i)
            for (int j = 0; j < k; ++j) {
                const double x = a[(size_t)i * k + j] * a[(size_t)i * k + j];
                tot += x;
                if (i != j) off += x;
            }
        if (off <= 1e-30 * tot || off == 0.0) break;
        for (int p = 0; p < k; ++p)
            for (int q = p + 1; q < k; ++q) {
                const double apq = a[(size_t)p * k + q];
                if (apq == 0.0) continue;
                const double app = a[(size_t)p * k + p], aqq = a[(size_t)q * k + q];
                const double th = (aqq - app) / (2.0 * apq);
                const double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
                const double cs = 1.0 / sqrt(t * t + 1.0), sn = t * cs;
                for (int r = 0; r < k; ++r) {   // columns p, q
                    const double arp = a[(size_t)r * k + p], arq = a[(size_t)r * k + q];
                    a[(size_t)r * k + p] = cs * arp - sn * arq;
                    a[(size_t)r * k + q] = sn * arp + cs * arq;
                }
                for (int r = 0; r < k; ++r) {   // rows p, q
                    const double apr = a[(size_t)p * k + r], aqr = a[(size_t)q * k + r];
                    a[(size_t)p * k + r] = cs * apr - sn * aqr;
                    a[(size_t)q * k + r] = sn * apr + cs * aqr;
                }
                if (vec)
                    for (int r = 0; r < k; ++r) {
                        const double vrp = v[(size_t)r * k + p], vrq = v[(size_t)r * k + q];
                        v[(size_t)r * k + p] = cs * vrp - sn * vrq;
                        v[(size_t)r * k + q] = sn * vrp + cs * vrq;
                    }
            }
    }
    ev.assign(k, 0.0);
    for (int i = 0; i < k; ++i) ev[i] = a[(size_t)i * k + i];
    if (vec) *vec = v;
}

// [[G, -1], [-1, 0]] for the k x k Gram matrix G.
std::vector<double> bordered(const std::vector<double>& g, int k)
{
    const int m = k + 1;
    std::vector<double> B((size_t)m * m, -1.0);
    for (int i = 0; i < k; ++i)
        for (int j = 0; j < k; ++j) B[(size_t)i * m + j] = g[(size_t)i * SCF_MAXV + j];
    B[(size_t)k * m + k] = 0.0;
    return B;
}

double cond(const std::vector<double>& B, int m)
{
    std::vector<double> ev;
    sym_eig(B, m, ev, nullptr);
    double lo = INFINITY, hi = 0.0;
    for (double e : ev) {
        lo = std::min(lo, fabs(e));
        hi = std::max(hi, fabs(e));
    }
    return lo == 0.0 ? INFINITY : hi / lo;
}

// Solve B c = rhs (m x m): LU with partial pivoting; an exactly singular
// B takes the minimum-norm least-squares solution (eigenvalues below
// m * eps * max |eigenvalue| dropped).
std::vector<double> solve(std::vector<double> B, int m, std::vector<double> rhs)
{
    std::vector<double> a = B, x = rhs;
    bool singular = false;
    for (int c = 0; c < m && !singular; ++c) {
        int p = c;
        for (int r = c + 1; r < m; ++r)
            if (fabs(a[(size_t)r * m + c]) > fabs(a[(size_t)p * m + c])) p = r;
        if (a[(size_t)p * m + c] == 0.0) {
            singular = true;
            break;
        }
        if (p != c) {
            for (int j = 0; j < m; ++j) std::swap(a[(size_t)p * m + j], a[(size_t)c * m + j]);
            std::swap(x[p], x[c]);
        }
        for (int r = c + 1; r < m; ++r) {
            const double f = a[(size_t)r * m + c] / a[(size_t)c * m + c];
            if (f == 0.0) continue;
            for (int j = c; j < m; ++j) a[(size_t)r * m + j] -= f * a[(size_t)c * m + j];
            x[r] -= f * x[c];
        }
    }
    if (!singular) {
        for (int r = m - 1; r >= 0; --r) {
            double s = x[r];
            for (int j = r + 1; j < m; ++j) s -= a[(size_t)r * m + j] * x[j];
            x[r] = s / a[(size_t)r * m + r];
        }
        return x;
    }
    std::vector<double> ev, V;
    sym_eig(B, m, ev, &V);
    double mx = 0.0;
    for (double e : ev) mx = std::max(mx, fabs(e));
    const double cut = m * 2.220446049250313e-16 * mx;
    std::vector<double> out(m, 0.0);
    for (int k = 0; k < m; ++k) {
        if (fabs(ev[k]) <= cut) continue;
        double d = 0.0;
        for (int i = 0; i < m; ++i) d += V[(size_t)i * m + k] * rhs[i];
        d /= ev[k];
        for (int i = 0; i < m; ++i) out[i] += d * V[(size_t)i * m + k];
    }
    return out;
}

// Pulay DIIS state of one spin: the stored (residual, F, F') slots oldest
// first and their Gram matrix.
struct Diis {
    int k = 0;
    int slot[SCF_MAXV];
    double gram[SCF_MAXV * SCF_MAXV];
    std::vector<double> coef;   // coefficients of the last extrapolation (empty: none)
    int free_slot(int nslot) const
    {
        for (int v = 0; v < nslot; ++v) {
            bool used = false;
            for (int i = 0; i < k; ++i) used = used || slot[i] == v;
            if (!used) return v;
        }
        return -1;
    }
};

}   // namespace

extern "C" long long pq_scf_work(int n, int nx, int ncart, int unrestricted, int diis_size)
{
    if (n <= 0 || nx <= 0 || nx > n || ncart < n || diis_size < 1 || diis_size >= SCF_MAXV) return -1;
    Bufs b;
    return lay_out(b, nullptr, n, nx, ncart, unrestricted ? 2 : 1, diis_size + 1);
}

extern "C" int pq_scf_run(pq_basis* basis, const double* H, const double* S, const double* X, int n, int nx,
                          const double* T, int ncart, int nalpha, int nbeta, int unrestricted, double conv,
                          int maxiter, int use_diis, int diis_start, double diis_threshold, double diis_max_cond,
                          int diis_size, double* Da, double* Db, double* Fa, double* Fb, double* work,
                          double* energies, double* jk_ms, double* result, void* stream)
{
    if (!basis || !H || !S || !X || !Da || !Fa || !work || !result) return bad("pq_scf_run: null argument");
    if (n <= 0 || nx <= 0 || nx > n) return bad("pq_scf_run: need 0 < nx <= n");
    if ((ncart != n) != (T != nullptr) || ncart < n) return bad("pq_scf_run: T (n x ncart) exactly when ncart != n");
    if (diis_size < 1 || diis_size >= SCF_MAXV) return bad("pq_scf_run: diis_size must be 1..15");
    if ((unrestricted ? 4 + 2 * (diis_size + 1) : 2 + diis_size + 1) > SCF_MAXDOT)
        return bad("pq_scf_run: diis_size too large for one scalar reduction (restricted <= 15, unrestricted <= 9)");
    if (maxiter < 1 || conv <= 0.0) return bad("pq_scf_run: need maxiter >= 1 and conv > 0");
    if (nalpha < 0 || nbeta < 0 || nalpha > nx || nbeta > nx) return bad("pq_scf_run: bad occupations");
    if (!unrestricted && nalpha != nbeta) return bad("pq_scf_run: restricted needs nalpha == nbeta");
    if (unrestricted && (!Db || !Fb || Db == Da || Fb == Fa))
        return bad("pq_scf_run: unrestricted needs separate Db and Fb");
    if ((uintptr_t)work % 256) return bad("pq_scf_run: work must be 256-byte aligned");
    hipStream_t st = (hipStream_t)stream;
    const int nsp = unrestricted ? 2 : 1, nslot = diis_size + 1;
    const long long n2 = (long long)n * n, x2 = (long long)nx * nx;
    Bufs b;
    lay_out(b, work, n, nx, ncart, nsp, nslot);
    double* D[2] = {Da, unrestricted ? Db : Da};
    double* Fout[2] = {Fa, unrestricted ? Fb : Fa};
    const int nocc[2] = {nalpha, nbeta};
    const bool sph = ncart != n;
    Diis diis[2];
    bool have_p[2] = {false, false};   // D_s = X P_s X^T holds
    int cur[2] = {0, 0};               // this cycle's slot per spin
    double E = 0.0, dE = 1.0;
    int it = 0, steps_total = 0;
    const double gtol = sqrt(conv);

    // Cartesian density -> J/K input; J, K -> final basis
    auto to_cart = [&](const double* Dn, double* Dc) -> int {   // Dc = T^T D T
        SCFCALL(gemm(0, 0, n, ncart, n, Dn, n, T, ncart, b.Tmp, ncart, st));       // D T (n x nc)
        return gemm(1, 0, ncart, ncart, n, T, ncart, b.Tmp, ncart, Dc, ncart, st);  // T^T (D T)
    };
    auto to_final = [&](const double* Mc, double* Mn) -> int {  // Mn = T M T^T
        SCFCALL(gemm(0, 1, ncart, n, ncart, Mc, ncart, T, ncart, b.Tmp, n, st));   // M T^T (nc x n)
        return gemm(0, 0, n, n, ncart, T, ncart, b.Tmp, n, Mn, n, st);             // T (M T^T)
    };

    while (conv < fabs(dE)) {
        ++it;
        if (it > maxiter) break;
        // ---- Fock matrices into this cycle's slots
        for (int s = 0; s < nsp; ++s) {
            cur[s] = diis[s].free_slot(nslot);
            if (cur[s] < 0) return bad("pq_scf_run: no free DIIS slot");
        }
        if (!unrestricted) {
            const double* Dj = D[0];
            if (sph) {
                SCFCALL(to_cart(D[0], b.Dca));
                Dj = b.Dca;
            }
            SCFCALL(pq_jk(basis, Dj, Dj, b.J, b.Ka, st));   // J(2 D) = 2 J(D) exactly
            double* F = b.F[0][cur[0]];
            if (sph) {   // J, K to the final basis first (A, B as scratch)
                SCFCALL(to_final(b.J, b.A));
                SCFCALL(to_final(b.Ka, b.B));
                hipLaunchKernelGGL(scf_fock_kernel, dim3(blocks_for(n2)), dim3(256), 0, st, H, b.A, 2.0, b.B, F, n2);
            } else {
                hipLaunchKernelGGL(scf_fock_kernel, dim3(blocks_for(n2)), dim3(256), 0, st, H, b.J, 2.0, b.Ka, F, n2);
            }
        } else {
            hipLaunchKernelGGL(scf_add_kernel, dim3(blocks_for(n2)), dim3(256), 0, st, D[0], D[1], b.Dt, n2);
            const double *Djt = b.Dt, *Dka = D[0], *Dkb = D[1];
            if (sph) {
                SCFCALL(to_cart(b.Dt, b.Dct));
                SCFCALL(to_cart(D[0], b.Dca));
                SCFCALL(to_cart(D[1], b.Dcb));
                Djt = b.Dct;
                Dka = b.Dca;
                Dkb = b.Dcb;
            }
            SCFCALL(pq_jk(basis, Djt, Dka, b.J, b.Ka, st));
            if (jk_ms) {
                double ms = -1.0;
                if (pq_get_timings(basis, nullptr, &ms, nullptr) == 0) jk_ms[2 * (it - 1)] = ms;
            }
            SCFCALL(pq_jk(basis, Djt, Dkb, b.J2, b.Kb, st));
            if (jk_ms) {
                double ms = -1.0;
                if (pq_get_timings(basis, nullptr, &ms, nullptr) == 0) jk_ms[2 * (it - 1) + 1] = ms;
            }
            const double *Jf = b.J, *Kfa = b.Ka, *Kfb = b.Kb;
            if (sph) {
                SCFCALL(to_final(b.J, b.A));
                SCFCALL(to_final(b.Ka, b.SD));
                SCFCALL(to_final(b.Kb, b.B));
                Jf = b.A;
                Kfa = b.SD;
                Kfb = b.B;
            }
            hipLaunchKernelGGL(scf_fock_kernel, dim3(blocks_for(n2)), dim3(256), 0, st, H, Jf, 1.0, Kfa,
                               b.F[0][cur[0]], n2);
            hipLaunchKernelGGL(scf_fock_kernel, dim3(blocks_for(n2)), dim3(256), 0, st, H, Jf, 1.0, Kfb,
                               b.F[1][cur[1]], n2);
        }
        SCFCHK(hipGetLastError());
        // ---- F' = X^T F X and the orbital gradient per spin
        for (int s = 0; s < nsp; ++s) {
            double* F = b.F[s][cur[s]];
            double* Fo = b.Fo[s][cur[s]];
            SCFCALL(gemm(0, 0, n, nx, n, F, n, X, nx, b.T1, nx, st));    // F X
            SCFCALL(gemm(1, 0, nx, nx, n, X, nx, b.T1, nx, Fo, nx, st));  // X^T (F X)
            if (have_p[s]) {
                SCFCALL(gemm(0, 0, nx, nx, nx, b.P[s], nx, Fo, nx, b.A, nx, st));   // P F'
            } else {
                SCFCALL(gemm(0, 0, n, n, n, S, n, D[s], n, b.SD, n, st));          // S D
                SCFCALL(gemm(0, 0, n, n, n, b.SD, n, F, n, b.B, n, st));           // (S D) F
                SCFCALL(gemm(0, 0, n, nx, n, b.B, n, X, nx, b.T1, nx, st));        // (S D F) X
                SCFCALL(gemm(1, 0, nx, nx, n, X, nx, b.T1, nx, b.A, nx, st));      // X^T (S D F X)
            }
            hipLaunchKernelGGL(scf_antisym_kernel, dim3(blocks_for(n2)), dim3(256), 0, st, b.A, nx,
                               b.R[s][cur[s]], n);
        }
        SCFCHK(hipGetLastError());
        // ---- scalars: energy terms, DIIS overlaps, residual maxima
        const bool diis_now = use_diis && it > diis_start;
        std::vector<const double*> xs, ys, rs;
        if (!unrestricted) {
            xs = {D[0], D[0]};
            ys = {H, b.F[0][cur[0]]};
        } else {
            xs = {D[0], D[1], D[0], D[1]};
            ys = {H, H, b.F[0][cur[0]], b.F[1][cur[1]]};
        }
        const int ne = (int)xs.size();
        int nd[2] = {0, 0};
        for (int s = 0; s < nsp; ++s) {
            if (!diis_now) continue;
            const double* r = b.R[s][cur[s]];
            for (int i = 0; i < diis[s].k; ++i) {
                xs.push_back(r);
                ys.push_back(b.R[s][diis[s].slot[i]]);
            }
            xs.push_back(r);
            ys.push_back(r);
            nd[s] = diis[s].k + 1;
        }
        for (int s = 0; s < nsp; ++s) rs.push_back(b.R[s][cur[s]]);
        double vals[SCF_MAXDOT + 4];
        SCFCALL(pq_scf_dots(xs.data(), ys.data(), (int)xs.size(), rs.data(), nsp, n2, b.dots, vals, st));
        if (jk_ms && !unrestricted) {   // this cycle's J/K has finished: its events are complete
            double ms = -1.0;
            if (pq_get_timings(basis, nullptr, &ms, nullptr) == 0) jk_ms[2 * (it - 1)] = ms;
        }
        const double Ev = unrestricted ? 0.5 * (vals[0] + vals[1] + vals[2] + vals[3]) : vals[0] + vals[1];
        const double* mx = vals + xs.size();   // per spin: max, max |.|
        const double gmax = unrestricted ? std::max(mx[1], mx[3]) : mx[1];
        dE = it == 1 ? 1.0 : Ev - E;
        E = Ev;
        if (fabs(dE) <= conv && gmax > gtol) dE = dE >= 0 ? 2.0 * conv : -2.0 * conv;
        if (energies) energies[it - 1] = E;
        if (fabs(dE) <= conv) {
            for (int s = 0; s < nsp; ++s)
                SCFCHK(hipMemcpyAsync(Fout[s], b.F[s][cur[s]], sizeof(double) * n2, hipMemcpyDeviceToDevice, st));
            break;
        }
        // ---- DIIS, then the new density per spin
        int off = ne;
        for (int s = 0; s < nsp; ++s) {
            const double* Fuse = b.Fo[s][cur[s]];   // purification input
            Diis& d = diis[s];
            d.coef.clear();
            const double err = mx[2 * s], amax = mx[2 * s + 1];
            const double* dots = vals + off;
            off += nd[s];
            if (diis_now && !(fabs(err) >= fabs(diis_threshold)) && !(amax < 1e-14)) {
                // grow the Gram matrix by the new row, store the pair
                const int k = d.k;
                for (int i = 0; i < k; ++i) d.gram[i * SCF_MAXV + k] = d.gram[k * SCF_MAXV + i] = dots[i];
                d.gram[k * SCF_MAXV + k] = dots[k];
                d.slot[k] = cur[s];
                d.k = k + 1;
                if (d.k >= 2) {
                    std::vector<double> g(d.gram, d.gram + SCF_MAXV * SCF_MAXV);
                    std::vector<double> Bm = bordered(g, d.k);
                    while (d.k > diis_size || cond(Bm, d.k + 1) > diis_max_cond) {
                        if (d.k == 1) break;
                        // drop the oldest
                        for (int i = 0; i + 1 < d.k; ++i) {
                            d.slot[i] = d.slot[i + 1];
                            for (int j = 0; j + 1 < d.k; ++j)
                                d.gram[i * SCF_MAXV + j] = d.gram[(i + 1) * SCF_MAXV + (j + 1)];
                        }
                        --d.k;
                        g.assign(d.gram, d.gram + SCF_MAXV * SCF_MAXV);
                        Bm = bordered(g, d.k);
                    }
                    std::vector<double> rhs(d.k + 1, 0.0);
                    rhs[d.k] = -1.0;
                    std::vector<double> c = solve(Bm, d.k + 1, rhs);
                    c.resize(d.k);
                    d.coef = c;
                    const double* mats[SCF_MAXCOMB];
                    for (int i = 0; i < d.k; ++i) mats[i] = b.Fo[s][d.slot[i]];
                    SCFCALL(pq_lincomb(mats, c.data(), d.k, x2, b.Fx[s], st));
                    Fuse = b.Fx[s];
                }
            }
            // P = purify(F'), D = X P X^T
            if (nocc[s] == 0 || nocc[s] >= nx) {
                hipLaunchKernelGGL(scf_diag_kernel, dim3(blocks_for(x2)), dim3(256), 0, st, b.P[s], nx,
                                   nocc[s] == 0 ? 0.0 : 1.0);
                SCFCHK(hipGetLastError());
            } else {
                int steps = 0;
                double idem = 0.0;
                SCFCALL(pq_purify(Fuse, nx, nocc[s], b.P[s], b.pur, 1e-11, SCF_MAX_STEPS, &steps, &idem, st));
                steps_total += steps;
            }
            SCFCALL(gemm(0, 0, n, nx, nx, X, nx, b.P[s], nx, b.T1, nx, st));   // X P
            SCFCALL(gemm(0, 1, n, n, nx, b.T1, nx, X, nx, D[s], n, st));       // (X P) X^T
            have_p[s] = true;
        }
        if (it == maxiter) {   // the loop ends here: hand back the (extrapolated) Fock matrices
            for (int s = 0; s < nsp; ++s) {
                const Diis& d = diis[s];
                if (d.coef.empty()) {
                    SCFCHK(hipMemcpyAsync(Fout[s], b.F[s][cur[s]], sizeof(double) * n2, hipMemcpyDeviceToDevice,
                                          st));
                } else {
                    const double* mats[SCF_MAXCOMB];
                    for (int i = 0; i < d.k; ++i) mats[i] = b.F[s][d.slot[i]];
                    SCFCALL(pq_lincomb(mats, d.coef.data(), d.k, n2, Fout[s], st));
                }
            }
        }
    }
    if (!unrestricted && Db && Db != Da) SCFCHK(hipMemcpyAsync(Db, Da, sizeof(double) * n2, hipMemcpyDeviceToDevice, st));
    if (!unrestricted && Fb && Fb != Fa) SCFCHK(hipMemcpyAsync(Fb, Fa, sizeof(double) * n2, hipMemcpyDeviceToDevice, st));
    SCFCHK(hipStreamSynchronize(st));
    result[0] = E;
    result[1] = dE;
    result[2] = (double)it;
    result[3] = (double)steps_total;
    return 0;
}
