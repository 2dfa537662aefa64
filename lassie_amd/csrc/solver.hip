// Persistent single-model solver: the whole LASSIE integration loop
// (Simulation.cuh:632-925) in one cooperative launch.
//
//   workgroup 0   sample check, RKF45 attempt, step control, Newton loop
//   all groups    dense Jacobian assembly of M = I - hJ and the blocked LU
//
// Every floating-point operation follows the order written in the reference
// kernels (kernels.cuh) with no contraction (built with -ffp-contract=off),
// so the trajectory is bit-identical to the CPU restatement used as oracle.
#include "device_common.h"
#include "solver.h"

namespace lassie {

struct Shared {
    double red[NT];
    int ired[NT];
    double cand_row[2][NWAVES][16];
    double cand_v[2][NWAVES];
    int cand_p[2][NWAVES];
    double U12[16][16];
    double L11[16][16];
    double oldP[16][16];
    int iscal[4];
    double b[kMaxDenseN];
};

// ---------------------------------------------------------------------------
// workgroup reductions (all threads receive the result)

// sum over lanes in the fixed "tid strided, then halving tree" order used for
// the Newton norm (Simulation.cuh:883-888 with a float-returning square).
__device__ double wg_tree_sum(double v, Shared &sh)
{
    const int tid = threadIdx.x;
    sh.red[tid] = v;
    __syncthreads();
    for (int stride = NT / 2; stride > 0; stride >>= 1) {
        if (tid < stride) sh.red[tid] = sh.red[tid] + sh.red[tid + stride];
        __syncthreads();
    }
    const double r = sh.red[0];
    __syncthreads();
    return r;
}

__device__ double wg_fmin(double v, Shared &sh)
{
    const int tid = threadIdx.x;
    sh.red[tid] = v;
    __syncthreads();
    for (int stride = NT / 2; stride > 0; stride >>= 1) {
        if (tid < stride) sh.red[tid] = fmin(sh.red[tid], sh.red[tid + stride]);
        __syncthreads();
    }
    const double r = sh.red[0];
    __syncthreads();
    return r;
}

__device__ int wg_isum(int v, Shared &sh)
{
    const int tid = threadIdx.x;
    sh.ired[tid] = v;
    __syncthreads();
    for (int stride = NT / 2; stride > 0; stride >>= 1) {
        if (tid < stride) sh.ired[tid] = sh.ired[tid] + sh.ired[tid + stride];
        __syncthreads();
    }
    const int r = sh.ired[0];
    __syncthreads();
    return r;
}

// ---------------------------------------------------------------------------
// RKF45 attempt (Simulation.cuh:692-740; kernels.cuh:68-199), workgroup-wide.
// Returns the number of species passing the test; *mind = min of the step
// factors with thrust::min_element semantics (a NaN first element wins,
// other NaNs are ignored).
__device__ int rkf_attempt_wg(const SolveArgs &A, double dt, double *mind, Shared &sh)
{
    const DevModel &m = A.m;
    const int S = m.S;
    const int tid = threadIdx.x;
    double *__restrict__ y1 = A.y1;
    double *__restrict__ ua = A.u;
    double *__restrict__ ub = A.u2;
    double *k1 = A.k1, *k2 = A.k2, *k3 = A.k3, *k4 = A.k4, *k5 = A.k5;
    const double *fd = m.feed;

    for (int s = tid; s < S; s += NT) {
        const double f = rhs_species(m, y1, s);
        if (fd[s] != 0.0) {
            k1[s] = fd[s];
            ua[s] = fd[s];
        } else {
            const double kk = f * dt;
            k1[s] = kk;
            ua[s] = y1[s] + kk / 4.0;
        }
    }
    __syncthreads();
    for (int s = tid; s < S; s += NT) {
        const double f = rhs_species(m, ua, s);
        if (fd[s] != 0.0) {
            k2[s] = fd[s];
            ub[s] = fd[s];
        } else {
            const double kk = f * dt;
            k2[s] = kk;
            ub[s] = y1[s] + (3.0 * k1[s]) / 32.0 + (9.0 * kk) / 32.0;
        }
    }
    __syncthreads();
    for (int s = tid; s < S; s += NT) {
        const double f = rhs_species(m, ub, s);
        if (fd[s] != 0.0) {
            k3[s] = fd[s];
            ua[s] = fd[s];
        } else {
            const double kk = f * dt;
            k3[s] = kk;
            ua[s] = y1[s] + (1932.0 * k1[s]) / 2197.0 - (7200.0 * k2[s]) / 2197.0 + (7296.0 * kk) / 2197.0;
        }
    }
    __syncthreads();
    for (int s = tid; s < S; s += NT) {
        const double f = rhs_species(m, ua, s);
        if (fd[s] != 0.0) {
            k4[s] = fd[s];
            ub[s] = fd[s];
        } else {
            const double kk = f * dt;
            k4[s] = kk;
            ub[s] = y1[s] + (439.0 * k1[s]) / 216.0 - (8.0 * k2[s]) + (3680.0 * k3[s]) / 513.0
                    - (845.0 * kk) / 4104.0;
        }
    }
    __syncthreads();
    for (int s = tid; s < S; s += NT) {
        const double f = rhs_species(m, ub, s);
        if (fd[s] != 0.0) {
            k5[s] = fd[s];
            ua[s] = fd[s];
        } else {
            const double kk = f * dt;
            k5[s] = kk;
            ua[s] = y1[s] - (8.0 * k1[s]) / 27.0 + (2.0 * k2[s]) - (3544.0 * k3[s]) / 2565.0
                    + (1859.0 * k4[s]) / 4104.0 - (11.0 * kk) / 40.0;
        }
    }
    __syncthreads();
    int nok = 0;
    double dmin = INFINITY;
    for (int s = tid; s < S; s += NT) {
        const double f = rhs_species(m, ua, s);
        double w1, w2;
        if (fd[s] != 0.0) {
            w1 = fd[s];
            w2 = fd[s];
        } else {
            const double kk6 = f * dt;
            w1 = y1[s] + (25.0 * k1[s]) / 216.0 + (1408.0 * k3[s]) / 2565.0 + (2197.0 * k4[s]) / 4104.0
                 - k5[s] / 5.0;
            w2 = y1[s] + (16.0 * k1[s]) / 135.0 + (6656.0 * k3[s]) / 12825.0
                 + (28561.0 * k4[s]) / 56430.0 - (9.0 * k5[s] / 50.0) + (2.0 * kk6) / 55.0;
        }
        A.w1[s] = w1;
        const double R = fabs(w1 - w2) / dt;
        const double tol = (A.rtol > 0.0) ? m.atol[s] + A.rtol * fabs(w1) : m.atol[s];
        const double d = 0.84 * sqrt(sqrt(tol / R));
        if (s == 0) sh.iscal[0] = isnan(d) ? 1 : 0;
        dmin = fmin(dmin, d);
        if (R <= tol && !isnan(R) && !isinf(R) && w1 > 0) nok++;
    }
    nok = wg_isum(nok, sh);
    double r = wg_fmin(dmin, sh);
    if (sh.iscal[0]) r = NAN;
    *mind = r;
    return nok;
}

// ---------------------------------------------------------------------------
// Jacobian (createJac_, kernels.cuh:219-271) and M = I - hJ
// (aggiorna_Jacobiano1, kernels.cuh:393-402) for row i, into the column-major
// dense M that was zero-filled beforehand.  Rows are owned by one thread, so
// the accumulation order per entry is the reference's (reaction ascending).
__device__ void jac_row(const DevModel &m, const double *__restrict__ x, double *__restrict__ M, int n,
                        int i, double h)
{
    const int e0 = m.hptr[i], e1 = m.hptr[i + 1];
    for (int e = e0; e < e1; e++) {
        const int r = m.hrx[e];
        const double sk = m.hk[e];
        const int a0 = m.lptr[r], a1 = m.lptr[r + 1];
        for (int j = a0; j < a1; j++) {
            const int g = m.lsp[j];
            const int ex = m.lexp[j];
            double temp = 0.0;
            if (ex == 1)
                temp = temp + sk;
            else if (ex > 1)
                temp = temp + (sk * (double)ex) * mypow(x[g], ex - 1);
            else
                continue;
            for (int jj = a0; jj < a1; jj++)
                if (m.lsp[jj] != g) temp = temp * x[m.lsp[jj]];
            double *p = M + (size_t)g * n + i;
            *p = *p + temp;
        }
    }
    double *d = M + (size_t)i * n + i;
    *d = 1.0 - h * *d;
    for (int e = e0; e < e1; e++) {
        const int r = m.hrx[e];
        const int a0 = m.lptr[r], a1 = m.lptr[r + 1];
        for (int j = a0; j < a1; j++) {
            const int g = m.lsp[j];
            if (m.lexp[j] < 1 || g == i) continue;
            bool dup = false;
            for (int e2 = e0; e2 <= e && !dup; e2++) {
                const int r2 = m.hrx[e2];
                const int b0 = m.lptr[r2], b1 = m.lptr[r2 + 1];
                for (int j2 = b0; j2 < b1; j2++) {
                    if (e2 == e && j2 >= j) break;
                    if (m.lsp[j2] == g && m.lexp[j2] >= 1) {
                        dup = true;
                        break;
                    }
                }
            }
            if (dup) continue;
            double *p = M + (size_t)g * n + i;
            *p = 0.0 - h * *p;
        }
    }
}

// ---------------------------------------------------------------------------
// Panel factorization with partial pivoting of columns [r0, r0+nbp), rows
// [r0, n), held in registers: RPT rows per thread, NBC columns.  Pivot: the
// largest |a| among remaining rows, first (lowest position) on ties, a NaN at
// the diagonal position wins (LAPACK idamax / dgetf2 semantics).  Row swaps
// are tracked as positions and applied when the panel is written back.
// Publishes piv[k], perm_src[pos] (original row of the pivot now at pos) and
// moved_to[r0+j] (final position of the row that started at r0+j).
template <int RPT, int NBC>
__device__ void panel_factor(double *__restrict__ M, int n, int p, int *piv, int *perm_src, int *moved_to,
                             int *info, Shared &sh)
{
    const int r0 = p * NBC;
    const int nbp = min(NBC, n - r0);
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    double a[RPT][NBC];
    int pos[RPT];
    bool act[RPT];
#pragma unroll
    for (int q = 0; q < RPT; q++) {
        const int r = r0 + tid + NT * q;
        pos[q] = r;
        act[q] = r < n;
#pragma unroll
        for (int c = 0; c < NBC; c++) a[q][c] = (r < n && c < nbp) ? M[(size_t)(r0 + c) * n + r] : 0.0;
    }
    int buf = 0;
    bool stop = false;
#pragma unroll
    for (int j = 0; j < NBC; j++) {
        if (j < nbp && !stop) {
        const int k = r0 + j;
        double bv = -2.0;
        int bp = 0x7fffffff;
        int bq = -1;
#pragma unroll
        for (int q = 0; q < RPT; q++) {
            if (act[q]) {
                const double xv = a[q][j];
                double v = fabs(xv);
                if (isnan(xv)) v = (pos[q] == k) ? INFINITY : -1.0;
                if (v > bv || (v == bv && pos[q] < bp)) {
                    bv = v;
                    bp = pos[q];
                    bq = q;
                }
            }
        }
        const int mybp = bp;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const double ov = __shfl_xor(bv, off);
            const int op = __shfl_xor(bp, off);
            if (ov > bv || (ov == bv && op < bp)) {
                bv = ov;
                bp = op;
            }
        }
        if (lane == 0) {
            sh.cand_v[buf][wave] = bv;
            sh.cand_p[buf][wave] = bp;
        }
        if (bq >= 0 && mybp == bp) {
#pragma unroll
            for (int q = 0; q < RPT; q++)
                if (q == bq) {
#pragma unroll
                    for (int c = 0; c < NBC; c++) sh.cand_row[buf][wave][c] = a[q][c];
                }
        }
        __syncthreads();
        int w = 0;
        bv = sh.cand_v[buf][0];
        bp = sh.cand_p[buf][0];
#pragma unroll
        for (int ww = 1; ww < NWAVES; ww++) {
            const double ov = sh.cand_v[buf][ww];
            const int op = sh.cand_p[buf][ww];
            if (ov > bv || (ov == bv && op < bp)) {
                bv = ov;
                bp = op;
                w = ww;
            }
        }
        const double *prow = sh.cand_row[buf][w];
        const double ukk = prow[j];
        if (ukk == 0.0) {
            if (tid == 0) st_rel(info, k + 1);
            stop = true;
        } else {
        if (tid == 0) piv[k] = bp;
#pragma unroll
        for (int q = 0; q < RPT; q++) {
            const int old = pos[q];
            if (act[q] && old == bp) {
                act[q] = false;
                pos[q] = k;
            } else if (act[q] && old == k) {
                pos[q] = bp;
            }
        }
#pragma unroll
        for (int q = 0; q < RPT; q++) {
            if (act[q]) {
                const double l = a[q][j] / ukk;
                a[q][j] = l;
#pragma unroll
                for (int c = 0; c < NBC; c++)
                    if (c > j) a[q][c] = fma(-l, prow[c], a[q][c]);
            }
        }
        buf ^= 1;
        }
        }
    }
    if (stop) return;
#pragma unroll
    for (int q = 0; q < RPT; q++) {
        const int orig = r0 + tid + NT * q;
        if (orig < n) {
#pragma unroll
            for (int c = 0; c < NBC; c++)
                if (c < nbp) M[(size_t)(r0 + c) * n + pos[q]] = a[q][c];
            if (!act[q]) perm_src[pos[q]] = orig;
            if (orig < r0 + nbp) moved_to[orig] = pos[q];
        }
    }
}

// Apply panel p (swaps, triangular solve for the U block row, trailing update)
// to strip q (columns [q*NBC, q*NBC + ncols)), q > p.  Owned by one workgroup.
template <int NBC>
__device__ void strip_update(double *__restrict__ M, int n, int p, int q, const int *perm_src,
                             const int *moved_to, Shared &sh)
{
    const int tid = threadIdx.x;
    const int r0 = p * NBC;
    const int nbp = min(NBC, n - r0);
    const int c0 = q * NBC;
    const int ncols = min(NBC, n - c0);
    // gather the permuted panel rows and the old panel-range rows; stage L11
    for (int e = tid; e < NBC * NBC; e += NT) {
        const int kk = e / NBC, c = e % NBC;
        if (kk < nbp && c < ncols) {
            sh.U12[kk][c] = M[(size_t)(c0 + c) * n + perm_src[r0 + kk]];
            sh.oldP[kk][c] = M[(size_t)(c0 + c) * n + r0 + kk];
        }
        if (kk < nbp && c < nbp) sh.L11[kk][c] = M[(size_t)(r0 + c) * n + r0 + kk];
    }
    __syncthreads();
    // displaced panel-range rows go to the positions their pivots vacated
    for (int e = tid; e < NBC * NBC; e += NT) {
        const int jj = e / NBC, c = e % NBC;
        if (jj < nbp && c < ncols) {
            const int dst = moved_to[r0 + jj];
            if (dst >= r0 + nbp) M[(size_t)(c0 + c) * n + dst] = sh.oldP[jj][c];
        }
    }
    // U12 = L11^-1 * A12 (unit lower), one column per thread, k ascending
    if (tid < ncols) {
        const int c = tid;
        for (int i = 1; i < nbp; i++) {
            double v = sh.U12[i][c];
            for (int kk = 0; kk < i; kk++) v = fma(-sh.L11[i][kk], sh.U12[kk][c], v);
            sh.U12[i][c] = v;
        }
    }
    __syncthreads();
    for (int e = tid; e < NBC * NBC; e += NT) {
        const int kk = e / NBC, c = e % NBC;
        if (kk < nbp && c < ncols) M[(size_t)(c0 + c) * n + r0 + kk] = sh.U12[kk][c];
    }
    // trailing update A22 -= L21 * U12, fma chain with k ascending.  A wave
    // walks 64 consecutive rows (coalesced in the column-major layout) and
    // owns CPW columns of the strip, whose U12 entries stay in registers.
    constexpr int CPW = NBC / NWAVES;
    const int lane = tid & 63;
    const int cbase = (tid >> 6) * CPW;
    double u[NBC][CPW];
#pragma unroll
    for (int kk = 0; kk < NBC; kk++)
#pragma unroll
        for (int cc = 0; cc < CPW; cc++) u[kk][cc] = (kk < nbp) ? sh.U12[kk][cbase + cc] : 0.0;
    for (int i = r0 + nbp + lane; i < n; i += 64) {
        double acc[CPW];
#pragma unroll
        for (int cc = 0; cc < CPW; cc++)
            acc[cc] = (cbase + cc < ncols) ? M[(size_t)(c0 + cbase + cc) * n + i] : 0.0;
#pragma unroll
        for (int kk = 0; kk < NBC; kk++) {
            if (kk < nbp) {
                const double l = M[(size_t)(r0 + kk) * n + i];
#pragma unroll
                for (int cc = 0; cc < CPW; cc++) acc[cc] = fma(-l, u[kk][cc], acc[cc]);
            }
        }
#pragma unroll
        for (int cc = 0; cc < CPW; cc++)
            if (cbase + cc < ncols) M[(size_t)(c0 + cbase + cc) * n + i] = acc[cc];
    }
    __syncthreads();
}

// Blocked right-looking LU over all workgroups; strips are owned round-robin.
// The owner of strip p+1 updates it first and factors panel p+1 at once
// (look-ahead), so each panel costs one grid barrier.  Returns false on abort.
template <int RPT, int NBC>
__device__ bool lu_factor_grid(const SolveArgs &A, Shared &sh, unsigned int &gen_local, bool &singular)
{
    const int n = A.m.S;
    const int P = (n + NBC - 1) / NBC;
    const int G = A.G;
    const int wg = blockIdx.x;
    singular = false;
    // A zero pivot in panel p+1 may be recorded while slower groups are still
    // checking the flag for panel p, so each check only counts columns of
    // panels that were complete at the preceding barrier.
    auto failed_before = [&](int col_end) {
        const int inf = ld_acq(&A.ctl->info);
        return inf != 0 && inf - 1 < col_end;
    };
    if (wg == 0) panel_factor<RPT, NBC>(A.M, n, 0, A.piv, A.perm_src, A.moved_to, &A.ctl->info, sh);
    if (!grid_barrier(A.gs, G, gen_local, &A.ctl->abort)) return false;
    if (failed_before(NBC)) {
        singular = true;
        return true;
    }
    for (int p = 0; p < P; p++) {
        int q = p + 1 + ((wg - (p + 1)) % G + G) % G;
        for (; q < P; q += G) {
            strip_update<NBC>(A.M, n, p, q, A.perm_src, A.moved_to, sh);
            if (q == p + 1) {
                __syncthreads();
                panel_factor<RPT, NBC>(A.M, n, p + 1, A.piv, A.perm_src, A.moved_to, &A.ctl->info, sh);
                __syncthreads();
            }
        }
        if (p + 1 < P) {
            if (!grid_barrier(A.gs, G, gen_local, &A.ctl->abort)) return false;
            if (failed_before((p + 2) * NBC)) {
                singular = true;
                return true;
            }
        }
    }
    return true;
}

// getrs with one right-hand side (Simulation.cuh:1316), workgroup-wide, the
// vector staged in LDS: per panel the recorded row moves, then forward
// substitution (k ascending) and backward substitution (k descending).
template <int NBC>
__device__ void lu_solve_wg(const double *__restrict__ M, int n, const int *perm_src, const int *moved_to,
                            double *xio, Shared &sh)
{
    const int tid = threadIdx.x;
    const int P = (n + NBC - 1) / NBC;
    for (int i = tid; i < n; i += NT) sh.b[i] = xio[i];
    __syncthreads();
    for (int p = 0; p < P; p++) {
        const int r0 = p * NBC;
        const int nbp = min(NBC, n - r0);
        double v1 = 0.0, v2 = 0.0;
        int dst = -1;
        if (tid < nbp) {
            v1 = sh.b[perm_src[r0 + tid]];
            v2 = sh.b[r0 + tid];
            dst = moved_to[r0 + tid];
        }
        __syncthreads();
        if (tid < nbp) {
            sh.b[r0 + tid] = v1;
            if (dst >= r0 + nbp) sh.b[dst] = v2;
        }
        __syncthreads();
        if (tid < 64) {
            double bi = (tid < nbp) ? sh.b[r0 + tid] : 0.0;
            const double *col = M + (size_t)r0 * n + r0;
            for (int kk = 0; kk < nbp; kk++) {
                const double bk = __shfl(bi, kk);
                if (tid > kk && tid < nbp) bi = fma(-col[(size_t)kk * n + tid], bk, bi);
            }
            if (tid < nbp) sh.b[r0 + tid] = bi;
        }
        __syncthreads();
        for (int i = r0 + nbp + tid; i < n; i += NT) {
            double acc = sh.b[i];
            for (int kk = 0; kk < nbp; kk++) acc = fma(-M[(size_t)(r0 + kk) * n + i], sh.b[r0 + kk], acc);
            sh.b[i] = acc;
        }
        __syncthreads();
    }
    for (int p = P - 1; p >= 0; p--) {
        const int r0 = p * NBC;
        const int nbp = min(NBC, n - r0);
        if (tid < 64) {
            double bi = (tid < nbp) ? sh.b[r0 + tid] : 0.0;
            const double *col = M + (size_t)r0 * n + r0;
            for (int kk = nbp - 1; kk >= 0; kk--) {
                if (tid == kk) bi = bi / col[(size_t)kk * n + kk];
                const double xk = __shfl(bi, kk);
                if (tid < kk) bi = fma(-col[(size_t)kk * n + tid], xk, bi);
            }
            if (tid < nbp) sh.b[r0 + tid] = bi;
        }
        __syncthreads();
        for (int i = tid; i < r0; i += NT) {
            double acc = sh.b[i];
            for (int kk = nbp - 1; kk >= 0; kk--) acc = fma(-M[(size_t)(r0 + kk) * n + i], sh.b[r0 + kk], acc);
            sh.b[i] = acc;
        }
        __syncthreads();
    }
    for (int i = tid; i < n; i += NT) xio[i] = sh.b[i];
    __syncthreads();
}

__device__ __forceinline__ bool bad_value(double v) { return isnan(v) || isinf(v) || v < 0; }

// ---------------------------------------------------------------------------
template <int RPT, int NBC>
__global__ void __launch_bounds__(NT) solve_kernel(SolveArgs A)
{
    __shared__ Shared sh;
    const int tid = threadIdx.x;
    const int wg = blockIdx.x;
    const int G = A.G;
    const int S = A.m.S;
    const int n = S;
    unsigned int gen_local = 0;
    Ctl *ctl = A.ctl;

    // state of the control loop, held identically by every thread of group 0
    double t = 0.0, t1 = A.tv[0], dt = A.rkf_step, newdt = A.be_step;
    const double t_max = A.tv[A.nt - 1];
    int i_sample = 0;
    long long nrej = 0;
    long long st_loop = 0, st_att = 0, st_acc = 0, st_bdf = 0, st_newton = 0, st_lufail = 0, st_rep = 0,
              st_lu = 0;
    int rc = 0;
    // device-time attribution (group 0, 100 MHz wall clock): control/RKF,
    // Jacobian assembly, LU, Newton solves
    unsigned long long ph[4] = {0, 0, 0, 0};
    unsigned long long tmark = 0;

    if (wg == 0)
        for (int s = tid; s < S; s += NT) A.y1[s] = A.x0[s];
    __syncthreads();

    while (true) {
        // ---- control step (group 0) -----------------------------------------
        if (wg == 0) {
            tmark = wall_clock64();
            int mode = MODE_CONTINUE;
            double hb = 0.0;
            if (A.max_iters > 0 && st_loop >= A.max_iters) {
                mode = MODE_DONE;
                rc = 1;
            } else {
                if (tid == 0 && A.progress && st_loop < A.progress_cap) A.progress[st_loop] = t / t_max;
                st_loop++;
                bool finished = false;
                if (fabs(t1 - t) < 1e-5) {
                    for (int j = tid; j < A.ncs; j += NT)
                        A.out[(size_t)i_sample * A.ncs + j] = A.y1[A.cs[j]];
                    i_sample++;
                    if (i_sample == A.nt)
                        finished = true;
                    else
                        t1 = A.tv[i_sample];
                }
                if (finished) {
                    mode = MODE_DONE;
                } else {
                    st_att++;
                    double mind;
                    const int nok = rkf_attempt_wg(A, dt, &mind, sh);
                    if (nok == S) {
                        nrej = 0;
                        st_acc++;
                        for (int s = tid; s < S; s += NT) A.y1[s] = A.w1[s];
                        __syncthreads();
                        const double d_temp = dt * mind;
                        if ((t + d_temp) > t1)
                            dt = fabs(t1 - t);
                        else
                            dt = d_temp;
                        t = t + dt;
                    } else {
                        const double dt_old = dt;
                        dt = dt * mind;
                        nrej++;
                        if (dt < A.stiffness_tol || isnan(dt) || isinf(dt) || fabs(dt_old - dt) < 1e-3 ||
                            (A.max_rkf_retries > 0 && nrej >= A.max_rkf_retries)) {
                            nrej = 0;
                            st_bdf++;
                            dt = newdt;
                            if ((t + dt) > t1) dt = fabs(t1 - t);
                            hb = dt;
                            mode = (A.newton_iter > 0) ? MODE_BDF : MODE_BDF_NOOP;
                            // y_v = y1; Newton iteration 0 residual at y_v
                            for (int s = tid; s < S; s += NT) {
                                const double yv = A.y1[s];
                                A.yv[s] = yv;
                                const double f = rhs_species(A.m, A.y1, s);
                                A.x[s] = -(yv - A.y1[s] - hb * f);
                            }
                            __syncthreads();
                        }
                    }
                }
            }
            if (tid == 0) {
                st_rel_d(&ctl->h, hb);
                st_rel(&ctl->mode, mode);
                st_rel(&ctl->info, 0);
            }
        }
        if (wg == 0) {
            const unsigned long long now = wall_clock64();
            ph[0] += now - tmark;
            tmark = now;
        }
        if (G > 1) {
            if (!grid_barrier(A.gs, G, gen_local, &ctl->abort)) return;
        } else {
            __syncthreads();
        }
        const int mode = ld_acq(&ctl->mode);
        if (mode == MODE_DONE) break;
        if (mode == MODE_CONTINUE) continue;
        const double h = ld_acq_d(&ctl->h);

        bool singular = false;
        if (mode == MODE_BDF) {
            // ---- dense M = I - hJ(y_v), all groups -------------------------------
            {
                const size_t nn = (size_t)n * n;
                const size_t stride = (size_t)G * NT;
                for (size_t e = (size_t)wg * NT + tid; e < nn; e += stride) A.M[e] = 0.0;
            }
            if (G > 1) {
                if (!grid_barrier(A.gs, G, gen_local, &ctl->abort)) return;
            } else {
                __syncthreads();
            }
            for (int r = wg * NT + tid; r < n; r += G * NT) jac_row(A.m, A.yv, A.M, n, r, h);
            if (G > 1) {
                if (!grid_barrier(A.gs, G, gen_local, &ctl->abort)) return;
            } else {
                __syncthreads();
            }
            if (wg == 0) {
                const unsigned long long now = wall_clock64();
                ph[1] += now - tmark;
                tmark = now;
            }
            if (!lu_factor_grid<RPT, NBC>(A, sh, gen_local, singular)) return;
            if (wg == 0) {
                const unsigned long long now = wall_clock64();
                ph[2] += now - tmark;
                tmark = now;
            }
        }
        // ---- Newton loop, group 0 --------------------------------------------
        if (wg == 0) {
            st_lu += (mode == MODE_BDF) ? 1 : 0;
            bool luFail = false;
            if (mode == MODE_BDF) {
                int itN = 0;
                double errN = 1000;
                while (itN < A.newton_iter && errN > A.newton_tol) {
                    if (itN > 0) {
                        for (int s = tid; s < S; s += NT) {
                            const double f = rhs_species(A.m, A.yv, s);
                            A.x[s] = -(A.yv[s] - A.y1[s] - h * f);
                        }
                        __syncthreads();
                    }
                    if (singular) {
                        for (int s = tid; s < S; s += NT)
                            if (bad_value(A.yv[s])) A.yv[s] = A.y1[s];
                        __syncthreads();
                        newdt = h * 0.1;
                        luFail = true;
                        st_lufail++;
                        break;
                    }
                    lu_solve_wg<NBC>(A.M, n, A.perm_src, A.moved_to, A.x, sh);
                    double part = 0.0;
                    for (int s = tid; s < S; s += NT) {
                        const double yv1 = A.yv[s];
                        double yv;
                        if (A.m.feed[s] == 0)
                            yv = yv1 + A.x[s];
                        else
                            yv = A.m.feed[s];
                        A.yv[s] = yv;
                        const double d = yv - yv1;
                        part = part + (double)(float)(d * d);
                    }
                    errN = sqrt(wg_tree_sum(part, sh));
                    itN = itN + 1;
                }
                st_newton += itN;
            }
            if (!luFail) {
                int nbad = 0;
                for (int s = tid; s < S; s += NT) nbad += bad_value(A.yv[s]) ? 1 : 0;
                nbad = wg_isum(nbad, sh);
                if (nbad > 0) {
                    for (int s = tid; s < S; s += NT)
                        if (bad_value(A.yv[s])) A.yv[s] = A.y1[s];
                    newdt = h * 0.1;
                    st_rep++;
                } else if (A.bdf_regrow) {
                    newdt = fmin(newdt * 2.0, A.be_step);
                }
            }
            __syncthreads();
            for (int s = tid; s < S; s += NT) A.y1[s] = A.yv[s];
            __syncthreads();
            t = t + h;
            ph[3] += wall_clock64() - tmark;
        }
    }
    if (wg == 0 && tid == 0) {
        long long *o = A.stats_out;
        o[0] = st_loop;
        o[1] = st_att;
        o[2] = st_acc;
        o[3] = st_bdf;
        o[4] = st_newton;
        o[5] = st_lufail;
        o[6] = st_rep;
        o[7] = i_sample;
        o[8] = st_lu;
        o[9] = rc;
        double *d = A.dstats_out;
        d[0] = t;
        d[1] = dt;
        d[2] = newdt;
        for (int q = 0; q < 4; q++) d[3 + q] = (double)ph[q] * 1e-5;  // 100 MHz ticks -> ms
    }
}

// ---------------------------------------------------------------------------
template <int RPT, int NBC>
static hipError_t launch_t(const SolveArgs &args, hipStream_t stream)
{
    SolveArgs a = args;
    void *params[] = {&a};
    return hipLaunchCooperativeKernel(reinterpret_cast<const void *>(&solve_kernel<RPT, NBC>), dim3(a.G),
                                      dim3(NT), params, 0, stream);
}

int dense_panel_width(int n) { return n <= 2048 ? 16 : 8; }

int max_coop_groups(int n, int *out)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return -1;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return -1;
    int per = 0;
    hipError_t e;
    if (n <= 256)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void *)&solve_kernel<1, 16>, NT, 0);
    else if (n <= 512)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void *)&solve_kernel<2, 16>, NT, 0);
    else if (n <= 1024)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void *)&solve_kernel<4, 16>, NT, 0);
    else if (n <= 2048)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void *)&solve_kernel<8, 16>, NT, 0);
    else
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void *)&solve_kernel<16, 8>, NT, 0);
    if (e != hipSuccess) return -1;
    *out = per * cus;
    return 0;
}

hipError_t launch_solve(const SolveArgs &args, hipStream_t stream)
{
    const int n = args.m.S;
    if (n <= 256) return launch_t<1, 16>(args, stream);
    if (n <= 512) return launch_t<2, 16>(args, stream);
    if (n <= 1024) return launch_t<4, 16>(args, stream);
    if (n <= 2048) return launch_t<8, 16>(args, stream);
    return launch_t<16, 8>(args, stream);
}

}  // namespace lassie
