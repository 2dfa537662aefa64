// Workgroup-level building blocks shared by the single-model and batched
// solvers: LDS scratch, fixed-order reductions, the RKF45 attempt and the
// Jacobian row assembly.  Same operation order as the reference kernels.
#pragma once
#include "device_common.h"

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
    int rowcnt;
};

// Pointers used by the control group: LDS when the model fits, else global.
struct CtlView {
    const HRec *recs;
    const int *hptr;
    double *y1, *ua, *ub, *k1, *k2, *k3, *k4, *k5, *w1, *yv, *x;
};

// ---------------------------------------------------------------------------
// workgroup reductions (all threads receive the result)

// sum in the fixed "tid strided, then halving tree" order used for the Newton
// norm (Simulation.cuh:883-888 with a float-returning square).
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
__device__ int rkf_attempt_wg(const DevModel &m, double rtol, const CtlView &v, double dt, double *mind,
                              Shared &sh)
{
    const int S = m.S;
    const int tid = threadIdx.x;
    const double *y1 = v.y1;
    double *ua = v.ua, *ub = v.ub;
    double *k1 = v.k1, *k2 = v.k2, *k3 = v.k3, *k4 = v.k4, *k5 = v.k5;
    const double *fd = m.feed;

    for (int s = tid; s < S; s += NT) {
        const double f = rhs_species(m, v.recs, v.hptr, y1, s);
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
        const double f = rhs_species(m, v.recs, v.hptr, ua, s);
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
        const double f = rhs_species(m, v.recs, v.hptr, ub, s);
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
        const double f = rhs_species(m, v.recs, v.hptr, ua, s);
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
        const double f = rhs_species(m, v.recs, v.hptr, ub, s);
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
        const double f = rhs_species(m, v.recs, v.hptr, ua, s);
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
        v.w1[s] = w1;
        const double R = fabs(w1 - w2) / dt;
        const double tol = (rtol > 0.0) ? m.atol[s] + rtol * fabs(w1) : m.atol[s];
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
// (aggiorna_Jacobiano1, kernels.cuh:393-402) for row i, into the dense M
// (row- or column-major) that was zero-filled beforehand.  Rows are owned by one thread, so
// the accumulation order per entry is the reference's (reaction ascending).
template <bool ROW_MAJOR>
__device__ void jac_row(const DevModel &m, const double *__restrict__ x, double *__restrict__ M, int n,
                        int i, double h)
{
    auto at = [&](int row, int col) -> double * {
        return ROW_MAJOR ? M + (size_t)row * n + col : M + (size_t)col * n + row;
    };
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
            double *p = at(i, g);
            *p = *p + temp;
        }
    }
    double *d = at(i, i);
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
            double *p = at(i, g);
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
// Publishes piv[k], perm_src[pos] (original row of the pivot now at pos),
// moved_to[r0+j] (final position of the row that started at r0+j), and the
// positions below the panel whose multipliers are not all zero (rows, rowcnt):
// the only rows the trailing update changes.  T is the element type (double
// for the FP64 method, float for the precision form); RM selects row-major
// (else column-major) storage.
template <int RPT, int NBC, typename T, bool RM>
__device__ void panel_factor(T *__restrict__ M, int n, int p, int *piv, int *perm_src, int *moved_to,
                             int *rows, int *rowcnt, int *info, Shared &sh)
{
    auto at = [&](int row, int col) -> T & {
        return RM ? M[(size_t)row * n + col] : M[(size_t)col * n + row];
    };
    const int r0 = p * NBC;
    const int nbp = min(NBC, n - r0);
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    T a[RPT][NBC];
    int pos[RPT];
    bool act[RPT];
    if (tid == 0) sh.rowcnt = 0;
#pragma unroll
    for (int q = 0; q < RPT; q++) {
        const int r = r0 + tid + NT * q;
        pos[q] = r;
        act[q] = r < n;
#pragma unroll
        for (int c = 0; c < NBC; c++) a[q][c] = (r < n && c < nbp) ? at(r, r0 + c) : T(0);
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
                    const double xv = (double)a[q][j];
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
                        for (int c = 0; c < NBC; c++) sh.cand_row[buf][wave][c] = (double)a[q][c];
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
            const T ukk = (T)prow[j];
            if (ukk == T(0)) {
                if (tid == 0) st_rlx(info, k + 1);
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
                        const T l = a[q][j] / ukk;
                        a[q][j] = l;
#pragma unroll
                        for (int c = 0; c < NBC; c++)
                            if (c > j) a[q][c] = fma(-l, (T)prow[c], a[q][c]);
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
                if (c < nbp) at(pos[q], r0 + c) = a[q][c];
            if (!act[q]) perm_src[pos[q]] = orig;
            if (orig < r0 + nbp) moved_to[orig] = pos[q];
            if (act[q]) {
                bool nz = false;
#pragma unroll
                for (int c = 0; c < NBC; c++) nz = nz || (c < nbp && a[q][c] != T(0));
                if (nz) rows[atomicAdd(&sh.rowcnt, 1)] = pos[q];
            }
        }
    }
    __syncthreads();
    if (tid == 0) *rowcnt = sh.rowcnt;
}

// getrs with one right-hand side (Simulation.cuh:1316), one group, b in LDS.
// M holds L\U in final row order, so the panels' row moves are applied to b
// first (equal to LAPACK's sequential interchanges), then forward (k
// ascending) and backward (k descending) substitution as fma chains, in
// 64-row blocks: the diagonal block by one wave with register shuffles, the
// rows outside it by all waves with four rows and 16 matrix loads per row in
// flight.
// T: element type of the factors and of b; RM: row-major (else column-major).
template <int NBC, typename T, bool RM>
__device__ void lu_solve_wg(const T *__restrict__ M, int n, const int *perm_src, const int *moved_to, T *b)
{
    auto at = [&](int row, int col) -> T { return RM ? M[(size_t)row * n + col] : M[(size_t)col * n + row]; };
    constexpr int TB = 64;
    constexpr int RB = 4;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int P = (n + NBC - 1) / NBC;
    for (int p = 0; p < P; p++) {
        const int r0 = p * NBC;
        const int nbp = min(NBC, n - r0);
        T v1 = 0, v2 = 0;
        int dst = -1;
        if (tid < nbp) {
            v1 = b[perm_src[r0 + tid]];
            v2 = b[r0 + tid];
            dst = moved_to[r0 + tid];
        }
        __syncthreads();
        if (tid < nbp) {
            b[r0 + tid] = v1;
            if (dst >= r0 + nbp) b[dst] = v2;
        }
        __syncthreads();
    }
    const int NBk = (n + TB - 1) / TB;
    for (int B = 0; B < NBk; B++) {
        const int r0 = B * TB;
        const int nb = min(TB, n - r0);
        if (tid < 64) {
            T bi = (lane < nb) ? b[r0 + lane] : T(0);
#pragma unroll
            for (int sb = 0; sb < TB / 16; sb++) {
                T lsub[16];
#pragma unroll
                for (int kk = 0; kk < 16; kk++) {
                    const int k = sb * 16 + kk;
                    lsub[kk] = (k < lane && lane < nb) ? at(r0 + lane, r0 + k) : T(0);
                }
#pragma unroll
                for (int kk = 0; kk < 16; kk++) {
                    const int k = sb * 16 + kk;
                    if (k < nb) {
                        const T bk = __shfl(bi, k);
                        if (lane > k && lane < nb) bi = fma(-lsub[kk], bk, bi);
                    }
                }
            }
            if (lane < nb) b[r0 + lane] = bi;
        }
        __syncthreads();
        for (int i0 = r0 + nb + tid; i0 < n; i0 += NT * RB) {
            T acc[RB];
#pragma unroll
            for (int r = 0; r < RB; r++) acc[r] = (i0 + r * NT < n) ? b[i0 + r * NT] : T(0);
#pragma unroll
            for (int sb = 0; sb < TB / 16; sb++) {
                T l[RB][16];
#pragma unroll
                for (int r = 0; r < RB; r++)
#pragma unroll
                    for (int kk = 0; kk < 16; kk++) {
                        const int k = sb * 16 + kk;
                        l[r][kk] = (k < nb && i0 + r * NT < n) ? at(i0 + r * NT, r0 + k) : T(0);
                    }
#pragma unroll
                for (int kk = 0; kk < 16; kk++) {
                    const int k = sb * 16 + kk;
                    if (k < nb) {
                        const T bk = b[r0 + k];
#pragma unroll
                        for (int r = 0; r < RB; r++) acc[r] = fma(-l[r][kk], bk, acc[r]);
                    }
                }
            }
#pragma unroll
            for (int r = 0; r < RB; r++)
                if (i0 + r * NT < n) b[i0 + r * NT] = acc[r];
        }
        __syncthreads();
    }
    for (int B = NBk - 1; B >= 0; B--) {
        const int r0 = B * TB;
        const int nb = min(TB, n - r0);
        if (tid < 64) {
            T bi = (lane < nb) ? b[r0 + lane] : T(0);
#pragma unroll
            for (int sb = TB / 16 - 1; sb >= 0; sb--) {
                T usub[16];
#pragma unroll
                for (int kk = 0; kk < 16; kk++) {
                    const int k = sb * 16 + kk;
                    usub[kk] = (k >= lane && k < nb && lane < nb) ? at(r0 + lane, r0 + k) : T(1);
                }
#pragma unroll
                for (int kk = 15; kk >= 0; kk--) {
                    const int k = sb * 16 + kk;
                    if (k < nb) {
                        if (lane == k) bi = bi / usub[kk];
                        const T xk = __shfl(bi, k);
                        if (lane < k) bi = fma(-usub[kk], xk, bi);
                    }
                }
            }
            if (lane < nb) b[r0 + lane] = bi;
        }
        __syncthreads();
        for (int i0 = tid; i0 < r0; i0 += NT * RB) {
            T acc[RB];
#pragma unroll
            for (int r = 0; r < RB; r++) acc[r] = (i0 + r * NT < r0) ? b[i0 + r * NT] : T(0);
#pragma unroll
            for (int sb = TB / 16 - 1; sb >= 0; sb--) {
                T u[RB][16];
#pragma unroll
                for (int r = 0; r < RB; r++)
#pragma unroll
                    for (int kk = 0; kk < 16; kk++) {
                        const int k = sb * 16 + kk;
                        u[r][kk] = (k < nb && i0 + r * NT < r0) ? at(i0 + r * NT, r0 + k) : T(0);
                    }
#pragma unroll
                for (int kk = 15; kk >= 0; kk--) {
                    const int k = sb * 16 + kk;
                    if (k < nb) {
                        const T xk = b[r0 + k];
#pragma unroll
                        for (int r = 0; r < RB; r++) acc[r] = fma(-u[r][kk], xk, acc[r]);
                    }
                }
            }
#pragma unroll
            for (int r = 0; r < RB; r++)
                if (i0 + r * NT < r0) b[i0 + r * NT] = acc[r];
        }
        __syncthreads();
    }
}


__device__ __forceinline__ bool bad_value(double v) { return isnan(v) || isinf(v) || v < 0; }

}  // namespace lassie
