// Reverse pass of the differentiation form: the discrete adjoint of LASSIE's
// integration loop (Simulation.cuh:632-925) with respect to log k and x0.
//
// The forward pass (batch_kernel with a step record) stores the size and kind
// of every state-changing step and the state before every kCkptEvery-th step.
// Here one workgroup takes one instance from a work queue and walks its
// segments from the last: it replays the segment's steps from the checkpoint
// (same code as the forward, so the states are bit-identical), then runs the
// steps backwards, adding dL/d(sample) at each sampled state.
//
//   RKF45 step: y' = y + 25/216 k1 + 1408/2565 k3 + 2197/4104 k4 - 1/5 k5,
//     k_i = dt f(u_i) (rk_Fehlberg1..5 and _agg, kernels.cuh:68-182).
//   BDF1 step: modified Newton with M = I - h Jq(y), Jq the analytic Jacobian
//     of createJac_ (kernels.cuh:219-271) at the step's start state, run for
//     the recorded number of iterations: r_j = -(v_j - y - h f(v_j)),
//     x_j = M^-1 r_j, v_{j+1} = v_j + x_j, v_0 = y; entries ending NaN,
//     infinite or negative are reset to y (aggiornaNaN, kernels.cuh:512-521).
//     Its adjoint, per iteration from the last: z = M^-T vbar;
//     ybar += h d(z^T Jq(y) x_j)/dy + z; vbar += -z + h Jf(v_j)^T z;
//     log-k gradient += h d(z^T Jq x_j)/dlogk + h (df/dlogk)^T z.
//   Feed species hold the feed value in every stage and iterate.
//
// Mass action: rate_r = k_r prod x_a^e_a, f = H rate, df/dlogk_r = H[:,r] rate_r.
// Jf^T z and (df/dlogk)^T z go through c = H^T z (reaction-major H) and a
// species-major list of reactant entries, so every sum has one owner thread
// and a fixed order; there are no atomics.
#include <type_traits>

#include "adjoint.h"
#include "batch.h"
#include "device_common.h"
#include "wg_common.h"
#include "wg_lu.h"

namespace lassie {

__device__ __forceinline__ double rate_r(const AdjInst &I, const double *x, int r)
{
    double p = I.k[r];
    for (int j = I.lptr[r]; j < I.lptr[r + 1]; j++) p *= mypow(x[I.lsp[j]], I.lexp[j]);
    return p;
}

// d rate_r / d x_g for the reactant entry jg of reaction r
__device__ __forceinline__ double drate(const AdjInst &I, const double *x, int r, int jg)
{
    const int e = I.lexp[jg];
    double p = I.k[r] * (double)e * mypow(x[I.lsp[jg]], e - 1);
    for (int j = I.lptr[r]; j < I.lptr[r + 1]; j++)
        if (j != jg) p *= mypow(x[I.lsp[j]], I.lexp[j]);
    return p;
}

// createJac_'s term of reaction r in column lsp[jg], without the net
// coefficient: k e_g x_g^(e_g - 1) prod_{a != g} x_a
__device__ __forceinline__ double qterm(const AdjInst &I, const double *x, int r, int jg)
{
    const int e = I.lexp[jg];
    double p = I.k[r] * (double)e * mypow(x[I.lsp[jg]], e - 1);
    for (int j = I.lptr[r]; j < I.lptr[r + 1]; j++)
        if (j != jg) p *= x[I.lsp[j]];
    return p;
}

// d qterm(r, jg) / d x_{lsp[jb]}
__device__ __forceinline__ double dqterm(const AdjInst &I, const double *x, int r, int jg, int jb)
{
    const int e = I.lexp[jg];
    double p;
    if (jb == jg) {
        if (e < 2) return 0.0;
        p = I.k[r] * (double)e * (double)(e - 1) * mypow(x[I.lsp[jg]], e - 2);
        for (int j = I.lptr[r]; j < I.lptr[r + 1]; j++)
            if (j != jg) p *= x[I.lsp[j]];
    } else {
        p = I.k[r] * (double)e * mypow(x[I.lsp[jg]], e - 1);
        for (int j = I.lptr[r]; j < I.lptr[r + 1]; j++)
            if (j != jg && j != jb) p *= x[I.lsp[j]];
    }
    return p;
}

// c = H^T z, one thread per reaction
__device__ void wg_ht(const AdjInst &I, const double *z, double *c)
{
    for (int r = threadIdx.x; r < I.R; r += NT) {
        double a = 0.0;
        for (int e = I.htptr[r]; e < I.htptr[r + 1]; e++) a += (double)I.htcoef[e] * z[I.htsp[e]];
        c[r] = a;
    }
    __syncthreads();
}

// out += scale Jf(x)^T z ; gth += scale (df/dlogk)(x)^T z
__device__ void wg_vjp_f(const AdjInst &I, const double *x, const double *z, double scale, double *out, double *gth,
                         double *c)
{
    wg_ht(I, z, c);
    for (int g = threadIdx.x; g < I.S; g += NT) {
        double a = 0.0;
        for (int e = I.ltptr[g]; e < I.ltptr[g + 1]; e++) {
            const int jg = I.ltent[e];
            const int r = I.lrx[jg];
            if (c[r] != 0.0) a += c[r] * drate(I, x, r, jg);
        }
        out[g] += scale * a;
    }
    for (int r = threadIdx.x; r < I.R; r += NT)
        if (c[r] != 0.0) gth[r] += scale * c[r] * rate_r(I, x, r);
    __syncthreads();
}

// phi = z^T Jq(y) x: gy += scale dphi/dy ; gth += scale dphi/dlogk
__device__ void wg_vjp_jq(const AdjInst &I, const double *y, const double *z, const double *x, double scale,
                          double *gy, double *gth, double *c)
{
    wg_ht(I, z, c);
    for (int b = threadIdx.x; b < I.S; b += NT) {
        double a = 0.0;
        for (int e = I.ltptr[b]; e < I.ltptr[b + 1]; e++) {
            const int jb = I.ltent[e];
            const int r = I.lrx[jb];
            if (c[r] == 0.0) continue;
            double s2 = 0.0;
            for (int jg = I.lptr[r]; jg < I.lptr[r + 1]; jg++) s2 += x[I.lsp[jg]] * dqterm(I, y, r, jg, jb);
            a += c[r] * s2;
        }
        gy[b] += scale * a;
    }
    for (int r = threadIdx.x; r < I.R; r += NT) {
        if (c[r] == 0.0) continue;
        double th = 0.0;
        for (int jg = I.lptr[r]; jg < I.lptr[r + 1]; jg++) th += qterm(I, y, r, jg) * x[I.lsp[jg]];
        gth[r] += scale * c[r] * th;
    }
    __syncthreads();
}

// Solve M^T z = b with the factors of wg_lu_rm (P M = L U, rows of M in final
// order): U^T w = b forward, L^T u = w backward, in 64-row blocks (diagonal
// block by one wave with shuffles, the rest by all threads), then z = P^T u,
// the scatter z[pf[i]] = u[i] with pf the composed row order of compose_rows
// (the interchanges applied in reverse order, as one permutation).
// Arithmetic: every element's updates in the order of a plain column sweep
// (w_k = b_k / U_kk, then b_i = b_i - U_ki w_k for i > k, k ascending; u
// likewise with L, k descending inside a block, blocks descending, and the
// rows outside a block take its columns k ascending).  The factors are read
// with raw buffer loads (row in the vector offset, column in the scalar
// offset, an out-of-range offset reading zero where masked) in chunks whose
// successor is requested before the current one is used; full blocks run
// without per-column branches.
__device__ void lu_solve_t_wg(const double *__restrict__ M, int n, const int *pf, double *b)
{
    constexpr int TB = 64, CD = 16, CO = 8;
    M = uniform_ptr(M);
    n = uniform_int(n);
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const __amdgpu_buffer_rsrc_t rs = buf_rsrc(M, n * n * 8);
    const int NB = (n + TB - 1) / TB;
    auto ld = [&](int row, int voff) { return BufIO<double>::ld(rs, voff, uniform_int(row * n * 8)); };
    // the diagonal block of U^T (FWD) or L^T: lane l holds row k0 + kk's
    // entry in column k0 + l, chunk by chunk
    auto diag = [&](int k0, int nb, auto fwd, auto full) {
        constexpr bool FWD = decltype(fwd)::value, FULL = decltype(full)::value;
        double bi = lane < nb ? b[k0 + lane] : 0.0;
        const int vo = lane < nb ? (k0 + lane) * 8 : kBufOob;
        double cur[CD], nxt[CD];
        auto fetch = [&](double(&dst)[CD], int c) {
#pragma unroll
            for (int j = 0; j < CD; j++) {
                const int kk = FWD ? c * CD + j : TB - 1 - (c * CD + j);
                dst[j] = ld(k0 + kk, (FULL || kk < nb) ? vo : kBufOob);
            }
        };
        fetch(cur, 0);
#pragma unroll
        for (int c = 0; c < TB / CD; c++) {
            if (c + 1 < TB / CD) fetch(nxt, c + 1);
#pragma unroll
            for (int j = 0; j < CD; j++) {
                const int kk = FWD ? c * CD + j : TB - 1 - (c * CD + j);
                if (FULL || kk < nb) {
                    if (FWD) {
                        const double w = readlane_d(bi, kk) / readlane_d(cur[j], kk);
                        if (lane == kk) bi = w;
                        if (lane > kk && lane < nb) bi -= cur[j] * w;
                    } else {
                        const double u = readlane_d(bi, kk);
                        if (lane < kk) bi -= cur[j] * u;
                    }
                }
            }
            if (c + 1 < TB / CD) {
#pragma unroll
                for (int j = 0; j < CD; j++) cur[j] = nxt[j];
            }
        }
        if (lane < nb) b[k0 + lane] = bi;
    };
    // rows i in [lo, hi) outside the block: b_i -= sum over the block's
    // columns k ascending of M[k0 + kk][i] b[k0 + kk], two rows per thread
    auto outer = [&](int k0, int nb, int lo, int hi, auto full) {
        constexpr bool FULL = decltype(full)::value;
        for (int base = lo; base < hi; base += 2 * NT) {
            const int i0 = base + tid, i1 = i0 + NT;
            const int v0 = i0 < hi ? i0 * 8 : kBufOob, v1 = i1 < hi ? i1 * 8 : kBufOob;
            double a0 = i0 < hi ? b[i0] : 0.0, a1 = i1 < hi ? b[i1] : 0.0;
            double c0[CO], c1[CO], n0[CO], n1[CO];
            auto fetch = [&](double(&d0)[CO], double(&d1)[CO], int c) {
#pragma unroll
                for (int j = 0; j < CO; j++) {
                    const int kk = c * CO + j;
                    const bool ok = FULL || kk < nb;
                    d0[j] = ld(k0 + kk, ok ? v0 : kBufOob);
                    d1[j] = ld(k0 + kk, ok ? v1 : kBufOob);
                }
            };
            fetch(c0, c1, 0);
#pragma unroll
            for (int c = 0; c < TB / CO; c++) {
                if (c + 1 < TB / CO) fetch(n0, n1, c + 1);
#pragma unroll
                for (int j = 0; j < CO; j++) {
                    const int kk = c * CO + j;
                    if (FULL || kk < nb) {
                        const double bk = b[k0 + kk];
                        a0 -= c0[j] * bk;
                        a1 -= c1[j] * bk;
                    }
                }
                if (c + 1 < TB / CO) {
#pragma unroll
                    for (int j = 0; j < CO; j++) {
                        c0[j] = n0[j];
                        c1[j] = n1[j];
                    }
                }
            }
            if (i0 < hi) b[i0] = a0;
            if (i1 < hi) b[i1] = a1;
        }
    };
    for (int B = 0; B < NB; B++) {
        const int k0 = B * TB, nb = min(TB, n - k0);
        if (tid < 64) {
            if (nb == TB)
                diag(k0, nb, std::true_type{}, std::true_type{});
            else
                diag(k0, nb, std::true_type{}, std::false_type{});
        }
        __syncthreads();
        if (k0 + nb < n) {
            if (nb == TB)
                outer(k0, nb, k0 + nb, n, std::true_type{});
            else
                outer(k0, nb, k0 + nb, n, std::false_type{});
        }
        __syncthreads();
    }
    for (int B = NB - 1; B >= 0; B--) {
        const int k0 = B * TB, nb = min(TB, n - k0);
        if (tid < 64) {
            if (nb == TB)
                diag(k0, nb, std::false_type{}, std::true_type{});
            else
                diag(k0, nb, std::false_type{}, std::false_type{});
        }
        __syncthreads();
        if (k0 > 0) {
            if (nb == TB)
                outer(k0, nb, 0, k0, std::true_type{});
            else
                outer(k0, nb, 0, k0, std::false_type{});
        }
        __syncthreads();
    }
    // z = P^T u: z[pf[i]] = u[i]
    double u0 = tid < n ? b[tid] : 0.0, u1 = tid + NT < n ? b[tid + NT] : 0.0;
    __syncthreads();
    if (tid < n) b[pf[tid]] = u0;
    if (tid + NT < n) b[pf[tid + NT]] = u1;
    __syncthreads();
}

struct AdjWork {
    double *seg, *M, *V, *X;
    double *ua, *ub, *k1, *k2, *k3, *k4, *k5, *w1;
    double *lam, *yb, *vb, *ubar, *kb[6], *u[6], *c;
    int *piv, *perm_src, *moved_to, *info, *rowcnt, *rows;
};

// BDF1 step of size h from y with `iters` modified-Newton iterations, the
// arithmetic of batch_kernel's FP64 path: iterates in V (v_0 = y), raw
// corrections in X, factors in W.M; the repaired final iterate in y_out.
// Returns false when the LU fails.
__device__ bool bdf_iterates(const DevModel &m, const AdjInst &I, const double *y, double h, int iters, double *y_out,
                             AdjWork &W, double *s_z, int *s_pf, Shared &sh)
{
    const int S = m.S;
    const int tid = threadIdx.x;
    for (int s = tid; s < S; s += NT) W.V[s] = y[s];
    const size_t nn = (size_t)S * S;
    for (size_t e = tid; e < nn; e += NT) W.M[e] = 0.0;
    __syncthreads();
    for (int r = tid; r < S; r += NT) jac_row<true>(m, y, W.M, S, r, h);
    __syncthreads();
    if (wg_lu_rm<double, false>(W.M, S, W.piv, W.perm_src, W.moved_to, W.rows, W.rowcnt, W.info, sh) != 0) return false;
    compose_rows<kBatchNB>(W.perm_src, W.moved_to, S, s_pf);
    for (int j = 0; j < iters; j++) {
        const double *vj = W.V + (size_t)j * S;
        double *vn = W.V + (size_t)(j + 1) * S;
        double *xj = W.X + (size_t)j * S;
        for (int s = tid; s < S; s += NT) {
            const double f = rhs_species(m, I.recs, I.hptr, vj, s);
            s_z[s] = -(vj[s] - y[s] - h * f);
        }
        __syncthreads();
        lu_solve_rows(W.M, S, s_z, s_pf);
        for (int s = tid; s < S; s += NT) {
            xj[s] = s_z[s];
            vn[s] = (m.feed[s] == 0) ? vj[s] + s_z[s] : m.feed[s];
        }
        __syncthreads();
    }
    const double *vm = W.V + (size_t)iters * S;
    for (int s = tid; s < S; s += NT) y_out[s] = bad_value(vm[s]) ? y[s] : vm[s];
    __syncthreads();
    return true;
}

__device__ void rkf_forward(const DevModel &m, const AdjInst &I, double rtol, const double *y, double dt, AdjWork &W,
                            Shared &sh)
{
    CtlView v;
    v.recs = I.recs;
    v.hptr = I.hptr;
    v.y1 = const_cast<double *>(y);
    v.ua = W.ua;
    v.ub = W.ub;
    v.k1 = W.k1;
    v.k2 = W.k2;
    v.k3 = W.k3;
    v.k4 = W.k4;
    v.k5 = W.k5;
    v.w1 = W.w1;
    v.yv = nullptr;
    v.x = nullptr;
    double mind;
    (void)rkf_attempt_wg(m, rtol, v, dt, &mind, sh);
}

__global__ void __launch_bounds__(NT) adjoint_kernel(AdjArgs A)
{
    __shared__ Shared sh;
    __shared__ double s_z[kBatchMaxS];
    __shared__ int s_pf[kBatchMaxS];  // row order of the current factors (compose_rows)
    // operands of the sparse products of a reversed BDF1 step, in LDS: the
    // step's state, the iterate or correction of the current iteration, the
    // two cotangent accumulators, and c = H^T z (when R fits)
    constexpr int kAdjLdsR = 2048;
    __shared__ double s_y[kBatchMaxS], s_x[kBatchMaxS], s_vb[kBatchMaxS], s_yb[kBatchMaxS];
    __shared__ double s_c[kAdjLdsR];
    __shared__ int s_idx;
    const int tid = threadIdx.x;
    const int slot = blockIdx.x;
    const int smax = A.s_max;
    double *ws = A.ws + (size_t)slot * A.ws_stride;
    AdjWork W;
    {
        double *p = ws;
        W.seg = p;
        p += (size_t)(kCkptEvery + 1) * smax;
        W.M = p;
        p += (size_t)smax * smax + kLuSink;
        W.V = p;
        p += (size_t)(A.it_max + 1) * smax;
        W.X = p;
        p += (size_t)(A.it_max + 1) * smax;
        double **vecs[] = {&W.ua, &W.ub, &W.k1, &W.k2, &W.k3, &W.k4, &W.k5, &W.w1, &W.lam, &W.yb, &W.vb, &W.ubar};
        for (double **q : vecs) {
            *q = p;
            p += smax;
        }
        for (int i = 1; i <= 5; i++) {
            W.kb[i] = p;
            p += smax;
            W.u[i] = p;
            p += smax;
        }
        W.c = p;
        int *wi = A.wsi + (size_t)slot * A.wsi_stride;
        const int Pmax = (smax + kBatchNB - 1) / kBatchNB;
        W.piv = wi;
        W.perm_src = wi + smax;
        W.moved_to = wi + 2 * smax;
        W.info = wi + 3 * smax;
        W.rowcnt = wi + 3 * smax + 16;
        W.rows = W.rowcnt + Pmax;
    }

    while (true) {
        if (tid == 0) s_idx = atomicAdd(A.next, 1);
        __syncthreads();
        const int idx = s_idx;
        __syncthreads();
        if (idx >= A.n_inst) return;
        const int inst = A.order[idx];
        const AdjInst I = A.inst[inst];
        const int S = I.S, R = I.R;
        DevModel m;
        m.S = S;
        m.R = R;
        m.lptr = I.lptr;
        m.lsp = I.lsp;
        m.lexp = I.lexp;
        m.hptr = I.hptr;
        m.hrx = I.hrx;
        m.hk = I.hk;
        m.feed = I.feed;
        m.atol = I.atol;
        const double *fd = I.feed;
        double *cbuf = R <= kAdjLdsR ? s_c : W.c;
        const double *tdt = A.tape_dt + I.tape_off;
        const int *tkind = A.tape_kind + I.tape_off;
        const int *samp = A.tape_sample + (size_t)inst * A.nt;
        const double *gout = A.grad_out + I.g_off;
        double *gk = A.grad_logk + I.gk_off;
        int status = 0;
        for (int r = tid; r < R; r += NT) gk[r] = 0.0;
        for (int s = tid; s < S; s += NT) W.lam[s] = 0.0;
        __syncthreads();

        const long long N = I.n_steps;
        int isamp = A.nt - 1;
        auto inject = [&](long long n) {
            while (isamp >= 0 && samp[isamp] == n) {
                for (int s = tid; s < S; s += NT) W.lam[s] += gout[(size_t)isamp * S + s];
                isamp--;
            }
            __syncthreads();
        };
        const long long nseg = (N + kCkptEvery - 1) / kCkptEvery;
        for (long long sg = nseg - 1; sg >= 0 && status == 0; sg--) {
            const long long n0 = sg * kCkptEvery;
            const long long n1 = min(N, n0 + kCkptEvery);
            const double *ck = A.tape_ckpt + I.ckpt_off + sg * S;
            for (int s = tid; s < S; s += NT) W.seg[s] = ck[s];
            __syncthreads();
            // the replayed BDF1 step with the most Newton iterations keeps its
            // iterates in the second buffer pair (tape kinds: -1 RKF45, -2 no
            // change, >= 0 BDF1 iterations)
            long long nkeep = -1;
            AdjWork W2 = W;
            if (A.vx2) {
                int best = 0;
                for (long long n = n0; n < n1 - 1; n++)
                    if (tkind[n] > best) {
                        best = tkind[n];
                        nkeep = n;
                    }
                W2.V = A.vx2 + (size_t)slot * A.vx2_stride;
                W2.X = W2.V + (size_t)(A.it_max + 1) * smax;
            }
            // states n0 .. n1-1 of this segment
            for (long long n = n0; n < n1 - 1; n++) {
                const double *y = W.seg + (size_t)(n - n0) * S;
                double *yn = W.seg + (size_t)(n - n0 + 1) * S;
                const int kind = tkind[n];
                const double dt = tdt[n];
                if (kind == -1) {
                    rkf_forward(m, I, A.rtol, y, dt, W, sh);
                    for (int s = tid; s < S; s += NT) yn[s] = W.w1[s];
                    __syncthreads();
                } else if (kind == -2) {
                    for (int s = tid; s < S; s += NT) yn[s] = y[s];
                    __syncthreads();
                } else if (!bdf_iterates(m, I, y, dt, kind, yn, n == nkeep ? W2 : W, s_z, s_pf, sh)) {
                    status = 1;
                    break;
                }
            }
            if (status) break;
            for (long long n = n1; n > n0; n--) {
                inject(n);
                const long long step = n - 1;
                const int kind = tkind[step];
                const double dt = tdt[step];
                const double *y = W.seg + (size_t)(step - n0) * S;
                if (kind == -2) continue;
                if (kind == -1) {
                    rkf_forward(m, I, A.rtol, y, dt, W, sh);
                    const double *k1 = W.k1, *k2 = W.k2, *k3 = W.k3, *k4 = W.k4;
                    for (int s = tid; s < S; s += NT) {
                        W.u[1][s] = y[s];
                        const bool fed = fd[s] != 0.0;
                        W.u[2][s] = fed ? fd[s] : y[s] + k1[s] / 4.0;
                        W.u[3][s] = fed ? fd[s] : y[s] + (3.0 * k1[s]) / 32.0 + (9.0 * k2[s]) / 32.0;
                        W.u[4][s] = fed ? fd[s]
                                        : y[s] + (1932.0 * k1[s]) / 2197.0 - (7200.0 * k2[s]) / 2197.0 +
                                              (7296.0 * k3[s]) / 2197.0;
                        W.u[5][s] = fed ? fd[s]
                                        : y[s] + (439.0 * k1[s]) / 216.0 - (8.0 * k2[s]) + (3680.0 * k3[s]) / 513.0 -
                                              (845.0 * k4[s]) / 4104.0;
                        const double l = fed ? 0.0 : W.lam[s];
                        W.yb[s] = l;
                        W.kb[1][s] = (25.0 / 216.0) * l;
                        W.kb[2][s] = 0.0;
                        W.kb[3][s] = (1408.0 / 2565.0) * l;
                        W.kb[4][s] = (2197.0 / 4104.0) * l;
                        W.kb[5][s] = (-1.0 / 5.0) * l;
                    }
                    __syncthreads();
                    for (int st = 5; st >= 1; st--) {
                        for (int s = tid; s < S; s += NT) W.ubar[s] = 0.0;
                        __syncthreads();
                        wg_vjp_f(I, W.u[st], W.kb[st], dt, W.ubar, gk, cbuf);
                        for (int s = tid; s < S; s += NT) {
                            if (fd[s] != 0.0) continue;
                            const double ub = W.ubar[s];
                            W.yb[s] += ub;
                            if (st == 5) {
                                W.kb[1][s] += (439.0 / 216.0) * ub;
                                W.kb[2][s] += -8.0 * ub;
                                W.kb[3][s] += (3680.0 / 513.0) * ub;
                                W.kb[4][s] += (-845.0 / 4104.0) * ub;
                            } else if (st == 4) {
                                W.kb[1][s] += (1932.0 / 2197.0) * ub;
                                W.kb[2][s] += (-7200.0 / 2197.0) * ub;
                                W.kb[3][s] += (7296.0 / 2197.0) * ub;
                            } else if (st == 3) {
                                W.kb[1][s] += (3.0 / 32.0) * ub;
                                W.kb[2][s] += (9.0 / 32.0) * ub;
                            } else if (st == 2) {
                                W.kb[1][s] += 0.25 * ub;
                            }
                        }
                        __syncthreads();
                    }
                    for (int s = tid; s < S; s += NT) W.lam[s] = W.yb[s];
                    __syncthreads();
                } else {
                    // the kept step needs only its factors again (0 iterations)
                    const bool kept = step == nkeep;
                    if (!bdf_iterates(m, I, y, dt, kept ? 0 : kind, W.w1, W, s_z, s_pf, sh)) {
                        status = 1;
                        break;
                    }
                    const AdjWork &Wi = kept ? W2 : W;
                    const double *vm = Wi.V + (size_t)kind * S;
                    for (int s = tid; s < S; s += NT) {
                        double vb = W.lam[s], yb = 0.0;
                        if (bad_value(vm[s])) {
                            yb = vb;
                            vb = 0.0;
                        }
                        s_vb[s] = vb;
                        s_yb[s] = yb;
                        s_y[s] = y[s];
                    }
                    __syncthreads();
                    for (int j = kind - 1; j >= 0; j--) {
                        const double *xj = Wi.X + (size_t)j * S, *vj = Wi.V + (size_t)j * S;
                        for (int s = tid; s < S; s += NT) {
                            if (fd[s] != 0.0) s_vb[s] = 0.0;
                            s_z[s] = s_vb[s];
                            s_x[s] = xj[s];
                        }
                        __syncthreads();
                        lu_solve_t_wg(W.M, S, s_pf, s_z);
                        wg_vjp_jq(I, s_y, s_z, s_x, dt, s_yb, gk, cbuf);
                        for (int s = tid; s < S; s += NT) {
                            s_vb[s] -= s_z[s];
                            s_yb[s] += s_z[s];
                            s_x[s] = vj[s];
                        }
                        __syncthreads();
                        wg_vjp_f(I, s_x, s_z, dt, s_vb, gk, cbuf);
                    }
                    for (int s = tid; s < S; s += NT) W.lam[s] = s_yb[s] + s_vb[s];
                    __syncthreads();
                }
            }
        }
        if (status == 0) inject(0);
        double *gx = A.grad_x0 + I.gx_off;
        for (int s = tid; s < S; s += NT) gx[s] = W.lam[s];
        if (tid == 0) A.status[inst] = status;
        __syncthreads();
    }
}

int adjoint_occupancy(int *per_cu)
{
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, reinterpret_cast<const void *>(&adjoint_kernel), NT,
                                                        0) == hipSuccess
               ? 0
               : -1;
}

hipError_t launch_adjoint(const AdjArgs &args, int slots, hipStream_t stream)
{
    hipLaunchKernelGGL(adjoint_kernel, dim3(slots), dim3(NT), 0, stream, args);
    return hipGetLastError();
}

}  // namespace lassie
