// Batched persistent solver: many independent models of uneven size in one
// launch.  Every workgroup takes instances from an atomic work queue (largest
// first) and runs LASSIE's whole integration loop (Simulation.cuh:632-925)
// for each, with its own step control, so no workgroup ever waits for
// another and the caches keep their contents (no grid-wide fences).
//
// Newton linear solve per instance, in the workgroup, row-major dense M:
//   precision FP64   partial-pivot LU in FP64, bit-identical to the oracle
//   FP32 / BF16      M is rounded to FP32 and factored with the same pivoting
//                    in FP32 (BF16: the trailing updates on the BF16 matrix
//                    cores, FP32 accumulation); each Newton correction is refined in
//                    FP64 against the FP64 M to a componentwise backward error
//                    of refine_tol, else the step falls back to the FP64 LU.
#include "batch.h"
#include "device_common.h"
#include "wg_common.h"
#include "wg_lu.h"

namespace lassie {

// Mixed-precision Newton correction: solve M d = rhs with the FP32 factors and
// FP64 residuals r = rhs - M64 x, until every component meets the
// componentwise backward-error test |r_i| <= tol * ((|M64| |x|)_i + |rhs_i| + atol_i)
// (Oettli-Prager with the species' absolute tolerance as floor), so small
// components are as accurate as large ones, while components far below atol
// (under FP32's normal range, which no FP32 correction can reach) do not hold
// refinement back.  Returns true on convergence; x (double, LDS) holds the result.
// max_it == 0: the first FP32 solve, x = 0 + (double)d, with no residual and
// no test (the modified-Newton iteration refines it).
__device__ bool refine_solve(const double *__restrict__ M64, const float *__restrict__ M32, int n, const int *perm_src,
                             const int *moved_to, const int *pf, const double *atol, const double *rhs, double *x,
                             float *b32, double tol, int max_it, long long &iters, Shared &sh)
{
    const int tid = threadIdx.x;
    for (int i = tid; i < n; i += NT) {
        x[i] = 0.0;
        b32[i] = (float)rhs[i];
    }
    __syncthreads();
    if (max_it == 0) {
        iters++;
        lu_solve_wg<kBatchNB, float, true>(M32, n, perm_src, moved_to, b32, pf);
        for (int i = tid; i < n; i += NT) x[i] = x[i] + (double)b32[i];
        __syncthreads();
        return true;
    }
    for (int it = 0; it < max_it; it++) {
        iters++;
        lu_solve_wg<kBatchNB, float, true>(M32, n, perm_src, moved_to, b32, pf);
        for (int i = tid; i < n; i += NT) x[i] = x[i] + (double)b32[i];
        __syncthreads();
        // r = rhs - M64 x, one wave per row, lanes over columns (coalesced)
        int bad = 0;
        const int lane = tid & 63, wave = tid >> 6;
        for (int i = wave; i < n; i += NWAVES) {
            const double *row = M64 + (size_t)i * n;
            double pa = 0.0, pm = 0.0;
            for (int j = lane; j < n; j += 64) {
                const double mv = row[j], xv = x[j];
                pa = fma(mv, xv, pa);
                pm = fma(fabs(mv), fabs(xv), pm);
            }
            const double acc = rhs[i] - wave_sum(pa);
            const double mag = (fabs(rhs[i]) + wave_sum(pm)) + atol[i];
            if (lane == 0) b32[i] = (float)acc;
            bad |= !(fabs(acc) <= tol * mag);
        }
        if (__syncthreads_or(bad) == 0) return true;
    }
    return false;
}

template <int RT>
__global__ void __launch_bounds__(NT) batch_kernel(BatchArgs A)
{
    __shared__ Shared sh;
    __shared__ double s_x[kBatchMaxS];  // solve vector / refined correction
    __shared__ float s_b32[kBatchMaxS];
    __shared__ int s_pf[kBatchMaxS];    // row order of the current factors (compose_rows)
    // the instance's state vectors (y1, two stage vectors, k1..k5, w1, yv, rhs):
    // LDS, so the RKF45 stages and Newton updates make no global round trips
    __shared__ double s_vec[11 * kBatchMaxS];
    static_assert(kBatchMaxS <= 2 * NT, "lu_solve_wg and lu_solve_rows hold at most two rows per thread");
    __shared__ int s_idx;
    const int tid = threadIdx.x;
    const int slot = blockIdx.x;
    double *ws = A.ws64 + (size_t)slot * A.ws64_stride;
    float *M32 = A.ws32 ? A.ws32 + (size_t)slot * A.ws32_stride : nullptr;
    int *wi = A.wsi + (size_t)slot * A.wsi_stride;
    const int smax = A.s_max;
    const int Pmax = (smax + kBatchNB - 1) / kBatchNB;
    int *piv = wi, *perm_src = wi + smax, *moved_to = wi + 2 * smax, *info = wi + 3 * smax;
    int *rowcnt = wi + 3 * smax + 16, *rows = rowcnt + Pmax;

    unsigned int n_done = 0;  // instances this slot has finished
    while (true) {
        if (tid == 0) s_idx = atomicAdd(A.next, 1);
        __syncthreads();
        const int idx = s_idx;
        __syncthreads();
        if (idx >= A.n_inst) return;
        const int inst = A.order[idx];
        const BatchInst I = A.inst[inst];
        const int S = I.S;
        DevModel m;
        m.S = S;
        m.R = I.R;
        m.lptr = I.lptr;
        m.lsp = I.lsp;
        m.lexp = I.lexp;
        m.hptr = I.hptr;
        m.hrx = I.hrx;
        m.hk = I.hk;
        m.feed = I.feed;
        m.atol = I.atol;
        CtlView v;
        v.recs = I.recs;
        v.hptr = I.hptr;
        // the right-hand-side records (and their per-species offsets) go to the
        // LU's staging buffer whenever it is idle (at the start and after each
        // factorization), so the RKF45 stages and Newton residuals walk them in
        // LDS; the Jacobian staging and the LU overwrite it in between, when no
        // right-hand side is evaluated.  The same values in the same order.
        const int nh_recs = I.hptr[S];
        const bool recs_fit = (size_t)nh_recs * sizeof(HRec) + sizeof(int) * (size_t)(S + 1) <= sizeof(lp_raw);
        auto stage_recs = [&]() {
            if (!recs_fit) return;
            HRec *lr = reinterpret_cast<HRec *>(lp_raw);
            int *lh = reinterpret_cast<int *>(lr + nh_recs);
            for (int e = tid; e < nh_recs; e += NT) lr[e] = I.recs[e];
            for (int e = tid; e <= S; e += NT) lh[e] = I.hptr[e];
            v.recs = lr;
            v.hptr = lh;
            __syncthreads();
        };
        stage_recs();
        v.y1 = s_vec;
        v.ua = s_vec + S;
        v.ub = s_vec + 2 * S;
        v.k1 = s_vec + 3 * S;
        v.k2 = s_vec + 4 * S;
        v.k3 = s_vec + 5 * S;
        v.k4 = s_vec + 6 * S;
        v.k5 = s_vec + 7 * S;
        v.w1 = s_vec + 8 * S;
        v.yv = s_vec + 9 * S;
        v.x = s_x;
        double *rhs = s_vec + 10 * S;
        double *M = ws + 12 * (size_t)smax;
        double *out = A.out + I.out_off;

        double t = 0.0, t1 = A.tv[0], dt = A.rkf_step, newdt = A.be_step;
        const double t_max = A.tv[A.nt - 1];
        (void)t_max;
        int i_sample = 0;
        long long nrej = 0;
        // device time by phase (100 MHz ticks): RKF45 attempts, Jacobian, LU, Newton
        unsigned long long ph_rkf = 0, ph_jac = 0, ph_lu = 0, ph_nw = 0, tph = 0;
        long long st_loop = 0, st_att = 0, st_acc = 0, st_bdf = 0, st_newton = 0, st_lufail = 0, st_rep = 0,
                  st_lu = 0, st_ref = 0, st_fb = 0;
        int rc = 0;
        for (int s = tid; s < S; s += NT) v.y1[s] = I.x0[s];
        __syncthreads();
        // step record: called by every thread before the state changes
        const bool rec = A.tape_dt != nullptr;
        long long nstep = 0, max_kind = 0;
        auto record = [&](double dstep, int kind) -> bool {
            if (!rec) return true;
            if (nstep >= A.tape_cap) return false;
            if (nstep % kCkptEvery == 0) {
                double *ck = A.tape_ckpt + I.ckpt_off + (nstep / kCkptEvery) * S;
                for (int s = tid; s < S; s += NT) ck[s] = v.y1[s];
            }
            if (tid == 0) {
                A.tape_dt[I.tape_off + nstep] = dstep;
                A.tape_kind[I.tape_off + nstep] = kind;
            }
            max_kind = max(max_kind, (long long)kind);
            nstep++;
            return true;
        };

        while (true) {
            if ((A.max_iters > 0 && st_loop >= A.max_iters) || (A.max_newton > 0 && st_newton >= A.max_newton)) {
                rc = 1;
                break;
            }
            st_loop++;
            if (fabs(t1 - t) < 1e-5) {
                for (int j = tid; j < S; j += NT) out[(size_t)i_sample * S + j] = v.y1[j];
                if (rec && tid == 0) A.tape_sample[(size_t)inst * A.nt + i_sample] = (int)nstep;
                i_sample++;
                if (i_sample == A.nt) break;
                t1 = A.tv[i_sample];
            }
            st_att++;
            double mind;
            tph = wall_clock64();
            const int nok = rkf_attempt_wg(m, A.rtol, v, dt, &mind, sh);
            ph_rkf += wall_clock64() - tph;
            if (nok == S) {
                nrej = 0;
                st_acc++;
                if (!record(dt, -1)) {
                    rc = 2;
                    break;
                }
                __syncthreads();
                for (int s = tid; s < S; s += NT) v.y1[s] = v.w1[s];
                __syncthreads();
                const double d_temp = dt * mind;
                if ((t + d_temp) > t1)
                    dt = fabs(t1 - t);
                else
                    dt = d_temp;
                t = t + dt;
                continue;
            }
            const double dt_old = dt;
            dt = dt * mind;
            nrej++;
            if (!(dt < A.stiffness_tol || isnan(dt) || isinf(dt) || fabs(dt_old - dt) < 1e-3 ||
                  (A.max_rkf_retries > 0 && nrej >= A.max_rkf_retries)))
                continue;
            // ---- BDF1 step with modified Newton (Simulation.cuh:780-919) ------
            nrej = 0;
            st_bdf++;
            dt = newdt;
            if ((t + dt) > t1) dt = fabs(t1 - t);
            const double h = dt;
            for (int s = tid; s < S; s += NT) v.yv[s] = v.y1[s];
            __syncthreads();
            bool luFail = false;
            int itN = 0;
            if (A.newton_iter > 0) {
                tph = wall_clock64();
                const size_t nn = (size_t)S * S;
                for (size_t e = tid; e < nn; e += NT) M[e] = 0.0;
                // the model arrays the Jacobian walks go to the LU's panel buffer
                // (idle until the factorization) when they fit, so its chains of
                // dependent index loads read LDS; the same values, the same order
                DevModel mj = m;
                {
                    const int nh = m.hptr[S], nl = m.lptr[m.R];
                    const size_t bytes =
                        8 * (size_t)nh + 4 * ((size_t)(S + 1) + nh + (size_t)(m.R + 1) + 2 * (size_t)nl);
                    if (bytes <= sizeof(lp_raw)) {
                        double *hk = lp_raw;
                        int *hp = reinterpret_cast<int *>(lp_raw + nh);
                        int *hx = hp + (S + 1), *lp = hx + nh, *ls = lp + (m.R + 1), *le = ls + nl;
                        for (int e = tid; e < nh; e += NT) {
                            hk[e] = m.hk[e];
                            hx[e] = m.hrx[e];
                        }
                        for (int e = tid; e <= S; e += NT) hp[e] = m.hptr[e];
                        for (int e = tid; e <= m.R; e += NT) lp[e] = m.lptr[e];
                        for (int e = tid; e < nl; e += NT) {
                            ls[e] = m.lsp[e];
                            le[e] = m.lexp[e];
                        }
                        mj.hk = hk;
                        mj.hptr = hp;
                        mj.hrx = hx;
                        mj.lptr = lp;
                        mj.lsp = ls;
                        mj.lexp = le;
                    }
                }
                __syncthreads();
                for (int r = tid; r < S; r += NT) jac_row<true>(mj, v.yv, M, S, r, h);
                __syncthreads();
                {
                    const unsigned long long now = wall_clock64();
                    ph_jac += now - tph;
                    tph = now;
                }
                bool fp64 = A.precision == 0;
                int inf = 0;
                if (!fp64) {
                    for (size_t e = tid; e < nn; e += NT) M32[e] = (float)M[e];
                    __syncthreads();
                    const int inf32 = (A.precision == 2)
                                          ? wg_lu_rm<float, true, RT>(M32, S, piv, perm_src, moved_to, rows, rowcnt, info, sh)
                                          : wg_lu_rm<float, false, RT>(M32, S, piv, perm_src, moved_to, rows, rowcnt, info, sh);
                    if (inf32) {
                        fp64 = true;
                        st_fb++;
                    }
                }
                if (fp64) inf = wg_lu_rm<double, false, RT>(M, S, piv, perm_src, moved_to, rows, rowcnt, info, sh);
                if (!inf) compose_rows<kBatchNB>(perm_src, moved_to, S, s_pf);
                stage_recs();
                st_lu++;
                {
                    const unsigned long long now = wall_clock64();
                    ph_lu += now - tph;
                    tph = now;
                }
                double errN = 1000;
                while (itN < A.newton_iter && errN > A.newton_tol) {
                    for (int s = tid; s < S; s += NT) {
                        const double f = rhs_species(m, v.recs, v.hptr, v.yv, s);
                        rhs[s] = -(v.yv[s] - v.y1[s] - h * f);
                    }
                    __syncthreads();
                    if (inf) {
                        for (int s = tid; s < S; s += NT)
                            if (bad_value(v.yv[s])) v.yv[s] = v.y1[s];
                        __syncthreads();
                        newdt = h * 0.1;
                        luFail = true;
                        st_lufail++;
                        break;
                    }
                    if (fp64) {
                        for (int s = tid; s < S; s += NT) s_x[s] = rhs[s];
                        __syncthreads();
                        lu_solve_rows(M, S, s_x, s_pf);
                    } else if (!refine_solve(M, M32, S, perm_src, moved_to, s_pf, m.atol, rhs, s_x, s_b32, A.refine_tol,
                                             A.refine_max_iter, st_ref, sh)) {
                        // refinement did not converge: FP64 LU for the rest of this step
                        fp64 = true;
                        st_fb++;
                        inf = wg_lu_rm<double, false, RT>(M, S, piv, perm_src, moved_to, rows, rowcnt, info, sh);
                        stage_recs();
                        if (inf) continue;  // handled at the top of the loop as an LU failure
                        compose_rows<kBatchNB>(perm_src, moved_to, S, s_pf);
                        for (int s = tid; s < S; s += NT) s_x[s] = rhs[s];
                        __syncthreads();
                        lu_solve_rows(M, S, s_x, s_pf);
                    }
                    double part = 0.0;
                    for (int s = tid; s < S; s += NT) {
                        const double yv1 = v.yv[s];
                        double yv;
                        if (m.feed[s] == 0)
                            yv = yv1 + s_x[s];
                        else
                            yv = m.feed[s];
                        v.yv[s] = yv;
                        const double d = yv - yv1;
                        part = part + (double)(float)(d * d);
                    }
                    errN = sqrt(wg_tree_sum(part, sh));
                    itN = itN + 1;
                }
                st_newton += itN;
                ph_nw += wall_clock64() - tph;
            }
            if (!luFail) {
                int nbad = 0;
                for (int s = tid; s < S; s += NT) nbad += bad_value(v.yv[s]) ? 1 : 0;
                nbad = wg_isum(nbad, sh);
                if (nbad > 0) {
                    for (int s = tid; s < S; s += NT)
                        if (bad_value(v.yv[s])) v.yv[s] = v.y1[s];
                    newdt = h * 0.1;
                    st_rep++;
                } else if (A.bdf_regrow) {
                    newdt = fmin(newdt * 2.0, A.be_step);
                }
            }
            __syncthreads();
            if (!record(h, luFail ? -2 : itN)) {
                rc = 2;
                break;
            }
            __syncthreads();
            for (int s = tid; s < S; s += NT) v.y1[s] = v.yv[s];
            __syncthreads();
            t = t + h;
        }
        for (int s = tid; s < S; s += NT) A.y_final[I.y_off + s] = v.y1[s];
        if (tid == 0) {
            long long *o = A.stats + (size_t)inst * kBatchStats;
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
            o[10] = st_ref;
            o[11] = st_fb;
            o[12] = nstep;
            o[13] = max_kind;
            o[14] = (long long)ph_rkf;
            o[15] = (long long)ph_jac;
            o[16] = (long long)ph_lu;
            o[17] = (long long)ph_nw;
            double *d = A.dstats + (size_t)inst * 4;
            d[0] = t;
            d[1] = dt;
            d[2] = newdt;
            if (A.done_host) {
                n_done++;
                __hip_atomic_store(A.done_host + slot, n_done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            }
        }
        __syncthreads();
    }
}

static const void *batch_kernel_ptr(int variant)
{
    switch (variant) {
    case 0: return reinterpret_cast<const void *>(&batch_kernel<1>);
    case 1: return reinterpret_cast<const void *>(&batch_kernel<2>);
    default: return reinterpret_cast<const void *>(&batch_kernel<4>);
    }
}

int batch_variant_rt(int variant) { return variant == 0 ? 1 : (variant == 1 ? 2 : 4); }

int batch_occupancy(int *per_cu, int variant)
{
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, batch_kernel_ptr(variant), NT, 0) == hipSuccess ? 0
                                                                                                              : -1;
}

hipError_t launch_batch(const BatchArgs &args, int slots, hipStream_t stream, int variant)
{
    switch (variant) {
    case 0: hipLaunchKernelGGL(batch_kernel<1>, dim3(slots), dim3(NT), 0, stream, args); break;
    case 1: hipLaunchKernelGGL(batch_kernel<2>, dim3(slots), dim3(NT), 0, stream, args); break;
    default: hipLaunchKernelGGL(batch_kernel<4>, dim3(slots), dim3(NT), 0, stream, args); break;
    }
    return hipGetLastError();
}

// Resets the two work-queue heads and marks every output sample as not
// reached (NaN): a kernel, so that a captured graph holds no memset node.
__global__ void batch_reset_kernel(int *next, double *out, size_t n_out)
{
    const size_t i0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i0 < 2) next[i0] = 0;
    for (size_t i = i0; i < n_out; i += (size_t)gridDim.x * blockDim.x) out[i] = __builtin_nan("");
}

hipError_t launch_batch_reset(int *next, hipStream_t stream, double *out, size_t n_out)
{
    size_t blocks = (n_out + NT - 1) / NT;
    blocks = blocks < 1 ? 1 : (blocks > 2048 ? 2048 : blocks);
    hipLaunchKernelGGL(batch_reset_kernel, dim3((unsigned)blocks), dim3(NT), 0, stream, next, out, n_out);
    return hipGetLastError();
}

}  // namespace lassie
