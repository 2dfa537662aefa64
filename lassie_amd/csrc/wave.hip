// Wave-level batched solver: instances with at most 64 species, one
// wavefront (a 64-thread workgroup) per instance, lane s owning species s and
// row s of the Newton matrix.  Each wave takes instances from an atomic work
// queue and runs LASSIE's whole integration loop (Simulation.cuh:632-925) for
// them with its own step control.
//
// Everything an instance touches per step lives in the wave's LDS: the dense
// Newton matrix (rows padded to S+1 doubles, so row and column walks are both
// free of bank conflicts), the state vectors that other lanes read, and the
// model's right-hand-side records.  Per-species values (the RKF45 stage
// increments, w1, the Newton iterate) stay in registers.  A workgroup of one
// wave synchronises with a single s_barrier, so a Newton iteration of a small
// model costs a few microseconds instead of the tens a 256-thread workgroup
// spends on its barriers.
//
// The arithmetic is the FP64 method's, in the oracle's order, so results are
// bit-identical to the workgroup-level batched kernel and to the single-model
// solver:
//   RHS and Jacobian        rhs_species, jac_row (kernels.cuh:5-63, 205-271)
//   RKF45 attempt           kernels.cuh:68-199, same expressions
//   LU                      right-looking kij with explicit row swaps,
//                           first largest |a| (strict >), multipliers by
//                           division, a_ij = fma(-l_ik, u_kj, a_ij) with k
//                           ascending (cuSOLVER getrf seat, Simulation.cuh:1304)
//   triangular solves       forward k ascending, backward k descending, one
//                           fma chain per row (getrs seat, Simulation.cuh:1316)
//   Newton norm             float-rounded squares, the 256-lane halving tree
//                           (Simulation.cuh:58-64, 883-888)
//   step-factor minimum     thrust::min_element semantics (Simulation.cuh:753-754)
#include "device_common.h"
#include <utility>
#include "wave.h"
#include "wg_common.h"

namespace lassie {

namespace {

constexpr int WL = 64;  // lanes

// fmin over the wave.  On the step factors (never -0; NaNs ignored by fmin)
// the result does not depend on the order.
__device__ __forceinline__ double wave_fmin(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off));
    return v;
}

// The Newton norm's sum: the workgroup kernels and the oracle (orc_normsq)
// add the terms in 256 strided lanes and then halve the lanes pairwise.  With
// at most 64 species only lanes 0..63 hold a term; the strides 128 and 64 add
// +0 (every term is a non-negative square), and strides 32..1 remain, each
// step a[t] = a[t] + a[t + stride].
__device__ __forceinline__ double wave_tree_sum(double v)
{
#pragma unroll
    for (int stride = 32; stride > 0; stride >>= 1) v = v + __shfl_down(v, stride);
    return readlane_d(v, 0);
}

// One RKF45 attempt (Simulation.cuh:692-740; kernels.cuh:68-199) of the
// instance, lane s computing species s.  y1s, ua, ub are LDS vectors (y1s the
// current state, ua/ub the stage states other lanes read); y1, fv, at are this
// lane's state, feed value and absolute tolerance.  Returns the number of
// species passing calculate_delta's test; w1o receives this lane's
// fourth-order solution and *mind the smallest step factor.
__device__ int rkf_attempt_wave(const DevModel &m, const HRec *recs, const int *hptr, double rtol,
                                const double *y1s, double *ua, double *ub, double dt, double y1, double fv,
                                double at, bool act, double &w1o, double *mind)
{
    const int s = threadIdx.x;
    double k1 = 0.0, k2 = 0.0, k3 = 0.0, k4 = 0.0, k5 = 0.0;
    if (act) {
        const double f = rhs_species(m, recs, hptr, y1s, s);
        if (fv != 0.0) {
            k1 = fv;
            ua[s] = fv;
        } else {
            const double kk = f * dt;
            k1 = kk;
            ua[s] = y1 + kk / 4.0;
        }
    }
    __syncthreads();
    if (act) {
        const double f = rhs_species(m, recs, hptr, (const double *)ua, s);
        if (fv != 0.0) {
            k2 = fv;
            ub[s] = fv;
        } else {
            const double kk = f * dt;
            k2 = kk;
            ub[s] = y1 + (3.0 * k1) / 32.0 + (9.0 * kk) / 32.0;
        }
    }
    __syncthreads();
    if (act) {
        const double f = rhs_species(m, recs, hptr, (const double *)ub, s);
        if (fv != 0.0) {
            k3 = fv;
            ua[s] = fv;
        } else {
            const double kk = f * dt;
            k3 = kk;
            ua[s] = y1 + (1932.0 * k1) / 2197.0 - (7200.0 * k2) / 2197.0 + (7296.0 * kk) / 2197.0;
        }
    }
    __syncthreads();
    if (act) {
        const double f = rhs_species(m, recs, hptr, (const double *)ua, s);
        if (fv != 0.0) {
            k4 = fv;
            ub[s] = fv;
        } else {
            const double kk = f * dt;
            k4 = kk;
            ub[s] = y1 + (439.0 * k1) / 216.0 - (8.0 * k2) + (3680.0 * k3) / 513.0 - (845.0 * kk) / 4104.0;
        }
    }
    __syncthreads();
    if (act) {
        const double f = rhs_species(m, recs, hptr, (const double *)ub, s);
        if (fv != 0.0) {
            k5 = fv;
            ua[s] = fv;
        } else {
            const double kk = f * dt;
            k5 = kk;
            ua[s] = y1 - (8.0 * k1) / 27.0 + (2.0 * k2) - (3544.0 * k3) / 2565.0 + (1859.0 * k4) / 4104.0
                    - (11.0 * kk) / 40.0;
        }
    }
    __syncthreads();
    bool ok = false;
    double d = INFINITY;
    double w1 = 0.0;
    if (act) {
        const double f = rhs_species(m, recs, hptr, (const double *)ua, s);
        double w2;
        if (fv != 0.0) {
            w1 = fv;
            w2 = fv;
        } else {
            const double kk6 = f * dt;
            w1 = y1 + (25.0 * k1) / 216.0 + (1408.0 * k3) / 2565.0 + (2197.0 * k4) / 4104.0 - k5 / 5.0;
            w2 = y1 + (16.0 * k1) / 135.0 + (6656.0 * k3) / 12825.0 + (28561.0 * k4) / 56430.0 - (9.0 * k5 / 50.0)
                 + (2.0 * kk6) / 55.0;
        }
        const double R = fabs(w1 - w2) / dt;
        const double tol = (rtol > 0.0) ? (at + rtol * fabs(w1)) : at;
        d = 0.84 * sqrt(sqrt(tol / R));
        ok = R <= tol && !isnan(R) && !isinf(R) && w1 > 0;
    }
    w1o = w1;
    const int nok = __popcll(__ballot(ok));
    const bool nan0 = isnan(readlane_d(d, 0));  // thrust::min_element: a leading NaN wins
    const double r = wave_fmin(d);
    *mind = nan0 ? NAN : r;
    return nok;
}

// getrs with one right-hand side: b = rhs permuted by the factors' row order
// (lane i takes rhs[myorig]), forward substitution with the unit lower factor
// (k ascending), backward with the upper factor (k descending), lane i
// holding b_i; each step's pivot value is broadcast from its lane.  Returns,
// in lane s, component s of the solution.
__device__ double lu_solve_wave(const double *M, int ld, int S, int myorig, const double *rhs)
{
    constexpr int KB = 8;
    const int lane = threadIdx.x;
    double b = (lane < S) ? rhs[myorig] : 0.0;
    for (int k0 = 0; k0 < S; k0 += KB) {
        double l[KB];
#pragma unroll
        for (int q = 0; q < KB; q++) {
            const int k = k0 + q;
            l[q] = ((lane > k) & (lane < S)) ? M[lane * ld + k] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < KB; q++) {
            const int k = k0 + q;
            if (k < S) {
                const double bk = readlane_d(b, k);
                if ((lane > k) & (lane < S)) b = fma(-l[q], bk, b);
            }
        }
    }
    for (int k0 = S - 1; k0 >= 0; k0 -= KB) {
        double u[KB], dg[KB];
#pragma unroll
        for (int q = 0; q < KB; q++) {
            const int k = k0 - q;
            u[q] = ((k >= 0) & (lane < k)) ? M[lane * ld + k] : 0.0;
            dg[q] = (k >= 0) ? M[k * ld + k] : 1.0;
        }
#pragma unroll
        for (int q = 0; q < KB; q++) {
            const int k = k0 - q;
            if (k >= 0) {
                if (lane == k) b = b / dg[q];
                const double xk = readlane_d(b, k);
                if (lane < k) b = fma(-u[q], xk, b);
            }
        }
    }
    return b;
}

// LU with partial pivoting of the S x S row-major matrix M (leading dimension
// ld, LDS), getrf's row swaps, with the matrix in registers: lane i holds row
// i (NPC doubles) for the whole factorization and rows never move; each
// row's position (pos) follows the swaps, so the pivot search (the largest
// |a| among the rows at positions k.., the lowest position on ties; a NaN at
// position k keeps the diagonal, NaNs below never win) and every row's
// operations are those of the physical swaps.  Per column: the pivot's lane
// by ballot, the multipliers by division, the pivot row's later entries by
// readlane, one fma per entry; the factors go to M at their final positions.
// myorig receives, in lane i, the original row now at position i.  Returns
// 0, or k + 1 at the first zero pivot (the matrix is then singular for the
// loop, Simulation.cuh:1309).  FULL: S == NPC (no column checks).
template <typename F, int... K>
__device__ __forceinline__ void static_for_impl(F &&f, std::integer_sequence<int, K...>)
{
    (f(std::integral_constant<int, K>{}), ...);
}
// f(integral_constant<int, k>) for k = 0 .. N-1: every k a compile-time
// constant, so register arrays indexed by k stay in registers
template <int N, typename F>
__device__ __forceinline__ void static_for(F &&f)
{
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

template <int NPC, bool FULL>
__device__ int lu_wave_reg(double *M, int ld, int S, int &myorig)
{
    const int lane = threadIdx.x;
    const bool row = lane < S;
    double a[NPC];
#pragma unroll
    for (int c = 0; c < NPC; c++) a[c] = (row && (FULL || c < S)) ? M[lane * ld + c] : 0.0;
    int pos = lane;
    bool act = row;  // not yet a pivot row
    int info = 0;
    static_for<NPC>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        if (info != 0 || (!FULL && k >= S)) return;
        const double x = a[k];
        const double v = act ? (isnan(x) ? ((pos == k) ? INFINITY : -1.0) : fabs(x)) : -2.0;
        double wv;
        int wp;
        wave_argmax(v, act ? pos : 0x7fffffff, wv, wp);
        const int pl = __builtin_amdgcn_readfirstlane(__builtin_ctzll(__ballot(act & (pos == wp))));
        const double ukk = readlane_d(x, pl);
        if (ukk == 0.0) {
            info = k + 1;
            return;
        }
        const bool isp = lane == pl;
        pos = isp ? k : ((act & (pos == k)) ? wp : pos);
        act = act & !isp;
        const double l = x / ukk;
        a[k] = act ? l : x;
        if (act) {
#pragma unroll
            for (int c = k + 1; c < NPC; c++)
                if (FULL || c < S) a[c] = fma(-l, readlane_d(a[c], pl), a[c]);
        }
    });
    if (info != 0) return info;
    if (row) {
#pragma unroll
        for (int c = 0; c < NPC; c++)
            if (FULL || c < S) M[pos * ld + c] = a[c];
    }
    // lane i: the original row now at position i
    myorig = __builtin_amdgcn_ds_permute(pos * 4, lane);
    __syncthreads();
    return 0;
}

__device__ int lu_wave_any(double *M, int ld, int S, int &myorig)
{
    if (S == 16) return lu_wave_reg<16, true>(M, ld, S, myorig);
    if (S == 32) return lu_wave_reg<32, true>(M, ld, S, myorig);
    if (S == 64) return lu_wave_reg<64, true>(M, ld, S, myorig);
    if (S <= 16) return lu_wave_reg<16, false>(M, ld, S, myorig);
    if (S <= 32) return lu_wave_reg<32, false>(M, ld, S, myorig);
    return lu_wave_reg<64, false>(M, ld, S, myorig);
}

}  // namespace

size_t wave_lds_bytes(int smax, int nh_max)
{
    const size_t b = 8 * ((size_t)smax * (smax + 1) + 5 * (size_t)smax) + 16 * (size_t)nh_max + 4 * (size_t)(smax + 1);
    return (b + 15) & ~(size_t)15;
}

__global__ void __launch_bounds__(WL) wave_kernel(BatchArgs A)
{
    extern __shared__ double wdyn[];
    __shared__ int s_idx;
    const int lane = threadIdx.x;
    const int smax = A.s_max;
    double *Mw = wdyn;  // smax rows of smax + 1 (an instance uses S rows of S + 1)
    double *y1s = Mw + (size_t)smax * (smax + 1);
    double *ua = y1s + smax, *ub = ua + smax, *yvs = ub + smax, *rv = yvs + smax;
    HRec *recs = reinterpret_cast<HRec *>(rv + smax);
    int *hptr = reinterpret_cast<int *>(recs + A.nh_max);
    const int slot = blockIdx.x;
    unsigned int n_done = 0;

    while (true) {
        if (lane == 0) s_idx = atomicAdd(A.next, 1);
        __syncthreads();
        const int idx = s_idx;
        __syncthreads();
        if (idx >= A.n_inst) return;
        const int inst = A.order[idx];
        const BatchInst I = A.inst[inst];
        const int S = I.S;
        const int ld = S + 1;
        const bool act = lane < S;
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
        {
            const int nh = I.hptr[S];
            for (int e = lane; e < nh; e += WL) recs[e] = I.recs[e];
            for (int e = lane; e <= S; e += WL) hptr[e] = I.hptr[e];
        }
        const double fv = act ? I.feed[lane] : 0.0;
        const double atl = act ? I.atol[lane] : 0.0;
        double y1 = act ? I.x0[lane] : 0.0;
        if (act) y1s[lane] = y1;
        __syncthreads();
        double *out = A.out + I.out_off;

        double t = 0.0, t1 = A.tv[0], dt = A.rkf_step, newdt = A.be_step;
        int i_sample = 0;
        long long nrej = 0;
        // device time by phase (100 MHz ticks): RKF45 attempts, Jacobian, LU, Newton
        unsigned long long ph_rkf = 0, ph_jac = 0, ph_lu = 0, ph_nw = 0, tph = 0;
        long long st_loop = 0, st_att = 0, st_acc = 0, st_bdf = 0, st_newton = 0, st_lufail = 0, st_rep = 0,
                  st_lu = 0;
        int rc = 0;
        while (true) {
            if ((A.max_iters > 0 && st_loop >= A.max_iters) || (A.max_newton > 0 && st_newton >= A.max_newton)) {
                rc = 1;
                break;
            }
            st_loop++;
            if (fabs(t1 - t) < 1e-5) {
                if (act) out[(size_t)i_sample * S + lane] = y1;
                i_sample++;
                if (i_sample == A.nt) break;
                t1 = A.tv[i_sample];
            }
            st_att++;
            double mind, w1;
            tph = wall_clock64();
            const int nok = rkf_attempt_wave(m, recs, hptr, A.rtol, y1s, ua, ub, dt, y1, fv, atl, act, w1, &mind);
            ph_rkf += wall_clock64() - tph;
            if (nok == S) {
                nrej = 0;
                st_acc++;
                y1 = w1;
                if (act) y1s[lane] = y1;
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
            double yv = y1;
            if (act) yvs[lane] = yv;
            bool luFail = false;
            int itN = 0;
            if (A.newton_iter > 0) {
                tph = wall_clock64();
                for (int e = lane; e < S * ld; e += WL) Mw[e] = 0.0;
                __syncthreads();
                if (act) jac_row<true>(m, (const double *)yvs, Mw, ld, lane, h);
                __syncthreads();
                {
                    const unsigned long long now = wall_clock64();
                    ph_jac += now - tph;
                    tph = now;
                }
                int myorig;
                const int inf = lu_wave_any(Mw, ld, S, myorig);
                st_lu++;
                {
                    const unsigned long long now = wall_clock64();
                    ph_lu += now - tph;
                    tph = now;
                }
                double errN = 1000;
                while (itN < A.newton_iter && errN > A.newton_tol) {
                    if (act) {
                        const double f = rhs_species(m, recs, hptr, (const double *)yvs, lane);
                        rv[lane] = -(yv - y1 - h * f);
                    }
                    __syncthreads();
                    if (inf) {
                        if (act && bad_value(yv)) {
                            yv = y1;
                            yvs[lane] = yv;
                        }
                        __syncthreads();
                        newdt = h * 0.1;
                        luFail = true;
                        st_lufail++;
                        break;
                    }
                    const double x = lu_solve_wave(Mw, ld, S, myorig, rv);
                    double part = 0.0;
                    if (act) {
                        const double yv1 = yv;
                        double yn;
                        if (fv == 0)
                            yn = yv1 + x;
                        else
                            yn = fv;
                        yv = yn;
                        yvs[lane] = yn;
                        const double d = yn - yv1;
                        part = part + (double)(float)(d * d);
                    }
                    __syncthreads();
                    errN = sqrt(wave_tree_sum(part));
                    itN = itN + 1;
                }
                st_newton += itN;
                ph_nw += wall_clock64() - tph;
            }
            if (!luFail) {
                const bool bad = act && bad_value(yv);
                if (__ballot(bad) != 0ULL) {
                    if (bad) yv = y1;
                    newdt = h * 0.1;
                    st_rep++;
                } else if (A.bdf_regrow) {
                    newdt = fmin(newdt * 2.0, A.be_step);
                }
            }
            y1 = yv;
            if (act) y1s[lane] = y1;
            __syncthreads();
            t = t + h;
        }
        if (act) A.y_final[I.y_off + lane] = y1;
        if (lane == 0) {
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
            o[10] = 0;
            o[11] = 0;
            o[12] = 0;
            o[13] = 0;
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

hipError_t launch_wave(const BatchArgs &args, int blocks, size_t lds_bytes, hipStream_t stream)
{
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&wave_kernel),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(wave_kernel, dim3(blocks), dim3(WL), lds_bytes, stream, args);
    return hipGetLastError();
}

int wave_occupancy(size_t lds_bytes, int *per_cu)
{
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, reinterpret_cast<const void *>(&wave_kernel), WL,
                                                        lds_bytes) == hipSuccess
               ? 0
               : -1;
}

}  // namespace lassie
