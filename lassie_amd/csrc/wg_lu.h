// Workgroup-local dense LU of the batched and differentiation forms: one
// workgroup factors one instance's row-major Newton matrix (S <= 512).
#pragma once
#include <type_traits>
#include "batch.h"
#include "device_common.h"
#include "wg_common.h"

namespace lassie {

__device__ __forceinline__ float to_bf16_rne(float x)
{
    // round to the nearest BF16 value (ties to even), kept in an FP32 register
    unsigned int u = __float_as_uint(x);
    if ((u & 0x7f800000u) == 0x7f800000u) return x;  // inf / nan unchanged
    u += 0x7fffu + ((u >> 16) & 1u);
    return __uint_as_float(u & 0xffff0000u);
}

// Workgroup-local blocked right-looking LU of a row-major n x n matrix with
// partial pivoting (panel_factor semantics).  M must be followed by kLuSink
// writable elements.  RT: row tiles of the trailing update in flight per wave.  Row moves go to every column,
// so on return M holds L\U in final row order.  Returns 0 or k+1 of the
// first zero pivot.  BF16 = round the trailing-update operands to BF16.
constexpr int kLpStride = 17;  // LDS row stride of the staged L21 panel (elements)

// L21 of the current panel (at most kBatchMaxS - kBatchNB rows), shared by
// every precision's instantiation of wg_lu_rm in a kernel
static __shared__ double lp_raw[(kBatchMaxS - kBatchNB) * kLpStride];

// Trailing update A22 -= L21 U12 of an FP64 factorization with n % 16 == 0
// (every tile full, so no lane checks): the same four 16x16x4 MFMAs per
// output tile as the general form (k ascending, the FP64 instruction's
// sequential fma chain), with the per-lane address parts fixed for the whole
// update and the per-tile parts in scalar offsets.  A wave walks its tiles
// (column tiles ct = wave (mod 4), every row tile of each) through a ring of
// D sets, each holding one tile's output block and U12 operands in flight;
// the L21 operands come from the LDS panel (Lp, kLpStride per row).
template <int D, typename LT>
__device__ void trailing_update_full(double *M, int n, int r0, int c0, const LT *Lp)
{
    M = uniform_ptr(M);
    n = uniform_int(n);
    r0 = uniform_int(r0);
    c0 = uniform_int(c0);
    typedef typename Mfma<double>::acc Acc;
    const __amdgpu_buffer_rsrc_t rs = buf_rsrc(M, n * n * 8);
    const int tid = threadIdx.x, lane = tid & 63, l16 = lane & 15, lq = lane >> 4;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int nt = (n - c0) >> 4;
    const int W = wave < nt ? ((nt - wave + NWAVES - 1) / NWAVES) * nt : 0;
    // output element (c0 + lq + 4i + 16g, c0 + l16 + 16ct) at vc + 32n i + 128n g + 128 ct bytes
    const int vc = ((c0 + lq) * n + c0 + l16) * 8;
    // U12 element (r0 + lq + 4kc, c0 + l16 + 16ct) at vb + 32n kc + 128 ct
    const int vb = ((r0 + lq) * n + c0 + l16) * 8;
    // L21 element (16g + l16, lq + 4kc) of the LDS panel
    const LT *la = Lp + l16 * kLpStride + lq;
    struct Set {
        Acc d;
        double b[4];
    };
    auto issue = [&](Set &st, int g, int ct) {
#pragma unroll
        for (int kc = 0; kc < 4; kc++) st.b[kc] = BufIO<double>::ld(rs, vb, 128 * ct + 32 * n * kc);
#pragma unroll
        for (int i = 0; i < 4; i++) st.d[i] = BufIO<double>::ld(rs, vc, 128 * n * g + 128 * ct + 32 * n * i);
    };
    auto finish = [&](Set &st, int g, int ct) {
        const LT *ap = la + 16 * kLpStride * g;
        double a[4];
#pragma unroll
        for (int kc = 0; kc < 4; kc++) a[kc] = (double)ap[4 * kc];
        Acc d = st.d;
#pragma unroll
        for (int kc = 0; kc < 4; kc++) d = Mfma<double>::run(a[kc], st.b[kc], d);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const double v = d[i];
            BufIO<double>::st(rs, vc, v, 128 * n * g + 128 * ct + 32 * n * i);
        }
    };
    // tile w of the wave: column tile ct = wave + 4 (w / nt), row tile g = w % nt
    int gi = 0, ci = wave, gf = 0, cf = wave;
    auto next = [&](int &g, int &ct) {
        if (++g == nt) {
            g = 0;
            ct += NWAVES;
        }
    };
    // the scheduling barriers keep each set's memory operations in program
    // order, so the wait for one set leaves the later sets in flight
    Set S[D];
#pragma unroll
    for (int j = 0; j < D; j++) {
        if (j < W) {
            issue(S[j], gi, ci);
            next(gi, ci);
        }
        __builtin_amdgcn_sched_barrier(0);
    }
    for (int w = 0; w < W; w += D) {
#pragma unroll
        for (int j = 0; j < D; j++) {
            if (w + j < W) {
                finish(S[j], gf, cf);
                next(gf, cf);
            }
            __builtin_amdgcn_sched_barrier(0);
            if (w + j + D < W) {
                issue(S[j], gi, ci);
                next(gi, ci);
            }
            __builtin_amdgcn_sched_barrier(0);
        }
    }
}

template <typename T, bool BF16, int RT = 4>
__device__ int wg_lu_rm(T *__restrict__ M, int n, int *piv, int *perm_src, int *moved_to, int *rows, int *rowcnt,
                        int *info, Shared &sh)
{
    // uniform across the workgroup: scalar registers, exact buffer accesses
    M = uniform_ptr(M);
    n = uniform_int(n);
    constexpr int NBC = kBatchNB;
    constexpr int RPT = kBatchMaxS / NT;
    const int P = (n + NBC - 1) / NBC;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    if (tid == 0) st_rlx(info, 0);
    __syncthreads();
    for (int p = 0; p < P; p++) {
        const int r0 = p * NBC;
        const int nbp = min(NBC, n - r0);
        // the fewest rows per thread that cover the panel's rows [r0, n): each
        // row stays with the same thread, so the results do not change
        if (RPT > 1 && n - r0 <= NT)
            panel_factor<1, NBC, T, true, false>(M, n, p, piv, perm_src, moved_to, rows + (size_t)p * n, rowcnt + p,
                                                 info, sh);
        else
            panel_factor<RPT, NBC, T, true, false>(M, n, p, piv, perm_src, moved_to, rows + (size_t)p * n,
                                                   rowcnt + p, info, sh);
        __syncthreads();
        const int inf = ld_rlx(info);
        if (inf) return inf;
        for (int e = tid; e < NBC * NBC; e += NT) {
            const int kk = e / NBC, c = e % NBC;
            if (kk < nbp && c < kk) sh.L11[kk][c] = (double)M[(size_t)(r0 + kk) * n + r0 + c];
        }
        __syncthreads();
        // row moves in every other column; U block row by forward substitution.
        // A thread takes columns tid and tid + NT together and issues every
        // load of both before the first use (one memory round trip); masked
        // lanes use out-of-range buffer offsets instead of branches.
        {
            const __amdgpu_buffer_rsrc_t rs = buf_rsrc(M, n * n * (int)sizeof(T));
            int src[NBC], dst[NBC];
#pragma unroll
            for (int kk = 0; kk < NBC; kk++) {
                src[kk] = (kk < nbp) ? perm_src[r0 + kk] : 0;
                dst[kk] = (kk < nbp) ? moved_to[r0 + kk] : -1;
            }
            T vals[2][NBC], old[2][NBC];
#pragma unroll
            for (int j = 0; j < 2; j++) {
                const int c = tid + j * NT;
                const bool okc = c < n && (c < r0 || c >= r0 + nbp);
#pragma unroll
                for (int kk = 0; kk < NBC; kk++) {
                    const bool ok = okc && kk < nbp;
                    vals[j][kk] = BufIO<T>::ld(rs, ok ? (src[kk] * n + c) * (int)sizeof(T) : kBufOob);
                    old[j][kk] = BufIO<T>::ld(rs, ok ? ((r0 + kk) * n + c) * (int)sizeof(T) : kBufOob);
                }
            }
#pragma unroll
            for (int j = 0; j < 2; j++) {
                const int c = tid + j * NT;
                const bool okc = c < n && (c < r0 || c >= r0 + nbp);
#pragma unroll
                for (int kk = 0; kk < NBC; kk++) {
                    const bool ok = okc && kk < nbp && dst[kk] >= r0 + nbp;
                    BufIO<T>::st(rs, ok ? (dst[kk] * n + c) * (int)sizeof(T) : kBufOob, old[j][kk]);
                }
                if (okc && c >= r0 + nbp) {
#pragma unroll
                    for (int i = 1; i < NBC; i++) {
                        if (i < nbp) {
#pragma unroll
                            for (int k = 0; k < i; k++) vals[j][i] = fma(-(T)sh.L11[i][k], vals[j][k], vals[j][i]);
                        }
                    }
                }
#pragma unroll
                for (int kk = 0; kk < NBC; kk++) {
                    const bool ok = okc && kk < nbp;
                    BufIO<T>::st(rs, ok ? ((r0 + kk) * n + c) * (int)sizeof(T) : kBufOob, vals[j][kk]);
                }
            }
        }
        __syncthreads();
        // trailing update A22 -= L21 U12 on the matrix cores: 16 x 16 output
        // tiles, four 16x16x4 MFMAs each (k ascending; the FP64 instruction
        // accumulates as a sequential fma chain, the oracle's order).  Every
        // row is updated, as in the oracle.  A wave owns the column tiles
        // ct = wave (mod 4), keeps their U12 operands in registers and walks
        // the row tiles RT at a time, the next RT tiles' output tiles in
        // flight.  The L21 operands, which every column tile reuses, are
        // staged once in LDS (negated, and rounded for BF16), kLpStride
        // elements per row so that the 16 rows of a tile fall in different
        // banks.
        const int c0 = r0 + nbp;
        if (c0 < n) {
            typedef __attribute__((address_space(3))) T LT;
            LT *Lp = (LT *)lp_raw;
            {
                // every element's load in flight at once, then the LDS stores
                constexpr int kStage = ((kBatchMaxS - NBC) * NBC + NT - 1) / NT;
                const __amdgpu_buffer_rsrc_t rs = buf_rsrc(M, n * n * (int)sizeof(T));
                const int cnt = (n - c0) * NBC;
                T v[kStage];
#pragma unroll
                for (int j = 0; j < kStage; j++) {
                    const int e = tid + j * NT, rr = e / NBC, k = e % NBC;
                    const bool ok = e < cnt && k < nbp;
                    v[j] = BufIO<T>::ld(rs, ok ? ((c0 + rr) * n + r0 + k) * (int)sizeof(T) : kBufOob);
                }
#pragma unroll
                for (int j = 0; j < kStage; j++) {
                    const int e = tid + j * NT, rr = e / NBC, k = e % NBC;
                    if (e < cnt) {
                        T x = (k < nbp) ? -v[j] : T(-0.0);
                        if (BF16) x = (T)to_bf16_rne((float)x);
                        Lp[rr * kLpStride + k] = x;
                    }
                }
            }
            __syncthreads();
            if constexpr (std::is_same<T, double>::value && !BF16) {
                if ((n & 15) == 0) {
                    trailing_update_full<2 * RT>(M, n, r0, c0, Lp);
                    goto update_done;
                }
            }
            typedef typename Mfma<T>::acc Acc;
            // global-memory view of M (the workspace lives in device memory):
            // global loads/stores keep the memory counters precise, so the
            // next tiles' loads stay in flight across the current MFMAs
            typedef __attribute__((address_space(1))) T GT;
            GT *Mg = (GT *)M;
            const size_t last = (size_t)n * n - 1;
            const size_t sink = (size_t)n * n;  // discarded stores (kLuSink elements after M)
            const int nt = (n - c0 + 15) >> 4;
            const int l16 = lane & 15, lq = lane >> 4;
            for (int ct = wave; ct < nt; ct += NWAVES) {
                const int col = c0 + 16 * ct + l16;
                const bool okc = col < n;
                T b[4];
#pragma unroll
                for (int kc = 0; kc < 4; kc++) {
                    const int k = 4 * kc + lq;
                    const bool ok = okc && k < nbp;
                    T bv = Mg[ok ? (size_t)(r0 + k) * n + col : 0];
                    bv = ok ? bv : T(0);
                    if (BF16) bv = (T)to_bf16_rne((float)bv);
                    b[kc] = bv;
                }
                // branch-free loads: out-of-range lanes read a valid address and discard it
                auto load_group = [&](int g, Acc *dd, T (*aa)[4]) {
#pragma unroll
                    for (int u = 0; u < RT; u++) {
                        const int R = c0 + 16 * (g + u);
                        const bool okt = g + u < nt;
#pragma unroll
                        for (int i = 0; i < 4; i++) {
                            const int r = R + Mfma<T>::row(lq, i);
                            const bool ok = okt && r < n && okc;
                            const T v = Mg[ok ? (size_t)r * n + col : last];
                            dd[u][i] = ok ? v : T(0);
                        }
                        const int ra = R + l16;
#pragma unroll
                        for (int kc = 0; kc < 4; kc++) {
                            const int k = 4 * kc + lq;
                            const bool ok = okt && ra < n && k < nbp;
                            // padding products are (-0) * 0 = -0, which leaves every value unchanged
                            aa[u][kc] = ok ? (T)Lp[(ra - c0) * kLpStride + k] : T(-0.0);
                        }
                    }
                };
                Acc d[RT], dn[RT];
                T a[RT][4], an[RT][4];
                load_group(0, d, a);
                for (int g = 0; g < nt; g += RT) {
                    const bool more = g + RT < nt;
                    load_group(g + RT, dn, an);  // unconditional: no branch, exact memory counters
#pragma unroll
                    for (int u = 0; u < RT; u++) {
#pragma unroll
                        for (int kc = 0; kc < 4; kc++) d[u] = Mfma<T>::run(a[u][kc], b[kc], d[u]);
                    }
#pragma unroll
                    for (int u = 0; u < RT; u++) {
                        const int R = c0 + 16 * (g + u);
#pragma unroll
                        for (int i = 0; i < 4; i++) {
                            const int r = R + Mfma<T>::row(lq, i);
                            const bool ok = g + u < nt && r < n && okc;
                            Mg[ok ? (size_t)r * n + col : sink] = d[u][i];
                        }
                    }
                    if (more) {
#pragma unroll
                        for (int u = 0; u < RT; u++) {
                            d[u] = dn[u];
#pragma unroll
                            for (int kc = 0; kc < 4; kc++) a[u][kc] = an[u][kc];
                        }
                    }
                }
            }
        }
    update_done:
        __syncthreads();
    }
    return 0;
}

}  // namespace lassie
