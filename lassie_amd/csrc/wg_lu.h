// Workgroup-local dense LU of the batched and differentiation forms: one
// workgroup factors one instance's row-major Newton matrix (S <= 512).
#pragma once
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

template <typename T, bool BF16, int RT = 4>
__device__ int wg_lu_rm(T *__restrict__ M, int n, int *piv, int *perm_src, int *moved_to, int *rows, int *rowcnt,
                        int *info, Shared &sh)
{
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
        // row moves in every other column; U block row by forward substitution
        for (int c = tid; c < n; c += NT) {
            if (c >= r0 && c < r0 + nbp) continue;
            T vals[NBC], old[NBC];
#pragma unroll
            for (int kk = 0; kk < NBC; kk++) {
                vals[kk] = (kk < nbp) ? M[(size_t)perm_src[r0 + kk] * n + c] : T(0);
                old[kk] = (kk < nbp) ? M[(size_t)(r0 + kk) * n + c] : T(0);
            }
#pragma unroll
            for (int kk = 0; kk < NBC; kk++) {
                if (kk < nbp) {
                    const int d = moved_to[r0 + kk];
                    if (d >= r0 + nbp) M[(size_t)d * n + c] = old[kk];
                }
            }
            if (c >= r0 + nbp) {
#pragma unroll
                for (int i = 1; i < NBC; i++) {
                    if (i < nbp) {
#pragma unroll
                        for (int k = 0; k < i; k++) vals[i] = fma(-(T)sh.L11[i][k], vals[k], vals[i]);
                    }
                }
            }
#pragma unroll
            for (int kk = 0; kk < NBC; kk++)
                if (kk < nbp) M[(size_t)(r0 + kk) * n + c] = vals[kk];
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
            for (int e = tid; e < (n - c0) * NBC; e += NT) {
                const int rr = e / NBC, k = e % NBC;
                T v = (k < nbp) ? -M[(size_t)(c0 + rr) * n + r0 + k] : T(-0.0);
                if (BF16) v = (T)to_bf16_rne((float)v);
                Lp[rr * kLpStride + k] = v;
            }
            __syncthreads();
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
        __syncthreads();
    }
    return 0;
}

}  // namespace lassie
