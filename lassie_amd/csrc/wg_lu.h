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
// partial pivoting (panel_factor semantics).  Row moves go to every column,
// so on return M holds L\U in final row order.  Returns 0 or k+1 of the
// first zero pivot.  BF16 = round the trailing-update operands to BF16.
template <typename T, bool BF16>
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
        panel_factor<RPT, NBC, T, true>(M, n, p, piv, perm_src, moved_to, rows + (size_t)p * n, rowcnt + p, info,
                                        sh);
        __syncthreads();
        const int inf = ld_rlx(info);
        if (inf) return inf;
        for (int e = tid; e < NBC * NBC; e += NT) {
            const int kk = e / NBC, c = e % NBC;
            if (kk < nbp && c < kk) sh.L11[kk][c] = (double)M[(size_t)(r0 + kk) * n + r0 + c];
        }
        __syncthreads();
        // row moves in every other column; U block row by forward substitution
        int nonfinite = 0;
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
#pragma unroll
                for (int kk = 0; kk < NBC; kk++) nonfinite |= (kk < nbp) && !isfinite((double)vals[kk]);
            }
#pragma unroll
            for (int kk = 0; kk < NBC; kk++)
                if (kk < nbp) M[(size_t)(r0 + kk) * n + c] = vals[kk];
        }
        const bool all_rows = __syncthreads_or(nonfinite) != 0;
        // trailing update over rows with nonzero multipliers: lanes take 64
        // consecutive columns, waves take rows, two rows in flight per wave
        const int c0 = r0 + nbp;
        if (c0 < n) {
            const int cnt = all_rows ? n - c0 : rowcnt[p];
            const int *rl = rows + (size_t)p * n;
            for (int cb = c0; cb < n; cb += 64) {
                const int c = cb + lane;
                const bool okc = c < n;
                T u[NBC];
#pragma unroll
                for (int kk = 0; kk < NBC; kk++) {
                    T uv = (okc && kk < nbp) ? M[(size_t)(r0 + kk) * n + c] : T(0);
                    if (BF16) uv = (T)to_bf16_rne((float)uv);
                    u[kk] = uv;
                }
                for (int idx = wave; idx < cnt; idx += 2 * NWAVES) {
                    const int idx1 = idx + NWAVES;
                    const bool h1 = idx1 < cnt;
                    const int i0 = all_rows ? c0 + idx : rl[idx];
                    const int i1 = h1 ? (all_rows ? c0 + idx1 : rl[idx1]) : i0;
                    T l0[NBC], l1[NBC];
#pragma unroll
                    for (int kk = 0; kk < NBC; kk++) {
                        T a = (kk < nbp) ? M[(size_t)i0 * n + r0 + kk] : T(0);
                        T b = (kk < nbp && h1) ? M[(size_t)i1 * n + r0 + kk] : T(0);
                        if (BF16) {
                            a = (T)to_bf16_rne((float)a);
                            b = (T)to_bf16_rne((float)b);
                        }
                        l0[kk] = a;
                        l1[kk] = b;
                    }
                    T a0 = okc ? M[(size_t)i0 * n + c] : T(0);
                    T a1 = (okc && h1) ? M[(size_t)i1 * n + c] : T(0);
#pragma unroll
                    for (int kk = 0; kk < NBC; kk++) {
                        if (kk < nbp) {
                            a0 = fma(-l0[kk], u[kk], a0);
                            a1 = fma(-l1[kk], u[kk], a1);
                        }
                    }
                    if (okc) M[(size_t)i0 * n + c] = a0;
                    if (okc && h1) M[(size_t)i1 * n + c] = a1;
                }
            }
        }
        __syncthreads();
    }
    return 0;
}

}  // namespace lassie
