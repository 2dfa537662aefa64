// Workgroup-local dense LU of the batched and differentiation forms: one
// workgroup factors one instance's row-major Newton matrix (S <= 512).
#pragma once
#include <type_traits>
#include "batch.h"
#include "device_common.h"
#include "wg_common.h"

namespace lassie {

// Workgroup-local blocked right-looking LU of a row-major n x n matrix with
// partial pivoting (panel_factor semantics).  M must be followed by kLuSink
// writable elements.  RT: row tiles of the trailing update in flight per wave.  Row moves go to every column,
// so on return M holds L\U in final row order.  Returns 0 or k+1 of the
// first zero pivot.  BF16: the trailing update on the BF16 matrix cores
// (operands rounded to BF16, FP32 accumulation).
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

// ---------------------------------------------------------------------------
// Two-level blocked form of the FP64 factorization (n % 16 == 0): blocks of
// kLuBlock columns, each factored as four 16-column panels whose row moves,
// U blocks and trailing updates stay inside the block; then the block's row
// moves are applied once to the other columns, its U12 rows are solved, and
// the trailing matrix takes one rank-64 update (sixteen 16x16x4 MFMAs per
// output tile).  Every element still receives its updates as one fma chain
// fma(-l_ik, u_kj, a_ij) over k ascending, the pivots are chosen from the
// same values, and row moves are data movement only, so the factors are
// bit-identical to the 16-column form; the trailing matrix is read and
// written once per 64 columns instead of once per 16.
constexpr int kLuBlock = 64;
constexpr int kLuBlockedMin = 256;  // the blocked form factors n > kLuBlockedMin

// composed row order of the current block (position -> source position at
// the block's start) and the positions it changes (at most 2 kLuBlock)
static __shared__ int lu_pfb[kBatchMaxS];
static __shared__ int lu_chg[2 * kLuBlock];
static __shared__ int lu_nchg;

// -M[r][k0 + k] for rows rs <= r < re (re - rs <= MAXR), k < 16, into
// Lp[(r - rs) * kLpStride + k]: every load in flight at once, then the LDS
// stores.
template <int MAXR, typename LT>
__device__ __forceinline__ void stage_l16(const double *M, int n, int k0, int rs, int re, LT *Lp)
{
    constexpr int kPer = (MAXR * kBatchNB + NT - 1) / NT;
    const __amdgpu_buffer_rsrc_t rsc = buf_rsrc(M, n * n * 8);
    const int tid = threadIdx.x;
    const int cnt = (re - rs) * kBatchNB;
    double v[kPer];
#pragma unroll
    for (int j = 0; j < kPer; j++) {
        const int e = tid + j * NT, rr = e >> 4, k = e & 15;
        v[j] = BufIO<double>::ld(rsc, e < cnt ? ((rs + rr) * n + k0 + k) * 8 : kBufOob);
    }
#pragma unroll
    for (int j = 0; j < kPer; j++) {
        const int e = tid + j * NT, rr = e >> 4, k = e & 15;
        if (e < cnt) Lp[rr * kLpStride + k] = -v[j];
    }
}

// Row moves of the panel at r0 (16 rows) and the U-block forward
// substitution, in the columns c_lo <= c < c_hi outside the panel (at most
// 2 NT of them; thread t takes c_lo + t and c_lo + t + NT): with MOVES the
// panel's row moves in every such column, and in the columns right of the
// panel the substitution with the panel's L11 (in sh.L11), k ascending.
// Masked lanes use out-of-range buffer offsets instead of branches.
template <bool MOVES>
__device__ void panel_rows_update(double *M, int n, int r0, int c_lo, int c_hi, const int *perm_src,
                                  const int *moved_to, Shared &sh)
{
    constexpr int NBC = kBatchNB;
    const int tid = threadIdx.x;
    const __amdgpu_buffer_rsrc_t rs = buf_rsrc(M, n * n * 8);
    int src[NBC], dst[NBC];
#pragma unroll
    for (int kk = 0; kk < NBC; kk++) {
        src[kk] = MOVES ? perm_src[r0 + kk] : r0 + kk;
        dst[kk] = MOVES ? moved_to[r0 + kk] : -1;
    }
    double vals[2][NBC], old[2][NBC];
#pragma unroll
    for (int j = 0; j < 2; j++) {
        const int c = c_lo + tid + j * NT;
        const bool okc = c < c_hi && (c < r0 || c >= r0 + NBC);
#pragma unroll
        for (int kk = 0; kk < NBC; kk++) {
            vals[j][kk] = BufIO<double>::ld(rs, okc ? (src[kk] * n + c) * 8 : kBufOob);
            if (MOVES) old[j][kk] = BufIO<double>::ld(rs, okc ? ((r0 + kk) * n + c) * 8 : kBufOob);
        }
    }
#pragma unroll
    for (int j = 0; j < 2; j++) {
        const int c = c_lo + tid + j * NT;
        const bool okc = c < c_hi && (c < r0 || c >= r0 + NBC);
        if (MOVES) {
#pragma unroll
            for (int kk = 0; kk < NBC; kk++) {
                const bool ok = okc && dst[kk] >= r0 + NBC;
                BufIO<double>::st(rs, ok ? (dst[kk] * n + c) * 8 : kBufOob, old[j][kk]);
            }
        }
        if (okc && c >= r0 + NBC) {
#pragma unroll
            for (int i = 1; i < NBC; i++) {
#pragma unroll
                for (int k = 0; k < i; k++) vals[j][i] = fma(-sh.L11[i][k], vals[j][k], vals[j][i]);
            }
        }
#pragma unroll
        for (int kk = 0; kk < NBC; kk++) BufIO<double>::st(rs, okc ? ((r0 + kk) * n + c) * 8 : kBufOob, vals[j][kk]);
    }
}

// A -= L U over the rectangle of output tiles rows [rb, rb + 16 nrt) x
// columns [cb, cb + 16 nct): U rows [r0, r0 + 16) of the same columns, the
// negated L columns staged in Lp (row r at (r - rb) * kLpStride), four
// 16x16x4 MFMAs per tile, k ascending.  The tiles are dealt to the waves in
// turn; each wave keeps a ring of D tiles' loads in flight.
template <int D, typename LT>
__device__ void tile_update16(double *M, int n, int r0, int rb, int nrt, int cb, int nct, const LT *Lp)
{
    M = uniform_ptr(M);
    n = uniform_int(n);
    r0 = uniform_int(r0);
    rb = uniform_int(rb);
    nrt = uniform_int(nrt);
    cb = uniform_int(cb);
    nct = uniform_int(nct);
    typedef Mfma<double>::acc Acc;
    const __amdgpu_buffer_rsrc_t rs = buf_rsrc(M, n * n * 8);
    const int tid = threadIdx.x, lane = tid & 63, l16 = lane & 15, lq = lane >> 4;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int total = nrt * nct;
    const int W = wave < total ? (total - wave + NWAVES - 1) / NWAVES : 0;
    // output element (rb + 16g + lq + 4i, cb + 16ct + l16) at vc + 128n g + 128 ct + 32n i
    const int vc = ((rb + lq) * n + cb + l16) * 8;
    // U element (r0 + lq + 4kc, cb + 16ct + l16) at vb + 32n kc + 128 ct
    const int vb = ((r0 + lq) * n + cb + l16) * 8;
    const LT *la = Lp + l16 * kLpStride + lq;
    struct Set {
        Acc d;
        double b[4];
    };
    auto tile = [&](int w, int &g, int &ct) {
        const int idx = wave + NWAVES * w;
        g = idx / nct;
        ct = idx - g * nct;
    };
    auto issue = [&](Set &st, int w) {
        int g, ct;
        tile(w, g, ct);
#pragma unroll
        for (int kc = 0; kc < 4; kc++) st.b[kc] = BufIO<double>::ld(rs, vb, 128 * ct + 32 * n * kc);
#pragma unroll
        for (int i = 0; i < 4; i++) st.d[i] = BufIO<double>::ld(rs, vc, 128 * n * g + 128 * ct + 32 * n * i);
    };
    auto finish = [&](Set &st, int w) {
        int g, ct;
        tile(w, g, ct);
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
    Set S[D];
#pragma unroll
    for (int j = 0; j < D; j++) {
        if (j < W) issue(S[j], j);
        __builtin_amdgcn_sched_barrier(0);
    }
    for (int w = 0; w < W; w += D) {
#pragma unroll
        for (int j = 0; j < D; j++) {
            if (w + j < W) finish(S[j], w + j);
            __builtin_amdgcn_sched_barrier(0);
            if (w + j + D < W) issue(S[j], w + j + D);
            __builtin_amdgcn_sched_barrier(0);
        }
    }
}

// Rank-64 trailing update of a block: A22 -= L21 U12 for rows and columns
// [E, n), k = R0 .. R0 + 63 ascending (sixteen 16x16x4 MFMAs per output
// tile, one fma chain per element).  The negated L21 goes through LDS (lds,
// up to 128 rows x 64 at a time) in the MFMA's lane order, so each operand
// read is one contiguous 512-byte row.  Each wave owns the column tiles
// ct = wave (mod 4) and walks them in pairs: the U12 operands of the pair
// sit in two register banks, and every row tile of the chunk has one slot
// whose output tile is in flight, filled for the next column tile as soon
// as the current one is stored.  Slots, banks and loads are all fixed at
// compile time (masked lanes and tiles past the end use out-of-range buffer
// offsets), so the memory counters stay exact and only the tile being
// finished is waited for.
__device__ void tile_update64(double *M, int n, int R0, int E, double *lds)
{
    M = uniform_ptr(M);
    n = uniform_int(n);
    R0 = uniform_int(R0);
    E = uniform_int(E);
    typedef Mfma<double>::acc Acc;
    typedef __attribute__((address_space(3))) double LT;
    LT *La = (LT *)lds;
    const __amdgpu_buffer_rsrc_t rs = buf_rsrc(M, n * n * 8);
    const int tid = threadIdx.x, lane = tid & 63, l16 = lane & 15, lq = lane >> 4;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int nt = (n - E) >> 4;
    const int ncw = wave < nt ? (nt - wave + NWAVES - 1) / NWAVES : 0;  // column tiles of this wave
    // output element (E + 16g + lq + 4i, E + 16ct + l16) at vd + 128n g + 128 ct + 32n i
    const int vd = ((E + lq) * n + E + l16) * 8;
    // U12 element (R0 + lq + 4kc, E + 16ct + l16) at vb + 32n kc + 128 ct
    const int vb = ((R0 + lq) * n + E + l16) * 8;
    constexpr int kChunk = 8;  // row tiles staged at a time (LDS: kChunk * 16 * 64 doubles)
    const int nchunk = (nt + kChunk - 1) / kChunk;
    const int per = nchunk ? (nt + nchunk - 1) / nchunk : 0;
    for (int rc = 0; rc < nt; rc += per) {
        const int ng = min(per, nt - rc);
        __syncthreads();  // the previous chunk's operands are no longer read
        {
            // -L21 of row tile g, MFMA step kc, lane l at (16 g + kc) * 64 + l
            constexpr int kPer = kChunk * 16 * kLuBlock / NT;
            const int cnt = ng * 16 * kLuBlock;
            double v[kPer];
#pragma unroll
            for (int j = 0; j < kPer; j++) {
                const int e = tid + j * NT, rr = e >> 6, k = e & 63;
                v[j] = BufIO<double>::ld(rs, e < cnt ? ((E + 16 * rc + rr) * n + R0 + k) * 8 : kBufOob);
            }
#pragma unroll
            for (int j = 0; j < kPer; j++) {
                const int e = tid + j * NT, rr = e >> 6, k = e & 63;
                if (e < cnt) La[((rr >> 4) * 16 + (k >> 2)) * 64 + (rr & 15) + 16 * (k & 3)] = -v[j];
            }
        }
        __syncthreads();
        if (ncw == 0) continue;
        auto loadb = [&](double(&b)[16], int jc) {
            const int off = jc < ncw ? vb : kBufOob;
            const int ct = wave + NWAVES * jc;
#pragma unroll
            for (int kc = 0; kc < 16; kc++) b[kc] = BufIO<double>::ld(rs, off, 128 * ct + 32 * n * kc);
        };
        auto issue = [&](Acc &d, int jc, int g) {
            const int off = (jc < ncw && g < ng) ? vd : kBufOob;
            const int ct = wave + NWAVES * jc;
#pragma unroll
            for (int i = 0; i < 4; i++) d[i] = BufIO<double>::ld(rs, off, 128 * n * (rc + g) + 128 * ct + 32 * n * i);
        };
        auto finish = [&](const Acc &dset, int jc, int g, const double(&b)[16]) {
            const bool live = jc < ncw && g < ng;
            const int off = live ? vd : kBufOob;
            const int ct = wave + NWAVES * jc;
            Acc d = dset;
            if (live) {
                const LT *ap = La + g * 16 * 64 + lane;
#pragma unroll
                for (int kc = 0; kc < 16; kc++) d = Mfma<double>::run(ap[kc * 64], b[kc], d);
            }
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const double v = d[i];
                BufIO<double>::st(rs, off, v, 128 * n * (rc + g) + 128 * ct + 32 * n * i);
            }
        };
        double b0[16], b1[16];
        Acc S[kChunk];
        loadb(b0, 0);
        loadb(b1, 1);
#pragma unroll
        for (int g = 0; g < kChunk; g++) issue(S[g], 0, g);
        __builtin_amdgcn_sched_barrier(0);
        for (int jc = 0; jc < ncw; jc += 2) {
#pragma unroll
            for (int g = 0; g < kChunk; g++) {
                finish(S[g], jc, g, b0);
                __builtin_amdgcn_sched_barrier(0);
                issue(S[g], jc + 1, g);
                __builtin_amdgcn_sched_barrier(0);
            }
            loadb(b0, jc + 2);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int g = 0; g < kChunk; g++) {
                finish(S[g], jc + 1, g, b1);
                __builtin_amdgcn_sched_barrier(0);
                issue(S[g], jc + 2, g);
                __builtin_amdgcn_sched_barrier(0);
            }
            loadb(b1, jc + 3);
            __builtin_amdgcn_sched_barrier(0);
        }
    }
}

// One step of block_u12_sweep: row group G (rows 16G .. 16G + 15 of the
// block) takes the update of k = K, fma(-L[i][K], u_K, v_i) for every row
// i > K of the group, with the L values in cur; the next step's L values go
// into nxt first, so their LDS reads are in flight during these fmas.  The
// steps run G ascending and, within a group, K ascending: every element
// takes its updates in k order, and u_K is final when it is used (rows
// below 16G are done, and row K of the group has had every k < K).
template <int G, int K>
struct U12Step {
    typedef __attribute__((address_space(3))) double LT;
    static constexpr int kRows = 16;
    static constexpr int kLast = kRows * G + kRows - 1;  // the group's last k
    static __device__ __forceinline__ void load(double (&l)[kRows], const LT *Lc)
    {
#pragma unroll
        for (int q = 0; q < kRows; q++)
            if (kRows * G + q > K) l[q] = Lc[K * kLuBlock + kRows * G + q];
    }
    static __device__ __forceinline__ void run(double (&v)[kLuBlock], const LT *Lc, double (&cur)[kRows],
                                               double (&nxt)[kRows])
    {
        constexpr int Gn = K < kLast ? G : G + 1;
        constexpr int Kn = K < kLast ? K + 1 : 0;
        if constexpr (Gn < kLuBlock / kRows) U12Step<Gn, Kn>::load(nxt, Lc);
        __builtin_amdgcn_sched_barrier(0);
        const double u = v[K];
#pragma unroll
        for (int q = 0; q < kRows; q++)
            if (kRows * G + q > K) v[kRows * G + q] = fma(-cur[q], u, v[kRows * G + q]);
        __builtin_amdgcn_sched_barrier(0);
        if constexpr (Gn < kLuBlock / kRows) U12Step<Gn, Kn>::run(v, Lc, nxt, cur);
    }
};

// U12 of a 64-row block (rows [R0, R0 + 64)) in the columns [E, n): each
// thread takes whole columns, holds the column's 64 values in registers and
// runs the forward substitution with the block's unit lower L11 (in LDS,
// column-major: Lc[k * 64 + i] = L[R0 + i][R0 + k]), k ascending for every
// element, the order the 16-row steps (substitution in the diagonal 16 x 16
// block, MFMA update of the block's later rows) give.  No barrier: every
// column belongs to one thread.
__device__ __forceinline__ void block_u12_sweep(double *M, int n, int R0, int E, const double *lds)
{
    M = uniform_ptr(M);
    n = uniform_int(n);
    R0 = uniform_int(R0);
    E = uniform_int(E);
    typedef __attribute__((address_space(3))) double LT;
    const LT *Lc = (const LT *)lds;
    const __amdgpu_buffer_rsrc_t rs = buf_rsrc(M, n * n * 8);
    for (int c = E + (int)threadIdx.x; c < n; c += NT) {
        double v[kLuBlock];
#pragma unroll
        for (int i = 0; i < kLuBlock; i++) v[i] = BufIO<double>::ld(rs, ((R0 + i) * n + c) * 8);
        double la[16], lb[16];
        U12Step<0, 0>::load(la, Lc);
        U12Step<0, 0>::run(v, Lc, la, lb);
#pragma unroll
        for (int i = 0; i < kLuBlock; i++) BufIO<double>::st(rs, ((R0 + i) * n + c) * 8, v[i]);
    }
}

template <int RT>
__device__ int wg_lu_rm_blocked(double *__restrict__ M, int n, int *piv, int *perm_src, int *moved_to, int *rows,
                                int *rowcnt, int *info, Shared &sh)
{
    M = uniform_ptr(M);
    n = uniform_int(n);
    constexpr int NBC = kBatchNB;
    const int tid = threadIdx.x;
    typedef __attribute__((address_space(3))) double LT;
    LT *Lp = (LT *)lp_raw;
    if (tid == 0) st_rlx(info, 0);
    __syncthreads();
    for (int R0 = 0; R0 < n; R0 += kLuBlock) {
        const int E = min(n, R0 + kLuBlock);
        // the block's columns: its panels with row moves, U blocks and
        // trailing updates confined to columns [R0, E)
        for (int r0 = R0; r0 < E; r0 += NBC) {
            const int p = r0 / NBC;
            if (n - r0 <= NT)
                panel_factor<1, NBC, double, true, false>(M, n, p, piv, perm_src, moved_to, rows + (size_t)p * n,
                                                          rowcnt + p, info, sh);
            else
                panel_factor<kBatchMaxS / NT, NBC, double, true, false>(M, n, p, piv, perm_src, moved_to,
                                                                        rows + (size_t)p * n, rowcnt + p, info, sh);
            __syncthreads();
            const int inf = ld_rlx(info);
            if (inf) return inf;
            for (int e = tid; e < NBC * NBC; e += NT) {
                const int kk = e / NBC, c = e % NBC;
                if (c < kk) sh.L11[kk][c] = M[(size_t)(r0 + kk) * n + r0 + c];
            }
            __syncthreads();
            panel_rows_update<true>(M, n, r0, R0, E, perm_src, moved_to, sh);
            __syncthreads();
            const int c0 = r0 + NBC;
            if (c0 < E) {
                stage_l16<kBatchMaxS - NBC>(M, n, r0, c0, n, Lp);
                __syncthreads();
                tile_update16<2 * RT>(M, n, r0, c0, (n - c0) / NBC, c0, (E - c0) / NBC, Lp);
                __syncthreads();
            }
        }
        if (R0 == 0 && E == n) break;
        // the block's row moves in every other column: its panels' moves
        // composed (as compose_rows does) over positions [R0, n)
        for (int i = tid; i < n; i += NT) lu_pfb[i] = i;
        if (tid == 0) lu_nchg = 0;
        __syncthreads();
        for (int r0 = R0; r0 < E; r0 += NBC) {
            int v1 = 0, v2 = 0, dst = -1;
            if (tid < NBC) {
                v1 = lu_pfb[perm_src[r0 + tid]];
                v2 = lu_pfb[r0 + tid];
                dst = moved_to[r0 + tid];
            }
            __syncthreads();
            if (tid < NBC) {
                lu_pfb[r0 + tid] = v1;
                if (dst >= r0 + NBC) lu_pfb[dst] = v2;
            }
            __syncthreads();
        }
        for (int i = R0 + tid; i < n; i += NT) {
            if (lu_pfb[i] != i) {
                const int k = atomicAdd(&lu_nchg, 1);
                if (k < 2 * kLuBlock) lu_chg[k] = i;
            }
        }
        __syncthreads();
        {
            // 64 columns at a time: thread t takes column (t % 64) of the
            // chunk and the changed positions t / 64 + 4 j; every source is
            // read before any position is written
            const int nchg = min(lu_nchg, 2 * kLuBlock);
            const int ncol = n - (E - R0);
            const __amdgpu_buffer_rsrc_t rs = buf_rsrc(M, n * n * 8);
            const int cl = tid & 63, e0 = tid >> 6;
            constexpr int kPer = 2 * kLuBlock / NWAVES;
            int dpos[kPer], spos[kPer];
#pragma unroll
            for (int j = 0; j < kPer; j++) {
                const int e = e0 + NWAVES * j;
                dpos[j] = e < nchg ? lu_chg[e] : -1;
                spos[j] = e < nchg ? lu_pfb[dpos[j]] : 0;
            }
            for (int cs = 0; cs < ncol; cs += 64) {
                const int cc = cs + cl;
                const int c = cc < R0 ? cc : cc + (E - R0);
                const bool okc = cc < ncol;
                double v[kPer];
#pragma unroll
                for (int j = 0; j < kPer; j++)
                    v[j] = BufIO<double>::ld(rs, (okc && dpos[j] >= 0) ? (spos[j] * n + c) * 8 : kBufOob);
                __syncthreads();
#pragma unroll
                for (int j = 0; j < kPer; j++)
                    BufIO<double>::st(rs, (okc && dpos[j] >= 0) ? (dpos[j] * n + c) * 8 : kBufOob, v[j]);
            }
        }
        __syncthreads();
        if (E == n) break;
        // U12 of the block (rows [R0, E), columns [E, n); E = R0 + 64 here):
        // the block's L11 to LDS, then one register sweep per column
        {
            const __amdgpu_buffer_rsrc_t rs = buf_rsrc(M, n * n * 8);
            constexpr int kPer = kLuBlock * kLuBlock / NT;
            double v[kPer];
#pragma unroll
            for (int j = 0; j < kPer; j++) {
                const int e = tid + j * NT, i = e / kLuBlock, k = e % kLuBlock;
                v[j] = BufIO<double>::ld(rs, ((R0 + i) * n + R0 + k) * 8);
            }
#pragma unroll
            for (int j = 0; j < kPer; j++) {
                const int e = tid + j * NT, i = e / kLuBlock, k = e % kLuBlock;
                Lp[k * kLuBlock + i] = v[j];
            }
        }
        __syncthreads();
        block_u12_sweep(M, n, R0, E, lp_raw);
        __syncthreads();
        tile_update64(M, n, R0, E, lp_raw);
        __syncthreads();
    }
    return 0;
}

template <typename T, bool BF16, int RT = 4>
__device__ int wg_lu_rm(T *__restrict__ M, int n, int *piv, int *perm_src, int *moved_to, int *rows, int *rowcnt,
                        int *info, Shared &sh)
{
    if constexpr (std::is_same<T, double>::value && !BF16) {
        // the blocked form's extra phases and barriers cost more than the
        // trailing-matrix traffic they save up to S = 256 (measured: 541
        // against 506 us per LU at 256, 203 against 178 at 128; 1.76 against
        // 2.56 ms at 512)
        if ((n & 15) == 0 && n > kLuBlockedMin)
            return wg_lu_rm_blocked<RT>(M, n, piv, perm_src, moved_to, rows, rowcnt, info, sh);
    }
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
                if constexpr (BF16) {
                    // one v_mfma_f32_16x16x16_bf16 per tile (k = 0..15, the
                    // whole panel): lane l holds A[l%16][4(l/16)+j] and
                    // B[4(l/16)+j][l%16], j < 4, as BF16; FP32 accumulation
                    // with the FP32 instruction's output layout (measured)
                    typedef short S4 __attribute__((ext_vector_type(4)));
                    auto bits16 = [](float x) { return (short)(__float_as_uint(x) >> 16); };
                    S4 b4;
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        const int k = 4 * lq + j;
                        const bool ok = okc && k < nbp;
                        const T bv = Mg[ok ? (size_t)(r0 + k) * n + col : 0];
                        b4[j] = bits16(ok ? to_bf16_rne((float)bv) : 0.f);
                    }
                    for (int g = 0; g < nt; g += RT) {
                        Acc d[RT];
                        S4 a4[RT];
#pragma unroll
                        for (int u = 0; u < RT; u++) {
                            const int R = c0 + 16 * (g + u);
                            const bool okt = g + u < nt;
#pragma unroll
                            for (int i = 0; i < 4; i++) {
                                const int r = R + Mfma<T>::row(lq, i);
                                const bool ok = okt && r < n && okc;
                                const T v = Mg[ok ? (size_t)r * n + col : last];
                                d[u][i] = ok ? v : T(0);
                            }
                            const int ra = R + l16;
#pragma unroll
                            for (int j = 0; j < 4; j++) {
                                const int k = 4 * lq + j;
                                const bool ok = okt && ra < n && k < nbp;
                                // staged negated and rounded to BF16; padding -0 * 0 = -0
                                a4[u][j] = bits16(ok ? (float)Lp[(ra - c0) * kLpStride + k] : -0.f);
                            }
                        }
#pragma unroll
                        for (int u = 0; u < RT; u++) d[u] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a4[u], b4, d[u], 0, 0, 0);
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
                    }
                    continue;
                }
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
