// Persistent single-model solver: the whole LASSIE integration loop
// (Simulation.cuh:632-925) in one cooperative launch.
//
//   group 0      sample check, RKF45 attempt, step control, Newton loop; keeps
//                the model and the state vectors in LDS when they fit
//   all groups   dense assembly of M = I - hJ and the blocked LU
//
// Every floating-point operation follows the order written in the reference
// kernels (kernels.cuh) with no contraction (built with -ffp-contract=off),
// so the trajectory is bit-identical to the CPU restatement used as oracle.
//
// T is the state type: double for the reference's default, float for its
// `-p float` mode (Simulation<float>, simulator.cu:259-304), in which the
// state vectors, t, dt, the Newton matrix and its LU are float while the
// expressions with double literals and the per-term RHS/Jacobian
// accumulators stay double, as the reference's kernels are written.
#include "device_common.h"
#include "solver.h"
#include "wg_common.h"

namespace lassie {

// Row moves of panel p applied to a strip left of it (q < p), so that every
// L column ends in the final row order (LAPACK's dlaswp on the left part).
template <int NBC, typename T>
__device__ void strip_moves(T *__restrict__ M, int n, int p, int q, const int *perm_src, const int *moved_to,
                            Shared &sh)
{
    const int tid = threadIdx.x;
    const int r0 = p * NBC;
    const int nbp = min(NBC, n - r0);
    const int c0 = q * NBC;
    const int ncols = min(NBC, n - c0);
    for (int e = tid; e < NBC * NBC; e += NT) {
        const int kk = e / NBC, c = e % NBC;
        if (kk < nbp && c < ncols) {
            sh.U12[kk][c] = M[(size_t)(c0 + c) * n + perm_src[r0 + kk]];
            sh.oldP[kk][c] = M[(size_t)(c0 + c) * n + r0 + kk];
        }
    }
    __syncthreads();
    for (int e = tid; e < NBC * NBC; e += NT) {
        const int kk = e / NBC, c = e % NBC;
        if (kk < nbp && c < ncols) {
            M[(size_t)(c0 + c) * n + r0 + kk] = sh.U12[kk][c];
            const int dst = moved_to[r0 + kk];
            if (dst >= r0 + nbp) M[(size_t)(c0 + c) * n + dst] = sh.oldP[kk][c];
        }
    }
    __syncthreads();
}

// Apply panel p (row moves, triangular solve for the U block row, trailing
// update) to strip q (columns [q*NBC, q*NBC + ncols)), q > p.  One group.
// M must be followed by kLuSink writable elements.
//
// With 16-column panels the trailing update A22 -= L21 U12 runs on the matrix
// cores over every row below the panel, as the oracle does: 16-row tiles,
// computed transposed (A22^T -= U12^T L21^T) so that the lanes of a 16-lane
// group hold consecutive rows of the column-major matrix, four 16x16x4
// MFMAs per tile (FP64, or FP32 for the single-precision loop and the FP32
// factors of the precision form) with k ascending: both instructions
// accumulate as a sequential fma chain, the oracle's order (FP32: equal to
// an fmaf chain on 1048576 of 1048576 random elements over 20 binades).  A
// wave walks the row tiles wave, wave+4, ... RT at a time with the next RT
// tiles' operands in flight.
// Otherwise (8-column panels) the update uses vector fmas over the rows
// listed by the panel (all others have zero multipliers, so
// a_ij = fma(-0, u, a_ij) = a_ij), or every row when the U block holds a
// non-finite value.
template <int NBC, typename T, bool BF16 = false>
__device__ void strip_update(T *__restrict__ M, int n, int p, int q, const int *perm_src, const int *moved_to,
                             const int *rows, int rowcnt, Shared &sh)
{
    static_assert(!BF16 || sizeof(T) == 4, "BF16 updates of FP32 factors");
    n = uniform_int(n);
    p = uniform_int(p);
    q = uniform_int(q);
    const int tid = threadIdx.x;
    const int r0 = p * NBC;
    const int nbp = min(NBC, n - r0);
    const int c0 = q * NBC;
    const int ncols = min(NBC, n - c0);
    // gather the permuted panel rows and the old panel-range rows; stage L11.
    // Thread (kk, c) also fetches where row r0+kk was displaced to.
    static_assert(NBC * NBC <= NT, "one strip element per thread");
    const int ek = tid / NBC, ec = tid % NBC;
    const bool eok = tid < NBC * NBC && ek < nbp;
    int dst = -1;
    if (eok) {
        if (ec < ncols) {
            sh.U12[ek][ec] = M[(size_t)(c0 + ec) * n + perm_src[r0 + ek]];
            sh.oldP[ek][ec] = M[(size_t)(c0 + ec) * n + r0 + ek];
            dst = moved_to[r0 + ek];
        }
        if (ec < nbp) sh.L11[ek][ec] = M[(size_t)(r0 + ec) * n + r0 + ek];
    }
    __syncthreads();
    // displaced panel-range rows go to the positions their pivots vacated
    if (dst >= r0 + nbp) M[(size_t)(c0 + ec) * n + dst] = sh.oldP[ek][ec];
    // U12 = L11^-1 * A12 (unit lower), one column per thread in registers,
    // k ascending; written to M and back to LDS
    int nonfinite = 0;
    if (tid < ncols) {
        const int c = tid;
        T u[NBC];
#pragma unroll
        for (int kk = 0; kk < NBC; kk++) u[kk] = (kk < nbp) ? (T)sh.U12[kk][c] : (T)0.0;
#pragma unroll
        for (int i = 1; i < NBC; i++) {
            if (i < nbp) {
                T v = u[i];
#pragma unroll
                for (int kk = 0; kk < i; kk++) v = fma(-(T)sh.L11[i][kk], u[kk], v);
                u[i] = v;
            }
        }
#pragma unroll
        for (int kk = 0; kk < NBC; kk++) {
            if (kk < nbp) {
                sh.U12[kk][c] = u[kk];
                M[(size_t)(c0 + c) * n + r0 + kk] = u[kk];
                nonfinite |= !isfinite(u[kk]);
            }
        }
    }
    // the 8-column update visits every row when U12 holds a non-finite value
    constexpr bool kMfma = NBC == 16 && !BF16;
    bool all_rows = false;
    if constexpr (kMfma)
        __syncthreads();
    else
        all_rows = __syncthreads_or(nonfinite) != 0;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    if constexpr (BF16) {
        // every row below the panel on the BF16 matrix cores: one
        // v_mfma_f32_16x16x16_bf16 per 16 x 16 tile covers the whole panel
        // (k < 16; an 8-column panel pads k with zeros), the transposed
        // product of the FP64/FP32 form (A22^T -= U12^T L21^T) with both
        // operands rounded to BF16 (nearest even) and FP32 accumulation.
        // Lane l holds A[l%16][4(l/16)+j] and B[4(l/16)+j][l%16], j < 4; D item
        // i of lane l is D[4(l/16)+i][l%16].
        typedef short S4 __attribute__((ext_vector_type(4)));
        typedef float F4 __attribute__((ext_vector_type(4)));
        auto bits16 = [](float x) { return (short)(__float_as_uint(x) >> 16); };
        typedef __attribute__((address_space(1))) float GD;
        GD *Mg = (GD *)M;
        const size_t last = (size_t)n * n - 1;
        const size_t sink = (size_t)n * n;
        const int l16 = lane & 15, lq = lane >> 4;
        const int rb = r0 + nbp;
        // 8-column panels list the rows below with a nonzero multiplier: tile
        // rows come from the list (a zero multiplier row keeps its values
        // while U12 is finite, as in the vector update); else every row
        const bool listed = NBC != 16 && !all_rows;
        const int nrows = listed ? rowcnt : n - rb;
        const int nt = (nrows + 15) >> 4;
        S4 ua;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int k = 4 * lq + j;
            ua[j] = bits16((k < nbp && l16 < ncols) ? to_bf16_rne((float)sh.U12[k][l16]) : 0.f);
        }
        for (int t = wave; t < nt; t += NWAVES) {
            const int idx = 16 * t + l16;
            const int r = idx < nrows ? (listed ? rows[idx] : rb + idx) : n;
            const bool okr = r < n;
            F4 d;
            S4 bv;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int c = 4 * lq + i;
                const bool ok = okr && c < ncols;
                const float v = Mg[ok ? (size_t)(c0 + c) * n + r : last];
                d[i] = ok ? v : 0.f;
            }
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int k = 4 * lq + j;
                const bool ok = okr && k < nbp;
                const float v = Mg[ok ? (size_t)(r0 + k) * n + r : last];
                // padding products are 0 * (-0) = -0, which leaves every value unchanged
                bv[j] = bits16(ok ? to_bf16_rne(-v) : -0.f);
            }
            d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ua, bv, d, 0, 0, 0);
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int c = 4 * lq + i;
                const bool ok = okr && c < ncols;
                Mg[ok ? (size_t)(c0 + c) * n + r : sink] = d[i];
            }
        }
    } else if constexpr (kMfma) {
        (void)rows;
        (void)rowcnt;
        constexpr int RT = 4;
        typedef typename Mfma<T>::acc Acc;
        // global-memory view: precise memory counters keep the next tiles'
        // loads in flight across the current MFMAs
        typedef __attribute__((address_space(1))) T GD;
        GD *Mg = (GD *)M;
        const size_t last = (size_t)n * n - 1;
        const size_t sink = (size_t)n * n;
        const int l16 = lane & 15, lq = lane >> 4;
        const int rb = r0 + nbp;             // first row below the panel
        const int nt = (n - rb + 15) >> 4;   // row tiles
        // A operand: U12^T (A[c][k] = U12[k][c]); padding columns are 0
        T ua[4];
#pragma unroll
        for (int kc = 0; kc < 4; kc++) {
            const int k = 4 * kc + lq;
            ua[kc] = (k < nbp && l16 < ncols) ? (T)sh.U12[k][l16] : T(0);
        }
        // tile t of this wave: rows rb + 16 (wave + NWAVES t) + l16
        auto load_group = [&](int g, Acc *dd, T (*bb)[4]) {
#pragma unroll
            for (int u = 0; u < RT; u++) {
                const int r = rb + 16 * (wave + NWAVES * (g + u)) + l16;
                const bool okr = r < n;
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int c = Mfma<T>::row(lq, i);
                    const bool ok = okr && c < ncols;
                    const T v = Mg[ok ? (size_t)(c0 + c) * n + r : last];
                    dd[u][i] = ok ? v : T(0);
                }
#pragma unroll
                for (int kc = 0; kc < 4; kc++) {
                    const int k = 4 * kc + lq;
                    const bool ok = okr && k < nbp;
                    const T v = Mg[ok ? (size_t)(r0 + k) * n + r : last];
                    // padding products are 0 * (-0) = -0, which leaves every value unchanged
                    bb[u][kc] = ok ? -v : T(-0.0);
                }
            }
        };
        const int ntw = (nt - wave + NWAVES - 1) / NWAVES;  // tiles of this wave
        Acc d[RT], dn[RT];
        T b[RT][4], bn[RT][4];
        load_group(0, d, b);
        for (int g = 0; g < ntw; g += RT) {
            const bool more = g + RT < ntw;
            load_group(g + RT, dn, bn);  // unconditional: no branch, exact memory counters
#pragma unroll
            for (int u = 0; u < RT; u++) {
#pragma unroll
                for (int kc = 0; kc < 4; kc++) d[u] = Mfma<T>::run(ua[kc], b[u][kc], d[u]);
            }
#pragma unroll
            for (int u = 0; u < RT; u++) {
                const int r = rb + 16 * (wave + NWAVES * (g + u)) + l16;
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int c = Mfma<T>::row(lq, i);
                    const bool ok = g + u < ntw && r < n && c < ncols;
                    Mg[ok ? (size_t)(c0 + c) * n + r : sink] = d[u][i];
                }
            }
            if (more) {
#pragma unroll
                for (int u = 0; u < RT; u++) {
                    d[u] = dn[u];
#pragma unroll
                    for (int kc = 0; kc < 4; kc++) b[u][kc] = bn[u][kc];
                }
            }
        }
    } else {
        // vector update: a wave takes 64 rows at a time and owns CPW columns
        // of the strip, whose U12 entries stay in registers; two row blocks
        // are in flight per iteration.
        constexpr int CPW = NBC / NWAVES;
        const int cbase = wave * CPW;
        T u[NBC][CPW];
#pragma unroll
        for (int kk = 0; kk < NBC; kk++)
#pragma unroll
            for (int cc = 0; cc < CPW; cc++) u[kk][cc] = (kk < nbp) ? (T)sh.U12[kk][cbase + cc] : (T)0.0;
        const int nr = all_rows ? n - (r0 + nbp) : rowcnt;
        // raw buffer accesses: the row in the vector offset (out of range,
        // reading zero and dropping stores, for masked elements), the column
        // in the scalar offset; no branches, so every load of a row pair is
        // in flight at once.  Padding k (kk >= nbp) has l = u = +0, and
        // fma(-0, 0, acc) == acc bit for bit.
        constexpr int E = (int)sizeof(T);
        const __amdgpu_buffer_rsrc_t rs = buf_rsrc(M, n * n * E);
        const int cb = uniform_int(cbase);
        const int nbu = uniform_int(nbp), ncu = uniform_int(ncols);
        for (int x0 = lane; x0 < nr; x0 += 128) {
            const int x1 = x0 + 64;
            const bool has1 = x1 < nr;
            const int i0 = all_rows ? r0 + nbp + x0 : rows[x0];
            const int i1 = has1 ? (all_rows ? r0 + nbp + x1 : rows[x1]) : i0;
            const int v0 = i0 * E, v1 = has1 ? i1 * E : kBufOob;
            T acc0[CPW], acc1[CPW], l0[NBC], l1[NBC];
#pragma unroll
            for (int cc = 0; cc < CPW; cc++) {
                const bool okc = cb + cc < ncu;
                acc0[cc] = BufIO<T>::ld(rs, okc ? v0 : kBufOob, (c0 + cb + cc) * n * E);
                acc1[cc] = BufIO<T>::ld(rs, okc ? v1 : kBufOob, (c0 + cb + cc) * n * E);
            }
#pragma unroll
            for (int kk = 0; kk < NBC; kk++) {
                l0[kk] = BufIO<T>::ld(rs, kk < nbu ? v0 : kBufOob, (r0 + kk) * n * E);
                l1[kk] = BufIO<T>::ld(rs, kk < nbu ? v1 : kBufOob, (r0 + kk) * n * E);
            }
#pragma unroll
            for (int kk = 0; kk < NBC; kk++) {
#pragma unroll
                for (int cc = 0; cc < CPW; cc++) {
                    acc0[cc] = fma(-l0[kk], u[kk][cc], acc0[cc]);
                    acc1[cc] = fma(-l1[kk], u[kk][cc], acc1[cc]);
                }
            }
#pragma unroll
            for (int cc = 0; cc < CPW; cc++) {
                const bool okc = cb + cc < ncu;
                BufIO<T>::st(rs, okc ? v0 : kBufOob, acc0[cc], (c0 + cb + cc) * n * E);
                BufIO<T>::st(rs, okc ? v1 : kBufOob, acc1[cc], (c0 + cb + cc) * n * E);
            }
        }
    }
    __syncthreads();
}

// panel_factor with the fewest rows per thread that cover the panel's rows
// [p*NBC, n): later panels have fewer live rows, and the register panel is
// processed without per-row control flow, so a smaller RPT is less work.  The
// rows each thread holds, and so every result, are the same for any RPT that
// covers them.
template <int RPT, int NBC, typename T, typename Store>
__device__ void panel_fit(const Store st, int n, int p, int *piv, int *perm_src, int *moved_to, int *rows,
                          int *rowcnt, int *info, Shared &sh)
{
    if constexpr (RPT > 1) {
        if (n - p * NBC <= NT * (RPT / 2)) {
            panel_fit<RPT / 2, NBC, T>(st, n, p, piv, perm_src, moved_to, rows, rowcnt, info, sh);
            return;
        }
    }
    // row lists feed the vector strip update only
    constexpr bool kList = NBC != 16;  // row lists feed the vector (8-column) update only
    panel_factor_s<RPT, NBC, T, kList>(st, n, p, piv, perm_src, moved_to, rows, rowcnt, info, sh);
}

// ---------------------------------------------------------------------------
// LDS-resident strips (FP64, 16-column panels, one strip per workgroup).  The
// owner of strip q keeps it in LDS for the whole factorization: column-major,
// leading dimension ld, column c at Ls[c*ld].  Only the factored panel goes to
// the global matrix at once (the other groups read L11/L21 from there); the
// rest of the strip is written back when the factorization ends.
typedef __attribute__((address_space(3))) double LDouble;
__device__ __forceinline__ int tid_of() { return threadIdx.x; }

// strip_update on an LDS strip: same operations and order as strip_update's
// MFMA path; the accumulator tiles are read from and written to LDS.  The
// L21 operands come from the global matrix: every tile's operands are
// requested on entry (panel p is final), so their latency overlaps the
// gather and the U12 step.  Panels are never the last one here (q > p), so
// they have 16 columns; LDS strips are used for n <= 1024, so a wave has at
// most kMaxTiles row tiles.
constexpr int kMaxTiles = 16;

__device__ void strip_update_lds(const double *__restrict__ M, LDouble *Ls, int ld, int n, int p, int q,
                                 const int *perm_src, const int *moved_to, Shared &sh)
{
    constexpr int NBC = 16;
    typedef Mfma<double>::acc Acc;
    typedef __attribute__((address_space(1))) const double GD;
    GD *Mg = (GD *)M;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int l16 = lane & 15, lq = lane >> 4;
    const int r0 = p * NBC;
    const int c0 = q * NBC;
    const int ncols = min(NBC, n - c0);
    const int rb = r0 + NBC;
    const int nt = (n - rb + 15) >> 4;
    const int ntw = (nt - wave + NWAVES - 1) / NWAVES;  // tiles of this wave, <= kMaxTiles
    const size_t last = (size_t)n * n - 1;
    // global loads, in the order they are needed: L11, the pivot-row sources
    // and displaced-row targets, then the L21 operands
    const int ek = tid / NBC, ec = tid % NBC;
    const double l11 = M[(size_t)(r0 + ec) * n + r0 + ek];
    int src = 0, dst = -1;
    if (ec < ncols) {
        src = perm_src[r0 + ek];
        dst = moved_to[r0 + ek];
    }
    // B[k][row] = -L[row][r0 + k] for tile t (rows rb + 16 (wave + 4t) + l16);
    // padding products are 0 * (-0) = -0, which leaves every value unchanged
    double b[kMaxTiles][4];
#pragma unroll
    for (int t = 0; t < kMaxTiles; t++) {
        const int r = rb + 16 * (wave + NWAVES * t) + l16;
#pragma unroll
        for (int kc = 0; kc < 4; kc++) {
            const int k = 4 * kc + lq;
            const bool ok = r < n;
            const double v = Mg[ok ? (size_t)(r0 + k) * n + r : last];
            b[t][kc] = ok ? -v : -0.0;
        }
    }
    if (ec < ncols) {
        sh.U12[ek][ec] = Ls[ec * ld + src];
        sh.oldP[ek][ec] = Ls[ec * ld + r0 + ek];
    }
    sh.L11[ek][ec] = l11;
    __syncthreads();
    // displaced panel-range rows go to the positions their pivots vacated
    if (dst >= rb) Ls[ec * ld + dst] = sh.oldP[ek][ec];
    // U12 = L11^-1 * A12 (unit lower), one column per thread, k ascending
    if (tid < ncols) {
        const int c = tid;
        double u[NBC];
#pragma unroll
        for (int kk = 0; kk < NBC; kk++) u[kk] = sh.U12[kk][c];
#pragma unroll
        for (int i = 1; i < NBC; i++) {
            double v = u[i];
#pragma unroll
            for (int kk = 0; kk < i; kk++) v = fma(-sh.L11[i][kk], u[kk], v);
            u[i] = v;
        }
#pragma unroll
        for (int kk = 0; kk < NBC; kk++) {
            sh.U12[kk][c] = u[kk];
            Ls[c * ld + r0 + kk] = u[kk];
        }
    }
    __syncthreads();
    // trailing update A22 -= L21 U12 over every row below the panel, computed
    // transposed on the matrix cores (see strip_update), four tiles at a time
    double ua[4];
#pragma unroll
    for (int kc = 0; kc < 4; kc++) ua[kc] = (l16 < ncols) ? sh.U12[4 * kc + lq][l16] : 0.0;
#pragma unroll
    for (int g = 0; g < kMaxTiles; g += 4) {
        if (g < ntw) {
            Acc d[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int r = rb + 16 * (wave + NWAVES * (g + u)) + l16;
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int c = Mfma<double>::row(lq, i);
                    const bool ok = r < n && c < ncols;
                    const double v = Ls[ok ? c * ld + r : 0];
                    d[u][i] = ok ? v : 0.0;
                }
            }
#pragma unroll
            for (int kc = 0; kc < 4; kc++)
#pragma unroll
                for (int u = 0; u < 4; u++) d[u] = Mfma<double>::run(ua[kc], b[g + u][kc], d[u]);
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int r = rb + 16 * (wave + NWAVES * (g + u)) + l16;
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int c = Mfma<double>::row(lq, i);
                    if (r < n && c < ncols) Ls[c * ld + r] = d[u][i];
                }
            }
        }
    }
    __syncthreads();
}

// strip_moves on an LDS strip (row moves of panel p, q < p)
__device__ void strip_moves_lds(LDouble *Ls, int ld, int n, int p, int q, const int *perm_src,
                                const int *moved_to, Shared &sh)
{
    constexpr int NBC = 16;
    const int tid = threadIdx.x;
    const int r0 = p * NBC;
    const int nbp = min(NBC, n - r0);
    const int ncols = min(NBC, n - q * NBC);
    const int kk = tid / NBC, c = tid % NBC;
    const bool ok = kk < nbp && c < ncols;
    int src = 0, dst = -1;
    if (ok) {
        src = perm_src[r0 + kk];
        dst = moved_to[r0 + kk];
        sh.U12[kk][c] = Ls[c * ld + src];
        sh.oldP[kk][c] = Ls[c * ld + r0 + kk];
    }
    __syncthreads();
    if (ok) {
        Ls[c * ld + r0 + kk] = sh.U12[kk][c];
        if (dst >= r0 + nbp) Ls[c * ld + dst] = sh.oldP[kk][c];
    }
    __syncthreads();
}

// Blocked right-looking LU over all groups; strips are owned round-robin.
// The owner of strip p+1 updates it first and factors panel p+1 at once
// (look-ahead), so each panel costs one grid barrier; left strips receive the
// panel's row moves.  On return M holds L\U in final row order.  Returns
// false on abort.
template <int RPT, int NBC, typename T, bool BF16 = false>
__device__ bool lu_factor_grid(const SolveArgs &A, int wg, Shared &sh, unsigned int &gen_local,
                               bool &singular)
{
    T *M = reinterpret_cast<T *>(A.M);
    const int n = A.m.S;
    const int P = (n + NBC - 1) / NBC;
    const int G = A.G;
    singular = false;
    // A zero pivot in panel p+1 may be recorded while slower groups are still
    // checking the flag for panel p, so each check only counts columns of
    // panels that were complete at the preceding barrier.
    auto failed_before = [&](int col_end) {
        const int inf = ld_rlx(&A.ctl->info);
        return inf != 0 && inf - 1 < col_end;
    };
    // strip wg in LDS (one strip per group; group 0's strip 0 stays in the
    // global matrix: after the first panel it only receives row moves)
    constexpr bool kLdsCapable = NBC == 16 && sizeof(T) == 8;
    const bool ldsq = kLdsCapable && A.lds_strip && wg >= 1 && wg < P;
    const int ld = A.strip_ld;
    extern __shared__ __align__(16) unsigned char dyn[];
    LDouble *Ls = (LDouble *)dyn;
    if (ldsq) {
        const int c0 = wg * NBC;
        const int ncols = min(NBC, n - c0);
        for (int e = tid_of(); e < ncols * n; e += NT) {
            const int c = e / n, r = e - c * n;
            Ls[c * ld + r] = reinterpret_cast<const double *>(M)[(size_t)(c0 + c) * n + r];
        }
        __syncthreads();
    }
    if (wg == 0)
        panel_fit<RPT, NBC, T>(MatStore<T, false>{M, n}, n, 0, A.piv, A.perm_src, A.moved_to, A.rows, A.rowcnt,
                               &A.ctl->info, sh);
    if (!grid_barrier(A.gs, G, gen_local, &A.ctl->abort)) return false;
    if (failed_before(NBC)) {
        singular = true;
        return true;
    }
    for (int p = 0; p < P; p++) {
        const int cnt = NBC == 16 ? 0 : A.rowcnt[p];  // row lists feed the vector update only
        const int *rows = A.rows + (size_t)p * n;
        int q = p + 1 + ((wg - (p + 1)) % G + G) % G;
        for (; q < P; q += G) {
            if constexpr (kLdsCapable) {
                if (ldsq) {
                    const double *Md = reinterpret_cast<const double *>(M);
                    strip_update_lds(Md, Ls, ld, n, p, q, A.perm_src, A.moved_to, sh);
                    if (q == p + 1)
                        panel_fit<RPT, NBC, T>(LdsStripStore<T>{(LDouble *)Ls, ld, q * NBC, M, n}, n, p + 1, A.piv,
                                               A.perm_src, A.moved_to, A.rows + (size_t)(p + 1) * n,
                                               A.rowcnt + p + 1, &A.ctl->info, sh);
                    __syncthreads();
                    continue;
                }
            }
            strip_update<NBC, T, BF16>(M, n, p, q, A.perm_src, A.moved_to, rows, cnt, sh);
            if (q == p + 1) {
                __syncthreads();
                panel_fit<RPT, NBC, T>(MatStore<T, false>{M, n}, n, p + 1, A.piv, A.perm_src, A.moved_to,
                                       A.rows + (size_t)(p + 1) * n, A.rowcnt + p + 1, &A.ctl->info, sh);
                __syncthreads();
            }
        }
        for (int ql = wg; ql < p; ql += G) {
            if (ldsq)
                strip_moves_lds(Ls, ld, n, p, ql, A.perm_src, A.moved_to, sh);
            else
                strip_moves<NBC>(M, n, p, ql, A.perm_src, A.moved_to, sh);
        }
        if (ldsq && p == P - 1) {
            // the strip in its final row order, for the triangular solves
            double *Mw = reinterpret_cast<double *>(M);
            const int c0 = wg * NBC;
            const int ncols = min(NBC, n - c0);
            for (int e = tid_of(); e < ncols * n; e += NT) {
                const int c = e / n, r = e - c * n;
                Mw[(size_t)(c0 + c) * n + r] = Ls[c * ld + r];
            }
        }
        if (!grid_barrier(A.gs, G, gen_local, &A.ctl->abort)) return false;
        if (p + 1 < P && failed_before((p + 2) * NBC)) {
            singular = true;
            return true;
        }
    }
    return true;
}

// ---------------------------------------------------------------------------
// Distributed triangular solves of the Newton correction (getrs,
// Simulation.cuh:1316).  Row block j (kTriBlock rows) belongs to wave 0 of
// group j % G.  Blocks are chained by per-block stamps: the writer stores its
// block, then publishes the stamp after an agent-scope fence; a reader waits
// for the stamp, fences, then loads the block.  Each element's fma chain runs
// in the order of lu_solve_wg and the oracle: forward k ascending (earlier
// blocks first, then the diagonal block), backward k descending.

// Wait (every lane of the calling wave) until *flag == want; bounded.
__device__ __forceinline__ bool wait_stamp(const unsigned int *flag, unsigned int want, int *abort_word)
{
    const unsigned long long t0 = wall_clock64();
    unsigned int spins = 0;
    while (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != want) {
        if ((++spins & 255) == 0) {
            if (wall_clock64() - t0 > kWaitLimit) {
                st_rlx(abort_word, 1);
                return false;
            }
            if (ld_rlx(abort_word)) return false;
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return true;
}

// Every lane releases its own stores, then lane 0 publishes the stamp.
__device__ __forceinline__ void publish_stamp(unsigned int *flag, unsigned int v)
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    if ((threadIdx.x & 63) == 0) __hip_atomic_store(flag, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// L y = b for block j (wave 0 of its owner): b = tri[0..S), y -> tri[S..2S).
// With rsrc, b is read from there instead, in the factors' row order and
// rounded to T: b_i = (T)rsrc[perm_full[i]].  The factors are read with raw
// buffer loads: the lane's row in the 32-bit vector offset, the column in the
// scalar offset, and an out-of-range vector offset (which reads zero) for the
// lanes past the last row, so the loads carry no branches and no 64-bit
// address arithmetic, and all of a chunk's loads stay in flight together.
template <typename T>
__device__ bool tri_forward_block(const SolveArgs &A, int j, unsigned int stamp, const double *rsrc = nullptr)
{
    constexpr int TB = kTriBlock;
    constexpr int E = (int)sizeof(T);
    // wave-uniform (scalar registers: the loads' scalar offsets)
    const int n = uniform_int(A.m.S);
    j = uniform_int(j);
    const int lane = threadIdx.x & 63;
    const int R0 = j * TB;
    const int nb = min(TB, n - R0);
    const int i = R0 + lane;
    const bool row = lane < nb;
    const __amdgpu_buffer_rsrc_t rs = buf_rsrc(A.M, n * n * E);
    const int vo = row ? i * E : kBufOob;
    const T *bp = reinterpret_cast<const T *>(A.tri);
    T *yb = reinterpret_cast<T *>(A.tri) + n;
    unsigned int *fwd = A.tflag;
    // the diagonal block's L entries first: the chain below must not wait on memory
    T d[TB];
#pragma unroll
    for (int k = 0; k < TB; k++) d[k] = BufIO<T>::ld(rs, k < lane ? vo : kBufOob, (R0 + k) * n * E);
    T acc = (T)0.0;
    if (row) acc = rsrc ? (T)rsrc[A.perm_full[i]] : bp[i];
    for (int B = 0; B < j; B++) {
        const int c0 = B * TB;
        T l[16], ln[16];
#pragma unroll
        for (int kk = 0; kk < 16; kk++) l[kk] = BufIO<T>::ld(rs, vo, (c0 + kk) * n * E);
        if (!wait_stamp(fwd + (size_t)B * kFlagStride, stamp, &A.ctl->abort)) return false;
        const T yB = yb[c0 + lane];
#pragma unroll
        for (int sb = 0; sb < TB / 16; sb++) {
            if (sb + 1 < TB / 16) {
#pragma unroll
                for (int kk = 0; kk < 16; kk++) ln[kk] = BufIO<T>::ld(rs, vo, (c0 + (sb + 1) * 16 + kk) * n * E);
            }
#pragma unroll
            for (int kk = 0; kk < 16; kk++) acc = fma(-l[kk], readlane_d(yB, sb * 16 + kk), acc);
            if (sb + 1 < TB / 16) {
#pragma unroll
                for (int kk = 0; kk < 16; kk++) l[kk] = ln[kk];
            }
        }
    }
    // diagonal block, unit lower triangle
#pragma unroll
    for (int k = 0; k < TB; k++) {
        if (k < nb) {
            const T yk = readlane_d(acc, k);
            if (lane > k && row) acc = fma(-d[k], yk, acc);
        }
    }
    if (row) yb[i] = acc;
    publish_stamp(fwd + (size_t)j * kFlagStride, stamp);
    return true;
}

// U x = y for block j: y = tri[S..2S), x -> tri[2S..3S) (buffer loads as in
// tri_forward_block).
template <typename T>
__device__ bool tri_backward_block(const SolveArgs &A, int j, int nblk, unsigned int stamp)
{
    constexpr int TB = kTriBlock;
    constexpr int E = (int)sizeof(T);
    // wave-uniform (scalar registers: the loads' scalar offsets)
    const int n = uniform_int(A.m.S);
    j = uniform_int(j);
    const int lane = threadIdx.x & 63;
    const int R0 = j * TB;
    const int nb = min(TB, n - R0);
    const int i = R0 + lane;
    const bool row = lane < nb;
    const __amdgpu_buffer_rsrc_t rs = buf_rsrc(A.M, n * n * E);
    const int vo = row ? i * E : kBufOob;
    const T *yb = reinterpret_cast<const T *>(A.tri) + n;
    T *xb = reinterpret_cast<T *>(A.tri) + 2 * (size_t)n;
    nblk = uniform_int(nblk);
    unsigned int *bwd = A.tflag + (size_t)nblk * kFlagStride;
    // the diagonal block's U entries (with the diagonal) first
    T d[TB];
#pragma unroll
    for (int k = 0; k < TB; k++) {
        const bool use = k >= lane && k < nb;
        const T v = BufIO<T>::ld(rs, use ? vo : kBufOob, (R0 + k) * n * E);
        d[k] = (use && row) ? v : (T)1.0;
    }
    T acc = row ? yb[i] : (T)0.0;
    for (int B = nblk - 1; B > j; B--) {
        const int c0 = B * TB;
        const int nbB = min(TB, n - c0);
        T u[16], un[16];
#pragma unroll
        for (int kk = 0; kk < 16; kk++) {
            const int k = TB - 16 + kk;
            u[kk] = BufIO<T>::ld(rs, k < nbB ? vo : kBufOob, (c0 + k) * n * E);
        }
        if (!wait_stamp(bwd + (size_t)B * kFlagStride, stamp, &A.ctl->abort)) return false;
        const T xB = (lane < nbB) ? xb[c0 + lane] : (T)0.0;
#pragma unroll
        for (int sb = TB / 16 - 1; sb >= 0; sb--) {
            if (sb > 0) {
#pragma unroll
                for (int kk = 0; kk < 16; kk++) {
                    const int k = (sb - 1) * 16 + kk;
                    un[kk] = BufIO<T>::ld(rs, k < nbB ? vo : kBufOob, (c0 + k) * n * E);
                }
            }
            // columns past the last row (k >= nbB) hold +0 (out-of-range
            // loads) and xB is 0 there: fma(-0, 0, acc) == acc bit for bit,
            // so they need no branch (a branch here makes every outstanding
            // load complete at each join, which would cancel the prefetch)
#pragma unroll
            for (int kk = 15; kk >= 0; kk--) acc = fma(-u[kk], readlane_d(xB, sb * 16 + kk), acc);
            if (sb > 0) {
#pragma unroll
                for (int kk = 0; kk < 16; kk++) u[kk] = un[kk];
            }
        }
    }
    // diagonal block, upper triangle with the diagonal
#pragma unroll
    for (int k = TB - 1; k >= 0; k--) {
        if (k < nb) {
            if (lane == k) acc = acc / d[k];
            const T xk = readlane_d(acc, k);
            if (lane < k) acc = fma(-d[k], xk, acc);
        }
    }
    if (row) xb[i] = acc;
    publish_stamp(bwd + (size_t)j * kFlagStride, stamp);
    return true;
}

// ---------------------------------------------------------------------------
// Precision form, helpers (FP64 loop; REF instantiation of solve_kernel).

// One refinement sweep after the FP32 solve d = tri32[2S..3S): x_new = x_old
// + (double)d (x_old = 0 on the first sweep), r = rhs - M x_new and the
// Oettli-Prager test |r_i| <= tol * (|rhs_i| + (|M||x_new|)_i + atol_i), both
// sums as fma chains over j ascending from 0 (M is the unfactored FP64
// matrix, column-major); the species' absolute tolerance is the floor that
// keeps components far below it (under FP32's normal range) from holding
// refinement back.  Rows go in 64-row chunks to the waves of all groups, a lane
// per row; x_new arrives 64 entries at a time, broadcast from the lane that
// formed it.  A group with a row off target raises ctl->rbad to seq.
__device__ void refine_residual(const SolveArgs &A, int wg, const double *rhs, const double *xo, double *xn,
                                double *res, unsigned int seq)
{
    const int n = uniform_int(A.m.S);
    const int G = A.G;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    // M by raw buffer loads: the row in the vector offset (out of range for
    // rows past n, reading zero), the column in the scalar offset
    const __amdgpu_buffer_rsrc_t rs = buf_rsrc(A.M, n * n * 8);
    const float *d = reinterpret_cast<const float *>(A.tri) + 2 * (size_t)n;
    const double tol = A.refine_tol;
    const int nch = (n + 63) / 64;
    int bad = 0;
    for (int c = wg + G * wave; c < nch; c += G * NWAVES) {
        const int i = c * 64 + lane;
        const bool row = i < n;
        const int vo = row ? i * 8 : kBufOob;
        double pa = 0.0, pm = 0.0;
        for (int j0 = 0; j0 < n; j0 += 64) {
            const int j = j0 + lane;
            double xj = 0.0;
            if (j < n) xj = (xo ? xo[j] : 0.0) + (double)d[j];
            const int jn = min(64, n - j0);
#pragma unroll
            for (int sb = 0; sb < 4; sb++) {
                double mv[16];
#pragma unroll
                for (int kk = 0; kk < 16; kk++) {
                    const int k = sb * 16 + kk;
                    mv[kk] = BufIO<double>::ld(rs, k < jn ? vo : kBufOob, uniform_int((j0 + k) * n * 8));
                }
                // columns past n: mv = +0 and xj = 0, and pa, pm (sums from
                // +0) are never -0, so fma(+0, 0, pa) == pa: no branch
#pragma unroll
                for (int kk = 0; kk < 16; kk++) {
                    const double xk = readlane_d(xj, sb * 16 + kk);
                    pa = fma(mv[kk], xk, pa);
                    pm = fma(fabs(mv[kk]), fabs(xk), pm);
                }
            }
        }
        if (row) {
            xn[i] = (xo ? xo[i] : 0.0) + (double)d[i];
            const double b = rhs[i];
            const double r = b - pa;
            res[i] = r;
            const double mag = (fabs(b) + pm) + A.m.atol[i];
            if (!(fabs(r) <= tol * mag)) bad = 1;
        }
    }
    if (__syncthreads_or(bad) && tid == 0)
        __hip_atomic_fetch_max(&A.ctl->rbad, seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Distributed triangular solves of one right-hand side, b_i = (T)rsrc[perm_full[i]]
// (T = float: the FP32 factors in A.M; double: the FP64 ones).  Wave 0 of the
// owners of the row blocks; false on abort.
template <typename T>
__device__ bool tri_solve_all(const SolveArgs &A, int wg, unsigned int stamp, const double *rsrc)
{
    const int n = A.m.S;
    const int G = A.G;
    const int nblk = (n + kTriBlock - 1) / kTriBlock;
    __shared__ int s_ok;
    if (threadIdx.x < 64) {
        bool ok = true;
        for (int j = wg; j < nblk && ok; j += G) ok = tri_forward_block<T>(A, j, stamp, rsrc);
        int jl = wg;
        while (jl + G < nblk) jl += G;
        if (wg < nblk)
            for (int j = jl; j >= 0 && ok; j -= G) ok = tri_backward_block<T>(A, j, nblk, stamp);
        if (threadIdx.x == 0) s_ok = ok ? 1 : 0;
    }
    __syncthreads();
    const bool ok = s_ok != 0;
    __syncthreads();
    return ok;
}

// ---------------------------------------------------------------------------
// REF: precision form (T = double only): M is rounded to FP32 and factored on
// the grid; each Newton solve is refined in FP64 (refine_residual); a zero
// FP32 pivot or a refinement that misses its target within refine_max_iter
// sweeps switches the step to the FP64 LU of M.  refine_max_iter == 0: one
// FP32 solve per Newton iteration, x = 0 + (double)d with no residual (the
// modified-Newton iteration, whose residual is FP64, refines it), in the
// plain path's Newton loop (no grid barrier per iteration).
template <int RPT, int NBC, typename T, int REF = 0>
__global__ void __launch_bounds__(NT) solve_kernel(SolveArgs A)
{
    // REF 1: FP32 factors; REF 2: FP32 factors with the trailing updates on
    // the BF16 matrix cores
    static_assert(!REF || sizeof(T) == 8, "the precision form runs the FP64 loop");
    // panel width of the precision form's FP32 factors: NBC, as for the FP64
    // LU.  (16-column FP32 panels above n = 2048 hold 16 rows x 16 columns
    // per thread and spill: 52 ms per 4096 x 4096 factorization, against
    // 28 ms with 8-column panels and the vector update.)
    constexpr int kNbc32 = NBC;
    extern __shared__ __align__(16) unsigned char dyn[];
    __shared__ Shared sh;
    const int tid = threadIdx.x;
    const int wg = blockIdx.x;
    const int G = A.G;
    const int S = A.m.S;
    const int n = S;
    unsigned int gen_local = 0;
    Ctl *ctl = A.ctl;

    // ---- control-group views ---------------------------------------------
    CtlViewT<T> v;
    T *gv = reinterpret_cast<T *>(A.gvec);
    const size_t Sz = (size_t)S;
    if (A.lds_model) {
        const int nh = A.m.hptr[S];
        HRec *lrec = reinterpret_cast<HRec *>(dyn);
        int *lhptr = reinterpret_cast<int *>(dyn + sizeof(HRec) * (size_t)nh);
        T *lv = reinterpret_cast<T *>(dyn + ((sizeof(HRec) * (size_t)nh + 4 * (Sz + 1) + 15) & ~size_t(15)));
        v.recs = lrec;
        v.hptr = lhptr;
        v.y1 = lv;
        v.ua = lv + Sz;
        v.ub = lv + 2 * Sz;
        v.k1 = lv + 3 * Sz;
        v.k2 = lv + 4 * Sz;
        v.k3 = lv + 5 * Sz;
        v.k4 = lv + 6 * Sz;
        v.k5 = lv + 7 * Sz;
        v.w1 = lv + 8 * Sz;
        v.yv = lv + 9 * Sz;
        v.x = lv + 10 * Sz;
        if (wg == 0) {
            for (int e = tid; e < nh; e += NT) lrec[e] = A.recs[e];
            for (int s = tid; s <= S; s += NT) lhptr[s] = A.m.hptr[s];
        }
    } else {
        v.recs = A.recs;
        v.hptr = A.m.hptr;
        v.y1 = gv;
        v.ua = gv + Sz;
        v.ub = gv + 2 * Sz;
        v.k1 = gv + 3 * Sz;
        v.k2 = gv + 4 * Sz;
        v.k3 = gv + 5 * Sz;
        v.k4 = gv + 6 * Sz;
        v.k5 = gv + 7 * Sz;
        v.w1 = gv + 8 * Sz;
        v.yv = gv + 9 * Sz;
        v.x = reinterpret_cast<T *>(dyn);  // the solve vector lives in LDS
    }
    T *gyv = gv + 9 * Sz;  // y_v as the other groups read it
    T *gy1 = gv;

    // state of the control loop, held identically by every thread of group 0
    T t = 0.0, t1 = (T)A.tv[0], dt = (T)A.rkf_step, newdt = (T)A.be_step;
    const double t_max = (T)A.tv[A.nt - 1];
    int i_sample = 0;
    long long nrej = 0;
    long long st_loop = 0, st_att = 0, st_acc = 0, st_bdf = 0, st_newton = 0, st_lufail = 0, st_rep = 0,
              st_lu = 0, st_ref = 0, st_fb = 0;
    unsigned int rseq = 0;  // refinement sweeps seen (identical in every group)
    int rc = 0;
    // device-time attribution (group 0, 100 MHz wall clock): control/RKF,
    // Jacobian assembly, LU, Newton solves
    unsigned long long ph[4] = {0, 0, 0, 0};
    unsigned long long tmark = 0;

    if (wg == 0)
        for (int s = tid; s < S; s += NT) v.y1[s] = A.x0[s];
    __syncthreads();

    unsigned int parity = 0;
    unsigned int tri_stamp = 0;  // Newton iterations seen by this group (identical in every participant)
    while (true) {
        parity ^= 1;
        // ---- control step (group 0) -----------------------------------------
        if (wg == 0) {
            tmark = wall_clock64();
            int mode = MODE_CONTINUE;
            T hb = 0.0;
            if ((A.max_iters > 0 && st_loop >= A.max_iters) || (A.max_newton > 0 && st_newton >= A.max_newton)) {
                mode = MODE_DONE;
                rc = 1;
            } else {
                if (tid == 0 && A.progress && st_loop < A.progress_cap) A.progress[st_loop] = t / t_max;
                st_loop++;
                bool finished = false;
                if (fabs(t1 - t) < 1e-5) {
                    for (int j = tid; j < A.ncs; j += NT)
                        A.out[(size_t)i_sample * A.ncs + j] = v.y1[A.cs[j]];
                    i_sample++;
                    if (i_sample == A.nt)
                        finished = true;
                    else
                        t1 = (T)A.tv[i_sample];
                }
                if (finished) {
                    mode = MODE_DONE;
                } else {
                    st_att++;
                    double mind;  // the smallest step factor (a T value)
                    const int nok = rkf_attempt_wg(A.m, A.rtol, v, dt, &mind, sh);
                    if (nok == S) {
                        nrej = 0;
                        st_acc++;
                        for (int s = tid; s < S; s += NT) v.y1[s] = v.w1[s];
                        __syncthreads();
                        const T d_temp = dt * (T)mind;
                        if ((t + d_temp) > t1)
                            dt = fabs(t1 - t);
                        else
                            dt = d_temp;
                        t = t + dt;
                    } else {
                        const T dt_old = dt;
                        dt = dt * (T)mind;
                        nrej++;
                        if (dt < (T)A.stiffness_tol || isnan(dt) || isinf(dt) || fabs(dt_old - dt) < 1e-3 ||
                            (A.max_rkf_retries > 0 && nrej >= A.max_rkf_retries)) {
                            nrej = 0;
                            st_bdf++;
                            dt = newdt;
                            if ((t + dt) > t1) dt = fabs(t1 - t);
                            hb = dt;
                            mode = (A.newton_iter > 0) ? MODE_BDF : MODE_BDF_NOOP;
                            // y_v = y1; Newton iteration 0 residual at y_v
                            for (int s = tid; s < S; s += NT) {
                                const T yv = v.y1[s];
                                v.yv[s] = yv;
                                gyv[s] = yv;
                                const T f = rhs_species(A.m, v.recs, v.hptr, (const T *)v.y1, s);
                                v.x[s] = -(yv - v.y1[s] - hb * f);
                            }
                            __syncthreads();
                        }
                    }
                }
            }
            if (tid == 0) {
                st_rlx_d(&ctl->h[parity], hb);
                st_rlx(&ctl->mode[parity], mode);
                st_rlx(&ctl->info, 0);
            }
            const unsigned long long now = wall_clock64();
            ph[0] += now - tmark;
            tmark = now;
        }
        if (G > 1) {
            if (!grid_barrier(A.gs, G, gen_local, &ctl->abort)) return;
        } else {
            __syncthreads();
        }
        const int mode = ld_rlx(&ctl->mode[parity]);
        if (mode == MODE_DONE) break;
        if (mode == MODE_CONTINUE) continue;
        const T h = (T)ld_rlx_d(&ctl->h[parity]);

        bool singular = false;
        bool fp64 = true;  // REF: this step solves with the FP64 factors of M
        if (mode == MODE_BDF) {
            // ---- dense M = I - hJ(y_v), all groups -------------------------------
            {
                // zero fill: 16-byte stores, then the elements left over
                const size_t nn = (size_t)n * n;
                const size_t stride = (size_t)G * NT;
                const size_t n16 = nn * sizeof(T) / 16;
                uint4 *M16 = reinterpret_cast<uint4 *>(A.M);
                const uint4 z = make_uint4(0u, 0u, 0u, 0u);
                for (size_t e = ((size_t)wg * NT + tid); e < n16; e += stride) M16[e] = z;
                T *Mt = reinterpret_cast<T *>(A.M);
                if (wg == 0)
                    for (size_t e = n16 * 16 / sizeof(T) + tid; e < nn; e += NT) Mt[e] = (T)0.0;
            }
            if (G > 1) {
                if (!grid_barrier(A.gs, G, gen_local, &ctl->abort)) return;
            } else {
                __syncthreads();
            }
            for (int r = wg * NT + tid; r < n; r += G * NT)
                jac_row<false>(A.m, (const T *)gyv, reinterpret_cast<T *>(A.M), n, r, h);
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
            if constexpr (REF) {
                // M32 = (float)M, then the FP32 LU on the grid
                const size_t nn = (size_t)n * n;
                const double *Md = A.M;
                for (size_t e = (size_t)wg * NT + tid; e < nn; e += (size_t)G * NT) A.M32[e] = (float)Md[e];
                if (!grid_barrier(A.gs, G, gen_local, &ctl->abort)) return;
                SolveArgs A32 = A;
                A32.M = reinterpret_cast<double *>(A.M32);
                A32.lds_strip = 0;
                if (!lu_factor_grid<RPT, kNbc32, float, REF == 2>(A32, wg, sh, gen_local, singular)) return;
                fp64 = singular;
                if (singular) {
                    // zero FP32 pivot: the FP64 LU of M for this step.  Every
                    // group has read ctl->info before it is cleared.
                    if (!grid_barrier(A.gs, G, gen_local, &ctl->abort)) return;
                    if (wg == 0 && tid == 0) st_rlx(&ctl->info, 0);
                    if (!grid_barrier(A.gs, G, gen_local, &ctl->abort)) return;
                    st_fb++;
                    if (!lu_factor_grid<RPT, NBC, T>(A, wg, sh, gen_local, singular)) return;
                }
            } else {
                if (!lu_factor_grid<RPT, NBC, T>(A, wg, sh, gen_local, singular)) return;
            }
            if (wg == 0) {
                const unsigned long long now = wall_clock64();
                ph[2] += now - tmark;
                tmark = now;
            }
        }
        // ---- Newton loop -------------------------------------------------------
        // Group 0 forms each residual and updates the iterate; the triangular
        // solves are spread over the owners of the row blocks (tri_*_block).
        bool luFail = false;
        // REF with refine_max_iter == 0 runs the plain loop below with the
        // step's factors (FP32, or FP64 after a zero FP32 pivot)
        bool refined = false;
        if constexpr (REF) refined = A.refine_max_iter != 0;
        if constexpr (REF) {
            if (mode == MODE_BDF && refined) {
                // Every group takes part: group 0 forms the residual and updates
                // the iterate; the FP32 solves, the refinement sweeps and any
                // FP64 solves are spread over all groups.
                double *rhs64 = A.rvec;
                double *rres = A.rvec + Sz;
                double *xbuf[2] = {A.rvec + 2 * Sz, A.rvec + 3 * Sz};
                SolveArgs A32 = A;
                A32.M = reinterpret_cast<double *>(A.M32);
                if (wg == 0) {
                    st_lu++;
                    // the row order of the current factors: FP32 (kNbc32-column
                    // panels), or FP64 after a zero FP32 pivot
                    if (!singular) {
                        if (fp64)
                            compose_rows<NBC>(A.perm_src, A.moved_to, n, A.perm_full);
                        else
                            compose_rows<kNbc32>(A.perm_src, A.moved_to, n, A.perm_full);
                    }
                }
                int itN = 0;
                T errN = 1000;
                while (true) {
                    if (wg == 0) {
                        int cont = (itN < A.newton_iter && errN > (T)A.newton_tol) ? 1 : 0;
                        if (cont && itN > 0) {
                            for (int s = tid; s < S; s += NT) {
                                const T f = rhs_species(A.m, v.recs, v.hptr, (const T *)v.yv, s);
                                v.x[s] = -(v.yv[s] - v.y1[s] - h * f);
                            }
                            __syncthreads();
                        }
                        if (cont && singular) {
                            for (int s = tid; s < S; s += NT)
                                if (bad_value(v.yv[s])) v.yv[s] = v.y1[s];
                            __syncthreads();
                            newdt = h * 0.1;
                            luFail = true;
                            st_lufail++;
                            cont = 0;
                        }
                        if (cont)
                            for (int s = tid; s < S; s += NT) rhs64[s] = v.x[s];
                        if (tid == 0) st_rlx(&ctl->ncont, cont);
                    }
                    if (!grid_barrier(A.gs, G, gen_local, &ctl->abort)) return;
                    if (!ld_rlx(&ctl->ncont)) break;
                    const double *xsol = nullptr;
                    if (!fp64) {
                        for (int it = 0; it < A.refine_max_iter; it++) {
                            tri_stamp++;
                            rseq++;
                            st_ref++;
                            if (!tri_solve_all<float>(A32, wg, tri_stamp, it == 0 ? rhs64 : rres)) return;
                            if (!grid_barrier(A.gs, G, gen_local, &ctl->abort)) return;
                            double *xn = xbuf[(it + 1) & 1];
                            refine_residual(A, wg, rhs64, it == 0 ? nullptr : xbuf[it & 1], xn, rres, rseq);
                            if (!grid_barrier(A.gs, G, gen_local, &ctl->abort)) return;
                            if (__hip_atomic_load(&ctl->rbad, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != rseq) {
                                xsol = xn;
                                break;
                            }
                        }
                        if (!xsol) {
                            // no convergence: the FP64 LU of M for the rest of the step
                            fp64 = true;
                            st_fb++;
                            if (!lu_factor_grid<RPT, NBC, T>(A, wg, sh, gen_local, singular)) return;
                            if (singular) {
                                if (wg == 0) {
                                    for (int s = tid; s < S; s += NT)
                                        if (bad_value(v.yv[s])) v.yv[s] = v.y1[s];
                                    __syncthreads();
                                    newdt = h * 0.1;
                                    luFail = true;
                                    st_lufail++;
                                }
                                break;
                            }
                            if (wg == 0) compose_rows<NBC>(A.perm_src, A.moved_to, n, A.perm_full);
                            if (!grid_barrier(A.gs, G, gen_local, &ctl->abort)) return;
                        }
                    }
                    if (fp64) {
                        // the last block of the backward chain belongs to group 0,
                        // which has seen every other block's stamp when it ends
                        tri_stamp++;
                        if (!tri_solve_all<T>(A, wg, tri_stamp, rhs64)) return;
                        xsol = reinterpret_cast<const double *>(A.tri) + 2 * Sz;
                    }
                    if (wg == 0) {
                        __threadfence();
                        T part = 0.0;
                        for (int s = tid; s < S; s += NT) {
                            const T yv1 = v.yv[s];
                            const T fv = (T)A.m.feed[s];
                            T yv;
                            if (fv == 0)
                                yv = yv1 + xsol[s];
                            else
                                yv = fv;
                            v.yv[s] = yv;
                            const T d = yv - yv1;
                            part = part + norm_square(d);
                        }
                        errN = sqrt_t(wg_tree_sum(part, sh));
                        itN = itN + 1;
                    }
                }
                if (wg == 0) st_newton += itN;
            }
        }
        if (mode == MODE_BDF && !refined) {
            const int nblk = (n + kTriBlock - 1) / kTriBlock;
            // REF (refine_max_iter == 0): the FP32 factors in A32, b = (float)
            // of the permuted residual, x = 0 + (double)(FP32 solution)
            bool use32 = false;
            SolveArgs A32 = A;
            if constexpr (REF) {
                use32 = !fp64;
                A32.M = reinterpret_cast<double *>(A.M32);
            }
            if (wg == 0) {
                st_lu++;
                // final row order of the factors: the panels' row moves composed
                if (!singular) compose_rows<NBC>(A.perm_src, A.moved_to, n, A.perm_full);
            }
            if (wg < nblk) {
                __shared__ int s_cont;
                int itN = 0;
                T errN = 1000;
                while (true) {
                    tri_stamp++;
                    if (wg == 0) {
                        int cont = (itN < A.newton_iter && errN > (T)A.newton_tol) ? 1 : 0;
                        if (cont && itN > 0) {
                            for (int s = tid; s < S; s += NT) {
                                const T f = rhs_species(A.m, v.recs, v.hptr, (const T *)v.yv, s);
                                v.x[s] = -(v.yv[s] - v.y1[s] - h * f);
                            }
                            __syncthreads();
                        }
                        if (cont && singular) {
                            for (int s = tid; s < S; s += NT)
                                if (bad_value(v.yv[s])) v.yv[s] = v.y1[s];
                            __syncthreads();
                            newdt = h * 0.1;
                            luFail = true;
                            st_lufail++;
                            cont = 0;
                        }
                        if (cont) {
                            if (use32)
                                for (int s = tid; s < S; s += NT)
                                    reinterpret_cast<float *>(A.tri)[s] = (float)v.x[A.perm_full[s]];
                            else
                                for (int s = tid; s < S; s += NT) reinterpret_cast<T *>(A.tri)[s] = v.x[A.perm_full[s]];
                        }
                        __syncthreads();
                        if (tid == 0) {
                            __threadfence();
                            st_rlx(&ctl->ncont, cont);
                            __threadfence();
                            __hip_atomic_store(&ctl->ngo, tri_stamp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                            s_cont = cont;
                        }
                    } else if (tid == 0) {
                        int cont = -1;
                        if (wait_stamp(&ctl->ngo, tri_stamp, &ctl->abort)) cont = ld_rlx(&ctl->ncont);
                        s_cont = cont;
                    }
                    __syncthreads();
                    const int cont = s_cont;
                    __syncthreads();
                    if (cont < 0) return;
                    if (cont == 0) break;
                    __shared__ int s_ok;
                    if (tid < 64) {
                        bool ok = true;
                        int jl = wg;
                        while (jl + G < nblk) jl += G;
                        bool done = false;
                        if constexpr (REF) {
                            if (use32) {
                                for (int j = wg; j < nblk && ok; j += G) ok = tri_forward_block<float>(A32, j, tri_stamp);
                                for (int j = jl; j >= 0 && ok; j -= G)
                                    ok = tri_backward_block<float>(A32, j, nblk, tri_stamp);
                                done = true;
                            }
                        }
                        if (!done) {
                            for (int j = wg; j < nblk && ok; j += G) ok = tri_forward_block<T>(A, j, tri_stamp);
                            for (int j = jl; j >= 0 && ok; j -= G) ok = tri_backward_block<T>(A, j, nblk, tri_stamp);
                        }
                        if (tid == 0) s_ok = ok ? 1 : 0;
                    }
                    __syncthreads();
                    if (!s_ok) return;
                    if (use32) st_ref++;
                    if (wg == 0) {
                        __threadfence();
                        const T *xs = reinterpret_cast<const T *>(A.tri) + 2 * (size_t)n;
                        const float *xs32 = reinterpret_cast<const float *>(A.tri) + 2 * (size_t)n;
                        T part = 0.0;
                        for (int s = tid; s < S; s += NT) {
                            const T yv1 = v.yv[s];
                            const T fv = (T)A.m.feed[s];
                            const T xv = use32 ? (T)(0.0 + (double)xs32[s]) : xs[s];
                            T yv;
                            if (fv == 0)
                                yv = yv1 + xv;
                            else
                                yv = fv;
                            v.yv[s] = yv;
                            const T d = yv - yv1;
                            part = part + norm_square(d);
                        }
                        errN = sqrt_t(wg_tree_sum(part, sh));
                        itN = itN + 1;
                    }
                }
                if (wg == 0) st_newton += itN;
            }
        }
        if (wg == 0) {
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
                    newdt = (T)fmin((double)newdt * 2.0, (double)(T)A.be_step);
                }
            }
            __syncthreads();
            for (int s = tid; s < S; s += NT) v.y1[s] = v.yv[s];
            __syncthreads();
            t = t + h;
            ph[3] += wall_clock64() - tmark;
        }
    }
    if (wg == 0) {
        if (A.lds_model)
            for (int s = tid; s < S; s += NT) gy1[s] = v.y1[s];
        if (sizeof(T) == 4)
            for (int s = tid; s < S; s += NT) A.yfin[s] = v.y1[s];
        if (tid == 0) {
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
            o[10] = st_ref;
            o[11] = st_fb;
            double *d = A.dstats_out;
            d[0] = t;
            d[1] = dt;
            d[2] = newdt;
            for (int q = 0; q < 4; q++) d[3 + q] = (double)ph[q] * 1e-5;  // 100 MHz ticks -> ms
        }
    }
}

// ---------------------------------------------------------------------------
size_t lds_model_bytes(int S, int nnz_h)
{
    const size_t head = (sizeof(HRec) * (size_t)nnz_h + 4 * ((size_t)S + 1) + 15) & ~size_t(15);
    return head + sizeof(double) * kNumVecLds * (size_t)S;
}

template <int RPT, int NBC>
static const void *kernel_ptr(int fp32, int refine)
{
    if (fp32) return reinterpret_cast<const void *>(&solve_kernel<RPT, NBC, float>);
    if (refine == 2) return reinterpret_cast<const void *>(&solve_kernel<RPT, NBC, double, 2>);
    if (refine) return reinterpret_cast<const void *>(&solve_kernel<RPT, NBC, double, 1>);
    return reinterpret_cast<const void *>(&solve_kernel<RPT, NBC, double>);
}

static const void *pick_kernel(int n, int fp32, int refine)
{
    if (n <= 256) return kernel_ptr<1, 16>(fp32, refine);
    if (n <= 512) return kernel_ptr<2, 16>(fp32, refine);
    if (n <= 1024) return kernel_ptr<4, 16>(fp32, refine);
    if (n <= 2048) return kernel_ptr<8, 16>(fp32, refine);
    return kernel_ptr<16, 8>(fp32, refine);
}

int dense_panel_width(int n) { return n <= 2048 ? 16 : 8; }

int solver_occupancy(int n, int fp32, int refine, size_t dyn_lds, int *per_cu)
{
    const void *k = pick_kernel(n, fp32, refine);
    if (dyn_lds > 64 * 1024 &&
        hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn_lds) != hipSuccess)
        return -1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, k, NT, dyn_lds) != hipSuccess) return -1;
    return 0;
}

hipError_t launch_solve(const SolveArgs &args, size_t dyn_lds, hipStream_t stream)
{
    SolveArgs a = args;
    void *params[] = {&a};
    return hipLaunchCooperativeKernel(pick_kernel(a.m.S, a.fp32, a.refine), dim3(a.G), dim3(NT), params,
                                      (unsigned int)dyn_lds, stream);
}

}  // namespace lassie
