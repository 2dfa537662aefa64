// Workgroup-level building blocks shared by the single-model and batched
// solvers: LDS scratch, fixed-order reductions, the RKF45 attempt and the
// Jacobian row assembly.  Same operation order as the reference kernels.
#pragma once
#include <type_traits>
#include "device_common.h"

namespace lassie {

struct Shared {
    double red[NT];
    int ired[NT];
    double cand_row[2][NWAVES][16];  // pivot candidates, one per wave
    double cand_v[2][NWAVES];
    int cand_p[2][NWAVES];
    int qcount[16][NWAVES];  // panel row-list counts per (row slot, wave)
    double U12[16][16];
    double L11[16][16];
    double oldP[16][16];
    int iscal[4];
    int rowcnt;
    double prep[16];  // panel: the representative row of the all-zero rows
};

__device__ __forceinline__ float to_bf16_rne(float x)
{
    // round to the nearest BF16 value (ties to even), kept in an FP32 register
    unsigned int u = __float_as_uint(x);
    if ((u & 0x7f800000u) == 0x7f800000u) return x;  // inf / nan unchanged
    u += 0x7fffu + ((u >> 16) & 1u);
    return __uint_as_float(u & 0xffff0000u);
}

// Pointers used by the control group: LDS when the model fits, else global.
// T is the state type (double; float for the single-precision loop).
template <typename T>
struct CtlViewT {
    const HRec *recs;
    const int *hptr;
    T *y1, *ua, *ub, *k1, *k2, *k3, *k4, *k5, *w1, *yv, *x;
};
typedef CtlViewT<double> CtlView;

// square root in the state type (correctly rounded for double and float)
__device__ __forceinline__ double sqrt_t(double x) { return sqrt(x); }
__device__ __forceinline__ float sqrt_t(float x) { return sqrtf(x); }

// 16x16x4 matrix-core step D = A B + C for FP64 and FP32 operands.
template <typename T>
struct Mfma;
// Both take A lane l -> A[l%16][l/16] and B lane l -> B[l/16][l%16]; item i
// of lane l holds D[row(l/16, i)][l%16], whose row order differs between the
// two instructions (measured on gfx950).
template <>
struct Mfma<double> {
    typedef double acc __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ acc run(double a, double b, acc c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lq, int i) { return lq + 4 * i; }
};
template <>
struct Mfma<float> {
    typedef float acc __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ acc run(float a, float b, acc c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lq, int i) { return 4 * lq + i; }
};

// Raw buffer access to one matrix: a load whose byte offset is kBufOob (at
// or past the buffer's size) returns zero and a store to it is dropped, so
// masked lanes need no branches, and the memory counters stay exact.  The
// range check covers the per-lane offset only (the scalar offset is added
// unchecked), so every out-of-range access puts kBufOob in the lane offset.
constexpr int kBufOob = 0x7ffffff0;
typedef unsigned int u32x2_t __attribute__((ext_vector_type(2)));
// A pointer every lane of the wave holds, moved to scalar registers.  Inside
// a called function arguments arrive in vector registers, and a buffer
// resource or scalar offset the compiler cannot prove uniform is wrapped in a
// loop over the distinct lane values; these make the uniformity explicit.
template <typename P>
__device__ __forceinline__ P *uniform_ptr(P *p)
{
    const uintptr_t u = reinterpret_cast<uintptr_t>(p);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)u);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(u >> 32));
    return reinterpret_cast<P *>(((uintptr_t)hi << 32) | lo);
}
__device__ __forceinline__ int uniform_int(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void *p, int bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(uniform_ptr(p)), (short)0, uniform_int(bytes),
                                             0x00020000);
}
template <typename T>
struct BufIO;
template <>
struct BufIO<double> {
    static __device__ __forceinline__ double ld(__amdgpu_buffer_rsrc_t rs, int off, int soff = 0)
    {
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rs, off, soff, 0));
    }
    // v is a named double: __builtin_bit_cast of a vector element (d[i]) reads
    // element 0 in this compiler, so callers pass elements through a value
    static __device__ __forceinline__ void st(__amdgpu_buffer_rsrc_t rs, int off, double v, int soff = 0)
    {
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2_t, v), rs, off, soff, 0);
    }
};
template <>
struct BufIO<float> {
    static __device__ __forceinline__ float ld(__amdgpu_buffer_rsrc_t rs, int off, int soff = 0)
    {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, off, soff, 0));
    }
    static __device__ __forceinline__ void st(__amdgpu_buffer_rsrc_t rs, int off, float v, int soff = 0)
    {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), rs, off, soff, 0);
    }
};

// ---------------------------------------------------------------------------
// workgroup reductions (all threads receive the result)

// sum in the fixed "tid strided, then halving tree" order used for the Newton
// norm (Simulation.cuh:883-888 with a float-returning square), in the
// arithmetic of T (thrust::plus<T>).
template <typename T>
__device__ T wg_tree_sum(T v, Shared &sh)
{
    const int tid = threadIdx.x;
    sh.red[tid] = v;
    __syncthreads();
    for (int stride = NT / 2; stride > 0; stride >>= 1) {
        if (tid < stride) sh.red[tid] = (T)sh.red[tid] + (T)sh.red[tid + stride];
        __syncthreads();
    }
    const T r = (T)sh.red[0];
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
// other NaNs are ignored).  T is the state type: stage values, k_i, w1, w2,
// R and the step factors are T; the stage combinations have double literals
// and are evaluated in double, as in the reference kernels.
template <typename T>
__device__ int rkf_attempt_wg(const DevModel &m, double rtol, const CtlViewT<T> &v, T dt, double *mind,
                              Shared &sh)
{
    const int S = m.S;
    const int tid = threadIdx.x;
    const T *y1 = v.y1;
    T *ua = v.ua, *ub = v.ub;
    T *k1 = v.k1, *k2 = v.k2, *k3 = v.k3, *k4 = v.k4, *k5 = v.k5;
    const double *fd = m.feed;

    for (int s = tid; s < S; s += NT) {
        const T f = rhs_species(m, v.recs, v.hptr, y1, s);
        const T fv = (T)fd[s];
        if (fv != 0.0) {
            k1[s] = fv;
            ua[s] = fv;
        } else {
            const T kk = f * dt;
            k1[s] = kk;
            ua[s] = y1[s] + kk / 4.0;
        }
    }
    __syncthreads();
    for (int s = tid; s < S; s += NT) {
        const T f = rhs_species(m, v.recs, v.hptr, (const T *)ua, s);
        const T fv = (T)fd[s];
        if (fv != 0.0) {
            k2[s] = fv;
            ub[s] = fv;
        } else {
            const T kk = f * dt;
            k2[s] = kk;
            ub[s] = y1[s] + (3.0 * k1[s]) / 32.0 + (9.0 * kk) / 32.0;
        }
    }
    __syncthreads();
    for (int s = tid; s < S; s += NT) {
        const T f = rhs_species(m, v.recs, v.hptr, (const T *)ub, s);
        const T fv = (T)fd[s];
        if (fv != 0.0) {
            k3[s] = fv;
            ua[s] = fv;
        } else {
            const T kk = f * dt;
            k3[s] = kk;
            ua[s] = y1[s] + (1932.0 * k1[s]) / 2197.0 - (7200.0 * k2[s]) / 2197.0 + (7296.0 * kk) / 2197.0;
        }
    }
    __syncthreads();
    for (int s = tid; s < S; s += NT) {
        const T f = rhs_species(m, v.recs, v.hptr, (const T *)ua, s);
        const T fv = (T)fd[s];
        if (fv != 0.0) {
            k4[s] = fv;
            ub[s] = fv;
        } else {
            const T kk = f * dt;
            k4[s] = kk;
            ub[s] = y1[s] + (439.0 * k1[s]) / 216.0 - (8.0 * k2[s]) + (3680.0 * k3[s]) / 513.0
                    - (845.0 * kk) / 4104.0;
        }
    }
    __syncthreads();
    for (int s = tid; s < S; s += NT) {
        const T f = rhs_species(m, v.recs, v.hptr, (const T *)ub, s);
        const T fv = (T)fd[s];
        if (fv != 0.0) {
            k5[s] = fv;
            ua[s] = fv;
        } else {
            const T kk = f * dt;
            k5[s] = kk;
            ua[s] = y1[s] - (8.0 * k1[s]) / 27.0 + (2.0 * k2[s]) - (3544.0 * k3[s]) / 2565.0
                    + (1859.0 * k4[s]) / 4104.0 - (11.0 * kk) / 40.0;
        }
    }
    __syncthreads();
    int nok = 0;
    double dmin = INFINITY;
    for (int s = tid; s < S; s += NT) {
        const T f = rhs_species(m, v.recs, v.hptr, (const T *)ua, s);
        const T fv = (T)fd[s];
        T w1, w2;
        if (fv != 0.0) {
            w1 = fv;
            w2 = fv;
        } else {
            const T kk6 = f * dt;
            w1 = y1[s] + (25.0 * k1[s]) / 216.0 + (1408.0 * k3[s]) / 2565.0 + (2197.0 * k4[s]) / 4104.0
                 - k5[s] / 5.0;
            w2 = y1[s] + (16.0 * k1[s]) / 135.0 + (6656.0 * k3[s]) / 12825.0
                 + (28561.0 * k4[s]) / 56430.0 - (9.0 * k5[s] / 50.0) + (2.0 * kk6) / 55.0;
        }
        v.w1[s] = w1;
        const T R = fabs(w1 - w2) / dt;
        const T at = (T)m.atol[s];
        const T tol = (rtol > 0.0) ? (T)(at + rtol * fabs(w1)) : at;
        const T d = 0.84 * sqrt_t(sqrt_t(tol / R));
        if (s == 0) sh.iscal[0] = isnan(d) ? 1 : 0;
        dmin = fmin(dmin, (double)d);
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
// T is the state and matrix type; each term is accumulated in double and
// added to the T entry, and m.hk holds sign * k formed in T.
template <bool ROW_MAJOR, typename T>
__device__ void jac_row_rmw(const DevModel &m, const T *__restrict__ x, T *__restrict__ M, int n, int i, T h)
{
    auto at = [&](int row, int col) -> T * {
        return ROW_MAJOR ? M + (size_t)row * n + col : M + (size_t)col * n + row;
    };
    const int e0 = m.hptr[i], e1 = m.hptr[i + 1];
    for (int e = e0; e < e1; e++) {
        const int r = m.hrx[e];
        const T sk = (T)m.hk[e];
        const int a0 = m.lptr[r], a1 = m.lptr[r + 1];
        for (int j = a0; j < a1; j++) {
            const int g = m.lsp[j];
            const int ex = m.lexp[j];
            double temp = 0.0;
            if (ex == 1)
                temp = temp + sk;
            else if (ex > 1)
                temp = temp + (sk * (T)ex) * mypow_t(x[g], ex - 1);
            else
                continue;
            for (int jj = a0; jj < a1; jj++)
                if (m.lsp[jj] != g) temp = temp * x[m.lsp[jj]];
            T *p = at(i, g);
            *p = (T)(*p + temp);
        }
    }
    T *d = at(i, i);
    *d = (T)1.0 - h * *d;
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
            T *p = at(i, g);
            *p = (T)0.0 - h * *p;
        }
    }
}

// Row i of M = I - h J (createJac_ and aggiorna_Jacobiano1, kernels.cuh:205-271,
// 393-402) with the row's entries accumulated in registers: each column
// takes the same additions in the same order as jac_row_rmw's global
// read-modify-writes (starting from the zeroed matrix's +0), then one store
// per entry: 1 - h v on the diagonal (always stored), 0 - h v elsewhere.  M
// must be zero in row i.  A row with more than kJacCols distinct columns
// takes jac_row_rmw.
constexpr int kJacCols = 16;

template <bool ROW_MAJOR, typename T>
__device__ void jac_row(const DevModel &m, const T *__restrict__ x, T *__restrict__ M, int n, int i, T h)
{
    auto at = [&](int row, int col) -> T * {
        return ROW_MAJOR ? M + (size_t)row * n + col : M + (size_t)col * n + row;
    };
    int col[kJacCols];
    T val[kJacCols];
#pragma unroll
    for (int k = 0; k < kJacCols; k++) {
        col[k] = -1;
        val[k] = T(0);
    }
    int cnt = 0;
    bool over = false;
    const int e0 = m.hptr[i], e1 = m.hptr[i + 1];
    for (int e = e0; e < e1 && !over; e++) {
        const int r = m.hrx[e];
        const T sk = (T)m.hk[e];
        const int a0 = m.lptr[r], a1 = m.lptr[r + 1];
        for (int j = a0; j < a1; j++) {
            const int g = m.lsp[j];
            const int ex = m.lexp[j];
            double temp = 0.0;
            if (ex == 1)
                temp = temp + sk;
            else if (ex > 1)
                temp = temp + (sk * (T)ex) * mypow_t(x[g], ex - 1);
            else
                continue;
            for (int jj = a0; jj < a1; jj++)
                if (m.lsp[jj] != g) temp = temp * x[m.lsp[jj]];
            int hit = -1;
#pragma unroll
            for (int k = 0; k < kJacCols; k++) hit = (col[k] == g) ? k : hit;
            if (hit < 0) {
                if (cnt == kJacCols) {
                    over = true;
                    break;
                }
                hit = cnt++;
            }
            T cur = T(0);
#pragma unroll
            for (int k = 0; k < kJacCols; k++) cur = (k == hit) ? val[k] : cur;
            const T nv = (T)(cur + temp);
#pragma unroll
            for (int k = 0; k < kJacCols; k++) {
                val[k] = (k == hit) ? nv : val[k];
                col[k] = (k == hit) ? g : col[k];
            }
        }
    }
    if (over) {
        jac_row_rmw<ROW_MAJOR>(m, x, M, n, i, h);
        return;
    }
    T dv = T(0);
#pragma unroll
    for (int k = 0; k < kJacCols; k++) {
        if (k < cnt && col[k] != i) *at(i, col[k]) = (T)0.0 - h * val[k];
        dv = (k < cnt && col[k] == i) ? val[k] : dv;
    }
    *at(i, i) = (T)1.0 - h * dv;
}

// ---------------------------------------------------------------------------
// Cross-lane helpers without LDS traffic.  DPP moves within 16-lane rows
// (quad_perm, row_half_mirror, row_mirror) and readlane from a uniform lane.
template <int CTRL>
__device__ __forceinline__ double dpp_mov_d(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Pivot choice of one wave: the largest v, then the smallest position p among
// equal v (a total order, so the result does not depend on the reduction
// tree).  Values first (DPP within 16-lane rows, then the four rows by
// readlane), then the smallest position among the lanes holding that value.
// Branch-free; the results are wave-uniform.
template <int CTRL>
__device__ __forceinline__ int dpp_mov_i(int v)
{
    return __builtin_amdgcn_mov_dpp(v, CTRL, 0xf, 0xf, false);
}

__device__ __forceinline__ double readlane_d(double v, int lane)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ void wave_argmax(double v, int p, double &wv, int &wp)
{
    double m = v;
    m = fmax(m, dpp_mov_d<0xB1>(m));   // quad_perm [1,0,3,2]
    m = fmax(m, dpp_mov_d<0x4E>(m));   // quad_perm [2,3,0,1]
    m = fmax(m, dpp_mov_d<0x141>(m));  // row_half_mirror
    m = fmax(m, dpp_mov_d<0x140>(m));  // row_mirror
    m = fmax(fmax(readlane_d(m, 0), readlane_d(m, 16)), fmax(readlane_d(m, 32), readlane_d(m, 48)));
    int q = (v == m) ? p : 0x7fffffff;
    q = min(q, dpp_mov_i<0xB1>(q));
    q = min(q, dpp_mov_i<0x4E>(q));
    q = min(q, dpp_mov_i<0x141>(q));
    q = min(q, dpp_mov_i<0x140>(q));
    q = min(min(__builtin_amdgcn_readlane(q, 0), __builtin_amdgcn_readlane(q, 16)),
            min(__builtin_amdgcn_readlane(q, 32), __builtin_amdgcn_readlane(q, 48)));
    wv = m;
    wp = q;
}

// Sum over the wave in a fixed order (DPP within 16-lane rows, then the four
// row sums by readlane); every lane receives the result.
__device__ __forceinline__ double wave_sum(double v)
{
    v = v + dpp_mov_d<0xB1>(v);
    v = v + dpp_mov_d<0x4E>(v);
    v = v + dpp_mov_d<0x141>(v);
    v = v + dpp_mov_d<0x140>(v);
    return (readlane_d(v, 0) + readlane_d(v, 16)) + (readlane_d(v, 32) + readlane_d(v, 48));
}

// v of a uniform lane, to every lane (float overload; the double one is above)
__device__ __forceinline__ float readlane_d(float v, int lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// ---------------------------------------------------------------------------
// Panel factorization with partial pivoting of columns [r0, r0+nbp), rows
// [r0, n), held in registers: RPT rows per thread, NBC columns.  Pivot: the
// largest |a| among remaining rows, first (lowest position) on ties, a NaN at
// the diagonal position wins (LAPACK idamax / dgetf2 semantics).  Row swaps
// are tracked as positions and applied when the panel is written back.
// Publishes piv[k], perm_src[pos] (original row of the pivot now at pos),
// moved_to[r0+j] (final position of the row that started at r0+j), and the
// positions below the panel whose multipliers are not all zero (rows, rowcnt;
// LIST only): the only rows the trailing update changes.  T is the element type (double
// for the FP64 method, float for the precision form); RM selects row-major
// (else column-major) storage.
// Where the panel's columns live.  MatStore: the n x n matrix itself (row- or
// column-major).  LdsStripStore: an LDS copy of the column strip [c0, c0+NBC)
// (column-major, leading dimension ld); stores also go to the global
// column-major matrix, where the other workgroups read the factored panel.
template <typename T, bool RM>
struct MatStore {
    T *M;
    int n;
    __device__ __forceinline__ T load(int row, int col) const
    {
        return RM ? M[(size_t)row * n + col] : M[(size_t)col * n + row];
    }
    __device__ __forceinline__ void store(int row, int col, T v) const
    {
        if (RM)
            M[(size_t)row * n + col] = v;
        else
            M[(size_t)col * n + row] = v;
    }
};

template <typename T>
struct LdsStripStore {
    typedef __attribute__((address_space(3))) T LT;
    LT *L;
    int ld, c0;
    T *M;
    int n;
    __device__ __forceinline__ T load(int row, int col) const { return L[(col - c0) * ld + row]; }
    __device__ __forceinline__ void store(int row, int col, T v) const
    {
        L[(col - c0) * ld + row] = v;
        M[(size_t)col * n + row] = v;
    }
};

template <int RPT, int NBC, typename T, bool LIST, typename Store>
__device__ void panel_factor_s(const Store st, int n, int p, int *piv, int *perm_src, int *moved_to, int *rows,
                               int *rowcnt, int *info, Shared &sh);

template <int RPT, int NBC, typename T, bool RM, bool LIST = true>
__device__ void panel_factor(T *__restrict__ M, int n, int p, int *piv, int *perm_src, int *moved_to,
                             int *rows, int *rowcnt, int *info, Shared &sh)
{
    panel_factor_s<RPT, NBC, T, LIST>(MatStore<T, RM>{M, n}, n, p, piv, perm_src, moved_to, rows, rowcnt, info,
                                      sh);
}

template <int RPT, int NBC, typename T, bool LIST, typename Store>
__device__ void panel_factor_body(const Store st, int n, int p, int *piv, int *perm_src, int *moved_to, int *rows,
                                  int *rowcnt, int *info, Shared &sh)
{
    const int r0 = p * NBC;
    const int nbp = min(NBC, n - r0);
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    T a[RPT][NBC];
    int pos[RPT];
    bool act[RPT];
#pragma unroll
    for (int q = 0; q < RPT; q++) {
        const int r = r0 + tid + NT * q;
        pos[q] = r;
        act[q] = r < n;
#pragma unroll
        for (int c = 0; c < NBC; c++) a[q][c] = (r < n && c < nbp) ? st.load(r, r0 + c) : T(0);
    }
    int buf = 0;
    bool stop = false;
    int stopk = 0;
    int pk[NBC];  // pivot rows, stored once the panel is done
    // Columns run without control flow around them (a branch per column made
    // the compiler copy the whole register panel at every merge); conditions
    // combine with bitwise operators so they stay lane masks rather than
    // branches.  Padding columns (j >= nbp) and columns after a zero pivot
    // compute values that are never used: they change no position, pivot or
    // stored entry.
#pragma unroll
    for (int j = 0; j < NBC; j++) {
        const bool live = (j < nbp) & !stop;
        const int k = r0 + j;
        double bv = -2.0;
        int bp = 0x7fffffff;
        int bq = -1;
#pragma unroll
        for (int q = 0; q < RPT; q++) {
            const double xv = (double)a[q][j];
            const double v = isnan(xv) ? ((pos[q] == k) ? INFINITY : -1.0) : fabs(xv);
            const bool take = act[q] & ((v > bv) | ((v == bv) & (pos[q] < bp)));
            bv = take ? v : bv;
            bp = take ? pos[q] : bp;
            bq = take ? q : bq;
        }
        double wv;
        int wp;
        wave_argmax(bv, bp, wv, wp);
        if (lane == 0) {
            sh.cand_v[buf][wave] = wv;
            sh.cand_p[buf][wave] = wp;
        }
        if ((bq >= 0) & (bp == wp)) {
#pragma unroll
            for (int q = 0; q < RPT; q++)
                if (q == bq) {
#pragma unroll
                    for (int c = 0; c < NBC; c++) sh.cand_row[buf][wave][c] = (double)a[q][c];
                }
        }
        __syncthreads();
        // all candidates are read before any comparison (one LDS round trip)
        double cv[NWAVES];
        int cp[NWAVES];
#pragma unroll
        for (int ww = 0; ww < NWAVES; ww++) {
            cv[ww] = sh.cand_v[buf][ww];
            cp[ww] = sh.cand_p[buf][ww];
        }
        int w = 0;
        bv = cv[0];
        bp = cp[0];
#pragma unroll
        for (int ww = 1; ww < NWAVES; ww++) {
            const bool take = (cv[ww] > bv) | ((cv[ww] == bv) & (cp[ww] < bp));
            bv = take ? cv[ww] : bv;
            bp = take ? cp[ww] : bp;
            w = take ? ww : w;
        }
        const double *prow = sh.cand_row[buf][w];
        const T ukk = (T)prow[j];
        const bool zero = live & (ukk == T(0));
        stopk = (zero & !stop) ? k : stopk;
        stop = stop | zero;
        const bool go = live & (ukk != T(0));
        pk[j] = go ? bp : -1;
#pragma unroll
        for (int q = 0; q < RPT; q++) {
            const int old = pos[q];
            const bool isp = go & act[q] & (old == bp);
            const bool isk = go & act[q] & (old == k);
            pos[q] = isp ? k : (isk ? bp : old);
            act[q] = act[q] & !isp;
        }
        T pr[NBC];
#pragma unroll
        for (int c = 0; c < NBC; c++) pr[c] = (c > j) ? (T)prow[c] : T(0);
        // multipliers of all rows at once (independent divisions overlap);
        // rows already used as pivots keep their values
#pragma unroll
        for (int q = 0; q < RPT; q++) {
            const T l = a[q][j] / ukk;
            a[q][j] = (go & act[q]) ? l : a[q][j];
        }
#pragma unroll
        for (int q = 0; q < RPT; q++) {
            if (go & act[q]) {
                const T l = a[q][j];
#pragma unroll
                for (int c = 0; c < NBC; c++)
                    if (c > j) a[q][c] = fma(-l, pr[c], a[q][c]);
            }
        }
        buf ^= 1;
    }
    if (tid == 0) {
#pragma unroll
        for (int j = 0; j < NBC; j++)
            if (pk[j] >= 0) piv[r0 + j] = pk[j];
        if (stop) st_rlx(info, stopk + 1);
    }
    if (stop) return;
    if (!LIST) {
#pragma unroll
        for (int q = 0; q < RPT; q++) {
            const int orig = r0 + tid + NT * q;
            if (orig < n) {
#pragma unroll
                for (int c = 0; c < NBC; c++)
                    if (c < nbp) st.store(pos[q], r0 + c, a[q][c]);
                if (!act[q]) perm_src[pos[q]] = orig;
                if (orig < r0 + nbp) moved_to[orig] = pos[q];
            }
        }
        __syncthreads();
        return;
    }
    // rows below the panel with a nonzero multiplier, listed by (q, wave, lane)
    bool nzq[RPT];
#pragma unroll
    for (int q = 0; q < RPT; q++) {
        const int orig = r0 + tid + NT * q;
        bool nz = false;
        if (orig < n && act[q]) {
#pragma unroll
            for (int c = 0; c < NBC; c++) nz = nz || (c < nbp && a[q][c] != T(0));
        }
        nzq[q] = nz;
        const unsigned long long m = __ballot(nz);
        if (lane == 0) sh.qcount[q][wave] = __popcll(m);
    }
    __syncthreads();
    int total = 0, base[RPT];
#pragma unroll
    for (int q = 0; q < RPT; q++)
#pragma unroll
        for (int w = 0; w < NWAVES; w++) {
            if (w == wave) base[q] = total;
            total += sh.qcount[q][w];
        }
#pragma unroll
    for (int q = 0; q < RPT; q++) {
        const int orig = r0 + tid + NT * q;
        const unsigned long long m = __ballot(nzq[q]);
        const int off = __builtin_amdgcn_mbcnt_hi((int)(m >> 32), __builtin_amdgcn_mbcnt_lo((int)m, 0));
        if (orig < n) {
#pragma unroll
            for (int c = 0; c < NBC; c++)
                if (c < nbp) st.store(pos[q], r0 + c, a[q][c]);
            if (!act[q]) perm_src[pos[q]] = orig;
            if (orig < r0 + nbp) moved_to[orig] = pos[q];
            if (nzq[q]) rows[base[q] + off] = pos[q];
        }
    }
    if (tid == 0) *rowcnt = total;
    __syncthreads();
}

__device__ __forceinline__ bool bits_nonzero(double x) { return __double_as_longlong(x) != 0; }
__device__ __forceinline__ bool bits_nonzero(float x) { return __float_as_uint(x) != 0u; }

// panel_factor_body with one row per thread for the rows that take part:
// the C < NT rows listed in sh.ired (every row with an entry other than +0 in
// the panel's columns, and the rows at the panel's diagonal positions), plus
// one representative (thread C) of the other rows.  Those rows hold +0 in
// every panel column; none is a pivot unless a column's largest |a| is 0,
// which stops the factorization, and each takes the same operations, so
// the representative's values are every such row's (it is a pivot candidate
// only where such rows exist, at a position that loses every tie).  Where a
// non-finite pivot row turns the representative into NaN the rows could
// differ: nothing is stored and 0 is returned (the caller factors the panel
// whole).  A zero pivot: piv and info as panel_factor_body leaves them, 2.
// Otherwise the listed rows are stored, the representative's values are
// left in sh.prep for the others, 1.
template <int NBC, typename T, typename Store>
__device__ int panel_factor_listed(const Store st, int n, int p, int C, int *piv, int *perm_src, int *moved_to,
                                    int *info, Shared &sh)
{
    const int r0 = p * NBC;
    const int nbp = min(NBC, n - r0);
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const bool real = tid < C;
    const bool repr = tid == C;
    const int orig = real ? sh.ired[tid] : -1;
    T a[NBC];
#pragma unroll
    for (int c = 0; c < NBC; c++) a[c] = (real && c < nbp) ? st.load(orig, r0 + c) : T(0);
    int pos = real ? orig : 0x7ffffffe;
    bool act = real | repr;
    bool cand = real | (repr & (C < n - r0));
    int buf = 0;
    bool stop = false;
    int stopk = 0;
    int pk[NBC];
#pragma unroll
    for (int j = 0; j < NBC; j++) {
        const bool live = (j < nbp) & !stop;
        const int k = r0 + j;
        const double xv = (double)a[j];
        const double v = isnan(xv) ? ((pos == k) ? INFINITY : -1.0) : fabs(xv);
        const double bv = cand ? v : -2.0;
        const int bp = cand ? pos : 0x7fffffff;
        double wv;
        int wp;
        wave_argmax(bv, bp, wv, wp);
        if (lane == 0) {
            sh.cand_v[buf][wave] = wv;
            sh.cand_p[buf][wave] = wp;
        }
        if (cand & (bp == wp)) {
#pragma unroll
            for (int c = 0; c < NBC; c++) sh.cand_row[buf][wave][c] = (double)a[c];
        }
        __syncthreads();
        double cv[NWAVES];
        int cp[NWAVES];
#pragma unroll
        for (int ww = 0; ww < NWAVES; ww++) {
            cv[ww] = sh.cand_v[buf][ww];
            cp[ww] = sh.cand_p[buf][ww];
        }
        int w = 0;
        double gv = cv[0];
        int gp = cp[0];
#pragma unroll
        for (int ww = 1; ww < NWAVES; ww++) {
            const bool take = (cv[ww] > gv) | ((cv[ww] == gv) & (cp[ww] < gp));
            gv = take ? cv[ww] : gv;
            gp = take ? cp[ww] : gp;
            w = take ? ww : w;
        }
        const double *prow = sh.cand_row[buf][w];
        const T ukk = (T)prow[j];
        const bool zero = live & (ukk == T(0));
        stopk = (zero & !stop) ? k : stopk;
        stop = stop | zero;
        const bool go = live & (ukk != T(0));
        pk[j] = go ? gp : -1;
        const bool isp = go & act & (pos == gp);
        const bool isk = go & act & (pos == k);
        pos = isp ? k : (isk ? gp : pos);
        act = act & !isp;
        cand = cand & !isp;
        T pr[NBC];
#pragma unroll
        for (int c = 0; c < NBC; c++) pr[c] = (c > j) ? (T)prow[c] : T(0);
        const T l = a[j] / ukk;
        a[j] = (go & act) ? l : a[j];
        if (go & act) {
#pragma unroll
            for (int c = 0; c < NBC; c++)
                if (c > j) a[c] = fma(-l, pr[c], a[c]);
        }
        buf ^= 1;
    }
    bool repnan = false;
    if (repr) {
#pragma unroll
        for (int c = 0; c < NBC; c++) {
            sh.prep[c] = (double)a[c];
            repnan = repnan | isnan(a[c]);
        }
    }
    if (__syncthreads_or(repnan)) return 0;
    if (tid == 0) {
#pragma unroll
        for (int j = 0; j < NBC; j++)
            if (pk[j] >= 0) piv[r0 + j] = pk[j];
        if (stop) st_rlx(info, stopk + 1);
    }
    if (stop) return 2;
    if (real) {
#pragma unroll
        for (int c = 0; c < NBC; c++)
            if (c < nbp) st.store(pos, r0 + c, a[c]);
        if (!act) perm_src[pos] = orig;
        if (orig < r0 + nbp) moved_to[orig] = pos;
    }
    return 1;
}

// Panel factorization (panel_factor_body semantics).  With several rows per
// thread (and no row lists), the rows that take part are gathered first:
// when fewer than NT of them remain they are factored one per thread
// (panel_factor_listed), which makes the per-column work that of a panel of
// at most NT rows; the rows of value 0 take the representative's values.
template <int RPT, int NBC, typename T, bool LIST, typename Store>
__device__ void panel_factor_s(const Store st, int n, int p, int *piv, int *perm_src, int *moved_to, int *rows,
                               int *rowcnt, int *info, Shared &sh)
{
    if constexpr (RPT > 1 && !LIST) {
        const int r0 = p * NBC;
        const int nbp = min(NBC, n - r0);
        const int tid = threadIdx.x;
        const int lane = tid & 63;
        const int wave = tid >> 6;
        bool part[RPT];
#pragma unroll
        for (int q = 0; q < RPT; q++) {
            const int r = r0 + tid + NT * q;
            const int rr = min(r, n - 1);  // clamped: loads without branches
            bool nz = r < r0 + nbp;
#pragma unroll
            for (int c = 0; c < NBC; c++) nz = nz | ((c < nbp) & bits_nonzero(st.load(rr, r0 + min(c, nbp - 1))));
            part[q] = (r < n) & nz;
            const unsigned long long m = __ballot(part[q]);
            if (lane == 0) sh.qcount[q][wave] = __popcll(m);
        }
        __syncthreads();
        int total = 0, base[RPT];
#pragma unroll
        for (int q = 0; q < RPT; q++)
#pragma unroll
            for (int w = 0; w < NWAVES; w++) {
                if (w == wave) base[q] = total;
                total += sh.qcount[q][w];
            }
        const int C = __builtin_amdgcn_readfirstlane(total);
        if (C < NT) {
#pragma unroll
            for (int q = 0; q < RPT; q++) {
                const unsigned long long m = __ballot(part[q]);
                const int off = __builtin_amdgcn_mbcnt_hi((int)(m >> 32), __builtin_amdgcn_mbcnt_lo((int)m, 0));
                if (part[q]) sh.ired[base[q] + off] = r0 + tid + NT * q;
            }
            __syncthreads();
            const int done =
                __builtin_amdgcn_readfirstlane(panel_factor_listed<NBC, T>(st, n, p, C, piv, perm_src, moved_to, info, sh));
            if (done == 2) return;  // a zero pivot: nothing stored, as panel_factor_body
            if (done == 1) {
#pragma unroll
                for (int q = 0; q < RPT; q++) {
                    const int r = r0 + tid + NT * q;
                    if ((r < n) & !part[q]) {
#pragma unroll
                        for (int c = 0; c < NBC; c++)
                            if (c < nbp) st.store(r, r0 + c, (T)sh.prep[c]);
                    }
                }
                __syncthreads();
                return;
            }
        }
    }
    panel_factor_body<RPT, NBC, T, LIST>(st, n, p, piv, perm_src, moved_to, rows, rowcnt, info, sh);
}

// Final row order of factors left by the panels' row moves (perm_src,
// moved_to of panel_factor): row i of L U is row pf[i] of M.  One group.
template <int NBC>
__device__ void compose_rows(const int *perm_src, const int *moved_to, int n, int *pf)
{
    const int tid = threadIdx.x;
    for (int s = tid; s < n; s += NT) pf[s] = s;
    __syncthreads();
    const int P = (n + NBC - 1) / NBC;
    for (int p = 0; p < P; p++) {
        const int r0 = p * NBC;
        const int nbp = min(NBC, n - r0);
        int v1 = 0, v2 = 0, dst = -1;
        if (tid < nbp) {
            v1 = pf[perm_src[r0 + tid]];
            v2 = pf[r0 + tid];
            dst = moved_to[r0 + tid];
        }
        __syncthreads();
        if (tid < nbp) {
            pf[r0 + tid] = v1;
            if (dst >= r0 + nbp) pf[dst] = v2;
        }
        __syncthreads();
    }
}

// getrs with one right-hand side (Simulation.cuh:1316), one group, b in LDS.
// M holds L\U in final row order, so the panels' row moves are applied to b
// first (equal to LAPACK's sequential interchanges; with pf, the composed
// order of compose_rows, n <= 2 NT, gathered in one pass), then forward (k
// ascending) and backward (k descending) substitution as fma chains, in
// 64-row blocks: the diagonal block by one wave with register shuffles, the
// rows outside it by all waves, two rows per thread, the next 16 matrix
// loads of each row in flight while the current 16 are applied.
// T: element type of the factors and of b; RM: row-major (else column-major).
template <int NBC, typename T, bool RM>
__device__ void lu_solve_wg(const T *__restrict__ M, int n, const int *perm_src, const int *moved_to, T *b,
                            const int *pf = nullptr)
{
    auto at = [&](int row, int col) -> T { return RM ? M[(size_t)row * n + col] : M[(size_t)col * n + row]; };
    constexpr int TB = 64;
    constexpr int RB = 2;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int P = pf ? 0 : (n + NBC - 1) / NBC;
    if (pf) {
        T g0 = 0, g1 = 0;
        if (tid < n) g0 = b[pf[tid]];
        if (tid + NT < n) g1 = b[pf[tid + NT]];
        __syncthreads();
        if (tid < n) b[tid] = g0;
        if (tid + NT < n) b[tid + NT] = g1;
        __syncthreads();
    }
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
            T lsub[16], lnx[16];
#pragma unroll
            for (int kk = 0; kk < 16; kk++) lsub[kk] = (kk < lane && lane < nb) ? at(r0 + lane, r0 + kk) : T(0);
#pragma unroll
            for (int sb = 0; sb < TB / 16; sb++) {
                if (sb + 1 < TB / 16) {
#pragma unroll
                    for (int kk = 0; kk < 16; kk++) {
                        const int k = (sb + 1) * 16 + kk;
                        lnx[kk] = (k < lane && lane < nb) ? at(r0 + lane, r0 + k) : T(0);
                    }
                }
#pragma unroll
                for (int kk = 0; kk < 16; kk++) {
                    const int k = sb * 16 + kk;
                    if (k < nb) {
                        const T bk = readlane_d(bi, k);
                        if (lane > k && lane < nb) bi = fma(-lsub[kk], bk, bi);
                    }
                }
                if (sb + 1 < TB / 16) {
#pragma unroll
                    for (int kk = 0; kk < 16; kk++) lsub[kk] = lnx[kk];
                }
            }
            if (lane < nb) b[r0 + lane] = bi;
        }
        __syncthreads();
        for (int i0 = r0 + nb + tid; i0 < n; i0 += NT * RB) {
            T acc[RB];
#pragma unroll
            for (int r = 0; r < RB; r++) acc[r] = (i0 + r * NT < n) ? b[i0 + r * NT] : T(0);
            T l[RB][16], ln[RB][16];
#pragma unroll
            for (int r = 0; r < RB; r++)
#pragma unroll
                for (int kk = 0; kk < 16; kk++)
                    l[r][kk] = (kk < nb && i0 + r * NT < n) ? at(i0 + r * NT, r0 + kk) : T(0);
#pragma unroll
            for (int sb = 0; sb < TB / 16; sb++) {
                if (sb + 1 < TB / 16) {
#pragma unroll
                    for (int r = 0; r < RB; r++)
#pragma unroll
                        for (int kk = 0; kk < 16; kk++) {
                            const int k = (sb + 1) * 16 + kk;
                            ln[r][kk] = (k < nb && i0 + r * NT < n) ? at(i0 + r * NT, r0 + k) : T(0);
                        }
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
                if (sb + 1 < TB / 16) {
#pragma unroll
                    for (int r = 0; r < RB; r++)
#pragma unroll
                        for (int kk = 0; kk < 16; kk++) l[r][kk] = ln[r][kk];
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
            T usub[16], unx[16];
#pragma unroll
            for (int kk = 0; kk < 16; kk++) {
                const int k = (TB / 16 - 1) * 16 + kk;
                usub[kk] = (k >= lane && k < nb && lane < nb) ? at(r0 + lane, r0 + k) : T(1);
            }
#pragma unroll
            for (int sb = TB / 16 - 1; sb >= 0; sb--) {
                if (sb > 0) {
#pragma unroll
                    for (int kk = 0; kk < 16; kk++) {
                        const int k = (sb - 1) * 16 + kk;
                        unx[kk] = (k >= lane && k < nb && lane < nb) ? at(r0 + lane, r0 + k) : T(1);
                    }
                }
#pragma unroll
                for (int kk = 15; kk >= 0; kk--) {
                    const int k = sb * 16 + kk;
                    if (k < nb) {
                        if (lane == k) bi = bi / usub[kk];
                        const T xk = readlane_d(bi, k);
                        if (lane < k) bi = fma(-usub[kk], xk, bi);
                    }
                }
                if (sb > 0) {
#pragma unroll
                    for (int kk = 0; kk < 16; kk++) usub[kk] = unx[kk];
                }
            }
            if (lane < nb) b[r0 + lane] = bi;
        }
        __syncthreads();
        for (int i0 = tid; i0 < r0; i0 += NT * RB) {
            T acc[RB];
#pragma unroll
            for (int r = 0; r < RB; r++) acc[r] = (i0 + r * NT < r0) ? b[i0 + r * NT] : T(0);
            T u[RB][16], un[RB][16];
#pragma unroll
            for (int r = 0; r < RB; r++)
#pragma unroll
                for (int kk = 0; kk < 16; kk++) {
                    const int k = (TB / 16 - 1) * 16 + kk;
                    u[r][kk] = (k < nb && i0 + r * NT < r0) ? at(i0 + r * NT, r0 + k) : T(0);
                }
#pragma unroll
            for (int sb = TB / 16 - 1; sb >= 0; sb--) {
                if (sb > 0) {
#pragma unroll
                    for (int r = 0; r < RB; r++)
#pragma unroll
                        for (int kk = 0; kk < 16; kk++) {
                            const int k = (sb - 1) * 16 + kk;
                            un[r][kk] = (k < nb && i0 + r * NT < r0) ? at(i0 + r * NT, r0 + k) : T(0);
                        }
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
                if (sb > 0) {
#pragma unroll
                    for (int r = 0; r < RB; r++)
#pragma unroll
                        for (int kk = 0; kk < 16; kk++) u[r][kk] = un[r][kk];
                }
            }
#pragma unroll
            for (int r = 0; r < RB; r++)
                if (i0 + r * NT < r0) b[i0 + r * NT] = acc[r];
        }
        __syncthreads();
    }
}


// getrs with one right-hand side for FP64 factors in row-major final row
// order (pf of compose_rows), n <= 2 NT, b in LDS: the same fma chains as
// lu_solve_wg (each row's forward chain k ascending from b, its backward
// chain k descending, then the division by the diagonal), so the results are
// bit-identical; organised for the memory system instead of the panels.
// Thread t owns rows t and t + NT for the whole solve and keeps their
// partial sums in registers.  Columns go in blocks of TB: the wave holding
// the block's diagonal rows solves them with register shuffles and publishes
// x through LDS, one barrier, then every thread applies the block to its
// rows.  Each thread's TB-column row segments come from raw buffer loads
// (out-of-range offsets read zeros, so unused segments cost no branches) into
// two register sets: the set for block c + 2 is requested as soon as block c
// is applied, so a block's loads have one block of work to arrive.
// VEC: doubles per load (2 needs 16-byte aligned rows: n even, M aligned).
template <int TB, int VEC>
__device__ void lu_solve_rows_impl(const double *__restrict__ M, int n, double *b, const int *pf)
{
    M = uniform_ptr(M);
    n = uniform_int(n);
    constexpr int kOob = kBufOob;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const __amdgpu_buffer_rsrc_t rs = buf_rsrc(M, n * n * 8);
    double acc[2];
    acc[0] = (tid < n) ? b[pf[tid]] : 0.0;
    acc[1] = (tid + NT < n) ? b[pf[tid + NT]] : 0.0;
    __syncthreads();
    const int nc = (n + TB - 1) / TB;
    struct Set {
        double v[2][TB];
    };
    auto ld = [&](double *dst, int r, int k, bool use) {
        const int off = use ? (r * n + k) * 8 : kOob;
        if (VEC == 2) {
            const double2 d = __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0));
            dst[0] = d.x;
            dst[1] = d.y;
        } else {
            dst[0] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rs, off, 0, 0));
        }
    };
    // forward: row r needs L[r][k] for k < r
    auto loadF = [&](Set &X, int c) {
        const int k0 = c * TB;
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const int r = tid + q * NT;
            const bool row_ok = c < nc && r < n;
#pragma unroll
            for (int j = 0; j < TB; j += VEC) ld(&X.v[q][j], r, k0 + j, row_ok && k0 + j < r);
        }
    };
    // backward: row r needs U[r][k] for r <= k < n
    auto loadB = [&](Set &X, int c) {
        const int k0 = c * TB;
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const int r = tid + q * NT;
            const bool row_ok = c >= 0 && r < n;
#pragma unroll
            for (int j = 0; j < TB; j += VEC)
                ld(&X.v[q][j], r, k0 + j, row_ok && k0 + j + VEC - 1 >= r && k0 + j < n);
        }
    };
    auto k0_full = [&](int c) { return c * TB + TB <= n; };
    auto stepF_ = [&](Set &X, int c, auto full) {
        constexpr bool FULL = decltype(full)::value;
        const int k0 = c * TB, kb = FULL ? TB : min(TB, n - k0);
        const int q = k0 / NT, dw = (k0 % NT) / 64, l0 = k0 % 64;
        if (wave == dw) {
            auto chain = [&](double &a, const double(&L)[TB]) {
#pragma unroll
                for (int kk = 0; kk < TB; kk++) {
                    if (FULL || kk < kb) {
                        const double xk = readlane_d(a, l0 + kk);
                        if (lane > l0 + kk && lane < l0 + kb) a = fma(-L[kk], xk, a);
                    }
                }
                if (lane >= l0 && lane < l0 + kb) b[k0 + lane - l0] = a;
            };
            if (q == 0)
                chain(acc[0], X.v[0]);
            else
                chain(acc[1], X.v[1]);
        }
        __syncthreads();
        double xs[TB];
#pragma unroll
        for (int kk = 0; kk < TB; kk++) xs[kk] = (kk < kb) ? b[k0 + kk] : 0.0;
#pragma unroll
        for (int q2 = 0; q2 < 2; q2++) {
            const int r = tid + q2 * NT;
            if (r < n && r >= k0 + kb) {
#pragma unroll
                for (int kk = 0; kk < TB; kk++)
                    if (FULL || kk < kb) acc[q2] = fma(-X.v[q2][kk], xs[kk], acc[q2]);
            }
        }
    };
    // a full block (every block when TB divides n) without per-column checks,
    // which otherwise become selects inside the fma chains
    auto stepF = [&](Set &X, int c) {
        if (k0_full(c))
            stepF_(X, c, std::true_type{});
        else
            stepF_(X, c, std::false_type{});
    };
    auto stepB_ = [&](Set &X, int c, auto full) {
        constexpr bool FULL = decltype(full)::value;
        const int k0 = c * TB, kb = FULL ? TB : min(TB, n - k0);
        const int q = k0 / NT, dw = (k0 % NT) / 64, l0 = k0 % 64;
        if (wave == dw) {
            auto chain = [&](double &a, const double(&U)[TB]) {
#pragma unroll
                for (int kk = TB - 1; kk >= 0; kk--) {
                    if (FULL || kk < kb) {
                        if (lane == l0 + kk) a = a / U[kk];
                        const double xk = readlane_d(a, l0 + kk);
                        if (lane >= l0 && lane < l0 + kk) a = fma(-U[kk], xk, a);
                    }
                }
                if (lane >= l0 && lane < l0 + kb) b[k0 + lane - l0] = a;
            };
            if (q == 0)
                chain(acc[0], X.v[0]);
            else
                chain(acc[1], X.v[1]);
        }
        __syncthreads();
        double xs[TB];
#pragma unroll
        for (int kk = 0; kk < TB; kk++) xs[kk] = (kk < kb) ? b[k0 + kk] : 0.0;
#pragma unroll
        for (int q2 = 0; q2 < 2; q2++) {
            const int r = tid + q2 * NT;
            if (r < k0) {
#pragma unroll
                for (int kk = TB - 1; kk >= 0; kk--)
                    if (FULL || kk < kb) acc[q2] = fma(-X.v[q2][kk], xs[kk], acc[q2]);
            }
        }
    };
    auto stepB = [&](Set &X, int c) {
        if (k0_full(c))
            stepB_(X, c, std::true_type{});
        else
            stepB_(X, c, std::false_type{});
    };
    // the scheduling barriers keep each set's loads together and in program
    // order, so the wait for set c leaves set c + 1 in flight
#define LASSIE_SB() __builtin_amdgcn_sched_barrier(0)
    {
        Set X0, X1;
        loadF(X0, 0);
        LASSIE_SB();
        loadF(X1, 1);
        LASSIE_SB();
        for (int c = 0; c < nc; c += 2) {
            stepF(X0, c);
            LASSIE_SB();
            loadF(X0, c + 2);
            LASSIE_SB();
            if (c + 1 < nc) {
                stepF(X1, c + 1);
                LASSIE_SB();
                loadF(X1, c + 3);
                LASSIE_SB();
            }
        }
    }
    __syncthreads();
    {
        Set X0, X1;
        loadB(X0, nc - 1);
        LASSIE_SB();
        loadB(X1, nc - 2);
        LASSIE_SB();
        for (int c = nc - 1; c >= 0; c -= 2) {
            stepB(X0, c);
            LASSIE_SB();
            loadB(X0, c - 2);
            LASSIE_SB();
            if (c - 1 >= 0) {
                stepB(X1, c - 1);
                LASSIE_SB();
                loadB(X1, c - 3);
                LASSIE_SB();
            }
        }
    }
#undef LASSIE_SB
    __syncthreads();
}

__device__ __forceinline__ void lu_solve_rows(const double *__restrict__ M, int n, double *b, const int *pf)
{
    if ((n & 1) == 0 && (reinterpret_cast<uintptr_t>(M) & 15) == 0)
        lu_solve_rows_impl<16, 2>(M, n, b, pf);
    else
        lu_solve_rows_impl<16, 1>(M, n, b, pf);
}

__device__ __forceinline__ bool bad_value(double v) { return isnan(v) || isinf(v) || v < 0; }

// One term of the Newton norm (Simulation.cuh:58-64): the square returned as
// float by the reference's functor; summed in double for the FP64 loop and
// in float (thrust::plus<float>) for the single-precision loop.
__device__ __forceinline__ double norm_square(double d) { return (double)(float)(d * d); }
__device__ __forceinline__ float norm_square(float d) { return d * d; }

}  // namespace lassie
