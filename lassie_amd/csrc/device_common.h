// Device helpers shared by the LASSIE kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

namespace lassie {

constexpr int NT = 256;               // threads per workgroup
constexpr int NWAVES = NT / 64;
constexpr int kLuSink = 64;           // elements after each dense matrix that take discarded stores

// myPow, kernels.cuh:52-63: repeated multiplication starting from 1.0.
__device__ __forceinline__ double mypow(double base, int ex)
{
    double p = 1.0;
    while (ex > 0) {
        p = p * base;
        ex--;
    }
    return p;
}

// Device view of a compiled model.  hk[e] = net_coef[e] * k[net_reaction[e]]
// is the product `signum * c_vector[k_pos]` that generate_ode forms first
// (kernels.cuh:33, 45); it is computed once on the host with the same single
// IEEE multiply.
struct DevModel {
    int S, R;
    const int *lptr, *lsp, *lexp;
    const int *hptr, *hrx;
    const double *hk;
    const double *feed, *atol;
};

// One net-stoichiometry entry with its reaction's reactants inlined, so a
// right-hand-side term needs one record read instead of a chain of dependent
// index loads.  kind 0: no reactant; 1, 2: that many reactants, species in
// sp01 (16 bits each), exponents in meta bits 0-7 and 8-15; 3: any other
// reaction, sp01 holds the reaction index for the CSR walk.
struct HRec {
    double hk;
    int sp01;
    int meta;
};

// myPow<T> (kernels.cuh:52-63): products in the state type T, returned as
// double.  For T = double this is mypow.
template <typename T>
__device__ __forceinline__ double mypow_t(T base, int ex)
{
    T p = 1.0;
    while (ex > 0) {
        p = p * base;
        ex--;
    }
    return (double)p;
}

// Reactant product of reaction r (any number of reactants), as a double
// term; sk is the coefficient product in the state type.
template <typename T>
__device__ __forceinline__ double rhs_species_generic(const DevModel &m, const T *x, int r, T sk)
{
    const int a0 = m.lptr[r], a1 = m.lptr[r + 1];
    if (a1 <= a0) return sk;
    double temp = 0.0;
    for (int j = a0; j < a1; j++) {
        const double p = mypow_t(x[m.lsp[j]], m.lexp[j]);
        if (j == a0)
            temp = temp + sk * p;
        else
            temp = temp * p;
    }
    return temp;
}

// f_s of generate_ode<T> (kernels.cuh:5-50) for one species, same operation
// order: acc = ((0 + t_0) + t_1) + ..., t = ((0 + sk*x_a^ea) * x_b^eb).  The
// term t is a double; the output, of the state type T, is rounded after
// every term (y1[gid] = y1[gid] + temp).  rc.hk holds sign * k formed in T.
template <typename T>
__device__ __forceinline__ T rhs_species(const DevModel &m, const HRec *recs, const int *hptr, const T *x, int s)
{
    T acc = 0.0;
    const int e1 = hptr[s + 1];
    for (int e = hptr[s]; e < e1; e++) {
        const HRec rc = recs[e];
        const int kind = rc.meta >> 16;
        const T hk = (T)rc.hk;
        if (kind == 0) {
            acc = acc + hk;
        } else if (kind <= 2) {
            double temp = 0.0;
            temp = temp + hk * mypow_t(x[rc.sp01 & 0xffff], rc.meta & 0xff);
            if (kind == 2) temp = temp * mypow_t(x[(unsigned)rc.sp01 >> 16], (rc.meta >> 8) & 0xff);
            acc = (T)(acc + temp);
        } else {
            acc = (T)(acc + rhs_species_generic(m, x, rc.sp01, hk));
        }
    }
    return acc;
}

// ---- control words shared between workgroups (relaxed, agent scope) -------
// Ordering of everything else is provided by grid_barrier.
__device__ __forceinline__ int ld_rlx(const int *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_rlx(int *p, int v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double ld_rlx_d(const double *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_rlx_d(double *p, double v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Grid-wide barrier for a cooperative launch (every workgroup resident).
// Release and acquire are agent-scope fences (L2 write-back / invalidate), the
// only scope that orders plain global stores of one CU before plain loads on
// another.  Each wait is bounded: when the bound is hit the abort word is
// raised and the caller leaves the kernel, so no wave can spin forever.
// Returns false on abort.
struct GridSync {
    unsigned int count;
    unsigned int pad0[31];
    unsigned int gen;
    unsigned int pad1[31];
};

// bound on one wait, in ticks of the 100 MHz wall clock (60 s)
constexpr unsigned long long kWaitLimit = 6000000000ULL;

__device__ __forceinline__ bool grid_barrier(GridSync *gs, unsigned int nwg, unsigned int &gen_local,
                                             int *abort_word)
{
    __shared__ int s_ok;
    __syncthreads();
    if (threadIdx.x == 0) {
        int ok = 1;
        const unsigned int my = gen_local;
        __threadfence();
        const unsigned int arrived =
            __hip_atomic_fetch_add(&gs->count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (arrived == nwg - 1) {
            __hip_atomic_store(&gs->count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __threadfence();
            __hip_atomic_store(&gs->gen, my + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            const unsigned long long t0 = wall_clock64();
            unsigned int spins = 0;
            while (__hip_atomic_load(&gs->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == my) {
                __builtin_amdgcn_s_sleep(2);
                if ((++spins & 255) == 0) {
                    if (wall_clock64() - t0 > kWaitLimit) {
                        st_rlx(abort_word, 1);
                        ok = 0;
                        break;
                    }
                    if (ld_rlx(abort_word)) {
                        ok = 0;
                        break;
                    }
                }
            }
        }
        __threadfence();
        gen_local = my + 1;
        if (ok && ld_rlx(abort_word)) ok = 0;
        s_ok = ok;
    }
    __syncthreads();
    return s_ok != 0;
}

}  // namespace lassie
