// Device helpers shared by the LASSIE kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

namespace lassie {

constexpr int NT = 256;               // threads per workgroup
constexpr int NWAVES = NT / 64;

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

// f_s of generate_ode (kernels.cuh:5-50) for one species, same operation order.
__device__ __forceinline__ double rhs_species(const DevModel &m, const double *__restrict__ x, int s)
{
    double acc = 0.0;
    const int e1 = m.hptr[s + 1];
    for (int e = m.hptr[s]; e < e1; e++) {
        const int r = m.hrx[e];
        const double sk = m.hk[e];
        const int a0 = m.lptr[r], a1 = m.lptr[r + 1];
        if (a1 > a0) {
            double temp = 0.0;
            for (int j = a0; j < a1; j++) {
                const double p = mypow(x[m.lsp[j]], m.lexp[j]);
                if (j == a0)
                    temp = temp + sk * p;
                else
                    temp = temp * p;
            }
            acc = acc + temp;
        } else {
            acc = acc + sk;
        }
    }
    return acc;
}

// ---- scoped atomics for control words shared between workgroups ----------
__device__ __forceinline__ int ld_acq(const int *p)
{
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_rel(int *p, int v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double ld_acq_d(const double *p)
{
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_rel_d(double *p, double v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// Grid-wide barrier for a cooperative launch (every workgroup resident).
// Sense by generation counter; each wait is bounded: when the bound is hit the
// abort word is raised and the caller leaves the kernel, so no wave can spin
// forever.  Returns false on abort.
struct GridSync {
    unsigned int count;
    unsigned int pad0[31];
    unsigned int gen;
    unsigned int pad1[31];
};

constexpr long long kSpinLimit = 1LL << 27;  // ~ several seconds of s_sleep(1)

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
            __hip_atomic_fetch_add(&gs->count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (arrived == nwg - 1) {
            __hip_atomic_store(&gs->count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&gs->gen, my + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            long long spins = 0;
            while (__hip_atomic_load(&gs->gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == my) {
                __builtin_amdgcn_s_sleep(1);
                if (++spins > kSpinLimit) {
                    st_rel(abort_word, 1);
                    ok = 0;
                    break;
                }
                if ((spins & 1023) == 0 && ld_acq(abort_word)) {
                    ok = 0;
                    break;
                }
            }
        }
        __threadfence();
        gen_local = my + 1;
        if (ok && ld_acq(abort_word)) ok = 0;
        s_ok = ok;
    }
    __syncthreads();
    return s_ok != 0;
}

}  // namespace lassie
