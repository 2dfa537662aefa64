// lassie_solve_batch_static: the batched form for callers that cannot
// allocate device memory.  Every device buffer of the call (instance table,
// work order, queue heads, sampling times, counters and outputs) is carved
// out of `scratch`, one fixed array per GPU; the wave-level kernel keeps each
// instance's matrix and state in LDS, so nothing else is needed.  The host
// side stages tables through small stack buffers, so the call performs no
// device or host allocation on its success path.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/lassie.h"
#include "batch.h"
#include "model.h"
#include "wave.h"

namespace {

constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char *g_scratch[kMaxDevices];  // per device: the array's device address, once set up

// The device address of this GPU's `scratch` (the current device must be
// `device`).  The symbol's host-side name is a stand-in, not a device
// address, so kernels and copies only ever see this pointer.
int scratch_address(int device, unsigned char **out)
{
    if (device < 0 || device >= kMaxDevices) return lassie::api_fail(LASSIE_ERR_UNSUPPORTED, "device index too large");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return lassie::api_fail(LASSIE_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[device] = static_cast<unsigned char *>(p);
    }
    *out = g_scratch[device];
    return LASSIE_OK;
}

inline size_t up256(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace

extern "C" int lassie_solve_batch_static(lassie_model *const *models, int32_t n, const lassie_opts *opts,
                                         const double *t, int32_t nt, double *out, double *y_final,
                                         lassie_stats *stats, double *elapsed_ms)
{
    using lassie::api_fail;
    if (!models || n < 1) return api_fail(LASSIE_ERR_INVALID, "need at least one model");
    if (!t || nt < 1 || !out) return api_fail(LASSIE_ERR_INVALID, "need sampling times and an output buffer");
    lassie_opts o;
    if (opts)
        o = *opts;
    else
        lassie_opts_default(&o);
    if (o.precision != LASSIE_PREC_FP64)
        return api_fail(LASSIE_ERR_UNSUPPORTED, "lassie_solve_batch_static runs the FP64 method only");
    long long out_tot = 0, y_tot = 0;
    int smax = 1, nh_max = 0;
    for (int i = 0; i < n; i++) {
        if (!models[i]) return api_fail(LASSIE_ERR_INVALID, "NULL model in batch");
        if (lassie::model_device(models[i]) != lassie::model_device(models[0]))
            return api_fail(LASSIE_ERR_INVALID, "models on different devices");
        const int S = lassie::model_species(models[i]), nh = lassie::model_nnz_net(models[i]);
        if (S > lassie::kWaveMaxS || nh > lassie::kWaveMaxNh)
            return api_fail(LASSIE_ERR_UNSUPPORTED, "lassie_solve_batch_static needs every model to have at most 64 "
                                                    "species and 2048 net-stoichiometry entries");
        out_tot += (long long)nt * S;
        y_tot += S;
        smax = S > smax ? S : smax;
        nh_max = nh > nh_max ? nh : nh_max;
    }
    // layout of the call's buffers in scratch
    const size_t o_inst = 0;
    const size_t o_order = o_inst + up256(sizeof(lassie::BatchInst) * (size_t)n);
    const size_t o_next = o_order + up256(sizeof(int) * (size_t)n);
    const size_t o_tv = o_next + 256;
    const size_t o_st = o_tv + up256(sizeof(double) * (size_t)nt);
    const size_t o_dst = o_st + up256(sizeof(long long) * lassie::kBatchStats * (size_t)n);
    const size_t o_out = o_dst + up256(sizeof(double) * 4 * (size_t)n);
    const size_t o_y = o_out + up256(sizeof(double) * (size_t)out_tot);
    const size_t need = o_y + up256(sizeof(double) * (size_t)y_tot);
    if (need > kScratchBytes)
        return api_fail(LASSIE_ERR_UNSUPPORTED, "lassie_solve_batch_static: this batch needs " + std::to_string(need) +
                                                    " bytes of the 67108864-byte buffer; use lassie_solve_batch");

    const int device = lassie::model_device(models[0]);
    hipError_t e;
#define STATIC_TRY(expr)                                                                         \
    do {                                                                                         \
        if ((e = (expr)) != hipSuccess)                                                          \
            return api_fail(LASSIE_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e));   \
    } while (0)
    STATIC_TRY(hipSetDevice(device));
    unsigned char *base = nullptr;
    if (const int rc = scratch_address(device, &base)) return rc;
    auto *d_inst = reinterpret_cast<lassie::BatchInst *>(base + o_inst);
    auto *d_order = reinterpret_cast<int *>(base + o_order);
    auto *d_next = reinterpret_cast<int *>(base + o_next);
    auto *d_tv = reinterpret_cast<double *>(base + o_tv);
    auto *d_st = reinterpret_cast<long long *>(base + o_st);
    auto *d_dst = reinterpret_cast<double *>(base + o_dst);
    auto *d_out = reinterpret_cast<double *>(base + o_out);
    auto *d_y = reinterpret_cast<double *>(base + o_y);
    hipStream_t st = lassie::model_stream(models[0]);
    hipEvent_t ev0, ev1;
    lassie::model_events(models[0], &ev0, &ev1);

    // instance table, through a stack buffer (synchronous copies, so the
    // buffer can be refilled at once)
    {
        constexpr int CH = 32;
        lassie::BatchInst buf[CH];
        long long oo = 0, yo = 0;
        for (int i0 = 0; i0 < n; i0 += CH) {
            const int c = (n - i0) < CH ? (n - i0) : CH;
            for (int j = 0; j < c; j++) {
                lassie::fill_batch_inst(models[i0 + j], &buf[j]);
                buf[j].out_off = oo;
                buf[j].y_off = yo;
                oo += (long long)nt * buf[j].S;
                yo += buf[j].S;
            }
            STATIC_TRY(hipMemcpy(d_inst + i0, buf, sizeof(lassie::BatchInst) * c, hipMemcpyHostToDevice));
        }
    }
    // work order, largest S first (a counting pass per size)
    {
        constexpr int CH = 256;
        int buf[CH];
        int filled = 0, written = 0;
        for (int S = lassie::kWaveMaxS; S >= 1; S--)
            for (int i = 0; i < n; i++) {
                if (lassie::model_species(models[i]) != S) continue;
                buf[filled++] = i;
                if (filled == CH) {
                    STATIC_TRY(hipMemcpy(d_order + written, buf, sizeof(int) * CH, hipMemcpyHostToDevice));
                    written += CH;
                    filled = 0;
                }
            }
        if (filled) STATIC_TRY(hipMemcpy(d_order + written, buf, sizeof(int) * filled, hipMemcpyHostToDevice));
    }
    STATIC_TRY(hipMemcpy(d_tv, t, sizeof(double) * nt, hipMemcpyHostToDevice));

    int cus = 0, per_cu = 0;
    const size_t lds = lassie::wave_lds_bytes(smax, nh_max);
    STATIC_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    if (lassie::wave_occupancy(lds, &per_cu) != 0 || per_cu < 1)
        return api_fail(LASSIE_ERR_HIP, "wave kernel occupancy query failed");
    const int blocks = n < cus * per_cu ? n : cus * per_cu;
    lassie::BatchArgs a{};
    a.inst = d_inst;
    a.n_inst = n;
    a.order = d_order;
    a.next = d_next;
    a.tv = d_tv;
    a.nt = nt;
    a.out = d_out;
    a.y_final = d_y;
    a.stats = d_st;
    a.dstats = d_dst;
    a.rkf_step = o.rkf_step;
    a.be_step = o.be_step;
    a.newton_tol = o.newton_tol;
    a.stiffness_tol = o.stiffness_tol;
    a.rtol = o.rtol;
    a.newton_iter = o.newton_iter;
    a.bdf_regrow = o.bdf_regrow;
    a.max_rkf_retries = o.max_rkf_retries;
    a.max_iters = o.max_iters;
    a.max_newton = o.max_newton;
    a.precision = LASSIE_PREC_FP64;
    a.s_max = smax;
    a.nh_max = nh_max;
    // the queue heads and the outputs start fresh: the array keeps the last call's contents
    STATIC_TRY(hipEventRecord(ev0, st));
    STATIC_TRY(lassie::launch_batch_reset(d_next, st, d_out, (size_t)out_tot));
    STATIC_TRY(lassie::launch_wave(a, blocks, lds, st));
    STATIC_TRY(hipEventRecord(ev1, st));
    STATIC_TRY(hipMemcpyAsync(out, d_out, sizeof(double) * out_tot, hipMemcpyDeviceToHost, st));
    if (y_final) STATIC_TRY(hipMemcpyAsync(y_final, d_y, sizeof(double) * y_tot, hipMemcpyDeviceToHost, st));
    STATIC_TRY(hipStreamSynchronize(st));
    float ms = 0.f;
    STATIC_TRY(hipEventElapsedTime(&ms, ev0, ev1));
    if (elapsed_ms) *elapsed_ms = ms;
    bool maxit = false;
    {
        constexpr int CH = 32;
        long long c[CH * lassie::kBatchStats];
        double d[CH * 4];
        for (int i0 = 0; i0 < n; i0 += CH) {
            const int cnt = (n - i0) < CH ? (n - i0) : CH;
            STATIC_TRY(hipMemcpy(c, d_st + (size_t)i0 * lassie::kBatchStats,
                                 sizeof(long long) * lassie::kBatchStats * cnt, hipMemcpyDeviceToHost));
            STATIC_TRY(hipMemcpy(d, d_dst + (size_t)i0 * 4, sizeof(double) * 4 * cnt, hipMemcpyDeviceToHost));
            for (int j = 0; j < cnt; j++) {
                if (stats) lassie::fill_batch_stats(c + j * lassie::kBatchStats, d + 4 * j, ms, blocks, &stats[i0 + j]);
                maxit |= c[j * lassie::kBatchStats + 9] == 1;
            }
        }
    }
#undef STATIC_TRY
    if (maxit)
        return api_fail(LASSIE_ERR_MAXITER,
                        "max_iters or max_newton reached before the last sample in at least one instance");
    return LASSIE_OK;
}
