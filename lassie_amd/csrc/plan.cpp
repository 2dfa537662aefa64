// Batched plan: the batched form prepared once and replayed.
//
// lassie_batch_plan_create allocates the workspaces, the instance table and
// pinned staging buffers once, chooses the kernel variant (row tiles of the
// LU trailing update in flight per wave) by timing a short trial of every
// variant on the plan's own instances, and captures the launch sequence (a
// queue-reset kernel, the persistent kernel, copies into the pinned buffers)
// as a hipGraph.  lassie_batch_plan_run replays the graph.  The first run
// hands instances out largest S first; every run records each instance's
// device time (the kernels' phase timers), and the next run hands them out
// longest first (LPT list scheduling on the measured costs), so repeated
// runs of one plan, e.g. a parameter sweep, finish their slowest instances
// earlier.  The same measurements split the CUs between the two kernels: the
// wave-level kernel's workgroups can share a CU only with each other (its
// LDS does not fit beside a workgroup-level one), so after a run the
// workgroup-level kernel is launched on k fewer CUs, k chosen to give both
// kernels' measured work the same finishing time, and the wave-level kernel
// starts on those k CUs at once instead of waiting for the other kernel to
// drain.  Rates changed
// with lassie_model_set_rates are used by the next run, since the graph reads
// the models' device arrays.  Each instance follows Simulation<T>::run
// (Simulation.cuh:632-925), as in lassie_solve_batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/lassie.h"
#include "batch.h"
#include "model.h"
#include "tune_cache.h"
#include "wave.h"

static_assert(lassie::kBatchVariants == LASSIE_BATCH_VARIANTS, "variant count of the C ABI");

namespace {

std::mutex g_tune_mu;
std::map<std::tuple<int, int, int, int>, int> g_tuned;  // (device, largest S, count, precision) -> variant

}  // namespace

struct lassie_batch_plan {
    int device = 0, n = 0, nt = 0, smax = 1, slots = 1, variant = lassie::kBatchDefaultVariant;
    int n_wg = 0, n_wave = 0, wave_blocks = 0;
    size_t wave_lds = 0;
    lassie_opts opts{};
    long long out_tot = 0, y_tot = 0;
    lassie::BatchInst *d_inst = nullptr;
    int *d_order = nullptr, *d_next = nullptr, *d_wsi = nullptr;
    double *d_tv = nullptr, *d_out = nullptr, *d_y = nullptr, *d_ws64 = nullptr, *d_dst = nullptr;
    float *d_ws32 = nullptr;
    long long *d_st = nullptr;
    double *h_out = nullptr, *h_y = nullptr, *h_dst = nullptr;
    long long *h_st = nullptr;
    unsigned int *h_done = nullptr;  // per slot: instances finished in the current run
    hipStream_t stream = nullptr;   // queue reset, workgroup-level kernel, copies out
    hipStream_t wstream = nullptr;  // wave-level kernel, concurrent with the workgroup-level one
    // ev[0] run start, ev[1] queues reset, ev[2] workgroup-level kernel done,
    // ev[3] wave-level kernel start, ev[4] its end, ev[5] run end
    hipEvent_t ev[6] = {};
    hipEvent_t &ev0 = ev[0], &ev1 = ev[5];
    // graphs: 0 queue reset, 1 workgroup-level kernel, 2 wave-level kernel, 3 copies out
    hipGraph_t graph[4] = {};
    hipGraphExec_t exec[4] = {};
    float tune_ms[lassie::kBatchVariants] = {0};
    int per_cu[lassie::kBatchVariants] = {0};  // resident workgroups per CU of each variant
    int cus = 1;
    // workgroups a variant launches: what it keeps resident, within the slots
    int slots_for(int v) const { return std::max(1, std::min(slots, cus * per_cu[v])); }
    float last_ms[2] = {0.f, 0.f};  // device time of the last run's two kernels
    // the work queues' order: workgroup-level instances [0, n_wg), then
    // wave-level ones; by S for the first run, then by measured device time
    std::vector<int> order;
    bool measured_order = true;   // reorder by the last run's device times
    int runs = 0;
    int wave_per_cu = 1;  // resident wave-level workgroups per CU
    int reserve = 0;      // CUs the workgroup-level kernel leaves to the wave-level one
    int wg_launch() const { return std::max(1, slots_for(variant) - reserve); }
    lassie::BatchArgs args{}, wargs{};
};

static void free_plan(lassie_batch_plan *p)
{
    for (int g = 0; g < 4; g++) {
        if (p->exec[g]) (void)hipGraphExecDestroy(p->exec[g]);
        if (p->graph[g]) (void)hipGraphDestroy(p->graph[g]);
    }
    void *dptrs[] = {p->d_inst, p->d_order, p->d_next, p->d_wsi, p->d_tv, p->d_out,
                     p->d_y,    p->d_ws64,  p->d_dst,  p->d_ws32, p->d_st};
    for (void *q : dptrs)
        if (q) (void)hipFree(q);
    void *hptrs[] = {p->h_out, p->h_y, p->h_dst, p->h_st, p->h_done};
    for (void *q : hptrs)
        if (q) (void)hipHostFree(q);
    for (hipEvent_t e : p->ev)
        if (e) (void)hipEventDestroy(e);
    if (p->stream) (void)hipStreamDestroy(p->stream);
    if (p->wstream) (void)hipStreamDestroy(p->wstream);
    delete p;
}

extern "C" int lassie_batch_plan_create(lassie_model *const *models, int32_t n_models, const lassie_opts *opts,
                                        const double *t, int32_t nt, lassie_batch_plan **out)
{
    using lassie::api_fail;
    if (!out) return api_fail(LASSIE_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (!models || n_models < 1) return api_fail(LASSIE_ERR_INVALID, "need at least one model");
    if (!t || nt < 1) return api_fail(LASSIE_ERR_INVALID, "need sampling times");
    lassie_batch_plan *P = new lassie_batch_plan();
    auto bail = [&](int code, const std::string &msg) {
        free_plan(P);
        return api_fail(code, msg);
    };
    if (opts)
        P->opts = *opts;
    else
        lassie_opts_default(&P->opts);
    const lassie_opts &o = P->opts;
    if (o.precision < LASSIE_PREC_FP64 || o.precision > LASSIE_PREC_BF16_REFINE)
        return bail(LASSIE_ERR_INVALID, "unknown precision");
    std::vector<lassie::BatchInst> inst(n_models);
    for (int i = 0; i < n_models; i++) {
        if (!models[i]) return bail(LASSIE_ERR_INVALID, "NULL model in batch");
        if (lassie::model_species(models[i]) > lassie::kBatchMaxS)
            return bail(LASSIE_ERR_UNSUPPORTED, "batched form supports S <= 512");
        if (lassie::model_device(models[i]) != lassie::model_device(models[0]))
            return bail(LASSIE_ERR_INVALID, "models on different devices");
        lassie::fill_batch_inst(models[i], &inst[i]);
        inst[i].out_off = P->out_tot;
        inst[i].y_off = P->y_tot;
        P->out_tot += (long long)nt * inst[i].S;
        P->y_tot += inst[i].S;
        P->smax = std::max(P->smax, inst[i].S);
    }
    P->device = lassie::model_device(models[0]);
    P->n = n_models;
    P->nt = nt;
    if (hipSetDevice(P->device) != hipSuccess) return bail(LASSIE_ERR_HIP, "hipSetDevice failed");
    const lassie::BatchSplit sp =
        lassie::split_batch(models, inst, o.precision == LASSIE_PREC_FP64 && !(o.reserved & 2));
    const std::vector<int> &order = sp.order;
    P->order = sp.order;
    // LASSIE_BATCH_SCHEDULE=size keeps the by-size order on every run
    if (const char *sch = std::getenv("LASSIE_BATCH_SCHEDULE"))
        if (std::strcmp(sch, "size") == 0) P->measured_order = false;
    P->n_wg = sp.n_wg;
    P->n_wave = sp.n_wave;
    P->smax = sp.smax_wg;
    P->wave_lds = sp.wave_lds;
    int cus = 0, per_cu_max = 0, wave_per_cu = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, P->device) != hipSuccess)
        return bail(LASSIE_ERR_HIP, "device attribute query failed");
    for (int v = 0; v < lassie::kBatchVariants; v++) {
        if (lassie::batch_occupancy(&P->per_cu[v], v) != 0 || P->per_cu[v] < 1)
            return bail(LASSIE_ERR_HIP, "batch occupancy query failed");
        per_cu_max = std::max(per_cu_max, P->per_cu[v]);
    }
    if (sp.n_wave && (lassie::wave_occupancy(sp.wave_lds, &wave_per_cu) != 0 || wave_per_cu < 1))
        return bail(LASSIE_ERR_HIP, "wave kernel occupancy query failed");
    // workspaces for the variant with the most resident workgroups; each
    // variant launches as many as it can keep resident
    P->slots = std::max(1, std::min(sp.n_wg, cus * per_cu_max));
    // LASSIE_BATCH_SLOTS caps the workgroup-level kernel's resident
    // workgroups (to leave CUs to other work on a shared GPU)
    if (const char *cap = std::getenv("LASSIE_BATCH_SLOTS"))
        if (std::atoi(cap) > 0) P->slots = std::min(P->slots, std::atoi(cap));
    P->cus = cus;
    P->wave_per_cu = std::max(1, wave_per_cu);
    P->wave_blocks = std::max(1, std::min(sp.n_wave, cus * wave_per_cu));
    const int smax = P->smax, slots = P->slots;
    const int n_done = slots + P->wave_blocks;  // finished-instance counters: workgroup slots, then waves
    const int Pmax = (smax + lassie::kBatchNB - 1) / lassie::kBatchNB;
    const size_t ws64_stride = 12 * (size_t)smax + (size_t)smax * smax + lassie::kLuSink;
    const size_t ws32_stride = (size_t)smax * smax + lassie::kLuSink;
    const size_t wsi_stride = 3 * (size_t)smax + 16 + Pmax + (size_t)Pmax * smax;
    hipError_t e;
    for (hipEvent_t &ev : P->ev)
        if ((e = hipEventCreate(&ev)) != hipSuccess)
            return bail(LASSIE_ERR_HIP, std::string("batch plan: ") + hipGetErrorString(e));
    if ((e = hipStreamCreate(&P->stream)) != hipSuccess || (e = hipStreamCreate(&P->wstream)) != hipSuccess ||
        (e = hipMalloc((void **)&P->d_inst, sizeof(lassie::BatchInst) * n_models)) != hipSuccess ||
        (e = hipMalloc((void **)&P->d_order, sizeof(int) * n_models)) != hipSuccess ||
        (e = hipMalloc((void **)&P->d_next, 2 * sizeof(int))) != hipSuccess ||
        (e = hipMalloc((void **)&P->d_wsi, sizeof(int) * wsi_stride * slots)) != hipSuccess ||
        (e = hipMalloc((void **)&P->d_tv, sizeof(double) * nt)) != hipSuccess ||
        (e = hipMalloc((void **)&P->d_out, sizeof(double) * P->out_tot)) != hipSuccess ||
        (e = hipMalloc((void **)&P->d_y, sizeof(double) * P->y_tot)) != hipSuccess ||
        (e = hipMalloc((void **)&P->d_ws64, sizeof(double) * ws64_stride * slots)) != hipSuccess ||
        (e = hipMalloc((void **)&P->d_dst, sizeof(double) * 4 * n_models)) != hipSuccess ||
        (e = hipMalloc((void **)&P->d_st, sizeof(long long) * lassie::kBatchStats * n_models)) != hipSuccess ||
        (o.precision != LASSIE_PREC_FP64 &&
         (e = hipMalloc((void **)&P->d_ws32, sizeof(float) * ws32_stride * slots)) != hipSuccess) ||
        (e = hipHostMalloc((void **)&P->h_out, sizeof(double) * P->out_tot)) != hipSuccess ||
        (e = hipHostMalloc((void **)&P->h_y, sizeof(double) * P->y_tot)) != hipSuccess ||
        (e = hipHostMalloc((void **)&P->h_dst, sizeof(double) * 4 * n_models)) != hipSuccess ||
        (e = hipHostMalloc((void **)&P->h_st, sizeof(long long) * lassie::kBatchStats * n_models)) != hipSuccess ||
        // coherent: the kernel's system-scope stores reach the host while it runs
        (e = hipHostMalloc((void **)&P->h_done, sizeof(unsigned int) * n_done, hipHostMallocCoherent)) !=
            hipSuccess)
        return bail(LASSIE_ERR_NOMEM, std::string("batch plan: ") + hipGetErrorString(e));
    if ((e = hipMemcpy(P->d_inst, inst.data(), sizeof(lassie::BatchInst) * n_models, hipMemcpyHostToDevice)) !=
            hipSuccess ||
        (e = hipMemcpy(P->d_order, order.data(), sizeof(int) * n_models, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(P->d_tv, t, sizeof(double) * nt, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemset(P->d_wsi, 0, sizeof(int) * wsi_stride * slots)) != hipSuccess)
        return bail(LASSIE_ERR_HIP, std::string("batch plan upload: ") + hipGetErrorString(e));
    lassie::BatchArgs &a = P->args;
    a.inst = P->d_inst;
    a.n_inst = sp.n_wg;
    a.order = P->d_order;
    a.next = P->d_next;
    a.tv = P->d_tv;
    a.nt = nt;
    a.out = P->d_out;
    a.y_final = P->d_y;
    a.stats = P->d_st;
    a.dstats = P->d_dst;
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
    a.precision = o.precision;
    a.refine_tol = o.refine_tol;
    a.refine_max_iter = o.refine_max_iter == 0 ? 0 : std::max(1, o.refine_max_iter);
    a.s_max = smax;
    a.ws64 = P->d_ws64;
    a.ws64_stride = ws64_stride;
    a.ws32 = P->d_ws32;
    a.ws32_stride = ws32_stride;
    a.wsi = P->d_wsi;
    a.wsi_stride = wsi_stride;
    std::memset(P->h_done, 0, sizeof(unsigned int) * n_done);
    lassie::BatchArgs &aw = P->wargs;
    aw = a;
    aw.n_inst = sp.n_wave;
    aw.order = P->d_order + sp.n_wg;
    aw.next = P->d_next + 1;
    aw.s_max = sp.smax_wave;
    aw.nh_max = sp.nh_max_wave;
    aw.done_host = P->h_done + slots;

    // autotuner: a short trial of every kernel variant on the plan's own
    // workgroup-level instances (8 loop iterations, at most 64 Newton
    // iterations each); the fastest is kept for this device and workload
    // shape, in this process and in the cache file (tune_cache.h)
    const auto key = std::make_tuple(P->device, smax, sp.n_wg, (int)o.precision);
    std::string fkey;
    {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, P->device) == hipSuccess) {
            std::string arch(prop.gcnArchName);
            for (char &c : arch)
                if (c == '"' || c == '\\') c = '_';
            fkey = "batch_variant/" + arch + "/cu" + std::to_string(prop.multiProcessorCount) + "/smax" +
                   std::to_string(smax) + "/n" + std::to_string(sp.n_wg) + "/prec" + std::to_string(o.precision);
        }
    }
    int chosen = -1;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tuned.find(key);
        if (it != g_tuned.end()) chosen = it->second;
    }
    if (chosen < 0 && sp.n_wg == 0) chosen = lassie::kBatchDefaultVariant;
    // LASSIE_BATCH_VARIANT forces a variant (no trials), for measurements
    if (const char *fv = std::getenv("LASSIE_BATCH_VARIANT")) {
        const int v = std::atoi(fv);
        if (v >= 0 && v < lassie::kBatchVariants && *fv) chosen = v;
    }
    if (chosen < 0 && !fkey.empty() && lassie::tune_lookup(fkey, &chosen) &&
        (chosen < 0 || chosen >= lassie::kBatchVariants))
        chosen = -1;
    if (chosen < 0) {
        lassie::BatchArgs trial = a;
        trial.max_iters = 8;
        trial.max_newton = 64;
        trial.done_host = nullptr;
        float best = 0.f;
        for (int v = 0; v < lassie::kBatchVariants; v++) {
            if ((e = lassie::launch_batch_reset(P->d_next, P->stream)) != hipSuccess ||
                (e = hipEventRecord(P->ev0, P->stream)) != hipSuccess ||
                (e = lassie::launch_batch(trial, P->slots_for(v), P->stream, v)) != hipSuccess ||
                (e = hipEventRecord(P->ev1, P->stream)) != hipSuccess ||
                (e = hipStreamSynchronize(P->stream)) != hipSuccess)
                return bail(LASSIE_ERR_HIP, std::string("autotune trial: ") + hipGetErrorString(e));
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, P->ev0, P->ev1);
            P->tune_ms[v] = ms;
            if (chosen < 0 || ms < best) {
                best = ms;
                chosen = v;
            }
        }
        if (!fkey.empty()) lassie::tune_store(fkey, chosen, P->tune_ms, lassie::kBatchVariants);
    }
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tuned[key] = chosen;
    }
    P->variant = chosen;
    a.done_host = P->h_done;

    // the launch sequence as four graphs (no memset nodes: the queue heads
    // are reset by a kernel).  The workgroup-level kernel starts first and
    // holds every CU it can; the wave-level kernel, on a second stream, takes
    // CUs as the large instances' queue drains; events between the graphs
    // time each kernel on its own stream.
    for (int g = 0; g < 4; g++) {
        hipStream_t s = (g == 2) ? P->wstream : P->stream;
        if ((e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal)) != hipSuccess)
            return bail(LASSIE_ERR_HIP, std::string("hipStreamBeginCapture: ") + hipGetErrorString(e));
        hipError_t ec = hipSuccess;
        if (g == 0) {
            ec = lassie::launch_batch_reset(P->d_next, s, P->d_out, (size_t)P->out_tot);
        } else if (g == 1) {
            if (sp.n_wg) ec = lassie::launch_batch(a, P->wg_launch(), s, P->variant);
        } else if (g == 2) {
            if (sp.n_wave) ec = lassie::launch_wave(aw, P->wave_blocks, sp.wave_lds, s);
        } else {
            ec = hipMemcpyAsync(P->h_out, P->d_out, sizeof(double) * P->out_tot, hipMemcpyDeviceToHost, s);
            if (ec == hipSuccess)
                ec = hipMemcpyAsync(P->h_y, P->d_y, sizeof(double) * P->y_tot, hipMemcpyDeviceToHost, s);
            if (ec == hipSuccess)
                ec = hipMemcpyAsync(P->h_st, P->d_st, sizeof(long long) * lassie::kBatchStats * n_models,
                                    hipMemcpyDeviceToHost, s);
            if (ec == hipSuccess)
                ec = hipMemcpyAsync(P->h_dst, P->d_dst, sizeof(double) * 4 * n_models, hipMemcpyDeviceToHost, s);
        }
        e = hipStreamEndCapture(s, &P->graph[g]);
        if (ec != hipSuccess || e != hipSuccess)
            return bail(LASSIE_ERR_HIP,
                        std::string("graph capture: ") + hipGetErrorString(ec != hipSuccess ? ec : e));
        if ((e = hipGraphInstantiate(&P->exec[g], P->graph[g], nullptr, nullptr, 0)) != hipSuccess)
            return bail(LASSIE_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    *out = P;
    return LASSIE_OK;
}

// Re-captures the workgroup-level kernel's graph with P->wg_launch() workgroups.
static int recapture_wg(lassie_batch_plan *P)
{
    using lassie::api_fail;
    hipError_t e;
    if (P->exec[1]) (void)hipGraphExecDestroy(P->exec[1]);
    if (P->graph[1]) (void)hipGraphDestroy(P->graph[1]);
    P->exec[1] = nullptr;
    P->graph[1] = nullptr;
    if ((e = hipStreamBeginCapture(P->stream, hipStreamCaptureModeThreadLocal)) != hipSuccess)
        return api_fail(LASSIE_ERR_HIP, std::string("hipStreamBeginCapture: ") + hipGetErrorString(e));
    const hipError_t ec = lassie::launch_batch(P->args, P->wg_launch(), P->stream, P->variant);
    e = hipStreamEndCapture(P->stream, &P->graph[1]);
    if (ec != hipSuccess || e != hipSuccess)
        return api_fail(LASSIE_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(ec != hipSuccess ? ec : e));
    if ((e = hipGraphInstantiate(&P->exec[1], P->graph[1], nullptr, nullptr, 0)) != hipSuccess)
        return api_fail(LASSIE_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    return LASSIE_OK;
}

extern "C" int lassie_batch_plan_run(lassie_batch_plan *P, double *out, double *y_final, lassie_stats *stats,
                                     double *elapsed_ms)
{
    using lassie::api_fail;
    if (!P || !out) return api_fail(LASSIE_ERR_INVALID, "plan and out are required");
    hipError_t e;
    // no launch of this plan is in flight here (runs are synchronous)
    std::memset(P->h_done, 0, sizeof(unsigned int) * (P->slots + P->wave_blocks));
    hipStream_t s0 = P->stream, s1 = P->wstream;
    if ((e = hipSetDevice(P->device)) != hipSuccess || (e = hipEventRecord(P->ev[0], s0)) != hipSuccess ||
        (e = hipGraphLaunch(P->exec[0], s0)) != hipSuccess || (e = hipEventRecord(P->ev[1], s0)) != hipSuccess ||
        (e = hipGraphLaunch(P->exec[1], s0)) != hipSuccess || (e = hipEventRecord(P->ev[2], s0)) != hipSuccess ||
        (e = hipStreamWaitEvent(s1, P->ev[1], 0)) != hipSuccess || (e = hipEventRecord(P->ev[3], s1)) != hipSuccess ||
        (e = hipGraphLaunch(P->exec[2], s1)) != hipSuccess || (e = hipEventRecord(P->ev[4], s1)) != hipSuccess ||
        (e = hipStreamWaitEvent(s0, P->ev[4], 0)) != hipSuccess || (e = hipGraphLaunch(P->exec[3], s0)) != hipSuccess ||
        (e = hipEventRecord(P->ev[5], s0)) != hipSuccess || (e = hipStreamSynchronize(s0)) != hipSuccess)
        return api_fail(LASSIE_ERR_HIP, std::string("plan run: ") + hipGetErrorString(e));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, P->ev[0], P->ev[5]);
    (void)hipEventElapsedTime(&P->last_ms[0], P->ev[3], P->ev[4]);
    (void)hipEventElapsedTime(&P->last_ms[1], P->ev[1], P->ev[2]);
    if (elapsed_ms) *elapsed_ms = ms;
    std::memcpy(out, P->h_out, sizeof(double) * P->out_tot);
    if (y_final) std::memcpy(y_final, P->h_y, sizeof(double) * P->y_tot);
    bool maxit = false;
    for (int i = 0; i < P->n; i++) {
        const long long *c = P->h_st + (size_t)i * lassie::kBatchStats;
        if (stats) lassie::fill_batch_stats(c, P->h_dst + 4 * (size_t)i, ms, P->wg_launch(), &stats[i]);
        maxit |= c[9] == 1;
    }
    P->runs++;
    if (P->measured_order) {
        // longest measured device time first within each kernel's queue (ties
        // keep the previous order); the next run's graph reads d_order
        auto cost = [&](int i) {
            const long long *c = P->h_st + (size_t)i * lassie::kBatchStats;
            return c[14] + c[15] + c[16] + c[17];
        };
        auto longest = [&](int a, int b) { return cost(a) > cost(b); };
        std::stable_sort(P->order.begin(), P->order.begin() + P->n_wg, longest);
        std::stable_sort(P->order.begin() + P->n_wg, P->order.end(), longest);
        if ((e = hipMemcpy(P->d_order, P->order.data(), sizeof(int) * P->n, hipMemcpyHostToDevice)) != hipSuccess)
            return api_fail(LASSIE_ERR_HIP, std::string("plan schedule: ") + hipGetErrorString(e));
        // CUs for the wave-level kernel: W_wg / (C - k) = W_wave / (w k)
        // with W the measured work of each kernel's instances (CU-time for
        // the workgroup-level kernel, wavefront-time for the wave-level one,
        // w of whose workgroups share a CU); at most a quarter of the CUs
        if (P->n_wg && P->n_wave) {
            double w_wg = 0.0, w_wave = 0.0;
            for (int j = 0; j < P->n; j++) (j < P->n_wg ? w_wg : w_wave) += (double)cost(P->order[j]);
            int k = (int)std::lround((double)P->cus * w_wave / (P->wave_per_cu * w_wg + w_wave));
            k = std::max(0, std::min(k, P->cus / 4));
            k = std::min(k, P->slots_for(P->variant) - 1);
            if (k != P->reserve) {
                P->reserve = k;
                int rc = recapture_wg(P);
                if (rc != LASSIE_OK) return rc;
            }
        }
    }
    if (maxit)
        return api_fail(LASSIE_ERR_MAXITER,
                        "max_iters or max_newton reached before the last sample in at least one instance");
    return LASSIE_OK;
}

extern "C" int lassie_batch_plan_timing(const lassie_batch_plan *P, int32_t *n_wave, int32_t *n_wg, float *ms)
{
    if (!P) return lassie::api_fail(LASSIE_ERR_INVALID, "plan is NULL");
    if (n_wave) *n_wave = P->n_wave;
    if (n_wg) *n_wg = P->n_wg;
    if (ms) {
        ms[0] = P->last_ms[0];
        ms[1] = P->last_ms[1];
    }
    return LASSIE_OK;
}

extern "C" int lassie_batch_plan_progress(const lassie_batch_plan *P, int64_t *done)
{
    using lassie::api_fail;
    if (!P || !done) return api_fail(LASSIE_ERR_INVALID, "plan and done are required");
    int64_t sum = 0;
    for (int s = 0; s < P->slots + P->wave_blocks; s++)
        sum += __atomic_load_n(P->h_done + s, __ATOMIC_RELAXED);
    *done = sum;
    return LASSIE_OK;
}

extern "C" int lassie_batch_plan_info(const lassie_batch_plan *P, int32_t *variant, int32_t *tile_rows,
                                      float *trial_ms)
{
    if (!P) return lassie::api_fail(LASSIE_ERR_INVALID, "plan is NULL");
    if (variant) *variant = P->variant;
    if (tile_rows) *tile_rows = lassie::batch_variant_rt(P->variant);
    if (trial_ms)
        for (int v = 0; v < lassie::kBatchVariants; v++) trial_ms[v] = P->tune_ms[v];
    return LASSIE_OK;
}

extern "C" int lassie_batch_plan_destroy(lassie_batch_plan *P)
{
    if (P) free_plan(P);
    return LASSIE_OK;
}
