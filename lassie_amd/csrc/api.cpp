// C ABI of liblassie (include/lassie.h): model upload, solve launch, helpers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/lassie.h"
#include "solver.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                            \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return fail(LASSIE_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));      \
    } while (0)

template <typename T>
int dev_upload(T **dst, const T *src, size_t n)
{
    *dst = nullptr;
    size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
    HIP_TRY(hipMalloc((void **)dst, bytes));
    if (n) HIP_TRY(hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice));
    return LASSIE_OK;
}

}  // namespace

struct lassie_model {
    int S = 0, R = 0;
    int device = 0;
    std::vector<int> lptr, lsp, lexp, hptr, hrx, hcoef;
    std::vector<double> k, x0, feed, atol, hk;
    int *d_lptr = nullptr, *d_lsp = nullptr, *d_lexp = nullptr, *d_hptr = nullptr, *d_hrx = nullptr;
    double *d_hk = nullptr, *d_x0 = nullptr, *d_feed = nullptr, *d_atol = nullptr;
    double *d_vec = nullptr;  // work vectors, 12 x S
    double *d_M = nullptr;    // dense S x S Newton matrix (column-major)
    int *d_piv = nullptr, *d_perm = nullptr, *d_moved = nullptr;
    lassie::Ctl *d_ctl = nullptr;
    lassie::GridSync *d_gs = nullptr;
    long long *d_stats = nullptr;
    double *d_dstats = nullptr;
    double *d_tv = nullptr, *d_out = nullptr, *d_prog = nullptr;
    int *d_cs = nullptr;
    size_t cap_tv = 0, cap_out = 0, cap_prog = 0, cap_cs = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

static void free_model(lassie_model *m)
{
    void *ptrs[] = {m->d_lptr, m->d_lsp, m->d_lexp, m->d_hptr, m->d_hrx, m->d_hk, m->d_x0, m->d_feed,
                    m->d_atol, m->d_vec, m->d_M, m->d_piv, m->d_perm, m->d_moved, m->d_ctl, m->d_gs,
                    m->d_stats, m->d_dstats, m->d_tv, m->d_out, m->d_prog, m->d_cs};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (m->ev0) (void)hipEventDestroy(m->ev0);
    if (m->ev1) (void)hipEventDestroy(m->ev1);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

static int validate_desc(const lassie_model_desc *d)
{
    if (!d) return fail(LASSIE_ERR_INVALID, "desc is NULL");
    const int S = d->n_species, R = d->n_reactions;
    if (S < 1 || R < 1) return fail(LASSIE_ERR_INVALID, "need n_species >= 1 and n_reactions >= 1");
    if (!d->react_ptr || !d->net_ptr || !d->k || !d->x0)
        return fail(LASSIE_ERR_INVALID, "react_ptr, net_ptr, k and x0 are required");
    if (d->react_ptr[0] != 0 || d->net_ptr[0] != 0) return fail(LASSIE_ERR_INVALID, "CSR pointers must start at 0");
    for (int r = 0; r < R; r++)
        if (d->react_ptr[r + 1] < d->react_ptr[r]) return fail(LASSIE_ERR_INVALID, "react_ptr not monotone");
    for (int s = 0; s < S; s++)
        if (d->net_ptr[s + 1] < d->net_ptr[s]) return fail(LASSIE_ERR_INVALID, "net_ptr not monotone");
    const int nl = d->react_ptr[R], nh = d->net_ptr[S];
    if ((nl > 0 && (!d->react_species || !d->react_exp)) || (nh > 0 && (!d->net_reaction || !d->net_coef)))
        return fail(LASSIE_ERR_INVALID, "CSR index arrays missing");
    for (int e = 0; e < nl; e++)
        if (d->react_species[e] < 0 || d->react_species[e] >= S)
            return fail(LASSIE_ERR_INVALID, "react_species out of range");
    for (int e = 0; e < nh; e++)
        if (d->net_reaction[e] < 0 || d->net_reaction[e] >= R)
            return fail(LASSIE_ERR_INVALID, "net_reaction out of range");
    return LASSIE_OK;
}

extern "C" {

void lassie_opts_default(lassie_opts *o)
{
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->rkf_step = 1e-3;
    o->be_step = 1e-2;
    o->newton_tol = 1e-6;
    o->stiffness_tol = 1e-6;
    o->rtol = 0.0;
    o->newton_iter = 10000;
    o->bdf_regrow = 0;
    o->max_rkf_retries = 0;
    o->precision = LASSIE_PREC_FP64;
    o->max_iters = 100000000LL;
    o->refine_tol = 1e-10;
    o->refine_max_iter = 8;
}

int lassie_abi_version(void) { return LASSIE_ABI_VERSION; }

const char *lassie_last_error(void) { return g_err.c_str(); }

int lassie_device_count(int32_t *count)
{
    if (!count) return fail(LASSIE_ERR_INVALID, "count is NULL");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        *count = 0;
        return fail(LASSIE_ERR_HIP, std::string("hipGetDeviceCount: ") + hipGetErrorString(e));
    }
    *count = n;
    return LASSIE_OK;
}

int lassie_csr_from_dense(int32_t S, int32_t R, const int16_t *left, const int16_t *right, int32_t *nnz_left,
                          int32_t *nnz_net, int32_t *react_ptr, int32_t *react_species, int32_t *react_exp,
                          int32_t *net_ptr, int32_t *net_reaction, int32_t *net_coef)
{
    if (S < 1 || R < 1 || !left || !right) return fail(LASSIE_ERR_INVALID, "bad dense stoichiometry");
    int32_t nl = 0, nh = 0;
    for (size_t e = 0; e < (size_t)S * R; e++) {
        nl += left[e] != 0;
        nh += (right[e] - left[e]) != 0;
    }
    if (nnz_left) *nnz_left = nl;
    if (nnz_net) *nnz_net = nh;
    if (!react_ptr) return LASSIE_OK;
    if (!react_species || !react_exp || !net_ptr || !net_reaction || !net_coef)
        return fail(LASSIE_ERR_INVALID, "output arrays missing");
    // create_sparse2 (Simulation.cuh:1107-1130): reaction rows, species ascending
    int32_t c = 0;
    for (int r = 0; r < R; r++) {
        react_ptr[r] = c;
        for (int s = 0; s < S; s++) {
            const int16_t v = left[(size_t)r * S + s];
            if (v != 0) {
                react_species[c] = s;
                react_exp[c] = v;
                c++;
            }
        }
    }
    react_ptr[R] = c;
    // createTransposte + create_sparse (Simulation.cuh:1061-1105): H = (right-left)^T
    c = 0;
    for (int s = 0; s < S; s++) {
        net_ptr[s] = c;
        for (int r = 0; r < R; r++) {
            const int v = (int)right[(size_t)r * S + s] - (int)left[(size_t)r * S + s];
            if (v != 0) {
                net_reaction[c] = r;
                net_coef[c] = v;
                c++;
            }
        }
    }
    net_ptr[S] = c;
    return LASSIE_OK;
}

int lassie_model_create(const lassie_model_desc *d, lassie_model **out)
{
    if (!out) return fail(LASSIE_ERR_INVALID, "out is NULL");
    *out = nullptr;
    int rc = validate_desc(d);
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(LASSIE_ERR_HIP, "no HIP device visible");
    const int S = d->n_species, R = d->n_reactions;
    lassie_model *m = new lassie_model();
    m->S = S;
    m->R = R;
    (void)hipGetDevice(&m->device);
    const int nl = d->react_ptr[R], nh = d->net_ptr[S];
    m->lptr.assign(d->react_ptr, d->react_ptr + R + 1);
    m->lsp.assign(d->react_species, d->react_species + nl);
    m->lexp.assign(d->react_exp, d->react_exp + nl);
    m->hptr.assign(d->net_ptr, d->net_ptr + S + 1);
    m->hrx.assign(d->net_reaction, d->net_reaction + nh);
    m->hcoef.assign(d->net_coef, d->net_coef + nh);
    m->k.assign(d->k, d->k + R);
    m->x0.assign(d->x0, d->x0 + S);
    m->feed.assign(S, 0.0);
    if (d->feed) m->feed.assign(d->feed, d->feed + S);
    m->atol.assign(S, 1e-12);
    if (d->atol) m->atol.assign(d->atol, d->atol + S);
    m->hk.resize(nh);
    for (int e = 0; e < nh; e++) m->hk[e] = (double)m->hcoef[e] * m->k[m->hrx[e]];
    auto bail = [&](int code) {
        free_model(m);
        return code;
    };
    if ((rc = dev_upload(&m->d_lptr, m->lptr.data(), m->lptr.size()))) return bail(rc);
    if ((rc = dev_upload(&m->d_lsp, m->lsp.data(), m->lsp.size()))) return bail(rc);
    if ((rc = dev_upload(&m->d_lexp, m->lexp.data(), m->lexp.size()))) return bail(rc);
    if ((rc = dev_upload(&m->d_hptr, m->hptr.data(), m->hptr.size()))) return bail(rc);
    if ((rc = dev_upload(&m->d_hrx, m->hrx.data(), m->hrx.size()))) return bail(rc);
    if ((rc = dev_upload(&m->d_hk, m->hk.data(), m->hk.size()))) return bail(rc);
    if ((rc = dev_upload(&m->d_x0, m->x0.data(), m->x0.size()))) return bail(rc);
    if ((rc = dev_upload(&m->d_feed, m->feed.data(), m->feed.size()))) return bail(rc);
    if ((rc = dev_upload(&m->d_atol, m->atol.data(), m->atol.size()))) return bail(rc);
    hipError_t e;
    if ((e = hipMalloc((void **)&m->d_vec, sizeof(double) * 12 * (size_t)S)) != hipSuccess ||
        (e = hipMalloc((void **)&m->d_piv, sizeof(int) * (size_t)S)) != hipSuccess ||
        (e = hipMalloc((void **)&m->d_perm, sizeof(int) * (size_t)S)) != hipSuccess ||
        (e = hipMalloc((void **)&m->d_moved, sizeof(int) * (size_t)S)) != hipSuccess ||
        (e = hipMalloc((void **)&m->d_ctl, sizeof(lassie::Ctl))) != hipSuccess ||
        (e = hipMalloc((void **)&m->d_gs, sizeof(lassie::GridSync))) != hipSuccess ||
        (e = hipMalloc((void **)&m->d_stats, sizeof(long long) * 16)) != hipSuccess ||
        (e = hipMalloc((void **)&m->d_dstats, sizeof(double) * 8)) != hipSuccess ||
        (e = hipStreamCreate(&m->stream)) != hipSuccess || (e = hipEventCreate(&m->ev0)) != hipSuccess ||
        (e = hipEventCreate(&m->ev1)) != hipSuccess) {
        fail(LASSIE_ERR_HIP, std::string("allocation: ") + hipGetErrorString(e));
        return bail(LASSIE_ERR_HIP);
    }
    // pivot bookkeeping starts as valid row indices
    if ((e = hipMemset(m->d_perm, 0, sizeof(int) * (size_t)S)) != hipSuccess ||
        (e = hipMemset(m->d_moved, 0, sizeof(int) * (size_t)S)) != hipSuccess) {
        fail(LASSIE_ERR_HIP, std::string("hipMemset: ") + hipGetErrorString(e));
        return bail(LASSIE_ERR_HIP);
    }
    if (S <= lassie::kMaxDenseN) {
        if ((e = hipMalloc((void **)&m->d_M, sizeof(double) * (size_t)S * S)) != hipSuccess) {
            fail(LASSIE_ERR_NOMEM, std::string("dense Newton matrix: ") + hipGetErrorString(e));
            return bail(LASSIE_ERR_NOMEM);
        }
    }
    *out = m;
    return LASSIE_OK;
}

int lassie_model_destroy(lassie_model *m)
{
    if (!m) return LASSIE_OK;
    free_model(m);
    return LASSIE_OK;
}

}  // extern "C"

template <typename T>
static int ensure(T **p, size_t *cap, size_t n)
{
    if (*cap >= n && *p) return LASSIE_OK;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    HIP_TRY(hipMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(T)));
    *cap = n;
    return LASSIE_OK;
}

extern "C" int lassie_solve(lassie_model *m, const lassie_opts *opts, const double *t, int32_t nt,
                            const int32_t *cs, int32_t ncs, double *out, double *y_final, double *progress,
                            int64_t progress_cap, lassie_stats *stats)
{
    if (!m) return fail(LASSIE_ERR_INVALID, "model is NULL");
    if (!t || nt < 1) return fail(LASSIE_ERR_INVALID, "need at least one sampling time");
    if (ncs < 0 || (ncs > 0 && (!cs || !out))) return fail(LASSIE_ERR_INVALID, "bad species selection");
    for (int j = 0; j < ncs; j++)
        if (cs[j] < 0 || cs[j] >= m->S) return fail(LASSIE_ERR_INVALID, "cs entry out of range");
    lassie_opts o;
    if (opts)
        o = *opts;
    else
        lassie_opts_default(&o);
    if (o.precision != LASSIE_PREC_FP64)
        return fail(LASSIE_ERR_UNSUPPORTED, "single-model solve supports LASSIE_PREC_FP64");
    if (m->S > lassie::kMaxDenseN)
        return fail(LASSIE_ERR_UNSUPPORTED, "dense Newton matrix limited to 4096 species");
    if (progress_cap < 0 || (progress_cap > 0 && !progress)) return fail(LASSIE_ERR_INVALID, "bad progress buffer");
    HIP_TRY(hipSetDevice(m->device));
    int rc;
    if ((rc = ensure(&m->d_tv, &m->cap_tv, (size_t)nt))) return rc;
    if ((rc = ensure(&m->d_cs, &m->cap_cs, (size_t)ncs))) return rc;
    if ((rc = ensure(&m->d_out, &m->cap_out, (size_t)nt * ncs))) return rc;
    if (progress_cap > 0 && (rc = ensure(&m->d_prog, &m->cap_prog, (size_t)progress_cap))) return rc;
    HIP_TRY(hipMemcpyAsync(m->d_tv, t, sizeof(double) * nt, hipMemcpyHostToDevice, m->stream));
    if (ncs) HIP_TRY(hipMemcpyAsync(m->d_cs, cs, sizeof(int) * ncs, hipMemcpyHostToDevice, m->stream));
    HIP_TRY(hipMemsetAsync(m->d_ctl, 0, sizeof(lassie::Ctl), m->stream));
    HIP_TRY(hipMemsetAsync(m->d_gs, 0, sizeof(lassie::GridSync), m->stream));
    HIP_TRY(hipMemsetAsync(m->d_stats, 0, sizeof(long long) * 16, m->stream));

    const int S = m->S;
    const int nb = lassie::dense_panel_width(S);
    const int P = (S + nb - 1) / nb;
    int maxg = 0;
    if (lassie::max_coop_groups(S, &maxg) != 0 || maxg < 1)
        return fail(LASSIE_ERR_HIP, "cannot size the cooperative launch");
    lassie::SolveArgs a{};
    a.m.S = S;
    a.m.R = m->R;
    a.m.lptr = m->d_lptr;
    a.m.lsp = m->d_lsp;
    a.m.lexp = m->d_lexp;
    a.m.hptr = m->d_hptr;
    a.m.hrx = m->d_hrx;
    a.m.hk = m->d_hk;
    a.m.feed = m->d_feed;
    a.m.atol = m->d_atol;
    a.x0 = m->d_x0;
    a.rkf_step = o.rkf_step;
    a.be_step = o.be_step;
    a.newton_tol = o.newton_tol;
    a.stiffness_tol = o.stiffness_tol;
    a.rtol = o.rtol;
    a.newton_iter = o.newton_iter;
    a.bdf_regrow = o.bdf_regrow;
    a.max_rkf_retries = o.max_rkf_retries;
    a.max_iters = o.max_iters;
    a.tv = m->d_tv;
    a.nt = nt;
    a.cs = m->d_cs;
    a.ncs = ncs;
    a.out = m->d_out;
    a.progress = progress_cap > 0 ? m->d_prog : nullptr;
    a.progress_cap = progress_cap;
    double *v = m->d_vec;
    a.y1 = v;
    a.u = v + (size_t)S;
    a.u2 = v + 2 * (size_t)S;
    a.k1 = v + 3 * (size_t)S;
    a.k2 = v + 4 * (size_t)S;
    a.k3 = v + 5 * (size_t)S;
    a.k4 = v + 6 * (size_t)S;
    a.k5 = v + 7 * (size_t)S;
    a.k6 = v + 8 * (size_t)S;
    a.w1 = v + 9 * (size_t)S;
    a.yv = v + 10 * (size_t)S;
    a.x = v + 11 * (size_t)S;
    a.M = m->d_M;
    a.piv = m->d_piv;
    a.perm_src = m->d_perm;
    a.moved_to = m->d_moved;
    a.ctl = m->d_ctl;
    a.gs = m->d_gs;
    a.stats_out = m->d_stats;
    a.dstats_out = m->d_dstats;
    a.G = std::max(1, std::min(P, maxg));

    HIP_TRY(hipEventRecord(m->ev0, m->stream));
    HIP_TRY(lassie::launch_solve(a, m->stream));
    HIP_TRY(hipEventRecord(m->ev1, m->stream));
    long long st[16];
    double dst[8];
    HIP_TRY(hipMemcpyAsync(st, m->d_stats, sizeof(st), hipMemcpyDeviceToHost, m->stream));
    HIP_TRY(hipMemcpyAsync(dst, m->d_dstats, sizeof(double) * 7, hipMemcpyDeviceToHost, m->stream));
    lassie::Ctl ctl;
    HIP_TRY(hipMemcpyAsync(&ctl, m->d_ctl, sizeof(ctl), hipMemcpyDeviceToHost, m->stream));
    if (ncs) HIP_TRY(hipMemcpyAsync(out, m->d_out, sizeof(double) * (size_t)nt * ncs, hipMemcpyDeviceToHost, m->stream));
    if (y_final) HIP_TRY(hipMemcpyAsync(y_final, a.y1, sizeof(double) * S, hipMemcpyDeviceToHost, m->stream));
    int64_t nprog = 0;
    HIP_TRY(hipStreamSynchronize(m->stream));
    if (ctl.abort) return fail(LASSIE_ERR_DEVICE, "device-side wait exceeded its bound");
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, m->ev0, m->ev1));
    if (progress_cap > 0) {
        nprog = std::min<int64_t>(progress_cap, st[0]);
        if (nprog > 0) HIP_TRY(hipMemcpy(progress, m->d_prog, sizeof(double) * nprog, hipMemcpyDeviceToHost));
    }
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->loop_iters = st[0];
        stats->rkf_attempts = st[1];
        stats->rkf_accepts = st[2];
        stats->bdf_steps = st[3];
        stats->newton_iters = st[4];
        stats->lu_failures = st[5];
        stats->nan_repairs = st[6];
        stats->samples = st[7];
        stats->lu_factorizations = st[8];
        stats->t_final = dst[0];
        stats->dt_final = dst[1];
        stats->newdt_final = dst[2];
        stats->kernel_ms = ms;
        stats->phase_ms[0] = dst[3];
        stats->phase_ms[1] = dst[4];
        stats->phase_ms[2] = dst[5];
        stats->phase_ms[3] = dst[6];
    }
    if (st[9]) return fail(LASSIE_ERR_MAXITER, "max_iters reached before the last sample");
    return LASSIE_OK;
}
