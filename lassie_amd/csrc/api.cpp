// C ABI of liblassie (include/lassie.h): model upload, solve launch, helpers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/lassie.h"
#include "adjoint.h"
#include "batch.h"
#include "model.h"
#include "solver.h"
#include "wave.h"

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
    // index maps of the reverse pass: H^T by reaction, reactant entries by species
    std::vector<int> htptr, htsp, htcoef, ltptr, ltent, lrx;
    int *d_htptr = nullptr, *d_htsp = nullptr, *d_htcoef = nullptr, *d_ltptr = nullptr, *d_ltent = nullptr,
        *d_lrx = nullptr;
    double *d_k = nullptr;
    unsigned long long version = 0;  // bumped by lassie_model_set_rates
    int *d_lptr = nullptr, *d_lsp = nullptr, *d_lexp = nullptr, *d_hptr = nullptr, *d_hrx = nullptr;
    double *d_hk = nullptr, *d_x0 = nullptr, *d_feed = nullptr, *d_atol = nullptr;
    lassie::HRec *d_recs = nullptr;
    std::vector<lassie::HRec> recs;
    // single-precision loop: sign * k formed in float (as the reference's
    // Simulation<float> does), built on first use after each rate change
    double *d_hk32 = nullptr;
    lassie::HRec *d_recs32 = nullptr;
    unsigned long long version32 = ~0ULL;
    // precision form: FP32 copy of the Newton matrix and the refinement
    // vectors, allocated on first use
    float *d_M32 = nullptr;
    double *d_rvec = nullptr;
    double *d_vec = nullptr;  // work vectors, 12 x S
    double *d_M = nullptr;    // dense S x S Newton matrix (column-major)
    int *d_piv = nullptr, *d_perm = nullptr, *d_moved = nullptr, *d_rows = nullptr, *d_rowcnt = nullptr;
    int *d_pfull = nullptr;             // final row order of the factors
    double *d_tri = nullptr;            // distributed triangular solves: 3 x S
    unsigned int *d_tflag = nullptr;    // their per-block stamps
    size_t tflag_words = 0;
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
                    m->d_atol, m->d_vec, m->d_M, m->d_piv, m->d_perm, m->d_moved, m->d_rows, m->d_rowcnt, m->d_ctl, m->d_gs,
                    m->d_stats, m->d_dstats, m->d_tv, m->d_out, m->d_prog, m->d_cs, m->d_recs,
                    m->d_htptr, m->d_htsp, m->d_htcoef, m->d_ltptr, m->d_ltent, m->d_lrx, m->d_k,
                    m->d_pfull, m->d_tri, m->d_tflag, m->d_hk32, m->d_recs32, m->d_M32, m->d_rvec};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (m->ev0) (void)hipEventDestroy(m->ev0);
    if (m->ev1) (void)hipEventDestroy(m->ev1);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

namespace lassie {

int api_fail(int code, const std::string &msg) { return fail(code, msg); }
int model_species(const lassie_model *m) { return m->S; }
int model_device(const lassie_model *m) { return m->device; }
int model_nnz_net(const lassie_model *m) { return m->hptr[m->S]; }
hipStream_t model_stream(const lassie_model *m) { return m->stream; }
void model_events(const lassie_model *m, hipEvent_t *ev0, hipEvent_t *ev1)
{
    *ev0 = m->ev0;
    *ev1 = m->ev1;
}

void fill_batch_inst(const lassie_model *m, BatchInst *b)
{
    b->S = m->S;
    b->R = m->R;
    b->lptr = m->d_lptr;
    b->lsp = m->d_lsp;
    b->lexp = m->d_lexp;
    b->hptr = m->d_hptr;
    b->hrx = m->d_hrx;
    b->recs = m->d_recs;
    b->hk = m->d_hk;
    b->x0 = m->d_x0;
    b->feed = m->d_feed;
    b->atol = m->d_atol;
    b->out_off = b->y_off = b->tape_off = b->ckpt_off = 0;
}

// c: kBatchStats counters of one instance; d: its t, dt, newdt
void fill_batch_stats(const long long *c, const double *d, float ms, int slots, lassie_stats *s)
{
    std::memset(s, 0, sizeof(*s));
    s->loop_iters = c[0];
    s->rkf_attempts = c[1];
    s->rkf_accepts = c[2];
    s->bdf_steps = c[3];
    s->newton_iters = c[4];
    s->lu_failures = c[5];
    s->nan_repairs = c[6];
    s->samples = c[7];
    s->lu_factorizations = c[8];
    s->refine_iters = c[10];
    s->refine_fallbacks = c[11];
    s->t_final = d[0];
    s->dt_final = d[1];
    s->newdt_final = d[2];
    s->kernel_ms = ms;
    s->groups = slots;
    for (int q = 0; q < 4; q++) s->phase_ms[q] = (double)c[14 + q] * 1e-5;
}

BatchSplit split_batch(lassie_model *const *models, const std::vector<BatchInst> &inst, bool allow_wave)
{
    BatchSplit sp;
    const int n = (int)inst.size();
    std::vector<int> wg, wv;
    for (int i = 0; i < n; i++) {
        const int nh = models[i]->hptr[models[i]->S];
        if (allow_wave && inst[i].S <= kWaveMaxS && nh <= kWaveMaxNh) {
            wv.push_back(i);
            sp.smax_wave = std::max(sp.smax_wave, inst[i].S);
            sp.nh_max_wave = std::max(sp.nh_max_wave, nh);
        } else {
            wg.push_back(i);
            sp.smax_wg = std::max(sp.smax_wg, inst[i].S);
        }
    }
    // most expensive first (dense LU work ~ S^3) so the tail of each queue is short
    auto by_size = [&](int a, int b) { return inst[a].S > inst[b].S; };
    std::stable_sort(wg.begin(), wg.end(), by_size);
    std::stable_sort(wv.begin(), wv.end(), by_size);
    sp.n_wg = (int)wg.size();
    sp.n_wave = (int)wv.size();
    sp.order = wg;
    sp.order.insert(sp.order.end(), wv.begin(), wv.end());
    if (sp.n_wave) sp.wave_lds = wave_lds_bytes(sp.smax_wave, sp.nh_max_wave);
    return sp;
}

}  // namespace lassie

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
    o->refine_tol = 1e-14;
    o->refine_max_iter = 4;
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
    // net entries with their reactions' reactants inlined (see HRec)
    m->recs.resize(nh);
    for (int e = 0; e < nh; e++) {
        const int r = m->hrx[e];
        const int a0 = m->lptr[r], na = m->lptr[r + 1] - a0;
        lassie::HRec rc;
        rc.hk = m->hk[e];
        bool inl = na <= 2 && S <= 32767;
        for (int j = 0; j < na && inl; j++) inl = m->lexp[a0 + j] >= 0 && m->lexp[a0 + j] <= 255;
        if (na == 0) {
            rc.sp01 = 0;
            rc.meta = 0;
        } else if (inl) {
            const int s0 = m->lsp[a0], e0 = m->lexp[a0];
            const int s1 = na == 2 ? m->lsp[a0 + 1] : 0, e1 = na == 2 ? m->lexp[a0 + 1] : 0;
            rc.sp01 = s0 | (s1 << 16);
            rc.meta = e0 | (e1 << 8) | (na << 16);
        } else {
            rc.sp01 = r;
            rc.meta = 3 << 16;
        }
        m->recs[e] = rc;
    }
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
    if ((rc = dev_upload(&m->d_recs, m->recs.data(), m->recs.size()))) return bail(rc);
    if ((rc = dev_upload(&m->d_x0, m->x0.data(), m->x0.size()))) return bail(rc);
    if ((rc = dev_upload(&m->d_feed, m->feed.data(), m->feed.size()))) return bail(rc);
    if ((rc = dev_upload(&m->d_atol, m->atol.data(), m->atol.size()))) return bail(rc);
    // reverse-pass maps (stable: entries keep the CSR order within a row)
    m->htptr.assign(R + 1, 0);
    for (int e = 0; e < nh; e++) m->htptr[m->hrx[e] + 1]++;
    for (int r = 0; r < R; r++) m->htptr[r + 1] += m->htptr[r];
    m->htsp.resize(nh);
    m->htcoef.resize(nh);
    {
        std::vector<int> fill(m->htptr.begin(), m->htptr.end() - 1);
        for (int sp = 0; sp < S; sp++)
            for (int e = m->hptr[sp]; e < m->hptr[sp + 1]; e++) {
                const int q = fill[m->hrx[e]]++;
                m->htsp[q] = sp;
                m->htcoef[q] = m->hcoef[e];
            }
    }
    m->ltptr.assign(S + 1, 0);
    m->lrx.resize(nl);
    for (int r = 0; r < R; r++)
        for (int j = m->lptr[r]; j < m->lptr[r + 1]; j++) {
            m->lrx[j] = r;
            m->ltptr[m->lsp[j] + 1]++;
        }
    for (int sp = 0; sp < S; sp++) m->ltptr[sp + 1] += m->ltptr[sp];
    m->ltent.resize(nl);
    {
        std::vector<int> fill(m->ltptr.begin(), m->ltptr.end() - 1);
        for (int j = 0; j < nl; j++) m->ltent[fill[m->lsp[j]]++] = j;
    }
    if ((rc = dev_upload(&m->d_k, m->k.data(), m->k.size()))) return bail(rc);
    if ((rc = dev_upload(&m->d_htptr, m->htptr.data(), m->htptr.size()))) return bail(rc);
    if ((rc = dev_upload(&m->d_htsp, m->htsp.data(), m->htsp.size()))) return bail(rc);
    if ((rc = dev_upload(&m->d_htcoef, m->htcoef.data(), m->htcoef.size()))) return bail(rc);
    if ((rc = dev_upload(&m->d_ltptr, m->ltptr.data(), m->ltptr.size()))) return bail(rc);
    if ((rc = dev_upload(&m->d_ltent, m->ltent.data(), m->ltent.size()))) return bail(rc);
    if ((rc = dev_upload(&m->d_lrx, m->lrx.data(), m->lrx.size()))) return bail(rc);
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
        const int P = (S + lassie::dense_panel_width(S) - 1) / lassie::dense_panel_width(S);
        if ((e = hipMalloc((void **)&m->d_rows, sizeof(int) * (size_t)P * S)) != hipSuccess ||
            (e = hipMalloc((void **)&m->d_rowcnt, sizeof(int) * (size_t)P)) != hipSuccess) {
            fail(LASSIE_ERR_NOMEM, std::string("panel row lists: ") + hipGetErrorString(e));
            return bail(LASSIE_ERR_NOMEM);
        }
        const int nblk = (S + lassie::kTriBlock - 1) / lassie::kTriBlock;
        m->tflag_words = 2 * (size_t)nblk * lassie::kFlagStride;
        if ((e = hipMalloc((void **)&m->d_pfull, sizeof(int) * (size_t)S)) != hipSuccess ||
            (e = hipMalloc((void **)&m->d_tri, sizeof(double) * 3 * (size_t)S)) != hipSuccess ||
            (e = hipMalloc((void **)&m->d_tflag, sizeof(unsigned int) * m->tflag_words)) != hipSuccess) {
            fail(LASSIE_ERR_NOMEM, std::string("triangular-solve buffers: ") + hipGetErrorString(e));
            return bail(LASSIE_ERR_NOMEM);
        }
        if ((e = hipMalloc((void **)&m->d_M, sizeof(double) * ((size_t)S * S + lassie::kLuSink))) != hipSuccess) {
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
    if (o.precision != LASSIE_PREC_FP64 && o.precision != LASSIE_PREC_FP32_LOOP &&
        o.precision != LASSIE_PREC_FP32_REFINE && o.precision != LASSIE_PREC_BF16_REFINE)
        return fail(LASSIE_ERR_UNSUPPORTED, "unknown precision");
    const int fp32 = o.precision == LASSIE_PREC_FP32_LOOP ? 1 : 0;
    const int refine = o.precision == LASSIE_PREC_FP32_REFINE ? 1 : (o.precision == LASSIE_PREC_BF16_REFINE ? 2 : 0);
    if (refine && !(o.refine_tol >= 0.0)) return fail(LASSIE_ERR_INVALID, "refine_tol must be >= 0");
    if (m->S > lassie::kMaxDenseN)
        return fail(LASSIE_ERR_UNSUPPORTED, "dense Newton matrix limited to 4096 species");
    if (progress_cap < 0 || (progress_cap > 0 && !progress)) return fail(LASSIE_ERR_INVALID, "bad progress buffer");
    HIP_TRY(hipSetDevice(m->device));
    int rc;
    const int nh = (int)m->hrx.size();
    if (fp32 && nh && m->version32 != m->version) {
        // (float)coef * (float)k: the float product `signum * c_vector[k_pos]`
        std::vector<double> hk32(nh);
        std::vector<lassie::HRec> recs32(m->recs);
        for (int e = 0; e < nh; e++) {
            hk32[e] = (double)((float)m->hcoef[e] * (float)m->k[m->hrx[e]]);
            recs32[e].hk = hk32[e];
        }
        if (!m->d_hk32 && (rc = dev_upload(&m->d_hk32, hk32.data(), hk32.size()))) return rc;
        if (!m->d_recs32 && (rc = dev_upload(&m->d_recs32, recs32.data(), recs32.size()))) return rc;
        HIP_TRY(hipMemcpy(m->d_hk32, hk32.data(), sizeof(double) * nh, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(m->d_recs32, recs32.data(), sizeof(lassie::HRec) * nh, hipMemcpyHostToDevice));
        m->version32 = m->version;
    }
    if (refine && !m->d_M32) {
        const size_t S2 = (size_t)m->S * m->S;
        if (hipMalloc((void **)&m->d_M32, sizeof(float) * (S2 + lassie::kLuSink)) != hipSuccess) {
            m->d_M32 = nullptr;
            return fail(LASSIE_ERR_NOMEM, "FP32 Newton matrix");
        }
    }
    if (refine && !m->d_rvec) {
        if (hipMalloc((void **)&m->d_rvec, sizeof(double) * 4 * (size_t)m->S) != hipSuccess) {
            m->d_rvec = nullptr;
            return fail(LASSIE_ERR_NOMEM, "refinement vectors");
        }
    }
    if ((rc = ensure(&m->d_tv, &m->cap_tv, (size_t)nt))) return rc;
    if ((rc = ensure(&m->d_cs, &m->cap_cs, (size_t)ncs))) return rc;
    if ((rc = ensure(&m->d_out, &m->cap_out, (size_t)nt * ncs))) return rc;
    if (progress_cap > 0 && (rc = ensure(&m->d_prog, &m->cap_prog, (size_t)progress_cap))) return rc;
    HIP_TRY(hipMemcpyAsync(m->d_tv, t, sizeof(double) * nt, hipMemcpyHostToDevice, m->stream));
    if (ncs) HIP_TRY(hipMemcpyAsync(m->d_cs, cs, sizeof(int) * ncs, hipMemcpyHostToDevice, m->stream));
    HIP_TRY(hipMemsetAsync(m->d_ctl, 0, sizeof(lassie::Ctl), m->stream));
    HIP_TRY(hipMemsetAsync(m->d_gs, 0, sizeof(lassie::GridSync), m->stream));
    HIP_TRY(hipMemsetAsync(m->d_stats, 0, sizeof(long long) * 16, m->stream));
    if (m->d_tflag) HIP_TRY(hipMemsetAsync(m->d_tflag, 0, sizeof(unsigned int) * m->tflag_words, m->stream));

    const int S = m->S;
    const int nb = lassie::dense_panel_width(S);
    const int P = (S + nb - 1) / nb;
    int cus = 0, max_lds = 0;
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, m->device));
    HIP_TRY(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, m->device));
    const size_t static_lds = 8 * 1024;
    const size_t model_lds = lassie::lds_model_bytes(S, m->hptr[S]);
    // opts.reserved bit 1 keeps the control group's model and vectors in
    // global memory (diagnostics)
    const bool lds_model = model_lds + static_lds <= (size_t)max_lds && !(o.reserved & 2);
    size_t dyn = lds_model ? model_lds : sizeof(double) * (size_t)S;
    int per_cu = 0;
    if (lassie::solver_occupancy(S, fp32, refine, dyn, &per_cu) != 0 || per_cu < 1)
        return fail(LASSIE_ERR_HIP, "cannot size the cooperative launch");
    // opts.reserved bits 8..15 cap the number of groups (diagnostics)
    const int gcap = (o.reserved >> 8) & 0xff;
    int G = std::max(1, std::min({P, cus * per_cu, gcap ? gcap : 64}));
    // FP64 with 16-column panels and one strip per group: groups >= 1 keep
    // their strip in LDS for the whole factorization (the dynamic LDS that
    // group 0 uses for the model).  opts.reserved bit 2 turns this off
    // (diagnostics).
    const int strip_ld = S + 2;
    bool lds_strip = false;
    if (!fp32 && nb == 16 && G == P && P > 1 && !(o.reserved & 4)) {
        const size_t dyn2 = std::max(dyn, sizeof(double) * 16 * (size_t)strip_ld);
        int pc = 0;
        if (lassie::solver_occupancy(S, fp32, refine, dyn2, &pc) == 0 && pc >= 1 && std::min(P, cus * pc) == P) {
            dyn = dyn2;
            per_cu = pc;
            lds_strip = true;
        }
    }
    lassie::SolveArgs a{};
    a.m.S = S;
    a.m.R = m->R;
    a.m.lptr = m->d_lptr;
    a.m.lsp = m->d_lsp;
    a.m.lexp = m->d_lexp;
    a.m.hptr = m->d_hptr;
    a.m.hrx = m->d_hrx;
    a.m.hk = fp32 ? m->d_hk32 : m->d_hk;
    a.m.feed = m->d_feed;
    a.m.atol = m->d_atol;
    a.recs = fp32 ? m->d_recs32 : m->d_recs;
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
    a.max_newton = o.max_newton;
    a.tv = m->d_tv;
    a.nt = nt;
    a.cs = m->d_cs;
    a.ncs = ncs;
    a.out = m->d_out;
    a.progress = progress_cap > 0 ? m->d_prog : nullptr;
    a.progress_cap = progress_cap;
    a.gvec = m->d_vec;
    a.M = m->d_M;
    a.piv = m->d_piv;
    a.perm_src = m->d_perm;
    a.moved_to = m->d_moved;
    a.rows = m->d_rows;
    a.rowcnt = m->d_rowcnt;
    a.perm_full = m->d_pfull;
    a.tri = m->d_tri;
    a.tflag = m->d_tflag;
    a.ctl = m->d_ctl;
    a.gs = m->d_gs;
    a.stats_out = m->d_stats;
    a.dstats_out = m->d_dstats;
    a.G = G;
    a.lds_model = lds_model ? 1 : 0;
    a.fp32 = fp32;
    a.yfin = m->d_vec + 6 * (size_t)S;  // past the 11 float vectors of the single-precision loop
    a.lds_strip = lds_strip ? 1 : 0;
    a.strip_ld = strip_ld;
    a.refine = refine;
    a.refine_max_iter = o.refine_max_iter == 0 ? 0 : std::max(1, o.refine_max_iter);
    a.refine_tol = o.refine_tol;
    a.M32 = m->d_M32;
    a.rvec = m->d_rvec;

    HIP_TRY(hipEventRecord(m->ev0, m->stream));
    HIP_TRY(lassie::launch_solve(a, dyn, m->stream));
    HIP_TRY(hipEventRecord(m->ev1, m->stream));
    long long st[16];
    double dst[8];
    HIP_TRY(hipMemcpyAsync(st, m->d_stats, sizeof(st), hipMemcpyDeviceToHost, m->stream));
    HIP_TRY(hipMemcpyAsync(dst, m->d_dstats, sizeof(double) * 7, hipMemcpyDeviceToHost, m->stream));
    lassie::Ctl ctl;
    HIP_TRY(hipMemcpyAsync(&ctl, m->d_ctl, sizeof(ctl), hipMemcpyDeviceToHost, m->stream));
    if (y_final)
        HIP_TRY(hipMemcpyAsync(y_final, fp32 ? a.yfin : a.gvec, sizeof(double) * S, hipMemcpyDeviceToHost,
                               m->stream));
    int64_t nprog = 0;
    HIP_TRY(hipStreamSynchronize(m->stream));
    // only the stored samples (rows past them are left untouched when max_iters stops the loop)
    const size_t rows = (size_t)std::min<long long>(std::max<long long>(st[7], 0), nt);
    if (ncs && rows) HIP_TRY(hipMemcpy(out, m->d_out, sizeof(double) * rows * ncs, hipMemcpyDeviceToHost));
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
        stats->refine_iters = st[10];
        stats->refine_fallbacks = st[11];
        stats->t_final = dst[0];
        stats->dt_final = dst[1];
        stats->newdt_final = dst[2];
        stats->kernel_ms = ms;
        stats->phase_ms[0] = dst[3];
        stats->phase_ms[1] = dst[4];
        stats->phase_ms[2] = dst[5];
        stats->phase_ms[3] = dst[6];
        stats->groups = G;
        stats->lds_model = lds_model ? 1 : 0;
    }
    if (st[9]) return fail(LASSIE_ERR_MAXITER, "max_iters reached before the last sample");
    return LASSIE_OK;
}

// ---------------------------------------------------------------------------
// Batched form: many uploaded models in one persistent launch; with a tape it
// is also the forward pass of the differentiation form.

struct lassie_tape {
    int device = 0;
    std::vector<lassie_model *> models;
    std::vector<unsigned long long> versions;
    lassie_opts opts{};
    int nt = 0;
    long long cap = 0;  // steps per instance
    std::vector<long long> n_steps, max_kind, tape_off, ckpt_off, out_off;
    std::vector<int> grad_finite;  // per model: last gradient finite (1), not (0), none yet (-1)
    long long out_tot = 0;
    double *d_dt = nullptr, *d_ckpt = nullptr;
    int *d_kind = nullptr, *d_sample = nullptr;
};

static void free_tape(lassie_tape *tp)
{
    void *ptrs[] = {tp->d_dt, tp->d_ckpt, tp->d_kind, tp->d_sample};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    delete tp;
}

static int batch_run(lassie_model *const *models, int32_t n_models, const lassie_opts *opts, const double *t,
                     int32_t nt, double *out, double *y_final, lassie_stats *stats, double *elapsed_ms,
                     lassie_tape *tape)
{
    if (!models || n_models < 1) return fail(LASSIE_ERR_INVALID, "need at least one model");
    if (!t || nt < 1 || !out) return fail(LASSIE_ERR_INVALID, "need sampling times and an output buffer");
    lassie_opts o;
    if (opts)
        o = *opts;
    else
        lassie_opts_default(&o);
    if (o.precision < LASSIE_PREC_FP64 || o.precision > LASSIE_PREC_BF16_REFINE)
        return fail(LASSIE_ERR_INVALID, "unknown precision");
    if (tape && o.precision != LASSIE_PREC_FP64)
        return fail(LASSIE_ERR_UNSUPPORTED, "the differentiation form runs the FP64 method");
    int smax = 1;
    for (int i = 0; i < n_models; i++) {
        if (!models[i]) return fail(LASSIE_ERR_INVALID, "NULL model in batch");
        if (models[i]->S > lassie::kBatchMaxS) return fail(LASSIE_ERR_UNSUPPORTED, "batched form supports S <= 512");
        if (models[i]->device != models[0]->device) return fail(LASSIE_ERR_INVALID, "models on different devices");
        smax = std::max(smax, models[i]->S);
    }
    lassie_model *m0 = models[0];
    HIP_TRY(hipSetDevice(m0->device));
    std::vector<lassie::BatchInst> inst(n_models);
    long long out_tot = 0, y_tot = 0, ckpt_tot = 0;
    const long long cap = tape ? tape->cap : 0;
    for (int i = 0; i < n_models; i++) {
        const lassie_model *m = models[i];
        lassie::BatchInst &b = inst[i];
        lassie::fill_batch_inst(m, &b);
        b.out_off = out_tot;
        b.y_off = y_tot;
        b.tape_off = (long long)i * cap;
        b.ckpt_off = ckpt_tot;
        out_tot += (long long)nt * m->S;
        y_tot += m->S;
        ckpt_tot += ((cap + lassie::kCkptEvery - 1) / lassie::kCkptEvery) * m->S;
    }
    const bool allow_wave = !tape && o.precision == LASSIE_PREC_FP64 && !(o.reserved & 2);
    const lassie::BatchSplit sp = lassie::split_batch(models, inst, allow_wave);
    const std::vector<int> &order = sp.order;
    smax = sp.smax_wg;

    int cus = 0, per_cu = 0, wave_per_cu = 0;
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, m0->device));
    if (lassie::batch_occupancy(&per_cu) != 0 || per_cu < 1) return fail(LASSIE_ERR_HIP, "batch occupancy query failed");
    if (sp.n_wave && (lassie::wave_occupancy(sp.wave_lds, &wave_per_cu) != 0 || wave_per_cu < 1))
        return fail(LASSIE_ERR_HIP, "wave kernel occupancy query failed");
    const int slots = std::max(1, std::min(sp.n_wg, cus * per_cu));
    const int wave_blocks = std::max(1, std::min(sp.n_wave, cus * wave_per_cu));
    const int Pmax = (smax + lassie::kBatchNB - 1) / lassie::kBatchNB;
    const size_t ws64_stride = 12 * (size_t)smax + (size_t)smax * smax + lassie::kLuSink;
    const size_t ws32_stride = (size_t)smax * smax + lassie::kLuSink;
    const size_t wsi_stride = 3 * (size_t)smax + 16 + Pmax + (size_t)Pmax * smax;

    lassie::BatchInst *d_inst = nullptr;
    int *d_order = nullptr, *d_next = nullptr, *d_wsi = nullptr;
    double *d_tv = nullptr, *d_out = nullptr, *d_y = nullptr, *d_ws64 = nullptr, *d_dst = nullptr;
    float *d_ws32 = nullptr;
    long long *d_st = nullptr;
    int rc = LASSIE_OK;
    auto cleanup = [&]() {
        void *ptrs[] = {d_inst, d_order, d_next, d_wsi, d_tv, d_out, d_y, d_ws64, d_dst, d_ws32, d_st};
        for (void *p : ptrs)
            if (p) (void)hipFree(p);
    };
    hipError_t e;
    if ((e = hipMalloc((void **)&d_inst, sizeof(lassie::BatchInst) * n_models)) != hipSuccess ||
        (e = hipMalloc((void **)&d_order, sizeof(int) * n_models)) != hipSuccess ||
        (e = hipMalloc((void **)&d_next, 2 * sizeof(int))) != hipSuccess ||
        (e = hipMalloc((void **)&d_wsi, sizeof(int) * wsi_stride * slots)) != hipSuccess ||
        (e = hipMalloc((void **)&d_tv, sizeof(double) * nt)) != hipSuccess ||
        (e = hipMalloc((void **)&d_out, sizeof(double) * out_tot)) != hipSuccess ||
        (e = hipMalloc((void **)&d_y, sizeof(double) * y_tot)) != hipSuccess ||
        (e = hipMalloc((void **)&d_ws64, sizeof(double) * ws64_stride * slots)) != hipSuccess ||
        (e = hipMalloc((void **)&d_dst, sizeof(double) * 4 * n_models)) != hipSuccess ||
        (e = hipMalloc((void **)&d_st, sizeof(long long) * lassie::kBatchStats * n_models)) != hipSuccess ||
        (o.precision != LASSIE_PREC_FP64 &&
         (e = hipMalloc((void **)&d_ws32, sizeof(float) * ws32_stride * slots)) != hipSuccess) ||
        (tape && ((e = hipMalloc((void **)&tape->d_dt, sizeof(double) * cap * n_models)) != hipSuccess ||
                  (e = hipMalloc((void **)&tape->d_kind, sizeof(int) * cap * n_models)) != hipSuccess ||
                  (e = hipMalloc((void **)&tape->d_sample, sizeof(int) * (size_t)nt * n_models)) != hipSuccess ||
                  (e = hipMalloc((void **)&tape->d_ckpt, sizeof(double) * std::max(1LL, ckpt_tot))) != hipSuccess))) {
        cleanup();
        return fail(LASSIE_ERR_NOMEM, std::string("batch workspace: ") + hipGetErrorString(e));
    }
    hipStream_t st = m0->stream;
#define BATCH_TRY(expr)                                                                          \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess) {                                                                  \
            cleanup();                                                                           \
            return fail(LASSIE_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));      \
        }                                                                                        \
    } while (0)
    BATCH_TRY(hipMemcpyAsync(d_inst, inst.data(), sizeof(lassie::BatchInst) * n_models, hipMemcpyHostToDevice, st));
    BATCH_TRY(hipMemcpyAsync(d_order, order.data(), sizeof(int) * n_models, hipMemcpyHostToDevice, st));
    BATCH_TRY(hipMemcpyAsync(d_tv, t, sizeof(double) * nt, hipMemcpyHostToDevice, st));
    BATCH_TRY(lassie::launch_batch_reset(d_next, st, d_out, (size_t)out_tot));
    BATCH_TRY(hipMemsetAsync(d_wsi, 0, sizeof(int) * wsi_stride * slots, st));
    if (tape) BATCH_TRY(hipMemsetAsync(tape->d_sample, 0, sizeof(int) * (size_t)nt * n_models, st));
    lassie::BatchArgs a{};
    a.inst = d_inst;
    a.n_inst = sp.n_wg;
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
    a.precision = o.precision;
    a.refine_tol = o.refine_tol;
    a.refine_max_iter = o.refine_max_iter == 0 ? 0 : std::max(1, o.refine_max_iter);
    a.s_max = smax;
    a.ws64 = d_ws64;
    a.ws64_stride = ws64_stride;
    a.ws32 = d_ws32;
    a.ws32_stride = ws32_stride;
    a.wsi = d_wsi;
    a.wsi_stride = wsi_stride;
    if (tape) {
        a.tape_dt = tape->d_dt;
        a.tape_kind = tape->d_kind;
        a.tape_sample = tape->d_sample;
        a.tape_ckpt = tape->d_ckpt;
        a.tape_cap = cap;
    }
    // the wave-level kernel takes the small instances first, then the
    // workgroup-level kernel the rest (two queues, same stream)
    lassie::BatchArgs aw = a;
    aw.n_inst = sp.n_wave;
    aw.order = d_order + sp.n_wg;
    aw.next = d_next + 1;
    aw.s_max = sp.smax_wave;
    aw.nh_max = sp.nh_max_wave;
    BATCH_TRY(hipEventRecord(m0->ev0, st));
    if (sp.n_wave) BATCH_TRY(lassie::launch_wave(aw, wave_blocks, sp.wave_lds, st));
    if (sp.n_wg) BATCH_TRY(lassie::launch_batch(a, slots, st));
    BATCH_TRY(hipEventRecord(m0->ev1, st));
    BATCH_TRY(hipMemcpyAsync(out, d_out, sizeof(double) * out_tot, hipMemcpyDeviceToHost, st));
    if (y_final) BATCH_TRY(hipMemcpyAsync(y_final, d_y, sizeof(double) * y_tot, hipMemcpyDeviceToHost, st));
    std::vector<long long> hst((size_t)lassie::kBatchStats * n_models);
    std::vector<double> hd(4 * (size_t)n_models);
    BATCH_TRY(hipMemcpyAsync(hst.data(), d_st, sizeof(long long) * hst.size(), hipMemcpyDeviceToHost, st));
    BATCH_TRY(hipMemcpyAsync(hd.data(), d_dst, sizeof(double) * hd.size(), hipMemcpyDeviceToHost, st));
    BATCH_TRY(hipStreamSynchronize(st));
    float ms = 0.f;
    BATCH_TRY(hipEventElapsedTime(&ms, m0->ev0, m0->ev1));
    if (elapsed_ms) *elapsed_ms = ms;
    bool maxit = false, full = false;
    for (int i = 0; i < n_models; i++) {
        const long long *c = &hst[(size_t)i * lassie::kBatchStats];
        if (stats) lassie::fill_batch_stats(c, &hd[4 * (size_t)i], ms, slots, &stats[i]);
        maxit |= c[9] == 1;
        full |= c[9] == 2;
        if (tape) {
            tape->n_steps[i] = c[12];
            tape->max_kind[i] = c[13];
            tape->tape_off[i] = inst[i].tape_off;
            tape->ckpt_off[i] = inst[i].ckpt_off;
            tape->out_off[i] = inst[i].out_off;
        }
    }
    cleanup();
    if (tape) tape->out_tot = out_tot;
    if (full) return fail(LASSIE_ERR_NOMEM, "step record full: raise max_steps");
    if (maxit) return fail(LASSIE_ERR_MAXITER, "max_iters reached before the last sample in at least one instance");
    return rc;
#undef BATCH_TRY
}

extern "C" int lassie_solve_batch(lassie_model *const *models, int32_t n_models, const lassie_opts *opts,
                                  const double *t, int32_t nt, double *out, double *y_final,
                                  lassie_stats *stats, double *elapsed_ms)
{
    return batch_run(models, n_models, opts, t, nt, out, y_final, stats, elapsed_ms, nullptr);
}

extern "C" int lassie_model_set_rates(lassie_model *m, const double *k)
{
    if (!m || !k) return fail(LASSIE_ERR_INVALID, "model and k are required");
    HIP_TRY(hipSetDevice(m->device));
    m->k.assign(k, k + m->R);
    const int nh = (int)m->hrx.size();
    for (int e = 0; e < nh; e++) {
        m->hk[e] = (double)m->hcoef[e] * m->k[m->hrx[e]];
        m->recs[e].hk = m->hk[e];
    }
    HIP_TRY(hipMemcpy(m->d_k, m->k.data(), sizeof(double) * m->R, hipMemcpyHostToDevice));
    if (nh) {
        HIP_TRY(hipMemcpy(m->d_hk, m->hk.data(), sizeof(double) * nh, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(m->d_recs, m->recs.data(), sizeof(lassie::HRec) * nh, hipMemcpyHostToDevice));
    }
    m->version++;
    return LASSIE_OK;
}

extern "C" int lassie_forward_batch(lassie_model *const *models, int32_t n_models, const lassie_opts *opts,
                                    const double *t, int32_t nt, int64_t max_steps, double *out,
                                    lassie_stats *stats, double *elapsed_ms, lassie_tape **tape_out)
{
    if (!tape_out) return fail(LASSIE_ERR_INVALID, "tape_out is NULL");
    *tape_out = nullptr;
    if (max_steps < 1) return fail(LASSIE_ERR_INVALID, "max_steps must be positive");
    if (!models || n_models < 1) return fail(LASSIE_ERR_INVALID, "need at least one model");
    for (int i = 0; i < n_models; i++)
        if (!models[i]) return fail(LASSIE_ERR_INVALID, "NULL model in batch");
    lassie_tape *tp = new lassie_tape();
    tp->device = models[0]->device;
    tp->models.assign(models, models + n_models);
    for (int i = 0; i < n_models; i++) tp->versions.push_back(models[i]->version);
    if (opts)
        tp->opts = *opts;
    else
        lassie_opts_default(&tp->opts);
    tp->nt = nt;
    tp->cap = max_steps;
    tp->n_steps.assign(n_models, 0);
    tp->max_kind.assign(n_models, 0);
    tp->grad_finite.assign(n_models, -1);
    tp->tape_off.assign(n_models, 0);
    tp->ckpt_off.assign(n_models, 0);
    tp->out_off.assign(n_models, 0);
    const int rc = batch_run(models, n_models, opts, t, nt, out, nullptr, stats, elapsed_ms, tp);
    if (rc != LASSIE_OK) {
        const std::string keep = g_err;
        free_tape(tp);
        g_err = keep;
        return rc;
    }
    *tape_out = tp;
    return LASSIE_OK;
}

extern "C" int lassie_tape_info(const lassie_tape *tp, int32_t model, int64_t *n_steps, int32_t *max_newton)
{
    if (!tp || model < 0 || model >= (int)tp->models.size()) return fail(LASSIE_ERR_INVALID, "bad tape or model index");
    if (n_steps) *n_steps = tp->n_steps[model];
    if (max_newton) *max_newton = (int32_t)tp->max_kind[model];
    return LASSIE_OK;
}

extern "C" int lassie_tape_read(const lassie_tape *tp, int32_t model, double *dt, int32_t *kind, int32_t *sample_step)
{
    if (!tp || model < 0 || model >= (int)tp->models.size()) return fail(LASSIE_ERR_INVALID, "bad tape or model index");
    HIP_TRY(hipSetDevice(tp->device));
    const long long n = tp->n_steps[model];
    if (dt && n) HIP_TRY(hipMemcpy(dt, tp->d_dt + tp->tape_off[model], sizeof(double) * n, hipMemcpyDeviceToHost));
    if (kind && n) HIP_TRY(hipMemcpy(kind, tp->d_kind + tp->tape_off[model], sizeof(int) * n, hipMemcpyDeviceToHost));
    if (sample_step)
        HIP_TRY(hipMemcpy(sample_step, tp->d_sample + (size_t)model * tp->nt, sizeof(int) * tp->nt,
                          hipMemcpyDeviceToHost));
    return LASSIE_OK;
}

extern "C" int lassie_adjoint_status(const lassie_tape *tp, int32_t model, int32_t *finite,
                                     int32_t *capped_newton_steps)
{
    if (!tp || model < 0 || model >= (int)tp->models.size()) return fail(LASSIE_ERR_INVALID, "bad tape or model index");
    if (finite) *finite = tp->grad_finite[model];
    if (capped_newton_steps) {
        *capped_newton_steps = 0;
        const long long n = tp->n_steps[model];
        if (n && tp->max_kind[model] >= tp->opts.newton_iter) {
            std::vector<int> kind(n);
            HIP_TRY(hipSetDevice(tp->device));
            HIP_TRY(hipMemcpy(kind.data(), tp->d_kind + tp->tape_off[model], sizeof(int) * n, hipMemcpyDeviceToHost));
            for (int k : kind) *capped_newton_steps += (k >= tp->opts.newton_iter) ? 1 : 0;
        }
    }
    return LASSIE_OK;
}

extern "C" int lassie_tape_destroy(lassie_tape *tp)
{
    if (tp) free_tape(tp);
    return LASSIE_OK;
}

extern "C" int lassie_adjoint_batch(lassie_tape *tp, const double *grad_out, double *grad_log_k, double *grad_x0,
                                    double *elapsed_ms)
{
    if (!tp || !grad_out || !grad_log_k) return fail(LASSIE_ERR_INVALID, "tape, grad_out and grad_log_k are required");
    const int n = (int)tp->models.size();
    for (int i = 0; i < n; i++)
        if (tp->models[i]->version != tp->versions[i])
            return fail(LASSIE_ERR_INVALID, "a model's rates changed after the forward pass");
    lassie_model *m0 = tp->models[0];
    HIP_TRY(hipSetDevice(tp->device));
    int smax = 1, rmax = 1, itmax = 0;
    long long gk_tot = 0, gx_tot = 0;
    std::vector<lassie::AdjInst> inst(n);
    for (int i = 0; i < n; i++) {
        const lassie_model *m = tp->models[i];
        lassie::AdjInst &b = inst[i];
        b.S = m->S;
        b.R = m->R;
        b.lptr = m->d_lptr;
        b.lsp = m->d_lsp;
        b.lexp = m->d_lexp;
        b.hptr = m->d_hptr;
        b.hrx = m->d_hrx;
        b.recs = m->d_recs;
        b.hk = m->d_hk;
        b.x0 = m->d_x0;
        b.feed = m->d_feed;
        b.atol = m->d_atol;
        b.k = m->d_k;
        b.htptr = m->d_htptr;
        b.htsp = m->d_htsp;
        b.htcoef = m->d_htcoef;
        b.ltptr = m->d_ltptr;
        b.ltent = m->d_ltent;
        b.lrx = m->d_lrx;
        b.tape_off = tp->tape_off[i];
        b.ckpt_off = tp->ckpt_off[i];
        b.n_steps = tp->n_steps[i];
        b.g_off = tp->out_off[i];
        b.gk_off = gk_tot;
        b.gx_off = gx_tot;
        gk_tot += m->R;
        gx_tot += m->S;
        smax = std::max(smax, m->S);
        rmax = std::max(rmax, m->R);
        itmax = std::max(itmax, (int)tp->max_kind[i]);
    }
    std::vector<int> order(n);
    for (int i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return inst[a].S > inst[b].S; });
    int cus = 0, per_cu = 0;
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, tp->device));
    if (lassie::adjoint_occupancy(&per_cu) != 0 || per_cu < 1) return fail(LASSIE_ERR_HIP, "adjoint occupancy query failed");
    const int slots = std::max(1, std::min(n, cus * per_cu));
    const int Pmax = (smax + lassie::kBatchNB - 1) / lassie::kBatchNB;
    const size_t ws_stride = (size_t)(lassie::kCkptEvery + 1) * smax + (size_t)smax * smax + lassie::kLuSink +
                             2 * (size_t)(itmax + 1) * smax + 22 * (size_t)smax + (size_t)rmax;
    const size_t wsi_stride = 3 * (size_t)smax + 16 + Pmax + (size_t)Pmax * smax;

    // optional second iterate/correction pair per slot (the segment's longest
    // BDF1 step keeps its replayed iterates for its reverse step)
    const size_t vx2_stride = 2 * (size_t)(itmax + 1) * smax;

    lassie::AdjInst *d_inst = nullptr;
    int *d_order = nullptr, *d_next = nullptr, *d_wsi = nullptr, *d_status = nullptr;
    double *d_gout = nullptr, *d_gk = nullptr, *d_gx = nullptr, *d_ws = nullptr, *d_vx2 = nullptr;
    auto cleanup = [&]() {
        void *ptrs[] = {d_inst, d_order, d_next, d_wsi, d_status, d_gout, d_gk, d_gx, d_ws, d_vx2};
        for (void *p : ptrs)
            if (p) (void)hipFree(p);
    };
    hipError_t e;
    if ((e = hipMalloc((void **)&d_inst, sizeof(lassie::AdjInst) * n)) != hipSuccess ||
        (e = hipMalloc((void **)&d_order, sizeof(int) * n)) != hipSuccess ||
        (e = hipMalloc((void **)&d_next, sizeof(int))) != hipSuccess ||
        (e = hipMalloc((void **)&d_status, sizeof(int) * n)) != hipSuccess ||
        (e = hipMalloc((void **)&d_wsi, sizeof(int) * wsi_stride * slots)) != hipSuccess ||
        (e = hipMalloc((void **)&d_gout, sizeof(double) * tp->out_tot)) != hipSuccess ||
        (e = hipMalloc((void **)&d_gk, sizeof(double) * gk_tot)) != hipSuccess ||
        (e = hipMalloc((void **)&d_gx, sizeof(double) * gx_tot)) != hipSuccess ||
        (e = hipMalloc((void **)&d_ws, sizeof(double) * ws_stride * slots)) != hipSuccess) {
        cleanup();
        return fail(LASSIE_ERR_NOMEM, std::string("adjoint workspace: ") + hipGetErrorString(e));
    }
    if (hipMalloc((void **)&d_vx2, sizeof(double) * vx2_stride * slots) != hipSuccess) {
        (void)hipGetLastError();
        d_vx2 = nullptr;
    }
    hipStream_t st = m0->stream;
#define ADJ_TRY(expr)                                                                            \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess) {                                                                  \
            cleanup();                                                                           \
            return fail(LASSIE_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));      \
        }                                                                                        \
    } while (0)
    ADJ_TRY(hipMemcpyAsync(d_inst, inst.data(), sizeof(lassie::AdjInst) * n, hipMemcpyHostToDevice, st));
    ADJ_TRY(hipMemcpyAsync(d_order, order.data(), sizeof(int) * n, hipMemcpyHostToDevice, st));
    ADJ_TRY(hipMemcpyAsync(d_gout, grad_out, sizeof(double) * tp->out_tot, hipMemcpyHostToDevice, st));
    ADJ_TRY(hipMemsetAsync(d_next, 0, sizeof(int), st));
    ADJ_TRY(hipMemsetAsync(d_wsi, 0, sizeof(int) * wsi_stride * slots, st));
    lassie::AdjArgs a{};
    a.inst = d_inst;
    a.n_inst = n;
    a.order = d_order;
    a.next = d_next;
    a.nt = tp->nt;
    a.tape_dt = tp->d_dt;
    a.tape_kind = tp->d_kind;
    a.tape_sample = tp->d_sample;
    a.tape_ckpt = tp->d_ckpt;
    a.grad_out = d_gout;
    a.grad_logk = d_gk;
    a.grad_x0 = d_gx;
    a.rtol = tp->opts.rtol;
    a.s_max = smax;
    a.r_max = rmax;
    a.it_max = itmax;
    a.ws = d_ws;
    a.ws_stride = ws_stride;
    a.wsi = d_wsi;
    a.wsi_stride = wsi_stride;
    a.status = d_status;
    a.vx2 = d_vx2;
    a.vx2_stride = vx2_stride;
    ADJ_TRY(hipEventRecord(m0->ev0, st));
    ADJ_TRY(lassie::launch_adjoint(a, slots, st));
    ADJ_TRY(hipEventRecord(m0->ev1, st));
    std::vector<int> status(n);
    ADJ_TRY(hipMemcpyAsync(grad_log_k, d_gk, sizeof(double) * gk_tot, hipMemcpyDeviceToHost, st));
    if (grad_x0) ADJ_TRY(hipMemcpyAsync(grad_x0, d_gx, sizeof(double) * gx_tot, hipMemcpyDeviceToHost, st));
    ADJ_TRY(hipMemcpyAsync(status.data(), d_status, sizeof(int) * n, hipMemcpyDeviceToHost, st));
    ADJ_TRY(hipStreamSynchronize(st));
    float ms = 0.f;
    ADJ_TRY(hipEventElapsedTime(&ms, m0->ev0, m0->ev1));
    if (elapsed_ms) *elapsed_ms = ms;
    cleanup();
    for (int i = 0; i < n; i++)
        if (status[i]) return fail(LASSIE_ERR_INVALID, "an LU that succeeded in the forward pass failed in the replay");
    int n_bad = 0;
    for (int i = 0; i < n; i++) {
        bool ok = true;
        for (long long r = 0; r < inst[i].R && ok; r++) ok = std::isfinite(grad_log_k[inst[i].gk_off + r]);
        if (grad_x0)
            for (long long s = 0; s < inst[i].S && ok; s++) ok = std::isfinite(grad_x0[inst[i].gx_off + s]);
        tp->grad_finite[i] = ok ? 1 : 0;
        n_bad += ok ? 0 : 1;
    }
    if (n_bad)
        return fail(LASSIE_ERR_NONFINITE, std::to_string(n_bad) +
                                              " gradient(s) not finite: a BDF1 step whose modified Newton iteration "
                                              "did not converge (see lassie_adjoint_status)");
    return LASSIE_OK;
#undef ADJ_TRY
}
