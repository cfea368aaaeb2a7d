// Plan management and the C ABI of include/mhx.h.
#include "mhx_internal.h"
#include <cstdint>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

namespace mhx {

namespace {
thread_local std::string g_last_error;
}

void set_error(const std::string &msg) { g_last_error = msg; }

int fail(int code, const std::string &msg)
{
    set_error(msg);
    return code;
}

bool is_device_ptr(const void *ptr)
{
    hipPointerAttribute_t attr;
    if (hipPointerGetAttributes(&attr, ptr) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

namespace {

// Gauss-Legendre colatitudes and weights, src/horizontal.f90:279-340
void gauleg(int n, std::vector<double> &theta, std::vector<double> &w)
{
    const double pi = 3.14159265358979323846;
    const double eps = 10.0 * 2.220446049250313e-16;
    theta.assign(n, 0.0);
    w.assign(n, 0.0);
    for (int i = 1; i <= (n + 1) / 2; ++i) {
        double z = std::cos(pi * ((double)i - 0.25) / ((double)n + 0.5));
        double z1 = z + 10.0 * eps, pp = 1.0;
        while (std::fabs(z - z1) > eps) {
            double p1 = 1.0, p2 = 0.0, p3;
            for (int j = 1; j <= n; ++j) {
                p3 = p2;
                p2 = p1;
                p1 = ((double)(2 * j - 1) * z * p2 - (double)(j - 1) * p3) / (double)j;
            }
            pp = (double)n * (z * p1 - p2) / (z * z - 1.0);
            z1 = z;
            z = z1 - p1 / pp;
        }
        theta[i - 1] = std::acos(z);
        theta[n - i] = std::acos(-z);
        w[i - 1] = 2.0 / ((1.0 - z * z) * pp * pp);
        w[n - i] = w[i - 1];
    }
}

bool factorize(int M, std::vector<int> &fac)
{
    fac.clear();
    while (M % 4 == 0) { fac.push_back(4); M /= 4; }
    while (M % 2 == 0) { fac.push_back(2); M /= 2; }
    while (M % 3 == 0) { fac.push_back(3); M /= 3; }
    while (M % 5 == 0) { fac.push_back(5); M /= 5; }
    return M == 1 && fac.size() <= 16;
}

template <typename T> void free_dev(T *&p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

int ensure_buf(double *&buf, size_t &have, size_t need)
{
    if (need <= have) return MHX_OK;
    free_dev(buf);
    have = 0;
    if (hipMalloc(&buf, need) != hipSuccess) {
        (void)hipGetLastError();
        return fail(MHX_E_NOMEM, "device allocation of " + std::to_string(need) + " bytes failed");
    }
    have = need;
    return MHX_OK;
}

// per-level workspaces of the batched paths
int ensure_ws(Plan *p, int nlev)
{
    // bytes per Fourier comp; theta padded to whole blocks of 16 rows (the
    // split form's theta-blocked forward side)
    size_t fc = (size_t)p->n_m * ((p->n_theta + 15) / 16 * 16) * 16;
    size_t gc = (size_t)p->n_phi * p->n_theta * 8;  // bytes per grid comp
    int rc;
    if ((rc = ensure_buf(p->ws_F, p->ws_F_bytes, fc * 18 * nlev))) return rc;
    // grids: the 3 du/dr components of the nonlinear step, or up to 3 grids
    // of a per-call transform
    if ((rc = ensure_buf(p->ws_G, p->ws_G_bytes, gc * 3 * nlev))) return rc;
    if ((rc = ensure_buf(p->ws_F2, p->ws_F2_bytes, fc * 10 * nlev))) return rc;
    if ((rc = ensure_buf(p->ws_C, p->ws_C_bytes, (size_t)p->lm_max * 16 * 8))) return rc;
    if (p->d_lcut_n < nlev) {
        free_dev(p->d_lcut);
        MHX_HIP_CHECK(hipMalloc(&p->d_lcut, sizeof(int) * nlev));
        p->d_lcut_n = nlev;
    }
    if (p->d_rad_n < nlev) {
        free_dev(p->d_rad);
        MHX_HIP_CHECK(hipMalloc(&p->d_rad, sizeof(double) * 3 * nlev));
        p->d_rad_n = nlev;
    }
    p->ws_levels = std::max(p->ws_levels, nlev);
    return MHX_OK;
}

int check_lcut(const Plan *p, int lcut)
{
    if (lcut < 0 || lcut > p->l_max)
        return fail(MHX_E_INVALID, "lcut=" + std::to_string(lcut) + " outside [0, l_max=" +
                                       std::to_string(p->l_max) + "]");
    return MHX_OK;
}

int upload_lcut(Plan *p, const int *lcut, int nlev)
{
    for (int i = 0; i < nlev; ++i) {
        int rc = check_lcut(p, lcut[i]);
        if (rc) return rc;
    }
    MHX_HIP_CHECK(hipMemcpyAsync(p->d_lcut, lcut, sizeof(int) * nlev, hipMemcpyHostToDevice,
                                 p->stream));
    return MHX_OK;
}

// stage a per-call input into device memory (slot of ws_C) unless it
// already is device memory
const double *stage_in(Plan *p, const double *src, int slot, size_t count_doubles)
{
    if (is_device_ptr(src)) return src;
    double *dst = p->ws_C + (size_t)slot * p->lm_max * 2;
    (void)hipMemcpyAsync(dst, src, count_doubles * sizeof(double), hipMemcpyHostToDevice, p->stream);
    return dst;
}

ColOp op(int s0, int g = G_ONE, int ipow = 0, int mpow = 0, double scale = 1.0, int s1 = -1,
         double h = 0.0)
{
    ColOp c;
    c.src0 = s0;
    c.src1 = s1;
    c.gkind = g;
    c.ipow = ipow;
    c.mpow = mpow;
    c.hcoef = h;
    c.scale = scale;
    return c;
}

SynJob syn_job(Plan *p, int kind, int nlev, size_t src_ld)
{
    SynJob j;
    std::memset(&j, 0, sizeof(j));
    j.kind = kind;
    j.nlev = nlev;
    j.src_ld = src_ld;
    j.fout = p->ws_F;
    j.fout_ld = (size_t)18 * p->n_m * p->n_theta;
    j.lcut = p->d_lcut;
    j.or2 = p->d_rad;
    for (int i = 0; i < MAX_OUT; ++i) j.out_comp[i] = i;
    return j;
}

// Fourier -> grid for ncomp comps starting at comp0 of ws_F into separate
// destinations (level stride ld_grid doubles)
int ifft_to(Plan *p, int comp0, int ncomp, int nlev, double *const *dst, size_t ld_grid)
{
    size_t Fld = (size_t)18 * p->n_m * p->n_theta;
    for (int c = 0; c < ncomp; ++c) {
        const double *F = p->ws_F + (size_t)(comp0 + c) * p->n_m * p->n_theta * 2;
        int rc = launch_ifft(p, F, Fld, 1, 1, nlev, dst[c], ld_grid, 0, p->stream);
        if (rc) return rc;
    }
    return MHX_OK;
}

// grids -> Fourier comps [0, ncomp) of ws_F2 (10-comp level stride)
int fft_from(Plan *p, int ncomp, int nlev, const double *const *src, size_t ld_grid,
             const bool *osin2)
{
    size_t Fld = (size_t)10 * p->n_m * p->n_theta;
    for (int c = 0; c < ncomp; ++c) {
        double *F = p->ws_F2 + (size_t)c * p->n_m * p->n_theta * 2;
        int rc = launch_fft(p, src[c], ld_grid, 0, 1, nlev, osin2[c] ? 1u : 0u, F, Fld, p->stream);
        if (rc) return rc;
    }
    return MHX_OK;
}

AnaJob ana_job(Plan *p, int kind, int nlev, size_t out_ld)
{
    AnaJob j;
    std::memset(&j, 0, sizeof(j));
    j.kind = kind;
    j.nlev = nlev;
    j.fin = p->ws_F2;
    j.fin_ld = (size_t)10 * p->n_m * p->n_theta;
    j.in_comp[0] = j.in_comp[1] = j.in_comp[2] = -1;
    j.out_ld = out_ld;
    j.lcut = p->d_lcut;
    return j;
}

// ---- per-call helpers: grids staged through ws_G, coefficients through ws_C
int finish_grids(Plan *p, int ncomp, double *const *out)
{
    size_t gbytes = (size_t)p->n_phi * p->n_theta * sizeof(double);
    for (int c = 0; c < ncomp; ++c)
        MHX_HIP_CHECK(hipMemcpyAsync(out[c], p->ws_G + (size_t)c * p->n_phi * p->n_theta, gbytes,
                                     hipMemcpyDefault, p->stream));
    MHX_HIP_CHECK(hipStreamSynchronize(p->stream));
    return MHX_OK;
}

int finish_coefs(Plan *p, int n, double *const *out)
{
    size_t bytes = (size_t)p->lm_max * 16;
    for (int c = 0; c < n; ++c)
        MHX_HIP_CHECK(hipMemcpyAsync(out[c], p->ws_C + (size_t)(4 + c) * p->lm_max * 2, bytes,
                                     hipMemcpyDefault, p->stream));
    MHX_HIP_CHECK(hipStreamSynchronize(p->stream));
    return MHX_OK;
}

// run a one-level synthesis job then inverse FFT of its ncomp outputs into ws_G
int synth_call(Plan *p, SynJob &j, int ncomp, int lcut, double or2, double *const *out)
{
    int rc = check_lcut(p, lcut);
    if (rc) return rc;
    MHX_HIP_CHECK(hipMemcpyAsync(p->d_lcut, &lcut, sizeof(int), hipMemcpyHostToDevice, p->stream));
    MHX_HIP_CHECK(hipMemcpyAsync(p->d_rad, &or2, sizeof(double), hipMemcpyHostToDevice, p->stream));
    if ((rc = launch_syn(p, j, p->stream))) return rc;
    double *dst[3] = {p->ws_G, p->ws_G + (size_t)p->n_phi * p->n_theta,
                      p->ws_G + (size_t)2 * p->n_phi * p->n_theta};
    if ((rc = ifft_to(p, 0, ncomp, 1, dst, 0))) return rc;
    return finish_grids(p, ncomp, out);
}

int analysis_call(Plan *p, int kind, int ngrid, const double *const *grids, const bool *os,
                  int lcut, int nout, double *const *out)
{
    int rc = check_lcut(p, lcut);
    if (rc) return rc;
    MHX_HIP_CHECK(hipMemcpyAsync(p->d_lcut, &lcut, sizeof(int), hipMemcpyHostToDevice, p->stream));
    size_t gsz = (size_t)p->n_phi * p->n_theta;
    const double *src[3];
    for (int c = 0; c < ngrid; ++c) {
        if (is_device_ptr(grids[c])) {
            src[c] = grids[c];
        } else {
            MHX_HIP_CHECK(hipMemcpyAsync(p->ws_G + c * gsz, grids[c], gsz * sizeof(double),
                                         hipMemcpyHostToDevice, p->stream));
            src[c] = p->ws_G + c * gsz;
        }
    }
    if ((rc = fft_from(p, ngrid, 1, src, 0, os))) return rc;
    AnaJob j = ana_job(p, kind, 1, 0);
    if (kind == ANA_SH) {
        j.in_comp[0] = 0;
    } else if (kind == ANA_SPHTOR) {
        j.in_comp[1] = 0;
        j.in_comp[2] = 1;
    } else {
        j.in_comp[0] = 0;
        j.in_comp[1] = 1;
        j.in_comp[2] = 2;
    }
    for (int c = 0; c < nout; ++c) j.out[c] = p->ws_C + (size_t)(4 + c) * p->lm_max * 2;
    if ((rc = launch_ana(p, j, p->stream))) return rc;
    return finish_coefs(p, nout, out);
}

__global__ void k_ic_preop(int lm_max, const int *__restrict__ lm2l, double rr, double ricb,
                           double r, int curl, const double *__restrict__ a0,
                           const double *__restrict__ a1, const double *__restrict__ a2,
                           const double *__restrict__ a3, double *__restrict__ Q,
                           double *__restrict__ S, double *__restrict__ T)
{
    int lm = blockIdx.x * blockDim.x + threadIdx.x;
    if (lm >= lm_max) return;
    int l = lm2l[lm];
    // rDep(l) = (r/r_ICB)^(l+1), rDep2(l) = (r/r_ICB)^l / r_ICB, built by
    // repeated multiplication as src/sht_native.f90:160-165
    double rd = rr, rd2 = 1.0 / ricb;
    for (int k = 1; k <= l; ++k) {
        rd *= rr;
        rd2 *= rr;
    }
    double ll = (double)l * (double)(l + 1);
    for (int c = 0; c < 2; ++c) {
        if (!curl) {
            // torpol_to_spat_IC (:209-228): a0=W a1=dW a2=Z
            Q[2 * lm + c] = rd * ll * a0[2 * lm + c];
            S[2 * lm + c] = rd2 * ((double)(l + 1) * a0[2 * lm + c] + r * a1[2 * lm + c]);
            T[2 * lm + c] = rd * a2[2 * lm + c];
        } else {
            // torpol_to_curl_spat_IC (:166-185): a0=dB a1=ddB a2=J a3=dJ
            Q[2 * lm + c] = rd * ll * a2[2 * lm + c];
            S[2 * lm + c] = rd2 * ((double)(l + 1) * a2[2 * lm + c] + r * a3[2 * lm + c]);
            T[2 * lm + c] = -rd2 * (2.0 * (double)(l + 1) * a0[2 * lm + c] + r * a1[2 * lm + c]);
        }
    }
}

// make the plan's stream wait for work already queued on a caller's stream
int order_after(Plan *p, void *stream)
{
    if (!stream) return MHX_OK;
    MHX_HIP_CHECK(hipEventRecord(p->ev_user, (hipStream_t)stream));
    MHX_HIP_CHECK(hipStreamWaitEvent(p->stream, p->ev_user, 0));
    return MHX_OK;
}

}  // namespace
}  // namespace mhx

using namespace mhx;

#define PLAN_OR_FAIL(pl)                                                       \
    if (!(pl)) return fail(MHX_E_INVALID, "null plan");                        \
    Plan *P = &(pl)->p;                                                        \
    {                                                                          \
        int rc_ = ensure_ws(P, std::max(1, P->ws_levels));                     \
        if (rc_) return rc_;                                                   \
    }

extern "C" {

const char *mhx_last_error(void) { return g_last_error.c_str(); }

int mhx_plan_create(int l_max, int m_max, int minc, int n_theta, int n_phi, int precision,
                    int max_levels, mhx_plan **plan)
{
    if (!plan) return fail(MHX_E_INVALID, "plan output pointer is null");
    *plan = nullptr;
    if (l_max < 1 || minc < 1) return fail(MHX_E_INVALID, "l_max must be >= 1 and minc >= 1");
    if (m_max <= 0 || m_max > l_max) m_max = (l_max / minc) * minc;
    m_max = (m_max / minc) * minc;
    if (n_theta < 4 || n_theta % 4) return fail(MHX_E_INVALID, "n_theta must be a multiple of 4");
    if (n_phi < 8 || n_phi % 4) return fail(MHX_E_INVALID, "n_phi must be a multiple of 4");
    if (precision != MHX_PREC_FP64 && precision != MHX_PREC_SPLIT16 && precision != MHX_PREC_SPLIT16_REC)
        return fail(MHX_E_INVALID, "unknown precision");
    const bool rec = precision == MHX_PREC_SPLIT16_REC;
    if (rec && (l_max <= 255 || l_max > 1023))
        return fail(MHX_E_UNSUPPORTED, "recurrence-generated tables need 255 < l_max <= 1023");
    if (rec) precision = MHX_PREC_SPLIT16;
    std::vector<int> fac;
    if (!factorize(n_phi / 2, fac)) return fail(MHX_E_INVALID, "n_phi/2 must factor into 2, 3, 5");
    int n_m = m_max / minc + 1;
    if (n_m > n_phi / 2) return fail(MHX_E_INVALID, "n_phi too small for m_max");
    if (n_phi / 2 > 4096) return fail(MHX_E_UNSUPPORTED, "n_phi above 8192");

    mhx_plan *h = new mhx_plan();
    Plan *p = &h->p;
    p->l_max = l_max;
    p->m_max = m_max;
    p->minc = minc;
    p->n_theta = n_theta;
    p->n_phi = n_phi;
    p->nh = n_theta / 2;
    p->n_m = n_m;
    p->precision = precision;
    p->rec = rec;
    p->nt16 = (p->nh + 15) / 16;
    p->nq4 = (p->nh + 3) / 4;
    p->M = n_phi / 2;
    p->fac = fac;
    p->lstart.resize(n_m);
    std::vector<int> lm2l, lm2m;
    int lm = 0;
    for (int mc = 0; mc < n_m; ++mc) {
        p->lstart[mc] = lm;
        for (int l = mc * minc; l <= l_max; ++l) lm2l.push_back(l), lm2m.push_back(mc * minc), ++lm;
    }
    p->lm_max = lm;

    std::vector<double> th, w;
    gauleg(n_theta, th, w);
    p->theta_ord.assign(th.begin(), th.begin() + p->nh);
    p->gauss.assign(w.begin(), w.begin() + p->nh);
    std::vector<double> osin2(n_theta);
    for (int k = 0; k < p->nh; ++k) {
        double s = std::sin(th[k]);
        osin2[2 * k] = osin2[2 * k + 1] = 1.0 / (s * s);
    }
    const double pi = 3.14159265358979323846;
    std::vector<double> tw(2 * p->M), twr(2 * (p->M + 1));
    for (int k = 0; k < p->M; ++k) {
        tw[2 * k] = std::cos(2.0 * pi * k / p->M);
        tw[2 * k + 1] = -std::sin(2.0 * pi * k / p->M);
    }
    for (int k = 0; k <= p->M; ++k) {
        twr[2 * k] = std::cos(2.0 * pi * k / n_phi);
        twr[2 * k + 1] = -std::sin(2.0 * pi * k / n_phi);
    }

    auto up = [&](auto *&dst, const auto &vec) -> int {
        size_t bytes = vec.size() * sizeof(vec[0]);
        MHX_HIP_CHECK(hipMalloc(&dst, bytes));
        MHX_HIP_CHECK(hipMemcpy(dst, vec.data(), bytes, hipMemcpyHostToDevice));
        return MHX_OK;
    };
    std::vector<float> tw32(tw.begin(), tw.end()), twr32(twr.begin(), twr.end());
    std::vector<float> twn32(2 * (size_t)n_phi);
    for (int k = 0; k < n_phi; ++k) {
        twn32[2 * k] = (float)std::cos(2.0 * pi * k / n_phi);
        twn32[2 * k + 1] = (float)-std::sin(2.0 * pi * k / n_phi);
    }
    {
        const char *gf = std::getenv("MHX_GRID_FORM");
        p->grid_form = (gf && std::strcmp(gf, "slots") == 0) ? 1 : 0;
    }
    int rc;
    int *d_lm2l = nullptr;
    if ((rc = up(p->d_lstart, p->lstart)) || (rc = up(p->d_theta, p->theta_ord)) ||
        (rc = up(p->d_gauss, p->gauss)) || (rc = up(p->d_osin2, osin2)) ||
        (rc = up(p->d_tw, tw)) || (rc = up(p->d_twr, twr)) || (rc = up(d_lm2l, lm2l)) ||
        (rc = up(p->d_tw32, tw32)) || (rc = up(p->d_twr32, twr32)) || (rc = up(p->d_lm2m, lm2m)) ||
        (rc = up(p->d_twn32, twn32))) {
        mhx_plan_destroy(h);
        return rc;
    }
    p->d_lm2l = d_lm2l;
    if (hipGetDevice(&p->device) != hipSuccess) {
        mhx_plan_destroy(h);
        return fail(MHX_E_HIP, "hipGetDevice failed");
    }
    if (hipStreamCreate(&p->stream) != hipSuccess) {
        mhx_plan_destroy(h);
        return fail(MHX_E_HIP, "hipStreamCreate failed");
    }
    // recurrence-table plans above l_max = 255 keep no FP64 tables: their
    // split kernels generate the tables in-kernel, and the per-call entries
    // that read FP64 tables build them on first use (ensure_fp64_tables)
    if ((rc = (rec && l_max > 255) ? table_layout(p) : build_tables(p)) ||
        (precision == MHX_PREC_SPLIT16 && (rc = build_split_tables(p))) ||
        (rc = ensure_ws(p, std::max(1, max_levels)))) {
        mhx_plan_destroy(h);
        return rc;
    }
    for (int i = 0; i < Plan::NEV; ++i) (void)hipEventCreate(&p->ev[i]);
    (void)hipEventCreateWithFlags(&p->ev_user, hipEventDisableTiming);
    (void)hipEventCreateWithFlags(&p->ev_done, hipEventDisableTiming);
    (void)hipEventCreateWithFlags(&p->ev_pin, hipEventDisableTiming);
    if (const char *e = std::getenv("MHX_BRANCHES")) p->nbranch = std::max(1, std::min(Plan::MAXBR, std::atoi(e)));
    (void)hipEventCreateWithFlags(&p->ev_fork, hipEventDisableTiming);
    for (int i = 0; i + 1 < p->nbranch; ++i) {
        if (hipStreamCreateWithFlags(&p->br[i], hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&p->ev_join[i], hipEventDisableTiming) != hipSuccess) {
            mhx_plan_destroy(h);
            return fail(MHX_E_HIP, "branch stream creation failed");
        }
    }
    *plan = h;
    return MHX_OK;
}

int mhx_plan_destroy(mhx_plan *h)
{
    if (!h) return MHX_OK;
    Plan *p = &h->p;
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    rl_free(p);
    free_dev(p->d_lstart);
    free_dev(p->d_syn_P);
    free_dev(p->d_syn_dP);
    free_dev(p->d_ana_wP);
    free_dev(p->d_ana_wdP);
    free_dev(p->d_syn_off);
    free_dev(p->d_ana_off);
    free_dev(p->d_osin2);
    free_dev(p->d_theta);
    free_dev(p->d_gauss);
    free_dev(p->d_tw);
    free_dev(p->d_twr);
    free_dev(p->d_lm2l);
    free_dev(p->d_lm2m);
    free_dev(p->d_tw32);
    free_dev(p->d_twr32);
    free_dev(p->d_twn32);
    free_dev(p->d_syn16_P);
    free_dev(p->d_syn16_dP);
    free_dev(p->d_ana16_P);
    free_dev(p->d_ana16_dP);
    free_dev(p->d_syn16_off);
    free_dev(p->d_ana16_off);
    free_dev(p->d_ana16_soff);
    free_dev(p->d_syn16_rs);
    free_dev(p->d_rec_seed);
    free_dev(p->d_cour_par);
    free_dev(p->d_cour_max);
    free_dev(p->d_td_rad);
    free_dev(p->d_ek);
    free_dev(p->d_td_lcut);
    free_dev(p->d_rec_aseed);
    free_dev(p->d_rec_aexp);
    free_dev(p->d_rec_segoff);
    free_dev(p->d_cost);
    free_dev(p->d_ana16_rs);
    free_dev(p->d_syn16x);
    free_dev(p->d_ana16x);
    free_dev(p->d_syn16x_off);
    free_dev(p->d_ana16x_off);
    free_dev(p->d_ana16x_soff);
    free_dev(p->d_syn16x_rs);
    free_dev(p->d_ana16x_rs);
    free_dev(p->d_crec);
    free_dev(p->ws_F);
    free_dev(p->ws_G);
    free_dev(p->ws_F2);
    free_dev(p->ws_C);
    free_dev(p->ws_H);
    free_dev(p->d_lcut);
    free_dev(p->d_rad);
    for (int i = 0; i < Plan::NEV; ++i)
        if (p->ev[i]) (void)hipEventDestroy(p->ev[i]);
    if (p->ev_user) (void)hipEventDestroy(p->ev_user);
    if (p->ev_done) (void)hipEventDestroy(p->ev_done);
    if (p->ev_pin) (void)hipEventDestroy(p->ev_pin);
    if (p->ev_fork) (void)hipEventDestroy(p->ev_fork);
    for (int i = 0; i + 1 < Plan::MAXBR; ++i) {
        if (p->ev_join[i]) (void)hipEventDestroy(p->ev_join[i]);
        if (p->br[i]) (void)hipStreamDestroy(p->br[i]);
    }
    if (p->gexec) (void)hipGraphExecDestroy(p->gexec);
    if (p->h_pin) (void)hipHostFree(p->h_pin);
    if (p->stream) (void)hipStreamDestroy(p->stream);
    delete h;
    return MHX_OK;
}

int mhx_plan_info(const mhx_plan *h, int *lm_max, int *n_m_max, int *n_theta, int *n_phi)
{
    if (!h) return fail(MHX_E_INVALID, "null plan");
    const Plan *p = &h->p;
    if (lm_max) *lm_max = p->lm_max;
    if (n_m_max) *n_m_max = p->n_m;
    if (n_theta) *n_theta = p->n_theta;
    if (n_phi) *n_phi = p->n_phi;
    return MHX_OK;
}

int mhx_plan_tables(const mhx_plan *h, double *theta_ord, double *gauss, double *plm, double *dplm)
{
    if (!h) return fail(MHX_E_INVALID, "null plan");
    const Plan *p = &h->p;
    if (theta_ord) std::memcpy(theta_ord, p->theta_ord.data(), sizeof(double) * p->nh);
    if (gauss) std::memcpy(gauss, p->gauss.data(), sizeof(double) * p->nh);
    if (plm || dplm) {
        int rc;
        if ((rc = ensure_fp64_tables(const_cast<Plan *>(p)))) return rc;
        std::vector<double> sP(p->syn_size), sdP(p->syn_size);
        MHX_HIP_CHECK(hipMemcpy(sP.data(), p->d_syn_P, p->syn_size * 8, hipMemcpyDeviceToHost));
        MHX_HIP_CHECK(hipMemcpy(sdP.data(), p->d_syn_dP, p->syn_size * 8, hipMemcpyDeviceToHost));
        for (int mc = 0; mc < p->n_m; ++mc) {
            int m = mc * p->minc;
            for (int l = m; l <= p->l_max; ++l) {
                int par = (l - m) & 1, kk = (l - m) >> 1;
                int lm = p->lstart[mc] + l - m;
                for (int th = 0; th < p->nh; ++th) {
                    size_t o = p->syn_off[2 * mc + par] +
                               ((size_t)(kk >> 2) * p->nt16 + (th >> 4)) * 64 + (th & 15) + 16 * (kk & 3);
                    if (plm) plm[(size_t)th * p->lm_max + lm] = sP[o];
                    if (dplm) dplm[(size_t)th * p->lm_max + lm] = sdP[o];
                }
            }
        }
    }
    return MHX_OK;
}

// ---------------------------------------------------------------- synthesis
int mhx_scal_to_spat(mhx_plan *h, const double *Slm, double *fieldc, int lcut)
{
    PLAN_OR_FAIL(h);
    SynJob j = syn_job(P, SYN_SCAL, 1, 0);
    j.src[0] = stage_in(P, Slm, 0, 2 * P->lm_max);
    j.col[0] = op(0);
    double *out[1] = {fieldc};
    return synth_call(P, j, 1, lcut, 0.0, out);
}

int mhx_scal_to_grad_spat(mhx_plan *h, const double *Slm, double *gradtc, double *gradpc, int lcut)
{
    PLAN_OR_FAIL(h);
    SynJob j = syn_job(P, SYN_GRAD, 1, 0);
    j.src[0] = stage_in(P, Slm, 0, 2 * P->lm_max);
    j.col[0] = op(0);
    j.col[1] = op(0, G_ONE, 1, 1);
    double *out[2] = {gradtc, gradpc};
    return synth_call(P, j, 2, lcut, 0.0, out);
}

int mhx_pol_to_grad_spat(mhx_plan *h, const double *Slm, double *gradtc, double *gradpc, int lcut)
{
    PLAN_OR_FAIL(h);
    SynJob j = syn_job(P, SYN_GRAD, 1, 0);
    j.src[0] = stage_in(P, Slm, 0, 2 * P->lm_max);
    j.col[0] = op(0, G_LL);
    j.col[1] = op(0, G_LL, 1, 1);
    double *out[2] = {gradtc, gradpc};
    return synth_call(P, j, 2, lcut, 0.0, out);
}

static void qst_cols(SynJob &j, ColOp Q, ColOp S, ColOp T)
{
    // columns Q, S, i m T, i m S, -T (T and S may be built from two sources)
    j.col[0] = Q;
    j.col[1] = S;
    ColOp miT = T;
    miT.ipow = (miT.ipow + 1) & 3;
    miT.mpow += 1;
    ColOp miS = S;
    miS.ipow = (miS.ipow + 1) & 3;
    miS.mpow += 1;
    ColOp mT = T;
    mT.scale = -mT.scale;
    j.col[2] = miT;
    j.col[3] = miS;
    j.col[4] = mT;
}

static void sphtor_cols(SynJob &j, ColOp S, ColOp T)
{
    // columns S, i m T, i m S, -T
    ColOp miT = T;
    miT.ipow = (miT.ipow + 1) & 3;
    miT.mpow += 1;
    ColOp miS = S;
    miS.ipow = (miS.ipow + 1) & 3;
    miS.mpow += 1;
    ColOp mT = T;
    mT.scale = -mT.scale;
    j.col[0] = S;
    j.col[1] = miT;
    j.col[2] = miS;
    j.col[3] = mT;
}

int mhx_torpol_to_spat(mhx_plan *h, const double *Wlm, const double *dWlm, const double *Zlm,
                       double *vrc, double *vtc, double *vpc, int lcut)
{
    PLAN_OR_FAIL(h);
    SynJob j = syn_job(P, SYN_QST, 1, 0);
    j.src[0] = stage_in(P, Wlm, 0, 2 * P->lm_max);
    j.src[1] = stage_in(P, dWlm, 1, 2 * P->lm_max);
    j.src[2] = stage_in(P, Zlm, 2, 2 * P->lm_max);
    qst_cols(j, op(0, G_LL), op(1), op(2));
    double *out[3] = {vrc, vtc, vpc};
    return synth_call(P, j, 3, lcut, 0.0, out);
}

int mhx_sphtor_to_spat(mhx_plan *h, const double *dWlm, const double *Zlm, double *vtc,
                       double *vpc, int lcut)
{
    PLAN_OR_FAIL(h);
    SynJob j = syn_job(P, SYN_SPHTOR, 1, 0);
    j.src[0] = stage_in(P, dWlm, 0, 2 * P->lm_max);
    j.src[1] = stage_in(P, Zlm, 1, 2 * P->lm_max);
    sphtor_cols(j, op(0), op(1));
    double *out[2] = {vtc, vpc};
    return synth_call(P, j, 2, lcut, 0.0, out);
}

static int ic_call(Plan *P, double r, double r_ICB, int curl, const double *a0, const double *a1,
                   const double *a2, const double *a3, double *o0, double *o1, double *o2)
{
    if (!(r_ICB != 0.0)) return fail(MHX_E_INVALID, "r_ICB must be nonzero");
    const double *s0 = stage_in(P, a0, 0, 2 * P->lm_max);
    const double *s1 = stage_in(P, a1, 1, 2 * P->lm_max);
    const double *s2 = stage_in(P, a2, 2, 2 * P->lm_max);
    const double *s3 = a3 ? stage_in(P, a3, 3, 2 * P->lm_max) : s2;
    double *Q = P->ws_C + (size_t)4 * P->lm_max * 2;
    double *S = P->ws_C + (size_t)5 * P->lm_max * 2;
    double *T = P->ws_C + (size_t)6 * P->lm_max * 2;
    k_ic_preop<<<(P->lm_max + 127) / 128, 128, 0, P->stream>>>(P->lm_max, P->d_lm2l, r / r_ICB, r_ICB,
                                                               r, curl, s0, s1, s2, s3, Q, S, T);
    MHX_HIP_CHECK(hipGetLastError());
    SynJob j = syn_job(P, SYN_QST, 1, 0);
    j.src[0] = Q;
    j.src[1] = S;
    j.src[2] = T;
    qst_cols(j, op(0), op(1), op(2));
    double *out[3] = {o0, o1, o2};
    return synth_call(P, j, 3, P->l_max, 0.0, out);
}

int mhx_torpol_to_curl_spat_IC(mhx_plan *h, double r, double r_ICB, const double *dBlm,
                               const double *ddBlm, const double *Jlm, const double *dJlm,
                               double *cbr, double *cbt, double *cbp)
{
    PLAN_OR_FAIL(h);
    return ic_call(P, r, r_ICB, 1, dBlm, ddBlm, Jlm, dJlm, cbr, cbt, cbp);
}

int mhx_torpol_to_spat_IC(mhx_plan *h, double r, double r_ICB, const double *Wlm,
                          const double *dWlm, const double *Zlm, double *Br, double *Bt, double *Bp)
{
    PLAN_OR_FAIL(h);
    return ic_call(P, r, r_ICB, 0, Wlm, dWlm, Zlm, nullptr, Br, Bt, Bp);
}

int mhx_torpol_to_dphspat(mhx_plan *h, const double *dWlm, const double *Zlm, double *dvtdp,
                          double *dvpdp, int lcut)
{
    PLAN_OR_FAIL(h);
    SynJob j = syn_job(P, SYN_SPHTOR, 1, 0);
    j.src[0] = stage_in(P, dWlm, 0, 2 * P->lm_max);
    j.src[1] = stage_in(P, Zlm, 1, 2 * P->lm_max);
    // S = i m dW, T = i m Z (src/sht_native.f90:249-262); outputs times
    // 1/sin^2(theta) (:266-270)
    sphtor_cols(j, op(0, G_ONE, 1, 1), op(1, G_ONE, 1, 1));
    j.row_scale[0] = j.row_scale[1] = 1;
    double *out[2] = {dvtdp, dvpdp};
    return synth_call(P, j, 2, lcut, 0.0, out);
}

int mhx_pol_to_curlr_spat(mhx_plan *h, const double *Qlm, double *cvrc, int lcut)
{
    PLAN_OR_FAIL(h);
    SynJob j = syn_job(P, SYN_SCAL, 1, 0);
    j.src[0] = stage_in(P, Qlm, 0, 2 * P->lm_max);
    j.col[0] = op(0, G_LL);
    double *out[1] = {cvrc};
    return synth_call(P, j, 1, lcut, 0.0, out);
}

int mhx_torpol_to_curl_spat(mhx_plan *h, double or2, const double *Blm, const double *ddBlm,
                            const double *Jlm, const double *dJlm, double *cvrc, double *cvtc,
                            double *cvpc, int lcut)
{
    PLAN_OR_FAIL(h);
    SynJob j = syn_job(P, SYN_QST, 1, 0);
    j.src[0] = stage_in(P, Blm, 0, 2 * P->lm_max);
    j.src[1] = stage_in(P, ddBlm, 1, 2 * P->lm_max);
    j.src[2] = stage_in(P, Jlm, 2, 2 * P->lm_max);
    j.src[3] = stage_in(P, dJlm, 3, 2 * P->lm_max);
    // Q = l(l+1) J, S = dJ, T = or2 l(l+1) B - ddB (src/sht_native.f90:320-330)
    qst_cols(j, op(2, G_LL), op(3), op(0, G_OR2LL, 0, 0, 1.0, 1, -1.0));
    double *out[3] = {cvrc, cvtc, cvpc};
    return synth_call(P, j, 3, lcut, or2, out);
}

// ----------------------------------------------------------------- analysis
int mhx_scal_to_SH(mhx_plan *h, const double *f, double *fLM, int lcut)
{
    PLAN_OR_FAIL(h);
    const double *g[1] = {f};
    bool os[1] = {false};
    double *out[1] = {fLM};
    return analysis_call(P, ANA_SH, 1, g, os, lcut, 1, out);
}

int mhx_spat_to_qst(mhx_plan *h, const double *f, const double *g, const double *hh, double *qLM,
                    double *sLM, double *tLM, int lcut)
{
    PLAN_OR_FAIL(h);
    const double *gr[3] = {f, g, hh};
    bool os[3] = {false, true, true};
    double *out[3] = {qLM, sLM, tLM};
    return analysis_call(P, ANA_QST, 3, gr, os, lcut, 3, out);
}

int mhx_spat_to_sphertor(mhx_plan *h, const double *f, const double *g, double *fLM, double *gLM,
                         int lcut)
{
    PLAN_OR_FAIL(h);
    const double *gr[2] = {f, g};
    bool os[2] = {true, true};
    double *out[2] = {fLM, gLM};
    return analysis_call(P, ANA_SPHTOR, 2, gr, os, lcut, 2, out);
}

// src/nonlinear_bcs.f90:58-64: br*vt*fac and br*(fac*vp - omega sin^2)
// with fac = O_r_E_2*O_rho.  a, b may alias br, vt (the staged grids): every
// thread reads its point's three inputs before writing the two outputs.
__global__ void k_br_v(size_t n, int nt, const double *br, const double *vt, const double *vp,
                       const double *__restrict__ osin2, double fac, double omega, double *a,
                       double *b)
{
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double s2 = 1.0 / osin2[i % nt];
    const double r = br[i], t = vt[i], f = vp[i];
    a[i] = fac * r * t;
    b[i] = r * (fac * f - omega * s2);
}

int mhx_get_br_v_bcs(mhx_plan *h, const double *br, const double *vt, const double *vp,
                     double omega, double O_r_E_2, double O_rho, double *br_vt_lm,
                     double *br_vp_lm)
{
    PLAN_OR_FAIL(h);
    if (!br || !vt || !vp || !br_vt_lm || !br_vp_lm) return fail(MHX_E_INVALID, "null array");
    const size_t gsz = (size_t)P->n_phi * P->n_theta;
    const double *in[3] = {br, vt, vp};
    const double *src[3];
    for (int c = 0; c < 3; ++c) {
        if (is_device_ptr(in[c])) {
            src[c] = in[c];
        } else {
            MHX_HIP_CHECK(hipMemcpyAsync(P->ws_G + c * gsz, in[c], gsz * sizeof(double),
                                         hipMemcpyHostToDevice, P->stream));
            src[c] = P->ws_G + c * gsz;
        }
    }
    double *a = P->ws_G, *b = P->ws_G + gsz;
    k_br_v<<<(unsigned)((gsz + 255) / 256), 256, 0, P->stream>>>(
        gsz, P->n_theta, src[0], src[1], src[2], P->d_osin2, O_r_E_2 * O_rho, omega, a, b);
    MHX_HIP_CHECK(hipGetLastError());
    const double *gr[2] = {a, b};
    bool os[2] = {true, true};
    double *out[2] = {br_vt_lm, br_vp_lm};
    return analysis_call(P, ANA_SPHTOR, 2, gr, os, P->l_max, 2, out);
}

// nonlinear boundary terms: -fac x per complex entry, lm = 1 (l = 0) zero
__global__ void k_b_nl_bcs(int n, const double *__restrict__ a, const double *__restrict__ b, double fac,
                           double *__restrict__ bo, double *__restrict__ jo)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool l0 = i < 2;
    if (bo) bo[i] = l0 ? 0.0 : -fac * a[i];
    jo[i] = l0 ? 0.0 : -fac * b[i];
}

int mhx_get_b_nl_bcs(mhx_plan *h, int bc, const double *br_vt_lm, const double *br_vp_lm,
                     double cond_or_sigma, double prmag, double *b_nl_bc, double *aj_nl_bc)
{
    PLAN_OR_FAIL(h);
    if (bc != MHX_BC_CMB && bc != MHX_BC_ICB) return fail(MHX_E_INVALID, "Wrong input of bc into get_b_nl_bcs");
    const bool cmb = bc == MHX_BC_CMB;
    if (!br_vp_lm || !aj_nl_bc || (cmb && (!br_vt_lm || !b_nl_bc))) return fail(MHX_E_INVALID, "null array");
    const size_t n = 2 * (size_t)P->lm_max;
    const double *a = cmb ? stage_in(P, br_vt_lm, 0, n) : nullptr;
    const double *b = stage_in(P, br_vp_lm, 1, n);
    double *out[2] = {aj_nl_bc, b_nl_bc};
    double *dst[2];
    for (int k = 0; k < (cmb ? 2 : 1); ++k)
        dst[k] = is_device_ptr(out[k]) ? out[k] : P->ws_C + (size_t)(4 + k) * n;
    const double fac = cond_or_sigma * prmag;  // src/nonlinear_bcs.f90:90, :104
    k_b_nl_bcs<<<(unsigned)((n + 255) / 256), 256, 0, P->stream>>>((int)n, a, b, fac, cmb ? dst[1] : nullptr,
                                                                   dst[0]);
    MHX_HIP_CHECK(hipGetLastError());
    for (int k = 0; k < (cmb ? 2 : 1); ++k)
        if (dst[k] != out[k])
            MHX_HIP_CHECK(hipMemcpyAsync(out[k], dst[k], n * sizeof(double), hipMemcpyDeviceToHost, P->stream));
    MHX_HIP_CHECK(hipStreamSynchronize(P->stream));
    return MHX_OK;
}

int mhx_axi_to_spat(mhx_plan *h, const double *fl_ax, double *f)
{
    PLAN_OR_FAIL(h);
    const double *fl = stage_in(P, fl_ax, 0, 2 * (P->l_max + 1));
    int rc = launch_axi(P, fl, P->ws_G, P->l_max, 0, nullptr, P->stream);
    if (rc) return rc;
    MHX_HIP_CHECK(hipMemcpyAsync(f, P->ws_G, P->n_theta * sizeof(double), hipMemcpyDefault, P->stream));
    MHX_HIP_CHECK(hipStreamSynchronize(P->stream));
    return MHX_OK;
}

int mhx_toraxi_to_spat(mhx_plan *h, const double *fl_ax, double *ft, double *fp, int lcut)
{
    PLAN_OR_FAIL(h);
    int rc = check_lcut(P, lcut);
    if (rc) return rc;
    const double *fl = stage_in(P, fl_ax, 0, 2 * (P->l_max + 1));
    rc = launch_axi(P, fl, P->ws_G, lcut, 1, P->ws_G + P->n_theta, P->stream);
    if (rc) return rc;
    MHX_HIP_CHECK(hipMemcpyAsync(ft, P->ws_G, P->n_theta * sizeof(double), hipMemcpyDefault, P->stream));
    MHX_HIP_CHECK(hipMemcpyAsync(fp, P->ws_G + P->n_theta, P->n_theta * sizeof(double),
                                 hipMemcpyDefault, P->stream));
    MHX_HIP_CHECK(hipStreamSynchronize(P->stream));
    return MHX_OK;
}

// ------------------------------------------------------------------ batched
int mhx_torpol_to_spat_batch(mhx_plan *h, int nlev, const double *W, const double *dW,
                             const double *Z, size_t ld_lm, double *vr, double *vt, double *vp,
                             size_t ld_grid, const int *lcut, void *stream)
{
    if (!h) return fail(MHX_E_INVALID, "null plan");
    Plan *P = &h->p;
    if (nlev <= 0) return MHX_OK;
    if (!W || !dW || !Z || !vr || !vt || !vp || !lcut) return fail(MHX_E_INVALID, "null pointer");
    int rc;
    if ((rc = order_after(P, stream)) || (rc = ensure_ws(P, nlev)) ||
        (rc = upload_lcut(P, lcut, nlev)))
        return rc;
    SynJob j = syn_job(P, SYN_QST, nlev, ld_lm);
    j.src[0] = W;
    j.src[1] = dW;
    j.src[2] = Z;
    qst_cols(j, op(0, G_LL), op(1), op(2));
    if ((rc = launch_syn(P, j, P->stream))) return rc;
    double *dst[3] = {vr, vt, vp};
    if ((rc = ifft_to(P, 0, 3, nlev, dst, ld_grid))) return rc;
    MHX_HIP_CHECK(hipStreamSynchronize(P->stream));
    return MHX_OK;
}

int mhx_spat_to_qst_batch(mhx_plan *h, int nlev, const double *f, const double *g, const double *hh,
                          size_t ld_grid, double *q, double *s, double *t, size_t ld_lm,
                          const int *lcut, void *stream)
{
    if (!h) return fail(MHX_E_INVALID, "null plan");
    Plan *P = &h->p;
    if (nlev <= 0) return MHX_OK;
    if (!f || !g || !hh || !q || !s || !t || !lcut) return fail(MHX_E_INVALID, "null pointer");
    int rc;
    if ((rc = order_after(P, stream)) || (rc = ensure_ws(P, nlev)) ||
        (rc = upload_lcut(P, lcut, nlev)))
        return rc;
    const double *src[3] = {f, g, hh};
    bool os[3] = {false, true, true};
    if ((rc = fft_from(P, 3, nlev, src, ld_grid, os))) return rc;
    AnaJob j = ana_job(P, ANA_QST, nlev, ld_lm);
    j.in_comp[0] = 0;
    j.in_comp[1] = 1;
    j.in_comp[2] = 2;
    j.out[0] = q;
    j.out[1] = s;
    j.out[2] = t;
    if ((rc = launch_ana(P, j, P->stream))) return rc;
    MHX_HIP_CHECK(hipStreamSynchronize(P->stream));
    return MHX_OK;
}

// ------------------------------------------------------------ nonlinear step
// zeroes n 64-bit words (a kernel, not a memset node, inside captured graphs)
__global__ void k_zero_u64(unsigned long long *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// the Courant inputs of mhx_nl_step_courant (host)
struct CourIn {
    double courfac, alffac;
    const double *valri2, *valhi2;
    double *vr2max, *vh2max;
};

// the lm-space assembly appended to the step (mhx_nl_step_td)
struct TdIn {
    double *const *out;  // dwdt dzdt dpdt dsdt dbdt djdt dVxBhLM
    const double *epsc;  // per level
    double CorFac;
};

static int nl_step_impl(mhx_plan *h, int nlev, const double *const *in, size_t ld_lm,
                        double *const *out, const double *or2, const double *or4,
                        const double *orho1, double LFfac, const int *lcut, double *grids,
                        void *stream, const CourIn *ci = nullptr, const TdIn *ti = nullptr)
{
    if (!h) return fail(MHX_E_INVALID, "null plan");
    Plan *P = &h->p;
    if (nlev <= 0) return MHX_OK;
    if (!in || !out || !or2 || !or4 || !orho1 || !lcut) return fail(MHX_E_INVALID, "null pointer");
    for (int i = 0; i < 11; ++i)
        if (!in[i]) return fail(MHX_E_INVALID, "null input array " + std::to_string(i));
    for (int i = 0; i < 10; ++i)
        if (!out[i]) return fail(MHX_E_INVALID, "null output array " + std::to_string(i));
    if (ti) {
        if (!ti->out || !ti->epsc) return fail(MHX_E_INVALID, "null pointer");
        for (int i = 0; i < 7; ++i)
            if (!ti->out[i]) return fail(MHX_E_INVALID, "null td output array " + std::to_string(i));
    }
    if (ld_lm < (size_t)P->lm_max) return fail(MHX_E_INVALID, "ld_lm below lm_max");
    int rc;
    if ((rc = order_after(P, stream)) || (rc = ensure_ws(P, nlev))) return rc;
    for (int i = 0; i < nlev; ++i)
        if ((rc = check_lcut(P, lcut[i]))) return rc;
    if (ti && P->td_nlev < nlev) {
        // per-level factors of the assembly (or2, or4, epsc), made before any
        // capture
        MHX_HIP_CHECK(hipStreamSynchronize(P->stream));
        free_dev(P->d_td_rad);
        free_dev(P->d_td_lcut);
        P->td_nlev = 0;
        MHX_HIP_CHECK(hipMalloc(&P->d_td_rad, 3 * (size_t)nlev * sizeof(double)));
        MHX_HIP_CHECK(hipMalloc(&P->d_td_lcut, (size_t)nlev * sizeof(int)));
        P->td_nlev = nlev;
    }
    // host arrays (the caller's *_Rloc fields) are staged through ws_H:
    // inputs copied in here, outputs copied back after the last launch
    const double *din[11];
    double *dout[10], *dtd[7] = {};
    bool hout[10], htd[7] = {};
    {
        const size_t arr = (size_t)nlev * ld_lm * 2;  // doubles per array
        int nh_in = 0, nh_out = 0;
        bool hin[11];
        for (int i = 0; i < 11; ++i) nh_in += (hin[i] = !is_device_ptr(in[i]));
        for (int i = 0; i < 10; ++i) nh_out += (hout[i] = !is_device_ptr(out[i]));
        if (ti)
            for (int i = 0; i < 7; ++i) nh_out += (htd[i] = !is_device_ptr(ti->out[i]));
        if (nh_in + nh_out &&
            (rc = ensure_buf(P->ws_H, P->ws_H_bytes, (size_t)(nh_in + nh_out) * arr * sizeof(double))))
            return rc;
        int slot = 0;
        for (int i = 0; i < 11; ++i) {
            din[i] = in[i];
            if (!hin[i]) continue;
            double *d = P->ws_H + (size_t)slot++ * arr;
            MHX_HIP_CHECK(hipMemcpyAsync(d, in[i], arr * sizeof(double), hipMemcpyHostToDevice,
                                         P->stream));
            din[i] = d;
        }
        for (int i = 0; i < 10; ++i) dout[i] = hout[i] ? P->ws_H + (size_t)slot++ * arr : out[i];
        if (ti)
            for (int i = 0; i < 7; ++i) dtd[i] = htd[i] ? P->ws_H + (size_t)slot++ * arr : ti->out[i];
        P->last_host_out = nh_out > 0;
    }
    std::vector<double> &rad = P->h_rad;
    rad.assign(3 * (size_t)nlev, 0.0);
    for (int i = 0; i < nlev; ++i) {
        rad[i] = or2[i];
        rad[nlev + i] = or4[i];
        rad[2 * nlev + i] = orho1[i];
    }
    {
        // per-level factors and lcut through a pinned buffer: the copies are
        // then asynchronous (the previous call's copies completed before its
        // end-of-call synchronisation, so the buffer is free to refill)
        const size_t ntd = ti ? 3 * (size_t)nlev : 0;
        const size_t need = (rad.size() + ntd) * sizeof(double) + (size_t)nlev * sizeof(int);
        if (P->h_pin_bytes < need) {
            if (P->h_pin) (void)hipHostFree(P->h_pin);
            P->h_pin = nullptr;
            P->h_pin_bytes = 0;
            MHX_HIP_CHECK(hipHostMalloc(&P->h_pin, need, hipHostMallocDefault));
            P->h_pin_bytes = need;
        }
        if (P->pin_pending) MHX_HIP_CHECK(hipEventSynchronize(P->ev_pin));  // previous copies read
        P->pin_pending = false;
        double *hr = static_cast<double *>(P->h_pin);
        int *hl = reinterpret_cast<int *>(hr + rad.size() + ntd);
        std::memcpy(hr, rad.data(), rad.size() * sizeof(double));
        std::memcpy(hl, lcut, (size_t)nlev * sizeof(int));
        MHX_HIP_CHECK(hipMemcpyAsync(P->d_rad, hr, rad.size() * sizeof(double), hipMemcpyHostToDevice,
                                     P->stream));
        if (ti) {
            double *ht = hr + rad.size();
            for (int i = 0; i < nlev; ++i) {
                ht[i] = or2[i];
                ht[nlev + i] = or4[i];
                ht[2 * nlev + i] = ti->epsc[i];
            }
            MHX_HIP_CHECK(hipMemcpyAsync(P->d_td_rad, ht, ntd * sizeof(double), hipMemcpyHostToDevice,
                                         P->stream));
        }
        MHX_HIP_CHECK(hipMemcpyAsync(P->d_lcut, hl, (size_t)nlev * sizeof(int), hipMemcpyHostToDevice,
                                     P->stream));
        MHX_HIP_CHECK(hipEventRecord(P->ev_pin, P->stream));
        P->pin_pending = true;
    }
    // Courant maxima: device parameters and per-level maxima (zeroed by a
    // kernel at the start of the step's launches)
    P->cour = CourArgs();
    if (ci) {
        if (!ci->valri2 || !ci->valhi2 || !ci->vr2max || !ci->vh2max)
            return fail(MHX_E_INVALID, "null Courant argument");
        if (P->cour_nlev < nlev) {
            free_dev(P->d_cour_par);
            free_dev(P->d_cour_max);
            P->cour_nlev = 0;
            MHX_HIP_CHECK(hipMalloc(&P->d_cour_par, (2 + 2 * (size_t)nlev) * sizeof(double)));
            MHX_HIP_CHECK(hipMalloc(&P->d_cour_max, 2 * (size_t)nlev * sizeof(unsigned long long)));
            P->cour_nlev = nlev;
        }
        MHX_HIP_CHECK(hipStreamSynchronize(P->stream));  // the previous upload has been read
        P->h_cour.assign(2 + 2 * (size_t)nlev, 0.0);
        P->h_cour[0] = ci->courfac * ci->courfac;
        P->h_cour[1] = ci->alffac * ci->alffac;
        for (int i = 0; i < nlev; ++i) {
            P->h_cour[2 + i] = ci->valri2[i];
            P->h_cour[2 + nlev + i] = ci->valhi2[i];
        }
        MHX_HIP_CHECK(hipMemcpyAsync(P->d_cour_par, P->h_cour.data(), P->h_cour.size() * sizeof(double),
                                     hipMemcpyHostToDevice, P->stream));
        P->cour.par = P->d_cour_par;
        P->cour.max = P->d_cour_max;
    }
    hipStream_t s = P->stream;
    // a split-precision plan runs the Legendre stage in scaled FP16 x3 and
    // keeps Fourier data and grids in FP32; the per-call API stays FP64
    const bool split = P->precision == MHX_PREC_SPLIT16;
    // du/dr grids: always when the caller asks for the grids, or when the
    // launches are timed one by one (their count assumes them)
    const bool dudr = P->dudr || grids || P->timing;
    int nev = 0;
    auto mark = [&]() -> int {
        if (P->timing && nev < Plan::NEV) MHX_HIP_CHECK(hipEventRecord(P->ev[nev++], s));
        return MHX_OK;
    };
    if ((rc = mark())) return rc;
    // independent launches dealt over the branch streams (one stream when
    // timing, so that the events between launches measure each launch)
    const int nbr = P->timing ? 1 : P->nbranch;
    auto fork = [&]() -> int {
        if (nbr < 2) return MHX_OK;
        MHX_HIP_CHECK(hipEventRecord(P->ev_fork, s));
        for (int i = 0; i + 1 < nbr; ++i) MHX_HIP_CHECK(hipStreamWaitEvent(P->br[i], P->ev_fork, 0));
        return MHX_OK;
    };
    auto join = [&]() -> int {
        for (int i = 0; i + 1 < nbr; ++i) {
            MHX_HIP_CHECK(hipEventRecord(P->ev_join[i], P->br[i]));
            MHX_HIP_CHECK(hipStreamWaitEvent(s, P->ev_join[i], 0));
        }
        return MHX_OK;
    };
    auto on = [&](int i) -> hipStream_t { return i % nbr == 0 ? s : P->br[i % nbr - 1]; };

    // inputs: 0 w, 1 dw, 2 ddw, 3 z, 4 dz, 5 s, 6 b, 7 db, 8 ddb, 9 aj, 10 dj
    enum { W = 0, DW, DDW, Z, DZ, S, B, DB, DDB, AJ, DJ };
    // s, ds/dtheta, ds/dphi -> comps 15..17
    {
        auto launch_split = [&]() -> int {
            int rc2;
            // FP32 Fourier data; F2 theta-blocked; compile-time column operations
            NSynArgs a;
            std::memset(&a, 0, sizeof(a));
            a.src_ld = ld_lm;
            a.fout = reinterpret_cast<float *>(P->ws_F);
            a.fout_ld = (size_t)18 * P->n_m * P->n_theta;
            a.lcut = P->d_lcut;
            a.or2 = P->d_rad;
            a.nlev = nlev;
            a.src[0] = din[S];
            a.comp0 = 15;
            if ((rc2 = fork())) return rc2;
            if ((rc2 = launch_nsyn16(P, NS_SG, a, on(5))) || (rc2 = mark())) return rc2;
            // u, curl u, du/dr, B, curl B (see the FP64 table below)
            const int qsrc[5][4] = {{W, DW, Z, -1}, {Z, DZ, W, DDW}, {DW, DDW, DZ, -1},
                                    {B, DB, AJ, -1}, {AJ, DJ, B, DDB}};
            for (int v = 0; v < 5; ++v) {
                if (v == 2 && !dudr) continue;
                for (int k = 0; k < 4; ++k) a.src[k] = qsrc[v][k] >= 0 ? din[qsrc[v][k]] : nullptr;
                a.comp0 = 3 * v;
                const int kind = qsrc[v][3] >= 0 ? NS_QSTC : NS_QST;
                if ((rc2 = launch_nsyn16(P, kind, a, on(v))) || (rc2 = mark())) return rc2;
            }
            if ((rc2 = join())) return rc2;
            if ((rc2 = launch_nl_fused(P, nlev, P->ws_F, P->ws_F2, P->d_rad, LFfac, dudr ? P->ws_G : nullptr,
                                       grids, s)) ||
                (rc2 = mark()))
                return rc2;
            NAnaArgs b;
            std::memset(&b, 0, sizeof(b));
            b.fin = reinterpret_cast<const float *>(P->ws_F2);
            b.fin_ld = (size_t)10 * P->n_m * ((P->n_theta + 15) / 16 * 16);
            b.out_ld = ld_lm;
            b.lcut = P->d_lcut;
            b.nlev = nlev;
            if ((rc2 = fork())) return rc2;
            for (int v = 0; v < 4; ++v) {
                const int nout = v < 3 ? 3 : 1;
                for (int k = 0; k < 3; ++k) {
                    b.in_comp[k] = k < nout ? 3 * v + k : -1;
                    b.out[k] = k < nout ? dout[3 * v + k] : nullptr;
                }
                if ((rc2 = launch_nana16(P, v < 3 ? NA_QST : NA_SH, b, on(v))) || (rc2 = mark()))
                    return rc2;
            }
            return join();
        };
        auto launch_fp64 = [&]() -> int {
            int rc2;
            {
                SynJob j = syn_job(P, SYN_SG, nlev, ld_lm);
                j.src[0] = din[S];
                j.col[0] = op(0);
                j.col[1] = op(0, G_ONE, 1, 1);
                j.out_comp[0] = 15; j.out_comp[1] = 16; j.out_comp[2] = 17;
                if ((rc2 = fork())) return rc2;
                if ((rc2 = launch_syn(P, j, on(5))) || (rc2 = mark())) return rc2;
            }
            struct QSpec { int q, sidx, t0, t1; bool curl; int comp0; };
            // torpol_to_spat(a, b, c): Q = l(l+1) a, S = b, T = c
            // torpol_to_curl_spat(or2, B, ddB, J, dJ): Q = l(l+1) J, S = dJ, T = or2 l(l+1) B - ddB
            const QSpec qs[5] = {
                {W, DW, Z, -1, false, 0},     // u
                {Z, DZ, W, DDW, true, 3},     // curl u
                {DW, DDW, DZ, -1, false, 6},  // du/dr
                {B, DB, AJ, -1, false, 9},    // B
                {AJ, DJ, B, DDB, true, 12},   // curl B
            };
            for (int v = 0; v < 5; ++v) {
                if (v == 2 && !dudr) continue;
                const QSpec &q = qs[v];
                SynJob j = syn_job(P, SYN_QST, nlev, ld_lm);
                j.src[0] = din[q.q];
                j.src[1] = din[q.sidx];
                j.src[2] = din[q.t0];
                j.src[3] = q.t1 >= 0 ? din[q.t1] : nullptr;
                ColOp T = q.curl ? op(2, G_OR2LL, 0, 0, 1.0, 3, -1.0) : op(2);
                qst_cols(j, op(0, G_LL), op(1), T);
                j.out_comp[0] = q.comp0; j.out_comp[1] = q.comp0 + 1; j.out_comp[2] = q.comp0 + 2;
                if ((rc2 = launch_syn(P, j, on(v))) || (rc2 = mark())) return rc2;
            }
            if ((rc2 = join())) return rc2;
            // fused: inverse FFTs, grid products, forward FFTs; du/dr grids to ws_G
            if (((rc2 = launch_nl_fused(P, nlev, P->ws_F, P->ws_F2, P->d_rad, LFfac, dudr ? P->ws_G : nullptr,
                                                 grids, s)) ||
                           (rc2 = mark())))
                return rc2;
            if ((rc2 = fork())) return rc2;
            for (int v = 0; v < 3; ++v) {
                AnaJob j = ana_job(P, ANA_QST, nlev, ld_lm);
                j.in_comp[0] = 3 * v;
                j.in_comp[1] = 3 * v + 1;
                j.in_comp[2] = 3 * v + 2;
                j.out[0] = dout[3 * v];
                j.out[1] = dout[3 * v + 1];
                j.out[2] = dout[3 * v + 2];
                if ((rc2 = launch_ana(P, j, on(v))) || (rc2 = mark())) return rc2;
            }
            {
                AnaJob j = ana_job(P, ANA_SH, nlev, ld_lm);
                j.in_comp[0] = 9;
                j.out[0] = dout[9];
                if ((rc2 = launch_ana(P, j, on(3))) || (rc2 = mark())) return rc2;
            }
            return join();
        };
        auto launch_all = [&]() -> int {
            if (P->cour.max) {
                k_zero_u64<<<(2 * nlev + 255) / 256, 256, 0, s>>>(P->cour.max, 2 * nlev);
                MHX_HIP_CHECK(hipGetLastError());
            }
            int rc2 = split ? launch_split() : launch_fp64();
            if (rc2 || !ti) return rc2;
            // the lm-space assembly on the step's outputs (src/get_td.f90)
            TdArgs a;
            std::memset(&a, 0, sizeof(a));
            const double *tin[10] = {dout[0], dout[1], dout[2], dout[4], dout[6], dout[7], dout[8],
                                     din[W], din[DW], din[Z]};
            for (int i = 0; i < 10; ++i) a.in[i] = tin[i];
            for (int i = 0; i < 7; ++i) a.out[i] = dtd[i];
            a.ld_lm = ld_lm;
            a.rad = P->d_td_rad;
            a.lcut = P->d_lcut;
            a.CorFac = ti->CorFac;
            a.nlev = nlev;
            if ((rc2 = launch_get_td(P, a, s)) || (rc2 = mark())) return rc2;
            return MHX_OK;
        };
        if (P->timing) {
            if ((rc = launch_all())) return rc;
        } else {
            std::vector<unsigned long long> key = {(unsigned long long)nlev, (unsigned long long)ld_lm,
                                                   (unsigned long long)(uintptr_t)grids};
            unsigned long long lf;
            std::memcpy(&lf, &LFfac, sizeof(lf));
            key.push_back(lf);
            for (int i = 0; i < 11; ++i) key.push_back((unsigned long long)(uintptr_t)din[i]);
            for (int i = 0; i < 10; ++i) key.push_back((unsigned long long)(uintptr_t)dout[i]);
            key.push_back(ti ? 1ull : 0ull);
            key.push_back(dudr ? 1ull : 0ull);
            if (ti) {
                for (int i = 0; i < 7; ++i) key.push_back((unsigned long long)(uintptr_t)dtd[i]);
                unsigned long long cf;
                std::memcpy(&cf, &ti->CorFac, sizeof(cf));
                key.push_back(cf);
                key.push_back((unsigned long long)(uintptr_t)P->d_td_rad);
            }
            // workspaces the launches address (reallocated when nlev grows)
            for (const void *w : {(const void *)P->ws_F, (const void *)P->ws_F2, (const void *)P->ws_G,
                                  (const void *)P->d_lcut, (const void *)P->d_rad,
                                  (const void *)P->cour.max, (const void *)P->cour.par})
                key.push_back((unsigned long long)(uintptr_t)w);
            if (P->gexec && key == P->gkey) {
                MHX_HIP_CHECK(hipGraphLaunch(P->gexec, s));
            } else if (key == P->gkey_seen) {
                // second call with these arrays: capture the launches once
                hipGraph_t g = nullptr;
                MHX_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                const int rcl = launch_all();
                const hipError_t ec = hipStreamEndCapture(s, &g);
                if (rcl) {
                    if (g) (void)hipGraphDestroy(g);
                    return rcl;
                }
                if (ec != hipSuccess) return fail(MHX_E_HIP, "graph capture of the nonlinear step failed");
                hipGraphExec_t ex = nullptr;
                const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                if (ei != hipSuccess) return fail(MHX_E_HIP, "graph instantiation failed");
                if (P->gexec) (void)hipGraphExecDestroy(P->gexec);
                P->gexec = ex;
                P->gkey = key;
                MHX_HIP_CHECK(hipGraphLaunch(P->gexec, s));
            } else {
                if ((rc = launch_all())) return rc;
                P->gkey_seen = key;
            }
        }
    }
    for (int i = 0; i < 10; ++i)
        if (hout[i])
            MHX_HIP_CHECK(hipMemcpyAsync(out[i], dout[i], (size_t)nlev * ld_lm * 2 * sizeof(double),
                                         hipMemcpyDeviceToHost, P->stream));
    if (ti)
        for (int i = 0; i < 7; ++i)
            if (htd[i])
                MHX_HIP_CHECK(hipMemcpyAsync(ti->out[i], dtd[i], (size_t)nlev * ld_lm * 2 * sizeof(double),
                                             hipMemcpyDeviceToHost, P->stream));
    if (P->timing) {
        // launches: 6 synthesis, fused grid stage, 4 analysis
        MHX_HIP_CHECK(hipEventSynchronize(P->ev[nev - 1]));
        P->n_launch = nev - 1;
        for (int i = 0; i + 1 < nev; ++i) {
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, P->ev[i], P->ev[i + 1]);
            P->launch_ms[i] = ms;
        }
        const int first[6] = {0, 6, 7, 7, 7, 11};
        for (int st = 0; st < 5; ++st) {
            double t = 0.0;
            for (int i = first[st]; i < first[st + 1] && i < P->n_launch; ++i) t += P->launch_ms[i];
            P->stage_ms[st] = t;
        }
    }
    return MHX_OK;
}

// end of a nonlinear step: with a caller's stream and device outputs the
// caller's stream is ordered after the step (no host wait); otherwise the
// host waits for the step
static int finish_step(mhx_plan *h, void *stream)
{
    Plan *P = &h->p;
    if (stream && !P->last_host_out) {
        MHX_HIP_CHECK(hipEventRecord(P->ev_done, P->stream));
        MHX_HIP_CHECK(hipStreamWaitEvent((hipStream_t)stream, P->ev_done, 0));
        return MHX_OK;
    }
    MHX_HIP_CHECK(hipStreamSynchronize(P->stream));
    return MHX_OK;
}

int mhx_nl_step(mhx_plan *h, int nlev, const double *const *in, size_t ld_lm, double *const *out,
                const double *or2, const double *or4, const double *orho1, double LFfac,
                const int *lcut, void *stream)
{
    int rc = nl_step_impl(h, nlev, in, ld_lm, out, or2, or4, orho1, LFfac, lcut, nullptr, stream);
    if (rc) return rc;
    return finish_step(h, stream);
}

int mhx_nl_step_td(mhx_plan *h, int nlev, const double *const *in, size_t ld_lm, double *const *out,
                   double *const *td_out, const double *or2, const double *or4, const double *orho1,
                   double LFfac, const double *epsc, double CorFac, const int *lcut, void *stream)
{
    TdIn ti{td_out, epsc, CorFac};
    int rc = nl_step_impl(h, nlev, in, ld_lm, out, or2, or4, orho1, LFfac, lcut, nullptr, stream,
                          nullptr, &ti);
    if (rc) return rc;
    return finish_step(h, stream);
}

// the per-level energy sums of src/kinetic_energy.f90:119-160 (mag = 0:
// w, dw, z, orho1) or src/magnetic_energy.f90:264-340 (mag = 1: b, db, aj)
static int energy_sums(mhx_plan *h, int nlev, const double *w, const double *dw, const double *z,
                       size_t ld_lm, const double *or2, const double *orho1, double *sums, int mag)
{
    if (!h) return fail(MHX_E_INVALID, "null plan");
    Plan *P = &h->p;
    if (nlev <= 0) return MHX_OK;
    if (!w || !dw || !z || !or2 || (!mag && !orho1) || !sums) return fail(MHX_E_INVALID, "null pointer");
    if (nlev > 65535) return fail(MHX_E_INVALID, "more than 65535 levels in one call");
    if (ld_lm < (size_t)P->lm_max) return fail(MHX_E_INVALID, "ld_lm below lm_max");
    const int nsum = mag ? 10 : 8;
    int rc;
    const size_t arr = (size_t)nlev * ld_lm * 2;
    const double *in[3] = {w, dw, z};
    bool hin[3];
    int nh = 0;
    for (int i = 0; i < 3; ++i) nh += (hin[i] = !is_device_ptr(in[i]));
    if (nh && (rc = ensure_buf(P->ws_H, P->ws_H_bytes, (size_t)nh * arr * sizeof(double)))) return rc;
    if (P->ek_nlev < nlev) {
        free_dev(P->d_ek);
        P->ek_nlev = 0;
        MHX_HIP_CHECK(hipMalloc(&P->d_ek, (12 + 10 * (size_t)EK_CHUNKS) * nlev * sizeof(double)));
        P->ek_nlev = nlev;
    }
    // the plan stream may still read the previous call's host staging
    MHX_HIP_CHECK(hipStreamSynchronize(P->stream));
    EkArgs a;
    std::memset(&a, 0, sizeof(a));
    const double *src[3];
    int slot = 0;
    for (int i = 0; i < 3; ++i) {
        src[i] = in[i];
        if (!hin[i]) continue;
        double *d = P->ws_H + (size_t)slot++ * arr;
        MHX_HIP_CHECK(hipMemcpyAsync(d, in[i], arr * sizeof(double), hipMemcpyHostToDevice, P->stream));
        src[i] = d;
    }
    std::vector<double> rad(2 * (size_t)nlev);
    for (int i = 0; i < nlev; ++i) {
        rad[i] = or2[i];
        rad[nlev + i] = mag ? 1.0 : orho1[i];
    }
    MHX_HIP_CHECK(hipMemcpyAsync(P->d_ek, rad.data(), 2 * (size_t)nlev * sizeof(double),
                                 hipMemcpyHostToDevice, P->stream));
    a.w = src[0];
    a.dw = src[1];
    a.z = src[2];
    a.ld_lm = ld_lm;
    a.or2 = P->d_ek;
    a.orho1 = mag ? nullptr : P->d_ek + nlev;
    a.out = P->d_ek + 2 * (size_t)nlev;
    a.part = P->d_ek + 12 * (size_t)nlev;
    a.nlev = nlev;
    a.mag = mag;
    a.nsum = nsum;
    if ((rc = launch_e_kin(P, a, P->stream))) return rc;
    MHX_HIP_CHECK(hipMemcpyAsync(sums, a.out, (size_t)nsum * nlev * sizeof(double), hipMemcpyDefault,
                                 P->stream));
    MHX_HIP_CHECK(hipStreamSynchronize(P->stream));
    return MHX_OK;
}

int mhx_get_e_kin(mhx_plan *h, int nlev, const double *w, const double *dw, const double *z,
                  size_t ld_lm, const double *or2, const double *orho1, double *sums)
{
    return energy_sums(h, nlev, w, dw, z, ld_lm, or2, orho1, sums, 0);
}

int mhx_get_e_mag(mhx_plan *h, int nlev, const double *b, const double *db, const double *aj,
                  size_t ld_lm, const double *or2, double *sums)
{
    return energy_sums(h, nlev, b, db, aj, ld_lm, or2, nullptr, sums, 1);
}

int mhx_get_td(mhx_plan *h, int nlev, const double *const *in, size_t ld_lm, double *const *out,
               const double *or2, const double *or4, const double *epsc, double CorFac,
               const int *lcut, void *stream)
{
    if (!h) return fail(MHX_E_INVALID, "null plan");
    Plan *P = &h->p;
    if (nlev <= 0) return MHX_OK;
    if (!in || !out || !or2 || !or4 || !epsc || !lcut) return fail(MHX_E_INVALID, "null pointer");
    for (int i = 0; i < 10; ++i)
        if (!in[i]) return fail(MHX_E_INVALID, "null input array " + std::to_string(i));
    for (int i = 0; i < 7; ++i)
        if (!out[i]) return fail(MHX_E_INVALID, "null output array " + std::to_string(i));
    if (ld_lm < (size_t)P->lm_max) return fail(MHX_E_INVALID, "ld_lm below lm_max");
    int rc;
    for (int i = 0; i < nlev; ++i)
        if ((rc = check_lcut(P, lcut[i]))) return rc;
    if ((rc = order_after(P, stream))) return rc;
    // host arrays staged through ws_H (inputs in, outputs back after the launch)
    const size_t arr = (size_t)nlev * ld_lm * 2;
    bool hin[10], hout[7];
    int nh = 0;
    for (int i = 0; i < 10; ++i) nh += (hin[i] = !is_device_ptr(in[i]));
    for (int i = 0; i < 7; ++i) nh += (hout[i] = !is_device_ptr(out[i]));
    if (nh && (rc = ensure_buf(P->ws_H, P->ws_H_bytes, (size_t)nh * arr * sizeof(double)))) return rc;
    TdArgs a;
    std::memset(&a, 0, sizeof(a));
    int slot = 0;
    for (int i = 0; i < 10; ++i) {
        a.in[i] = in[i];
        if (!hin[i]) continue;
        double *d = P->ws_H + (size_t)slot++ * arr;
        MHX_HIP_CHECK(hipMemcpyAsync(d, in[i], arr * sizeof(double), hipMemcpyHostToDevice, P->stream));
        a.in[i] = d;
    }
    for (int i = 0; i < 7; ++i) a.out[i] = hout[i] ? P->ws_H + (size_t)slot++ * arr : out[i];
    // per-level factors
    if (P->td_nlev < nlev) {
        free_dev(P->d_td_rad);
        free_dev(P->d_td_lcut);
        P->td_nlev = 0;
        MHX_HIP_CHECK(hipMalloc(&P->d_td_rad, 3 * (size_t)nlev * sizeof(double)));
        MHX_HIP_CHECK(hipMalloc(&P->d_td_lcut, (size_t)nlev * sizeof(int)));
        P->td_nlev = nlev;
    }
    // the host copy stays valid until the copies are done (synchronised
    // below, or by the next call's wait on the plan stream)
    MHX_HIP_CHECK(hipStreamSynchronize(P->stream));
    P->h_td.assign(3 * (size_t)nlev, 0.0);
    for (int i = 0; i < nlev; ++i) {
        P->h_td[i] = or2[i];
        P->h_td[nlev + i] = or4[i];
        P->h_td[2 * nlev + i] = epsc[i];
    }
    MHX_HIP_CHECK(hipMemcpyAsync(P->d_td_rad, P->h_td.data(), 3 * (size_t)nlev * sizeof(double),
                                 hipMemcpyHostToDevice, P->stream));
    P->h_td_lcut.assign(lcut, lcut + nlev);
    MHX_HIP_CHECK(hipMemcpyAsync(P->d_td_lcut, P->h_td_lcut.data(), (size_t)nlev * sizeof(int),
                                 hipMemcpyHostToDevice, P->stream));
    a.ld_lm = ld_lm;
    a.rad = P->d_td_rad;
    a.lcut = P->d_td_lcut;
    a.CorFac = CorFac;
    a.nlev = nlev;
    if ((rc = launch_get_td(P, a, P->stream))) return rc;
    for (int i = 0; i < 7; ++i)
        if (hout[i])
            MHX_HIP_CHECK(hipMemcpyAsync(out[i], a.out[i], arr * sizeof(double), hipMemcpyDeviceToHost,
                                         P->stream));
    P->last_host_out = nh > 0;
    if (stream && nh == 0) {
        MHX_HIP_CHECK(hipEventRecord(P->ev_done, P->stream));
        MHX_HIP_CHECK(hipStreamWaitEvent((hipStream_t)stream, P->ev_done, 0));
        return MHX_OK;
    }
    MHX_HIP_CHECK(hipStreamSynchronize(P->stream));
    return MHX_OK;
}

int mhx_nl_step_courant(mhx_plan *h, int nlev, const double *const *in, size_t ld_lm,
                        double *const *out, const double *or2, const double *or4,
                        const double *orho1, double LFfac, const int *lcut, double courfac,
                        double alffac, const double *valri2, const double *valhi2,
                        double *vr2max, double *vh2max, void *stream)
{
    CourIn ci{courfac, alffac, valri2, valhi2, vr2max, vh2max};
    int rc = nl_step_impl(h, nlev, in, ld_lm, out, or2, or4, orho1, LFfac, lcut, nullptr, stream, &ci);
    Plan *P = h ? &h->p : nullptr;
    if (rc) {
        if (P) P->cour = CourArgs();
        return rc;
    }
    // the maxima back to the host (the call then returns when the step is done)
    std::vector<unsigned long long> mx(2 * (size_t)nlev);
    MHX_HIP_CHECK(hipMemcpyAsync(mx.data(), P->d_cour_max, mx.size() * sizeof(unsigned long long),
                                 hipMemcpyDeviceToHost, P->stream));
    MHX_HIP_CHECK(hipStreamSynchronize(P->stream));
    P->cour = CourArgs();
    for (int i = 0; i < nlev; ++i) {
        double a, b;
        std::memcpy(&a, &mx[2 * i], sizeof(a));
        std::memcpy(&b, &mx[2 * i + 1], sizeof(b));
        vr2max[i] = a;
        vh2max[i] = b;
    }
    return finish_step(h, stream);
}

int mhx_nl_step_grids(mhx_plan *h, int nlev, const double *const *in, size_t ld_lm,
                      double *const *out, const double *or2, const double *or4,
                      const double *orho1, double LFfac, const int *lcut, double *grids,
                      void *stream)
{
    if (!grids) return fail(MHX_E_INVALID, "null grids pointer");
    int rc = nl_step_impl(h, nlev, in, ld_lm, out, or2, or4, orho1, LFfac, lcut, grids, stream);
    if (rc) return rc;
    return finish_step(h, stream);
}

int mhx_set_dudr(mhx_plan *h, int on)
{
    if (!h) return fail(MHX_E_INVALID, "null plan");
    h->p.dudr = on != 0;
    return MHX_OK;
}

int mhx_set_timing(mhx_plan *h, int on)
{
    if (!h) return fail(MHX_E_INVALID, "null plan");
    h->p.timing = on != 0;
    return MHX_OK;
}

int mhx_stage_times(const mhx_plan *h, double *ms5)
{
    if (!h || !ms5) return fail(MHX_E_INVALID, "null argument");
    for (int i = 0; i < 5; ++i) ms5[i] = h->p.stage_ms[i];
    return MHX_OK;
}

int mhx_launch_times(const mhx_plan *h, double *ms, int n)
{
    if (!h || !ms) return fail(MHX_E_INVALID, "null argument");
    int k = h->p.n_launch < n ? h->p.n_launch : n;
    for (int i = 0; i < k; ++i) ms[i] = h->p.launch_ms[i];
    return k;
}

}  // extern "C"
