// Tables of the split-precision form (MHX_PREC_SPLIT16): the FP64 P_lm and
// sin(theta) dP_lm/dtheta tables (mhx_tables.hip) scaled per row by a power
// of two and split into FP16 hi/lo pairs in the MFMA fragment layouts the
// split Legendre kernels of the nonlinear step read (mhx_nlsplit.hip), the
// c_l coefficients of the dP identity, and the FP64 seeds of the in-kernel
// recurrence (MHX_PREC_SPLIT16_REC).
//
// Scaling keeps every table operand inside FP16's normal range: one
// power-of-2 factor per output row (per (m, theta) for the synthesis, per
// (m, l) for the analysis), common to P and dP, chosen so that the row
// maximum lies in [2^14, 2^15); the kernels remove it in their epilogues.
//
// f16 16x16x32 lane layouts (checked on gfx950): A lane l holds row l%16,
// k = 8(l/16)..8(l/16)+7; B lane l holds column l%16, the same k; C element
// j of lane l is row 4(l/16)+j, column l%16.
#include "mhx_internal.h"

#include <cmath>

namespace mhx {

namespace {

__device__ inline void split_h(float v, _Float16 &hi, _Float16 &lo)
{
    hi = (_Float16)v;
    lo = (_Float16)(v - (float)hi);
}

// ------------------------------------------------------------ split tables
// per-row factors: synthesis rows (mc, theta): max over l of |P|, |dP|
__global__ void k_rowscale_syn(int l_max, int minc, int n_m, int nh, int nt16,
                               const double *__restrict__ sP, const double *__restrict__ sdP,
                               const size_t *__restrict__ off, float *__restrict__ rs)
{
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n_m * nh) return;
    int mc = idx / nh, th = idx - mc * nh;
    int m = mc * minc;
    double mx = 0.0;
    for (int l = m; l <= l_max; ++l) {
        int par = (l - m) & 1, kk = (l - m) >> 1;
        size_t o = off[2 * mc + par] + ((size_t)(kk >> 2) * nt16 + (th >> 4)) * 64 + (th & 15) + 16 * (kk & 3);
        mx = fmax(mx, fmax(fabs(sP[o]), fabs(sdP[o])));
    }
    rs[idx] = pow2_scale((float)mx);
}

__global__ void k_split_syn(int l_max, int minc, int n_m, int nh, int nt16,
                            const double *__restrict__ sP, const double *__restrict__ sdP,
                            const size_t *__restrict__ off, const size_t *__restrict__ off16,
                            const float *__restrict__ rs, _Float16 *__restrict__ hP,
                            _Float16 *__restrict__ hdP)
{
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n_m * nh) return;
    int mc = idx / nh, th = idx - mc * nh;
    int m = mc * minc;
    double s = (double)rs[idx];
    for (int l = m; l <= l_max; ++l) {
        int par = (l - m) & 1, kk = (l - m) >> 1;
        size_t o = off[2 * mc + par] + ((size_t)(kk >> 2) * nt16 + (th >> 4)) * 64 + (th & 15) + 16 * (kk & 3);
        int kc = kk >> 5, kin = kk & 31;
        int lane = (th & 15) + 16 * (kin >> 3);
        size_t o16 = off16[2 * mc + par] + (((size_t)kc * nt16 + (th >> 4)) * 64 + lane) * 16 + (kin & 7);
        _Float16 h, lo;
        split_h((float)(sP[o] * s), h, lo);
        hP[o16] = h;
        hP[o16 + 8] = lo;
        split_h((float)(sdP[o] * s), h, lo);
        hdP[o16] = h;
        hdP[o16 + 8] = lo;
    }
}

// analysis rows (mc, par, kk): max over theta of |wP|, |wdP|; one thread per row
__global__ void k_rowscale_ana(int l_max, int minc, int n_m, int nh, int nq4,
                               const double *__restrict__ aP, const double *__restrict__ adP,
                               const size_t *__restrict__ off, const size_t *__restrict__ soff,
                               float *__restrict__ rs, int nrows_total)
{
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nrows_total) return;
    // locate (mc, par, kk) from soff (row blocks of K[par] entries)
    int lo = 0, hi = 2 * n_m - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) / 2;
        if (soff[mid] <= (size_t)idx) lo = mid; else hi = mid - 1;
    }
    int blk = lo, mc = blk / 2, par = blk % 2;
    int kk = idx - (int)soff[blk];
    int m = mc * minc;
    if (m + par + 2 * kk > l_max) { rs[idx] = 1.f; return; }
    double mx = 0.0;
    for (int th = 0; th < nh; ++th) {
        size_t o = off[2 * mc + par] + ((size_t)(kk >> 4) * nq4 + (th >> 2)) * 64 + (kk & 15) + 16 * (th & 3);
        mx = fmax(mx, fmax(fabs(aP[o]), fabs(adP[o])));
    }
    rs[idx] = pow2_scale((float)mx);
}

__global__ void k_split_ana(int l_max, int minc, int n_m, int nh, int nq4, int nq32,
                            const double *__restrict__ aP, const double *__restrict__ adP,
                            const size_t *__restrict__ off, const size_t *__restrict__ off16,
                            const size_t *__restrict__ soff, const float *__restrict__ rs,
                            _Float16 *__restrict__ hP, _Float16 *__restrict__ hdP)
{
    // one thread per (mc, theta)
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n_m * nh) return;
    int mc = idx / nh, th = idx - mc * nh;
    int m = mc * minc;
    for (int l = m; l <= l_max; ++l) {
        int par = (l - m) & 1, kk = (l - m) >> 1;
        size_t o = off[2 * mc + par] + ((size_t)(kk >> 4) * nq4 + (th >> 2)) * 64 + (kk & 15) + 16 * (th & 3);
        double s = (double)rs[soff[2 * mc + par] + kk];
        int kc = th >> 5, kin = th & 31;
        int lane = (kk & 15) + 16 * (kin >> 3);
        size_t o16 = off16[2 * mc + par] + (((size_t)(kk >> 4) * nq32 + kc) * 64 + lane) * 16 + (kin & 7);
        _Float16 h, lo;
        split_h((float)(aP[o] * s), h, lo);
        hP[o16] = h;
        hP[o16 + 8] = lo;
        split_h((float)(adP[o] * s), h, lo);
        hdP[o16] = h;
        hdP[o16 + 8] = lo;
    }
}

// ---- P-only tables with l <= l_max + 1 --------------------------------
// c_l of the three-term recurrence cos(theta) P_l = c_{l+1} P_{l+1} + c_l P_{l-1}
// of orthonormal P_lm (src/plms.f90:81-97)
__device__ inline double rec_c(int l, int m)
{
    const double dl = (double)l, dm = (double)m;
    return sqrt((dl * dl - dm * dm) / (4.0 * dl * dl - 1.0));
}

// P_l(theta_h) for l in [m, l_max + 1] from the FP64 synthesis table, the
// last one by one step of the recurrence
__device__ inline double syn_p(const double *sP, const size_t *off, int mc, int m, int l,
                               int l_max, int th, int nt16, double ct)
{
    auto at = [&](int ll) {
        const int par = (ll - m) & 1, kk = (ll - m) >> 1;
        return sP[off[2 * mc + par] + ((size_t)(kk >> 2) * nt16 + (th >> 4)) * 64 + (th & 15) + 16 * (kk & 3)];
    };
    if (l <= l_max) return at(l);
    const double pl = at(l_max), plm1 = (l_max - 1 >= m) ? at(l_max - 1) : 0.0;
    return (ct * pl - rec_c(l_max, m) * plm1) / rec_c(l_max + 1, m);
}

__device__ inline double ana_p(const double *aP, const size_t *off, int mc, int m, int l,
                               int l_max, int th, int nq4, double ct)
{
    auto at = [&](int ll) {
        const int par = (ll - m) & 1, kk = (ll - m) >> 1;
        return aP[off[2 * mc + par] + ((size_t)(kk >> 4) * nq4 + (th >> 2)) * 64 + (kk & 15) + 16 * (th & 3)];
    };
    if (l <= l_max) return at(l);
    const double pl = at(l_max), plm1 = (l_max - 1 >= m) ? at(l_max - 1) : 0.0;
    return (ct * pl - rec_c(l_max, m) * plm1) / rec_c(l_max + 1, m);
}

// synthesis rows (mc, theta): factor from max over l of |P|, then split
__global__ void k_split_synx(int l_max, int minc, int n_m, int nh, int nt16,
                             const double *__restrict__ sP, const size_t *__restrict__ off,
                             const size_t *__restrict__ offx, const double *__restrict__ theta,
                             float *__restrict__ rs, _Float16 *__restrict__ hP)
{
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n_m * nh) return;
    int mc = idx / nh, th = idx - mc * nh;
    int m = mc * minc;
    const double ct = cos(theta[th]);
    double mx = 0.0;
    for (int l = m; l <= l_max + 1; ++l) mx = fmax(mx, fabs(syn_p(sP, off, mc, m, l, l_max, th, nt16, ct)));
    const float sc = pow2_scale((float)mx);
    rs[idx] = sc;
    for (int l = m; l <= l_max + 1; ++l) {
        int par = (l - m) & 1, kk = (l - m) >> 1;
        int kc = kk >> 5, kin = kk & 31;
        int lane = (th & 15) + 16 * (kin >> 3);
        // per 64-lane fragment: the 64 lanes' hi halves, then their lo halves
        size_t o16 = offx[2 * mc + par] + (((size_t)kc * nt16 + (th >> 4)) * 128 + lane) * 8 + (kin & 7);
        _Float16 h, lo;
        split_h((float)(syn_p(sP, off, mc, m, l, l_max, th, nt16, ct) * sc), h, lo);
        hP[o16] = h;
        hP[o16 + 512] = lo;
    }
}

// analysis rows (mc, par, kk): factor from max over theta of |wP|
__global__ void k_rowscale_anax(int l_max, int minc, int n_m, int nh, int nq4,
                                const double *__restrict__ aP, const size_t *__restrict__ off,
                                const size_t *__restrict__ soffx, const double *__restrict__ theta,
                                float *__restrict__ rs, int nrows_total)
{
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nrows_total) return;
    int lo = 0, hi = 2 * n_m - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) / 2;
        if (soffx[mid] <= (size_t)idx) lo = mid; else hi = mid - 1;
    }
    int blk = lo, mc = blk / 2, par = blk % 2;
    int kk = idx - (int)soffx[blk];
    int m = mc * minc;
    int l = m + par + 2 * kk;
    if (l > l_max + 1) { rs[idx] = 1.f; return; }
    double mx = 0.0;
    for (int th = 0; th < nh; ++th)
        mx = fmax(mx, fabs(ana_p(aP, off, mc, m, l, l_max, th, nq4, cos(theta[th]))));
    rs[idx] = pow2_scale((float)mx);
}

__global__ void k_split_anax(int l_max, int minc, int n_m, int nh, int nq4, int nq32,
                             const double *__restrict__ aP, const size_t *__restrict__ off,
                             const size_t *__restrict__ offx, const size_t *__restrict__ soffx,
                             const double *__restrict__ theta, const float *__restrict__ rs,
                             _Float16 *__restrict__ hP)
{
    // one thread per (mc, theta)
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n_m * nh) return;
    int mc = idx / nh, th = idx - mc * nh;
    int m = mc * minc;
    const double ct = cos(theta[th]);
    for (int l = m; l <= l_max + 1; ++l) {
        int par = (l - m) & 1, kk = (l - m) >> 1;
        double s = (double)rs[soffx[2 * mc + par] + kk];
        int kc = th >> 5, kin = th & 31;
        int lane = (kk & 15) + 16 * (kin >> 3);
        // per 64-lane fragment: the 64 lanes' hi halves, then their lo halves
        size_t o16 = offx[2 * mc + par] + (((size_t)(kk >> 4) * nq32 + kc) * 128 + lane) * 8 + (kin & 7);
        _Float16 h, lo;
        split_h((float)(ana_p(aP, off, mc, m, l, l_max, th, nq4, ct) * s), h, lo);
        hP[o16] = h;
        hP[o16 + 512] = lo;
    }
}

// c_l = sqrt((l^2 - m^2) / (4 l^2 - 1)) (0 for l = m) in FP64, stored FP32;
// one thread per (mc, l - m), l <= l_max + 2
__global__ void k_crec(int l_max, int minc, int n_m, const int *__restrict__ lstart, float *__restrict__ c)
{
    const int mc = blockIdx.y;
    const int m = mc * minc, d = blockIdx.x * blockDim.x + threadIdx.x;
    if (mc >= n_m || d > l_max + 2 - m) return;
    const int l = m + d;
    c[lstart[mc] + 2 * mc + d] = l > m ? (float)rec_c(l, m) : 0.f;
}

}  // namespace

// recurrence seeds: P_mm(theta) times the synthesis row scale, FP64 mantissa
// and power-of-two exponent (P_mm = sqrt(prod_{j=3,5..2m+1} j/(j-1)) / sqrt(4
// pi) sin^m theta, src/plms.f90:55-65, which leaves the double range for large
// m near the poles); one thread per (mc, theta_h)
__global__ void k_rec_seed(int minc, int n_m, int nh, const double *__restrict__ theta,
                           const float *__restrict__ rs, double2 *__restrict__ seed,
                           double *__restrict__ cost)
{
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n_m * nh) return;
    int mc = idx / nh, th = idx - mc * nh;
    int m = mc * minc;
    const double t = theta[th];
    const double st = sin(t);
    if (mc == 0) cost[th] = cos(t);
    double fac = 1.0;
    for (int j = 3; j <= 2 * m + 1; j += 2) fac *= (double)j / (double)(j - 1);
    // st^m = f^m 2^(k m) with st = f 2^k, f in [0.5, 1); f^m in steps of 256
    int k;
    const double f = frexp(st, &k);
    double pm = 0.28209479177387814 * sqrt(fac);
    long e = (long)k * m;
    int left = m;
    while (left > 0) {
        const int n = left < 256 ? left : 256;
        int k2;
        pm = frexp(pm * pow(f, (double)n), &k2);
        e += k2;
        left -= n;
    }
    // times the row scale (a power of two)
    int ks;
    frexpf(rs[idx], &ks);
    e += ks - 1;
    seed[idx] = make_double2(pm, (double)e);
}

// analysis seeds: for each (mc, theta_h), the FP64 recurrence from P_mm
// (src/plms.f90:55-97) with a separate power-of-two exponent, times the
// quadrature weight 2 pi w(theta) (src/shtransforms.f90:75-76); at every
// l = m + 32 s <= l_max the pair (P_{l-1}, P_l) is stored with its exponent
__global__ void k_rec_ana_seed(int l_max, int minc, int n_m, int nh, const double *__restrict__ theta,
                               const double *__restrict__ gauss, const size_t *__restrict__ segoff,
                               double2 *__restrict__ seed, int *__restrict__ sexp)
{
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n_m * nh) return;
    int mc = idx / nh, th = idx - mc * nh;
    int m = mc * minc;
    const double t = theta[th];
    const double st = sin(t), ct = cos(t);
    double fac = 1.0;
    for (int j = 3; j <= 2 * m + 1; j += 2) fac *= (double)j / (double)(j - 1);
    int k;
    const double f = frexp(st, &k);
    double q = 0.28209479177387814 * sqrt(fac) * 6.283185307179586 * gauss[th];
    long e = (long)k * m;
    for (int left = m; left > 0;) {
        const int n = left < 256 ? left : 256;
        int k2;
        q = frexp(q * pow(f, (double)n), &k2);
        e += k2;
        left -= n;
    }
    double pv = 0.0, eps = 0.0;
    const double dm = (double)m;
    for (int l = m; l <= l_max; ++l) {
        if (((l - m) & 31) == 0) {
            const size_t o = (segoff[mc] + (size_t)((l - m) >> 5)) * nh + th;
            seed[o] = make_double2(pv, q);
            sexp[o] = (int)(e < -2000000000L ? -2000000000L : e);
        }
        const double dl = (double)(l + 1);
        const double num = (dl - dm) * (dl + dm), den = (2.0 * dl - 1.0) * (2.0 * dl + 1.0);
        const double pn = sqrt(den / num) * (ct * q - eps * pv);
        pv = q;
        q = pn;
        eps = sqrt(num / den);
        const double v = fabs(q) + fabs(pv);
        if (v != 0.0 && v < 0x1p-256) {
            q *= 0x1p256;
            pv *= 0x1p256;
            e -= 256;
        } else if (v > 0x1p256) {
            q *= 0x1p-256;
            pv *= 0x1p-256;
            e += 256;
        }
    }
}

// c_l per (m, l <= l_max + 2) for the dP identity of the split kernels
static int crec_table(Plan *p)
{
    const size_t ncr = (size_t)p->lstart[p->n_m - 1] + 2 * (p->n_m - 1) +
                       (p->l_max + 3 - (p->n_m - 1) * p->minc);
    MHX_HIP_CHECK(hipMalloc(&p->d_crec, ncr * sizeof(float)));
    dim3 g((p->l_max + 3 + 127) / 128, p->n_m);
    k_crec<<<g, 128>>>(p->l_max, p->minc, p->n_m, p->d_lstart, p->d_crec);
    MHX_HIP_CHECK(hipGetLastError());
    return MHX_OK;
}

int build_split_tables(Plan *p)
{
    const int nh = p->nh;
    p->nq32 = (nh + 31) / 32;
    std::vector<size_t> off16(2 * p->n_m), aoff16(2 * p->n_m), soff(2 * p->n_m + 1);
    size_t so = 0, ao = 0, sr = 0;
    for (int mc = 0; mc < p->n_m; ++mc) {
        int m = mc * p->minc;
        int K[2] = {(p->l_max - m) / 2 + 1, (p->l_max - m + 1) / 2};
        int Kt0 = (K[0] + 15) / 16;
        for (int par = 0; par < 2; ++par) {
            off16[2 * mc + par] = so;
            so += (size_t)((K[par] + 31) / 32) * p->nt16 * 64 * 16;
            aoff16[2 * mc + par] = ao;
            ao += (size_t)Kt0 * p->nq32 * 64 * 16;
            soff[2 * mc + par] = sr;
            sr += (size_t)Kt0 * 16;
        }
    }
    soff[2 * p->n_m] = sr;
    p->syn16_size = so;
    p->ana16_size = ao;
    if (!p->rec) {
        MHX_HIP_CHECK(hipMalloc(&p->d_syn16_P, so * 2));
        MHX_HIP_CHECK(hipMalloc(&p->d_syn16_dP, so * 2));
        MHX_HIP_CHECK(hipMemset(p->d_syn16_P, 0, so * 2));
        MHX_HIP_CHECK(hipMemset(p->d_syn16_dP, 0, so * 2));
    }
    if (!p->rec) {
        MHX_HIP_CHECK(hipMalloc(&p->d_ana16_P, ao * 2));
        MHX_HIP_CHECK(hipMalloc(&p->d_ana16_dP, ao * 2));
        MHX_HIP_CHECK(hipMemset(p->d_ana16_P, 0, ao * 2));
        MHX_HIP_CHECK(hipMemset(p->d_ana16_dP, 0, ao * 2));
    }
    MHX_HIP_CHECK(hipMalloc(&p->d_syn16_off, off16.size() * sizeof(size_t)));
    MHX_HIP_CHECK(hipMalloc(&p->d_ana16_off, aoff16.size() * sizeof(size_t)));
    MHX_HIP_CHECK(hipMalloc(&p->d_ana16_soff, soff.size() * sizeof(size_t)));
    MHX_HIP_CHECK(hipMemcpy(p->d_syn16_off, off16.data(), off16.size() * sizeof(size_t), hipMemcpyHostToDevice));
    MHX_HIP_CHECK(hipMemcpy(p->d_ana16_off, aoff16.data(), aoff16.size() * sizeof(size_t), hipMemcpyHostToDevice));
    MHX_HIP_CHECK(hipMemcpy(p->d_ana16_soff, soff.data(), soff.size() * sizeof(size_t), hipMemcpyHostToDevice));
    MHX_HIP_CHECK(hipMalloc(&p->d_syn16_rs, (size_t)p->n_m * nh * sizeof(float)));
    MHX_HIP_CHECK(hipMalloc(&p->d_ana16_rs, sr * sizeof(float)));
    int n = p->n_m * nh;
    // plans without FP64 tables (recurrence tables, l_max > 255) take their
    // row factors from the recurrence and have no P-only tables (those serve
    // l_max <= 255 only)
    const bool have64 = p->d_syn_P != nullptr;
    int rc;
    if (have64)
        k_rowscale_syn<<<(n + 127) / 128, 128>>>(p->l_max, p->minc, p->n_m, nh, p->nt16, p->d_syn_P,
                                                 p->d_syn_dP, p->d_syn_off, p->d_syn16_rs);
    else if ((rc = split_rowscales_rec(p, p->d_syn16_rs, p->d_ana16_rs, p->d_ana16_soff)))
        return rc;
    if (p->rec) {
        // no stored synthesis tables: FP64 seeds for the in-kernel recurrence
        MHX_HIP_CHECK(hipMalloc(&p->d_rec_seed, (size_t)n * sizeof(double2)));
        MHX_HIP_CHECK(hipMalloc(&p->d_cost, (size_t)nh * sizeof(double)));
        k_rec_seed<<<(n + 127) / 128, 128>>>(p->minc, p->n_m, nh, p->d_theta, p->d_syn16_rs,
                                             p->d_rec_seed, p->d_cost);
    } else {
        k_split_syn<<<(n + 127) / 128, 128>>>(p->l_max, p->minc, p->n_m, nh, p->nt16, p->d_syn_P,
                                              p->d_syn_dP, p->d_syn_off, p->d_syn16_off, p->d_syn16_rs,
                                              p->d_syn16_P, p->d_syn16_dP);
    }
    if (have64)
        k_rowscale_ana<<<(int)((sr + 127) / 128), 128>>>(p->l_max, p->minc, p->n_m, nh, p->nq4,
                                                         p->d_ana_wP, p->d_ana_wdP, p->d_ana_off,
                                                         p->d_ana16_soff, p->d_ana16_rs, (int)sr);
    if (p->rec) {
        std::vector<size_t> segoff(p->n_m + 1);
        size_t ns = 0;
        for (int mc = 0; mc < p->n_m; ++mc) {
            segoff[mc] = ns;
            ns += (size_t)((p->l_max - mc * p->minc) / 32 + 1);
        }
        segoff[p->n_m] = ns;
        MHX_HIP_CHECK(hipMalloc(&p->d_rec_segoff, segoff.size() * sizeof(size_t)));
        MHX_HIP_CHECK(hipMemcpy(p->d_rec_segoff, segoff.data(), segoff.size() * sizeof(size_t),
                                hipMemcpyHostToDevice));
        MHX_HIP_CHECK(hipMalloc(&p->d_rec_aseed, ns * nh * sizeof(double2)));
        MHX_HIP_CHECK(hipMalloc(&p->d_rec_aexp, ns * nh * sizeof(int)));
        k_rec_ana_seed<<<(n + 127) / 128, 128>>>(p->l_max, p->minc, p->n_m, nh, p->d_theta, p->d_gauss,
                                                 p->d_rec_segoff, p->d_rec_aseed, p->d_rec_aexp);
    } else {
        k_split_ana<<<(n + 127) / 128, 128>>>(p->l_max, p->minc, p->n_m, nh, p->nq4, p->nq32,
                                              p->d_ana_wP, p->d_ana_wdP, p->d_ana_off, p->d_ana16_off,
                                              p->d_ana16_soff, p->d_ana16_rs, p->d_ana16_P,
                                              p->d_ana16_dP);
    }
    MHX_HIP_CHECK(hipGetLastError());

    if (!have64) {
        crec_table(p);
        MHX_HIP_CHECK(hipGetLastError());
        MHX_HIP_CHECK(hipDeviceSynchronize());
        return MHX_OK;
    }
    // P-only tables with l <= l_max + 1
    std::vector<size_t> offx(2 * p->n_m), aoffx(2 * p->n_m), soffx(2 * p->n_m + 1);
    size_t sox = 0, aox = 0, srx = 0;
    for (int mc = 0; mc < p->n_m; ++mc) {
        int m = mc * p->minc;
        int K[2] = {(p->l_max + 1 - m) / 2 + 1, (p->l_max + 2 - m) / 2};
        int Kt0 = (K[0] + 15) / 16;
        for (int par = 0; par < 2; ++par) {
            offx[2 * mc + par] = sox;
            sox += (size_t)((K[par] + 31) / 32) * p->nt16 * 64 * 16;
            aoffx[2 * mc + par] = aox;
            aox += (size_t)Kt0 * p->nq32 * 64 * 16;
            soffx[2 * mc + par] = srx;
            srx += (size_t)Kt0 * 16;
        }
    }
    soffx[2 * p->n_m] = srx;
    MHX_HIP_CHECK(hipMalloc(&p->d_syn16x, sox * 2));
    MHX_HIP_CHECK(hipMalloc(&p->d_ana16x, aox * 2));
    MHX_HIP_CHECK(hipMemset(p->d_syn16x, 0, sox * 2));
    MHX_HIP_CHECK(hipMemset(p->d_ana16x, 0, aox * 2));
    MHX_HIP_CHECK(hipMalloc(&p->d_syn16x_off, offx.size() * sizeof(size_t)));
    MHX_HIP_CHECK(hipMalloc(&p->d_ana16x_off, aoffx.size() * sizeof(size_t)));
    MHX_HIP_CHECK(hipMalloc(&p->d_ana16x_soff, soffx.size() * sizeof(size_t)));
    MHX_HIP_CHECK(hipMemcpy(p->d_syn16x_off, offx.data(), offx.size() * sizeof(size_t), hipMemcpyHostToDevice));
    MHX_HIP_CHECK(hipMemcpy(p->d_ana16x_off, aoffx.data(), aoffx.size() * sizeof(size_t), hipMemcpyHostToDevice));
    MHX_HIP_CHECK(hipMemcpy(p->d_ana16x_soff, soffx.data(), soffx.size() * sizeof(size_t), hipMemcpyHostToDevice));
    MHX_HIP_CHECK(hipMalloc(&p->d_syn16x_rs, (size_t)p->n_m * nh * sizeof(float)));
    MHX_HIP_CHECK(hipMalloc(&p->d_ana16x_rs, srx * sizeof(float)));
    k_split_synx<<<(n + 127) / 128, 128>>>(p->l_max, p->minc, p->n_m, nh, p->nt16, p->d_syn_P,
                                           p->d_syn_off, p->d_syn16x_off, p->d_theta, p->d_syn16x_rs,
                                           p->d_syn16x);
    k_rowscale_anax<<<(int)((srx + 127) / 128), 128>>>(p->l_max, p->minc, p->n_m, nh, p->nq4,
                                                       p->d_ana_wP, p->d_ana_off, p->d_ana16x_soff,
                                                       p->d_theta, p->d_ana16x_rs, (int)srx);
    k_split_anax<<<(n + 127) / 128, 128>>>(p->l_max, p->minc, p->n_m, nh, p->nq4, p->nq32,
                                           p->d_ana_wP, p->d_ana_off, p->d_ana16x_off,
                                           p->d_ana16x_soff, p->d_theta, p->d_ana16x_rs, p->d_ana16x);
    if ((rc = crec_table(p))) return rc;
    MHX_HIP_CHECK(hipGetLastError());
    MHX_HIP_CHECK(hipDeviceSynchronize());
    return MHX_OK;
}

}  // namespace mhx
