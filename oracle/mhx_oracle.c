/*
 * ORACLE - test infrastructure only (see mhx_oracle.h).
 *
 * FP64 C restatement of MagIC's native transform path.  Each routine cites
 * the reference lines it follows.  Pinned by tests/test_oracle.py against
 * golden vectors written by the reference itself (oracle/_ref/ref_harness),
 * against numpy's Gauss-Legendre nodes and real FFT, and against closed
 * forms (low-degree P_lm, orthonormality, round trips).
 */
#include "mhx_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef double _Complex cplx;

/* pi = 4 atan(1) in double, as src/constants.f90:17 */
#define ORC_PI 3.14159265358979323846

struct orc_plan {
    int l_max, m_max, minc, n_theta, n_phi, nh, n_m_max, lm_max;
    double *theta_ord; /* nh, northern colatitudes */
    double *gauss;     /* nh */
    double *plm, *dplm, *wplm, *wdplm; /* [nh][lm_max] */
    double *osin2;     /* n_theta, scrambled order */
    int *lstart;       /* n_m_max */
    int *lm2l, *lm2m;  /* lm_max */
};

/* ------------------------------------------------------------------ */
/* src/horizontal.f90:279-340 (interval [-1,1]) */
void orc_gauleg(int n, double *theta_ord, double *gauss)
{
    const double eps = 10.0 * 2.220446049250313e-16;
    int half = (n + 1) / 2;
    for (int i = 1; i <= half; ++i) {
        double z = cos(ORC_PI * ((double)i - 0.25) / ((double)n + 0.5));
        double z1 = z + 10.0 * eps, pp = 1.0;
        while (fabs(z - z1) > eps) {
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
        theta_ord[i - 1] = acos(z);
        theta_ord[n - i] = acos(-z);
        gauss[i - 1] = 2.0 / ((1.0 - z * z) * pp * pp);
        gauss[n - i] = gauss[i - 1];
    }
}

/* src/plms.f90:14-189, norm=2 (orthonormal, includes 1/sqrt(4 pi), no
   Condon-Shortley phase); dtheta_plm = sin(theta) dP/dtheta. */
void orc_plm_theta(double theta, int l_max, int m_min, int m_max, int minc,
                   double *plm_out, double *dplm_out)
{
    const double dnorm = 1.0 / sqrt(4.0 * ORC_PI);
    double st = sin(theta), ct = cos(theta);
    int pos = 0;
    for (int m = m_min; m <= m_max; m += minc) {
        double fac = 1.0;
        for (int j = 3; j <= 2 * m + 1; j += 2) fac *= (double)j / (double)(j - 1);
        double p = sqrt(fac);
        if (st != 0.0) p *= pow(st, (double)m);
        else if (m != 0) p = 0.0;
        plm_out[pos++] = dnorm * p;
        double p1 = 0.0, p2;
        for (int l = m + 1; l <= l_max; ++l) {
            p2 = p1;
            p1 = p;
            p = ct * sqrt((double)((2 * l - 1) * (2 * l + 1)) / (double)((l - m) * (l + m))) * p1
              - sqrt(((double)(2 * l + 1) * (double)(l + m - 1) * (double)(l - m - 1))
                     / ((double)(2 * l - 3) * (double)(l - m) * (double)(l + m))) * p2;
            plm_out[pos++] = dnorm * p;
        }
        /* degree l_max+1, parked in the derivative slot of l_max */
        int l = l_max + 1;
        p2 = p1;
        p1 = p;
        p = ct * sqrt((double)((2 * l - 1) * (2 * l + 1)) / (double)((l - m) * (l + m))) * p1
          - sqrt((double)((2 * l + 1) * (l + m - 1) * (l - m - 1))
                 / (double)((2 * l - 3) * (l - m) * (l + m))) * p2;
        dplm_out[pos - 1] = dnorm * p;
    }
    pos = 0;
    for (int m = m_min; m <= m_max; m += minc) {
        int l = m;
        if (m < l_max) dplm_out[pos] = l / sqrt((double)(2 * l + 3)) * plm_out[pos + 1];
        else           dplm_out[pos] = l / sqrt((double)(2 * l + 3)) * dplm_out[pos];
        pos++;
        for (l = m + 1; l <= l_max - 1; ++l) {
            dplm_out[pos] = l * sqrt((double)((l + m + 1) * (l - m + 1)) / (double)((2 * l + 1) * (2 * l + 3))) * plm_out[pos + 1]
                          - (l + 1) * sqrt((double)((l + m) * (l - m)) / (double)((2 * l - 1) * (2 * l + 1))) * plm_out[pos - 1];
            pos++;
        }
        if (m < l_max) {
            l = l_max;
            dplm_out[pos] = l * sqrt((double)((l + m + 1) * (l - m + 1)) / (double)((2 * l + 1) * (2 * l + 3))) * dplm_out[pos]
                          - (l + 1) * sqrt((double)((l + m) * (l - m)) / (double)((2 * l - 1) * (2 * l + 1))) * plm_out[pos - 1];
            pos++;
        }
    }
}

/* ------------------------------------------------------------------ */
orc_plan *orc_plan_create(int l_max, int m_max, int minc, int n_theta, int n_phi)
{
    orc_plan *p = (orc_plan *)calloc(1, sizeof(orc_plan));
    if (m_max > l_max) m_max = l_max;
    p->l_max = l_max; p->m_max = m_max; p->minc = minc;
    p->n_theta = n_theta; p->n_phi = n_phi; p->nh = n_theta / 2;
    p->n_m_max = m_max / minc + 1;
    int lm = 0;
    for (int m = 0; m <= m_max; m += minc) lm += l_max - m + 1;
    p->lm_max = lm;
    p->lstart = (int *)malloc(sizeof(int) * p->n_m_max);
    p->lm2l = (int *)malloc(sizeof(int) * lm);
    p->lm2m = (int *)malloc(sizeof(int) * lm);
    lm = 0;
    for (int mc = 0; mc < p->n_m_max; ++mc) {
        int m = mc * minc;
        p->lstart[mc] = lm;
        for (int l = m; l <= l_max; ++l) { p->lm2l[lm] = l; p->lm2m[lm] = m; lm++; }
    }
    double *th = (double *)malloc(sizeof(double) * n_theta);
    double *gw = (double *)malloc(sizeof(double) * n_theta);
    orc_gauleg(n_theta, th, gw);
    p->theta_ord = (double *)malloc(sizeof(double) * p->nh);
    p->gauss = (double *)malloc(sizeof(double) * p->nh);
    p->osin2 = (double *)malloc(sizeof(double) * n_theta);
    size_t tsz = (size_t)p->nh * p->lm_max;
    p->plm = (double *)malloc(sizeof(double) * tsz);
    p->dplm = (double *)malloc(sizeof(double) * tsz);
    p->wplm = (double *)malloc(sizeof(double) * tsz);
    p->wdplm = (double *)malloc(sizeof(double) * tsz);
    /* src/shtransforms.f90:67-77, src/horizontal.f90:141-146 */
    for (int k = 0; k < p->nh; ++k) {
        p->theta_ord[k] = th[k];
        p->gauss[k] = gw[k];
        double s = sin(th[k]);
        p->osin2[2 * k] = 1.0 / (s * s);
        p->osin2[2 * k + 1] = 1.0 / (s * s);
        double *P = p->plm + (size_t)k * p->lm_max;
        double *D = p->dplm + (size_t)k * p->lm_max;
        orc_plm_theta(th[k], l_max, 0, m_max, minc, P, D);
        for (int i = 0; i < p->lm_max; ++i) {
            p->wplm[(size_t)k * p->lm_max + i] = 2.0 * ORC_PI * gw[k] * P[i];
            p->wdplm[(size_t)k * p->lm_max + i] = 2.0 * ORC_PI * gw[k] * D[i];
        }
    }
    free(th);
    free(gw);
    return p;
}

void orc_plan_destroy(orc_plan *p)
{
    if (!p) return;
    free(p->theta_ord); free(p->gauss); free(p->plm); free(p->dplm);
    free(p->wplm); free(p->wdplm); free(p->osin2); free(p->lstart);
    free(p->lm2l); free(p->lm2m); free(p);
}

int orc_plan_lm_max(const orc_plan *p) { return p->lm_max; }
int orc_plan_n_m_max(const orc_plan *p) { return p->n_m_max; }

void orc_plan_tables(const orc_plan *p, double *theta_ord, double *gauss,
                     double *plm, double *dplm, double *wplm, double *wdplm,
                     double *osin2)
{
    size_t tsz = (size_t)p->nh * p->lm_max;
    if (theta_ord) memcpy(theta_ord, p->theta_ord, sizeof(double) * p->nh);
    if (gauss) memcpy(gauss, p->gauss, sizeof(double) * p->nh);
    if (plm) memcpy(plm, p->plm, sizeof(double) * tsz);
    if (dplm) memcpy(dplm, p->dplm, sizeof(double) * tsz);
    if (wplm) memcpy(wplm, p->wplm, sizeof(double) * tsz);
    if (wdplm) memcpy(wdplm, p->wdplm, sizeof(double) * tsz);
    if (osin2) memcpy(osin2, p->osin2, sizeof(double) * p->n_theta);
}

/* ------------------------------------------------------------------ */
/* generic mixed-radix complex DFT (recursive decimation in time) */
static void cfft_rec(int n, const cplx *in, int istride, cplx *out, int sign, cplx *tmp)
{
    if (n == 1) { out[0] = in[0]; return; }
    int r = 2;
    while (n % r) r++;
    int m = n / r;
    for (int s = 0; s < r; ++s) cfft_rec(m, in + (size_t)s * istride, istride * r, out + (size_t)s * m, sign, tmp);
    for (int k = 0; k < n; ++k) tmp[k] = out[k];
    for (int k = 0; k < m; ++k) {
        for (int q = 0; q < r; ++q) {
            int kk = k + q * m;
            cplx acc = 0.0;
            for (int s = 0; s < r; ++s) {
                long e = ((long)s * kk) % n;
                double ang = sign * 2.0 * ORC_PI * (double)e / (double)n;
                acc += tmp[s * m + k] * (cos(ang) + I * sin(ang));
            }
            out[kk] = acc;
        }
    }
}

static void cfft(int n, const cplx *in, cplx *out, int sign)
{
    cplx *tmp = (cplx *)malloc(sizeof(cplx) * n);
    cfft_rec(n, in, 1, out, sign, tmp);
    free(tmp);
}

/* src/fft.f90:164-209 with fft991(isign=-1): a_k=(1/n) sum x cos,
   b_k=-(1/n) sum x sin, b_0 = b_{n/2} = 0 */
void orc_fft_forward(const orc_plan *p, const double *grid, double *Fout)
{
    int n = p->n_phi, nt = p->n_theta, nk = n / 2 + 1;
    cplx *F = (cplx *)Fout;
    /* rows are independent: one thread per row, same arithmetic as serial */
#pragma omp parallel
    {
        cplx *x = (cplx *)malloc(sizeof(cplx) * n), *X = (cplx *)malloc(sizeof(cplx) * n);
#pragma omp for schedule(static)
        for (int t = 0; t < nt; ++t) {
            for (int j = 0; j < n; ++j) x[j] = grid[(size_t)j * nt + t];
            cfft(n, x, X, -1);
            for (int k = 0; k < nk; ++k) F[(size_t)k * nt + t] = X[k] / (double)n;
            F[t] = creal(F[t]);
            F[(size_t)(nk - 1) * nt + t] = creal(F[(size_t)(nk - 1) * nt + t]);
        }
        free(x); free(X);
    }
}

/* src/fft.f90:211-252 with fft991(isign=+1): x_j = sum_k c_k e^{2 pi i jk/n},
   c_{n-k} = conj(c_k); only the real parts of c_0 and c_{n/2} enter */
void orc_fft_inverse(const orc_plan *p, const double *Fin, double *grid)
{
    int n = p->n_phi, nt = p->n_theta, nk = n / 2 + 1;
    const cplx *F = (const cplx *)Fin;
#pragma omp parallel
    {
        cplx *X = (cplx *)malloc(sizeof(cplx) * n), *x = (cplx *)malloc(sizeof(cplx) * n);
#pragma omp for schedule(static)
        for (int t = 0; t < nt; ++t) {
            X[0] = creal(F[t]);
            X[n / 2] = creal(F[(size_t)(nk - 1) * nt + t]);
            for (int k = 1; k < n / 2; ++k) {
                X[k] = F[(size_t)k * nt + t];
                X[n - k] = conj(X[k]);
            }
            cfft(n, X, x, +1);
            for (int j = 0; j < n; ++j) grid[(size_t)j * nt + t] = creal(x[j]);
        }
        free(X); free(x);
    }
}

/* ------------------------------------------------------------------ */
#define PLM(p, k, lm) ((p)->plm[(size_t)(k) * (p)->lm_max + (lm)])
#define DPLM(p, k, lm) ((p)->dplm[(size_t)(k) * (p)->lm_max + (lm)])
#define WPLM(p, k, lm) ((p)->wplm[(size_t)(k) * (p)->lm_max + (lm)])
#define WDPLM(p, k, lm) ((p)->wdplm[(size_t)(k) * (p)->lm_max + (lm)])

static cplx *fourier_alloc(const orc_plan *p)
{
    return (cplx *)calloc((size_t)(p->n_phi / 2 + 1) * p->n_theta, sizeof(cplx));
}

/* src/shtransforms.f90:493-561 */
void orc_sph_to_spat(const orc_plan *p, const double *Slm_, double *sc, int lcut)
{
    const cplx *S = (const cplx *)Slm_;
    cplx *F = fourier_alloc(p);
    int nt = p->n_theta;
#pragma omp parallel for schedule(dynamic)
    for (int mc = 0; mc < p->n_m_max; ++mc) {
        int m = mc * p->minc;
        if (m > lcut) continue;
        int l0 = p->lstart[mc], lmS = l0 + lcut - m;
        for (int k = 0; k < p->nh; ++k) {
            cplx es = 0.0, ea = 0.0;
            for (int lm = l0; lm <= lmS; ++lm) {
                if (((lm - l0) & 1) == 0) es += S[lm] * PLM(p, k, lm);
                else                      ea += S[lm] * PLM(p, k, lm);
            }
            F[(size_t)mc * nt + 2 * k] = es + ea;
            F[(size_t)mc * nt + 2 * k + 1] = es - ea;
        }
    }
    orc_fft_inverse(p, (const double *)F, sc);
    free(F);
}

/* src/shtransforms.f90:563-646 */
void orc_sph_to_grad_spat(const orc_plan *p, const double *Slm_, double *gt, double *gp, int lcut)
{
    const cplx *S = (const cplx *)Slm_;
    cplx *Ft = fourier_alloc(p), *Fp = fourier_alloc(p);
    int nt = p->n_theta;
#pragma omp parallel for schedule(dynamic)
    for (int mc = 0; mc < p->n_m_max; ++mc) {
        int m = mc * p->minc;
        if (m > lcut) continue;
        double dm = (double)m;
        int l0 = p->lstart[mc], lmS = l0 + lcut - m;
        for (int k = 0; k < p->nh; ++k) {
            cplx ses = 0.0, sea = 0.0, ges = 0.0, gea = 0.0;
            for (int lm = l0; lm <= lmS; ++lm) {
                if (((lm - l0) & 1) == 0) { gea += S[lm] * DPLM(p, k, lm); ses += S[lm] * PLM(p, k, lm); }
                else                      { ges += S[lm] * DPLM(p, k, lm); sea += S[lm] * PLM(p, k, lm); }
            }
            Ft[(size_t)mc * nt + 2 * k] = ges + gea;
            Ft[(size_t)mc * nt + 2 * k + 1] = ges - gea;
            Fp[(size_t)mc * nt + 2 * k] = dm * I * (ses + sea);
            Fp[(size_t)mc * nt + 2 * k + 1] = dm * I * (ses - sea);
        }
    }
    orc_fft_inverse(p, (const double *)Ft, gt);
    orc_fft_inverse(p, (const double *)Fp, gp);
    free(Ft); free(Fp);
}

/* horizontal part shared by qst/sphtor: src/shtransforms.f90:191-219 */
static void sphtor_fourier(const orc_plan *p, const cplx *S, const cplx *T, cplx *Ft, cplx *Fp,
                           int lcut, int zero_l0)
{
    int nt = p->n_theta;
    cplx *bhG = (cplx *)malloc(sizeof(cplx) * p->lm_max);
    cplx *bhC = (cplx *)malloc(sizeof(cplx) * p->lm_max);
    double *PG = (double *)malloc(sizeof(double) * p->lm_max);
    double *PC = (double *)malloc(sizeof(double) * p->lm_max);
    for (int lm = 0; lm < p->lm_max; ++lm) { bhG[lm] = S[lm] - I * T[lm]; bhC[lm] = S[lm] + I * T[lm]; }
    if (zero_l0) { bhG[0] = 0.0; bhC[0] = 0.0; }
#pragma omp parallel for schedule(dynamic)
    for (int mc = 0; mc < p->n_m_max; ++mc) {
        int m = mc * p->minc;
        if (m > lcut) continue;
        double dm = (double)m;
        int l0 = p->lstart[mc], lmS = l0 + lcut - m;
        for (int k = 0; k < p->nh; ++k) {
            for (int lm = l0; lm <= lmS; ++lm) {
                PG[lm] = DPLM(p, k, lm) - dm * PLM(p, k, lm);
                PC[lm] = DPLM(p, k, lm) + dm * PLM(p, k, lm);
            }
            cplx n1 = 0.0, s1 = 0.0, n2 = 0.0, s2 = 0.0;
            for (int lm = l0; lm <= lmS; ++lm) {
                n1 += bhG[lm] * PG[lm];
                n2 += bhC[lm] * PC[lm];
                if (((lm - l0) & 1) == 0) { s1 -= bhG[lm] * PC[lm]; s2 -= bhC[lm] * PG[lm]; }
                else                      { s1 += bhG[lm] * PC[lm]; s2 += bhC[lm] * PG[lm]; }
            }
            n1 *= 0.5; s1 *= 0.5; n2 *= 0.5; s2 *= 0.5;
            Ft[(size_t)mc * nt + 2 * k] = n1 + n2;
            Ft[(size_t)mc * nt + 2 * k + 1] = s1 + s2;
            Fp[(size_t)mc * nt + 2 * k] = -I * (n1 - n2);
            Fp[(size_t)mc * nt + 2 * k + 1] = -I * (s1 - s2);
        }
    }
    free(bhG); free(bhC); free(PG); free(PC);
}

/* src/shtransforms.f90:103-245 */
void orc_qst_to_spat(const orc_plan *p, const double *Q, const double *S, const double *T,
                     double *vr, double *vt, double *vp, int lcut)
{
    cplx *Ft = fourier_alloc(p), *Fp = fourier_alloc(p);
    orc_sph_to_spat(p, Q, vr, lcut);
    sphtor_fourier(p, (const cplx *)S, (const cplx *)T, Ft, Fp, lcut, 1);
    orc_fft_inverse(p, (const double *)Ft, vt);
    orc_fft_inverse(p, (const double *)Fp, vp);
    free(Ft); free(Fp);
}

/* src/shtransforms.f90:247-372 */
void orc_sphtor_to_spat(const orc_plan *p, const double *S, const double *T,
                        double *vt, double *vp, int lcut)
{
    cplx *Ft = fourier_alloc(p), *Fp = fourier_alloc(p);
    sphtor_fourier(p, (const cplx *)S, (const cplx *)T, Ft, Fp, lcut, 0);
    orc_fft_inverse(p, (const double *)Ft, vt);
    orc_fft_inverse(p, (const double *)Fp, vp);
    free(Ft); free(Fp);
}

/* src/shtransforms.f90:374-403 (m=0 only, full l_max) */
void orc_axi_to_spat(const orc_plan *p, const double *Sl_, double *sc)
{
    const cplx *S = (const cplx *)Sl_;
    for (int k = 0; k < p->nh; ++k) {
        cplx es = 0.0, ea = 0.0;
        for (int l = 0; l <= p->l_max; ++l) {
            if ((l & 1) == 0) es += S[l] * PLM(p, k, l);
            else              ea += S[l] * PLM(p, k, l);
        }
        sc[2 * k] = creal(es + ea);
        sc[2 * k + 1] = creal(es - ea);
    }
}

/* src/shtransforms.f90:405-491 */
void orc_toraxi_to_spat(const orc_plan *p, const double *Tl_, double *vt, double *vp, int lcut)
{
    const cplx *T = (const cplx *)Tl_;
    for (int k = 0; k < p->nh; ++k) {
        cplx n1 = 0.0, s1 = 0.0, n2 = 0.0, s2 = 0.0;
        for (int l = 0; l <= p->l_max; ++l) {
            cplx g = (l > lcut) ? 0.0 : -I * T[l];
            cplx c = (l > lcut) ? 0.0 : I * T[l];
            double d = DPLM(p, k, l);
            n1 += g * d;
            n2 += c * d;
            if ((l & 1) == 0) { s1 -= g * d; s2 -= c * d; }
            else              { s1 += g * d; s2 += c * d; }
        }
        n1 *= 0.5; s1 *= 0.5; n2 *= 0.5; s2 *= 0.5;
        vt[2 * k] = creal(n1 + n2);
        vt[2 * k + 1] = creal(s1 + s2);
        vp[2 * k] = creal(-I * (n1 - n2));
        vp[2 * k + 1] = creal(-I * (s1 - s2));
    }
}

/* src/shtransforms.f90:648-729 */
void orc_spat_to_sph(const orc_plan *p, const double *f, double *flm_, int lcut)
{
    cplx *flm = (cplx *)flm_;
    cplx *F = fourier_alloc(p);
    int nt = p->n_theta;
    orc_fft_forward(p, f, (double *)F);
    memset(flm, 0, sizeof(cplx) * p->lm_max);
    int mlim = p->m_max < lcut ? p->m_max : lcut;
#pragma omp parallel for schedule(dynamic)
    for (int mc = 0; mc < p->n_m_max; ++mc) {
        int m = mc * p->minc;
        if (m > mlim) continue;
        int l0 = p->lstart[mc], lmS = l0 + lcut - m;
        for (int k = 0; k < p->nh; ++k) {
            cplx es = F[(size_t)mc * nt + 2 * k] + F[(size_t)mc * nt + 2 * k + 1];
            cplx ea = F[(size_t)mc * nt + 2 * k] - F[(size_t)mc * nt + 2 * k + 1];
            for (int lm = l0; lm <= lmS; ++lm)
                flm[lm] += (((lm - l0) & 1) == 0 ? es : ea) * WPLM(p, k, lm);
        }
    }
    free(F);
}

/* src/shtransforms.f90:731-872 */
void orc_spat_to_sph_tor(const orc_plan *p, const double *vt, const double *vp,
                         double *f1_, double *f2_, int lcut)
{
    cplx *f1 = (cplx *)f1_, *f2 = (cplx *)f2_;
    cplx *F1 = fourier_alloc(p), *F2 = fourier_alloc(p);
    int nt = p->n_theta, nk = p->n_phi / 2 + 1;
    orc_fft_forward(p, vt, (double *)F2);
    orc_fft_forward(p, vp, (double *)F1);
    for (int mc = 0; mc < nk; ++mc)
        for (int t = 0; t < nt; ++t) {
            F1[(size_t)mc * nt + t] *= p->osin2[t];
            F2[(size_t)mc * nt + t] *= p->osin2[t];
        }
    memset(f1, 0, sizeof(cplx) * p->lm_max);
    memset(f2, 0, sizeof(cplx) * p->lm_max);
    int mlim = p->m_max < lcut ? p->m_max : lcut;
#pragma omp parallel for schedule(dynamic)
    for (int mc = 0; mc < p->n_m_max; ++mc) {
        int m = mc * p->minc;
        if (m > mlim) continue;
        double dm = (double)m;
        int l0 = p->lstart[mc], lmS = l0 + lcut - m;
        for (int k = 0; k < p->nh; ++k) {
            cplx a = F1[(size_t)mc * nt + 2 * k], b = F1[(size_t)mc * nt + 2 * k + 1];
            cplx c = F2[(size_t)mc * nt + 2 * k], d = F2[(size_t)mc * nt + 2 * k + 1];
            cplx f1es = a + b, f1ea = a - b, f2es = c + d, f2ea = c - d;
            for (int lm = l0; lm <= lmS; ++lm) {
                double w = WPLM(p, k, lm), wd = WDPLM(p, k, lm);
                if (((lm - l0) & 1) == 0) {
                    f1[lm] += -I * dm * f1es * w + f2ea * wd;
                    f2[lm] += -f1ea * wd - I * dm * f2es * w;
                } else {
                    f1[lm] += -I * dm * f1ea * w + f2es * wd;
                    f2[lm] += -f1es * wd - I * dm * f2ea * w;
                }
            }
        }
    }
    for (int lm = 1; lm < p->lm_max; ++lm) {
        double l = (double)p->lm2l[lm];
        f1[lm] /= l * (l + 1.0);
        f2[lm] /= l * (l + 1.0);
    }
    free(F1); free(F2);
}

/* ------------------------------------------------------------------ */
/* src/sht_native.f90 wrappers */
static cplx *cbuf(const orc_plan *p) { return (cplx *)calloc(p->lm_max, sizeof(cplx)); }

void orc_scal_to_spat(const orc_plan *p, const double *S, double *f, int lcut)
{ orc_sph_to_spat(p, S, f, lcut); }

void orc_scal_to_grad_spat(const orc_plan *p, const double *S, double *gt, double *gp, int lcut)
{ orc_sph_to_grad_spat(p, S, gt, gp, lcut); }

/* :70-97 */
void orc_pol_to_grad_spat(const orc_plan *p, const double *S_, double *gt, double *gp, int lcut)
{
    const cplx *S = (const cplx *)S_;
    cplx *Q = cbuf(p);
    for (int lm = 0; lm < p->lm_max; ++lm) {
        int l = p->lm2l[lm];
        Q[lm] = (l <= lcut) ? (double)(l * (l + 1)) * S[lm] : 0.0;
    }
    orc_sph_to_grad_spat(p, (const double *)Q, gt, gp, lcut);
    free(Q);
}

/* :99-127 */
void orc_torpol_to_spat(const orc_plan *p, const double *W_, const double *dW, const double *Z,
                        double *vr, double *vt, double *vp, int lcut)
{
    const cplx *W = (const cplx *)W_;
    cplx *Q = cbuf(p);
    for (int lm = 0; lm < p->lm_max; ++lm) {
        int l = p->lm2l[lm];
        Q[lm] = (l <= lcut) ? (double)(l * (l + 1)) * W[lm] : 0.0;
    }
    orc_qst_to_spat(p, (const double *)Q, dW, Z, vr, vt, vp, lcut);
    free(Q);
}

/* :274-300 */
void orc_pol_to_curlr_spat(const orc_plan *p, const double *Q_, double *cvr, int lcut)
{
    const cplx *Qi = (const cplx *)Q_;
    cplx *Q = cbuf(p);
    for (int lm = 0; lm < p->lm_max; ++lm) {
        int l = p->lm2l[lm];
        Q[lm] = (l <= lcut) ? (double)(l * (l + 1)) * Qi[lm] : 0.0;
    }
    orc_sph_to_spat(p, (const double *)Q, cvr, lcut);
    free(Q);
}

/* :302-336 */
void orc_torpol_to_curl_spat(const orc_plan *p, double or2, const double *B_, const double *ddB_,
                             const double *J_, const double *dJ, double *cvr, double *cvt,
                             double *cvp, int lcut)
{
    const cplx *B = (const cplx *)B_, *ddB = (const cplx *)ddB_, *J = (const cplx *)J_;
    cplx *Q = cbuf(p), *T = cbuf(p);
    for (int lm = 0; lm < p->lm_max; ++lm) {
        int l = p->lm2l[lm];
        double dLh = (double)(l * (l + 1));
        if (l <= lcut) { Q[lm] = dLh * J[lm]; T[lm] = or2 * dLh * B[lm] - ddB[lm]; }
    }
    orc_qst_to_spat(p, (const double *)Q, dJ, (const double *)T, cvr, cvt, cvp, lcut);
    free(Q); free(T);
}

/* :231-272 */
void orc_torpol_to_dphspat(const orc_plan *p, const double *dW_, const double *Z_,
                           double *dvt, double *dvp, int lcut)
{
    const cplx *dW = (const cplx *)dW_, *Z = (const cplx *)Z_;
    cplx *S = cbuf(p), *T = cbuf(p);
    for (int lm = 0; lm < p->lm_max; ++lm) {
        int l = p->lm2l[lm];
        double m = (double)p->lm2m[lm];
        if (l <= lcut) { S[lm] = I * m * dW[lm]; T[lm] = I * m * Z[lm]; }
    }
    orc_sphtor_to_spat(p, (const double *)S, (const double *)T, dvt, dvp, lcut);
    int nt = p->n_theta;
    for (int j = 0; j < p->n_phi; ++j)
        for (int t = 0; t < nt; ++t) {
            dvt[(size_t)j * nt + t] *= p->osin2[t];
            dvp[(size_t)j * nt + t] *= p->osin2[t];
        }
    free(S); free(T);
}

/* :188-229 */
void orc_torpol_to_spat_IC(const orc_plan *p, double r, double r_icb, const double *W_,
                           const double *dW_, const double *Z_, double *br, double *bt, double *bp)
{
    const cplx *W = (const cplx *)W_, *dW = (const cplx *)dW_, *Z = (const cplx *)Z_;
    cplx *Q = cbuf(p), *S = cbuf(p), *T = cbuf(p);
    double *rD = (double *)malloc(sizeof(double) * (p->l_max + 1));
    double *rD2 = (double *)malloc(sizeof(double) * (p->l_max + 1));
    double rr = r / r_icb;
    rD[0] = rr; rD2[0] = 1.0 / r_icb;
    for (int l = 1; l <= p->l_max; ++l) { rD[l] = rD[l - 1] * rr; rD2[l] = rD2[l - 1] * rr; }
    for (int lm = 0; lm < p->lm_max; ++lm) {
        int l = p->lm2l[lm];
        Q[lm] = rD[l] * (double)(l * (l + 1)) * W[lm];
        S[lm] = rD2[l] * ((double)(l + 1) * W[lm] + r * dW[lm]);
        T[lm] = rD[l] * Z[lm];
    }
    orc_qst_to_spat(p, (const double *)Q, (const double *)S, (const double *)T, br, bt, bp, p->l_max);
    free(Q); free(S); free(T); free(rD); free(rD2);
}

/* :143-186 */
void orc_torpol_to_curl_spat_IC(const orc_plan *p, double r, double r_icb, const double *dB_,
                                const double *ddB_, const double *J_, const double *dJ_,
                                double *cbr, double *cbt, double *cbp)
{
    const cplx *dB = (const cplx *)dB_, *ddB = (const cplx *)ddB_, *J = (const cplx *)J_,
               *dJ = (const cplx *)dJ_;
    cplx *Q = cbuf(p), *S = cbuf(p), *T = cbuf(p);
    double *rD = (double *)malloc(sizeof(double) * (p->l_max + 1));
    double *rD2 = (double *)malloc(sizeof(double) * (p->l_max + 1));
    double rr = r / r_icb;
    rD[0] = rr; rD2[0] = 1.0 / r_icb;
    for (int l = 1; l <= p->l_max; ++l) { rD[l] = rD[l - 1] * rr; rD2[l] = rD2[l - 1] * rr; }
    for (int lm = 0; lm < p->lm_max; ++lm) {
        int l = p->lm2l[lm];
        Q[lm] = rD[l] * (double)(l * (l + 1)) * J[lm];
        S[lm] = rD2[l] * ((double)(l + 1) * J[lm] + r * dJ[lm]);
        T[lm] = -rD2[l] * (2.0 * (double)(l + 1) * dB[lm] + r * ddB[lm]);
    }
    orc_qst_to_spat(p, (const double *)Q, (const double *)S, (const double *)T, cbr, cbt, cbp, p->l_max);
    free(Q); free(S); free(T); free(rD); free(rD2);
}

/* :338-346 */
void orc_scal_to_SH(const orc_plan *p, const double *f, double *flm, int lcut)
{ orc_spat_to_sph(p, f, flm, lcut); }

/* :348-364 */
void orc_spat_to_qst(const orc_plan *p, const double *f, const double *g, const double *h,
                     double *q, double *s, double *t, int lcut)
{
    orc_spat_to_sph(p, f, q, lcut);
    orc_spat_to_sph_tor(p, g, h, s, t, lcut);
}

/* :366-379 */
void orc_spat_to_sphertor(const orc_plan *p, const double *f, const double *g,
                          double *flm, double *glm, int lcut)
{ orc_spat_to_sph_tor(p, f, g, flm, glm, lcut); }

/* ------------------------------------------------------------------ */
/* src/get_nl.f90:240-258 (Lorentz force), :262-273 (curl advection),
   :320-325 (V S), :376-388 (V x B), bulk level, nR > n_r_LCR */
void orc_get_nl(const orc_plan *p, double or2, double or4, double orho1, double LFfac,
                const double *const *in, double *const *out)
{
    int nt = p->n_theta;
    size_t npts = (size_t)nt * p->n_phi;
    const double *vr = in[0], *vt = in[1], *vp = in[2], *cvr = in[3], *cvt = in[4], *cvp = in[5];
    const double *br = in[6], *bt = in[7], *bp = in[8], *cbr = in[9], *cbt = in[10], *cbp = in[11];
    const double *s = in[12];
    for (size_t i = 0; i < npts; ++i) {
        double o2 = p->osin2[i % nt];
        out[3][i] = LFfac * o2 * (cbt[i] * bp[i] - cbp[i] * bt[i]);
        out[4][i] = LFfac * or4 * (cbp[i] * br[i] - cbr[i] * bp[i]);
        out[5][i] = LFfac * or4 * (cbr[i] * bt[i] - cbt[i] * br[i]);
        out[0][i] = -o2 * (cvt[i] * vp[i] - cvp[i] * vt[i]);
        out[1][i] = -or4 * (cvp[i] * vr[i] - cvr[i] * vp[i]);
        out[2][i] = -or4 * (cvr[i] * vt[i] - cvt[i] * vr[i]);
        out[9][i] = vr[i] * s[i];
        out[10][i] = or2 * vt[i] * s[i];
        out[11][i] = or2 * vp[i] * s[i];
        out[6][i] = orho1 * o2 * (vt[i] * bp[i] - vp[i] * bt[i]);
        out[7][i] = orho1 * or4 * (vp[i] * br[i] - vr[i] * bp[i]);
        out[8][i] = orho1 * or4 * (vr[i] * bt[i] - vt[i] * br[i]);
    }
}

/* ------------------------------------------------------------------ */
/* lm-space assembly of the explicit terms at an interior level (nBc = 0)
   of a rotating Boussinesq MHD run without the double-curl form:
   src/get_td.f90 get_dwdt (:133-203), get_dzdt (:379-444), get_dpdt
   (:314-377), get_dsdt (:446-500, epscProf = 1), get_dbdt (:530-589);
   operators of src/horizontal.f90:201-229; neighbours l -/+ 1 at the same m
   (src/blocking.f90:322-334).  or1 = sqrt(or2), r = 1/sqrt(or2). */
static double td_clm(int l, int m)
{
    if (l < m || l == 0) return 0.0;
    return sqrt((double)(l + m) * (double)(l - m) / ((double)(2 * l - 1) * (double)(2 * l + 1)));
}

void orc_get_td(const orc_plan *p, int lR, double or2, double or4, double CorFac, double epsc,
                const double *const *in, double *const *out)
{
    const cplx *Advr = (const cplx *)in[0], *Advt = (const cplx *)in[1], *Advp = (const cplx *)in[2];
    const cplx *VSt = (const cplx *)in[3], *VxBr = (const cplx *)in[4], *VxBt = (const cplx *)in[5];
    const cplx *VxBp = (const cplx *)in[6], *w = (const cplx *)in[7], *dw = (const cplx *)in[8];
    const cplx *z = (const cplx *)in[9];
    cplx *dwdt = (cplx *)out[0], *dzdt = (cplx *)out[1], *dpdt = (cplx *)out[2];
    cplx *dsdt = (cplx *)out[3], *dbdt = (cplx *)out[4], *djdt = (cplx *)out[5];
    cplx *dVxBh = (cplx *)out[6];
    const double or1 = sqrt(or2), r = 1.0 / sqrt(or2), c2 = 2.0 * CorFac;
    for (int lm = 0; lm < p->lm_max; ++lm) {
        const int l = p->lm2l[lm], m = p->lm2m[lm];
        const double dl = (double)l;
        const double dLh = dl * (dl + 1.0);
        const cplx dPhi = I * (double)m;
        const double cS = td_clm(l, m), cA = td_clm(l + 1, m);
        const double t2S = (dl - 1.0) * cS, t2A = (dl + 2.0) * cA;
        const double t3S = (dl - 1.0) * (dl + 1.0) * cS, t3A = dl * (dl + 2.0) * cA;
        const double t4S = (dl + 1.0) * cS * (dl - 1.0) * dl, t4A = dl * cA * (dl + 1.0) * (dl + 2.0);
        /* neighbours; the l = m "S" neighbour is multiplied by zero */
        const int lmS = (l > m) ? lm - 1 : lm;
        const int lmA = (l < p->l_max) ? lm + 1 : -1;
        const cplx zA = lmA >= 0 ? z[lmA] : 0.0, dwA = lmA >= 0 ? dw[lmA] : 0.0;
        const cplx wA = lmA >= 0 ? w[lmA] : 0.0;
        if (lm == 0) {
            /* l = 0, m = 0 */
            dwdt[0] = or2 * Advr[0] + c2 * or1 * t2A * zA;
            dzdt[0] = c2 * or2 * (t3A * dwA + or1 * t4A * wA);
            dpdt[0] = 0.0; /* get_dpdt starts at lm_min = 2 */
            dsdt[0] = epsc;
        } else {
            cplx cw = 0.0, cz = 0.0, cp_ = 0.0;
            if (l < lR) {
                cw = c2 * or1 * (dPhi * dw[lm] + t2A * zA - t2S * z[lmS]);
                cz = c2 * or2 * (dPhi * z[lm] + t3A * dwA + or1 * t4A * wA + t3S * dw[lmS] -
                                 or1 * t4S * w[lmS]);
                cp_ = c2 * or2 * (-dPhi * (dw[lm] + or1 * dLh * w[lm]) + t3A * zA + t3S * z[lmS]);
            } else if (l == lR) {
                cw = c2 * or1 * (dPhi * dw[lm] - t2S * z[lmS]);
                cz = c2 * or2 * (dPhi * z[lm] + t3S * dw[lmS] - or1 * t4S * w[lmS]);
                cp_ = c2 * or2 * (-dPhi * (dw[lm] + or1 * dLh * w[lm]) + t3S * z[lmS]);
            }
            dwdt[lm] = or2 * Advr[lm] + cw;
            dzdt[lm] = dLh * Advp[lm] + cz;
            dpdt[lm] = -dLh * Advt[lm] + cp_;
            dsdt[lm] = dLh * VSt[lm];
        }
        dbdt[lm] = dLh * VxBp[lm];
        dVxBh[lm] = -dLh * VxBt[lm] * r * r;
        djdt[lm] = dLh * or4 * VxBr[lm];
    }
}

/* ------------------------------------------------------------------ */
/* Courant maxima of one level (src/courant.f90:209-275, XSH_COURANT = 0,
   magnetic branch l_mag, l_mag_LF, not l_mag_kin): over the grid,
   vr2max = max or4 (cf^2 orho2 vr^2 + af^2 valr^2 / (valr + valri2)),
   valr = br^2 LFfac orho1, and vh2max = max or2 (cf^2 (vt^2 + vp^2)
   osin2 orho2 + af^2 valh^2 / (valh + valhi2)), valh = (bt^2 + bp^2) LFfac
   osin2 orho1; orho2 = orho1^2.  in[6] = vr vt vp br bt bp grids. */
void orc_courant(const orc_plan *p, double or2, double or4, double orho1, double LFfac,
                 double courfac, double alffac, double valri2, double valhi2,
                 const double *const *in, double *vr2max, double *vh2max)
{
    const int nt = p->n_theta, np_ = p->n_phi;
    const double cf2 = courfac * courfac, af2 = alffac * alffac, orho2 = orho1 * orho1;
    double mr = 0.0, mh = 0.0;
    for (int ip = 0; ip < np_; ++ip)
        for (int it = 0; it < nt; ++it) {
            const size_t g = (size_t)ip * nt + it;
            const double vr = in[0][g], vt = in[1][g], vp = in[2][g];
            const double br = in[3][g], bt = in[4][g], bp = in[5][g];
            const double o2 = p->osin2[it];
            const double vflr2 = orho2 * vr * vr;
            const double valr = br * br * LFfac * orho1;
            const double valr2 = valr * valr / (valr + valri2);
            const double a = or4 * (cf2 * vflr2 + af2 * valr2);
            if (a > mr) mr = a;
            const double vflh2 = (vt * vt + vp * vp) * o2 * orho2;
            const double valh2 = (bt * bt + bp * bp) * LFfac * o2 * orho1;
            const double valh2m = valh2 * valh2 / (valh2 + valhi2);
            const double b = or2 * (cf2 * vflh2 + af2 * valh2m);
            if (b > mh) mh = b;
        }
    *vr2max = mr;
    *vh2max = mh;
}

/* ------------------------------------------------------------------ */
/* src/nonlinear_bcs.f90:24-68 (get_br_v_bcs): the induction products at a
   conducting boundary, br*vt/(r^2 rho) and br*(vp/(r^2 rho) - omega sin^2),
   analysed by spat_to_sphertor at l_max.  sinTheta_E2 = sin(colat)**2 in
   the scrambled order (src/horizontal.f90:151-152). */
void orc_get_br_v_bcs(const orc_plan *p, const double *br, const double *vt, const double *vp,
                      double omega, double O_r_E_2, double O_rho, double *br_vt_lm,
                      double *br_vp_lm)
{
    int nt = p->n_theta;
    size_t npts = (size_t)nt * p->n_phi;
    double *a = (double *)malloc(sizeof(double) * npts);
    double *b = (double *)malloc(sizeof(double) * npts);
    const double fac = O_r_E_2 * O_rho;
    for (size_t i = 0; i < npts; ++i) {
        double s = sin(p->theta_ord[(i % nt) / 2]);
        double s2 = s * s;
        a[i] = fac * br[i] * vt[i];
        b[i] = br[i] * (fac * vp[i] - omega * s2);
    }
    orc_spat_to_sph_tor(p, a, b, br_vt_lm, br_vp_lm, p->l_max);
    free(a);
    free(b);
}

/* ------------------------------------------------------------------ */
/* src/nonlinear_bcs.f90:71-118 (get_b_nl_bcs): the nonlinear term of the
   magnetic boundary condition of a conducting mantle ('CMB', fac =
   conductance_ma prmag, :88-98) or inner core ('ICB', fac = sigma_ratio
   prmag, :100-110), lm = 1 (l = 0) zero. */
int orc_get_b_nl_bcs(int lm_max, int bc, const double *br_vt_lm, const double *br_vp_lm,
                     double cond_or_sigma, double prmag, double *b_nl_bc, double *aj_nl_bc)
{
    if (bc != 0 && bc != 1) return -1;
    const double fac = cond_or_sigma * prmag;
    for (int lm = 0; lm < lm_max; ++lm)
        for (int ri = 0; ri < 2; ++ri) {
            const size_t i = 2 * (size_t)lm + ri;
            if (bc == 0) b_nl_bc[i] = lm == 0 ? 0.0 : -fac * br_vt_lm[i];
            aj_nl_bc[i] = lm == 0 ? 0.0 : -fac * br_vp_lm[i];
        }
    return 0;
}

/* ------------------------------------------------------------------ */
/* The radial loop of any physics switches, one interior level (nBc = 0).
   Grid components in[ORC_RI_*] and products out[ORC_RO_*] are n_phi x
   n_theta grids (theta fastest, scrambled rows); absent components are NULL.
   flags[ORC_F_*], par[ORC_P_*], rad[ORC_R_*] as in mhx_oracle.h.
   src/get_nl.f90:240-436 per point, then the merge of
   src/rIter.f90:644-680 (Adv += LF, += precession, += centrifugal). */
void orc_rl_products(const orc_plan *p, const int *flags, const double *par, const double *rad,
                     const double *const *in, double *const *out)
{
    const int nt = p->n_theta, np_ = p->n_phi;
    const double r = rad[ORC_R_R], or1 = rad[ORC_R_OR1], or2 = rad[ORC_R_OR2];
    const double or4 = rad[ORC_R_OR4], orho1 = rad[ORC_R_ORHO1], beta = rad[ORC_R_BETA];
    const double otemp1 = rad[ORC_R_OTEMP1], visc = rad[ORC_R_VISC], lambda = rad[ORC_R_LAMBDA];
    const int lcr = flags[ORC_F_LCR];
    const double LFfac = par[ORC_P_LFFAC];
    /* posnalp = -2 oek po sin(prec_angle) (src/get_nl.f90:233) */
    const double posnalp = -2.0 * par[ORC_P_OEK] * par[ORC_P_PO] * sin(par[ORC_P_PREC_ANGLE]);
    const double dphi = 2.0 * ORC_PI / ((double)np_ * par[ORC_P_MINC]);
#define G(k) (in[k] ? in[k][g] : 0.0)
    for (int ip = 0; ip < np_; ++ip) {
        const double phi = (double)ip * dphi;
        for (int it = 0; it < nt; ++it) {
            const size_t g = (size_t)ip * nt + it;
            /* src/horizontal.f90:141-157 */
            const double colat = p->theta_ord[it / 2];
            const double sn = sin(colat), sgn = (it & 1) ? -1.0 : 1.0;
            const double o2 = 1.0 / (sn * sn), cosn2 = sgn * cos(colat) / sn / sn;
            const double sinT = sn, cosT = sgn * cos(colat), osinT = 1.0 / sn;
            const double vr = G(ORC_RI_VR), vt = G(ORC_RI_VT), vp = G(ORC_RI_VP);
            const double s = G(ORC_RI_S);
            double LFr = 0.0, LFt = 0.0, LFp = 0.0, Ar = 0.0, At = 0.0, Ap = 0.0;
            const int lf = flags[ORC_F_MAG_LF] && !lcr;
            if (lf) { /* :242-258 */
                const double br = G(ORC_RI_BR), bt = G(ORC_RI_BT), bp = G(ORC_RI_BP);
                const double cbr = G(ORC_RI_CBR), cbt = G(ORC_RI_CBT), cbp = G(ORC_RI_CBP);
                LFr = LFfac * o2 * (cbt * bp - cbp * bt);
                LFt = LFfac * or4 * (cbp * br - cbr * bp);
                LFp = LFfac * or4 * (cbr * bt - cbt * br);
            }
            if (flags[ORC_F_CONV_NL]) {
                if (flags[ORC_F_ADV_CURL]) { /* :262-273 */
                    const double cvr = G(ORC_RI_CVR), cvt = G(ORC_RI_CVT), cvp = G(ORC_RI_CVP);
                    Ar = -o2 * (cvt * vp - cvp * vt);
                    At = -or4 * (cvp * vr - cvr * vp);
                    Ap = -or4 * (cvr * vt - cvt * vr);
                } else { /* :274-308 */
                    const double dvrdr = G(ORC_RI_DVRDR), dvtdr = G(ORC_RI_DVTDR);
                    const double dvpdr = G(ORC_RI_DVPDR), dvrdt = G(ORC_RI_DVRDT);
                    const double dvrdp = G(ORC_RI_DVRDP), dvtdp = G(ORC_RI_DVTDP);
                    const double dvpdp = G(ORC_RI_DVPDP), cvr = G(ORC_RI_CVR);
                    Ar = -or2 * orho1 * (vr * (dvrdr - (2.0 * or1 + beta) * vr) +
                                         o2 * (vt * (dvrdt - r * vt) + vp * (dvrdp - r * vp)));
                    At = or4 * orho1 * (-vr * (dvtdr - beta * vt) + vt * (cosn2 * vt + dvpdp + dvrdr) +
                                        vp * (cosn2 * vp - dvtdp));
                    Ap = or4 * orho1 * (-vr * (dvpdr - beta * vp) - vt * (dvtdp + cvr) - vp * dvpdp);
                }
            }
            if (flags[ORC_F_PHASE]) { /* :333-344 */
                const double eps = par[ORC_P_EPSPHASE], pen = par[ORC_P_PENALTY];
                const double pc = G(ORC_RI_PHI);
                Ar = Ar - pc * vr / (eps * eps) / (pen * pen);
                At = At - or2 * pc * vt / (eps * eps) / (pen * pen);
                Ap = Ap - or2 * pc * vp / (eps * eps) / (pen * pen);
                out[ORC_RO_PHI][g] = -1.0 / (eps * eps) * pc * (1.0 - pc) *
                                     (par[ORC_P_PHASEDIFF] * (1.0 - 2.0 * pc) + s - par[ORC_P_TMELT]);
            }
            if (flags[ORC_F_CONV_NL]) {
                if (lf) { Ar += LFr; At += LFt; Ap += LFp; }
            } else if (flags[ORC_F_MAG_LF]) {
                Ar = LFr; At = LFt; Ap = LFp;
            }
            if (flags[ORC_F_PREC]) { /* :346-356 */
                const double a = par[ORC_P_OEK] * par[ORC_P_TIME] + phi;
                Ar += posnalp * osinT * r * (cos(a) * vp * cosT + sin(a) * vt);
                At += -posnalp * sinT * or2 * (cos(a) * vp + sin(a) * or1 * vr);
                Ap += posnalp * sinT * cos(a) * or2 * (vt - or1 * vr * cosT);
            }
            if (flags[ORC_F_CENTR]) { /* :358-374 */
                const double f = par[ORC_P_DILUTION] * par[ORC_P_RA] * par[ORC_P_OPR];
                Ar += -f * r * pow(sinT, 4) * s;
                At += -f * r * pow(sinT, 3) * cosT * s;
            }
            if (out[ORC_RO_ADVR]) {
                out[ORC_RO_ADVR][g] = Ar;
                out[ORC_RO_ADVT][g] = At;
                out[ORC_RO_ADVP][g] = Ap;
            }
            if (out[ORC_RO_VSR]) { /* :320-325 */
                const int on = flags[ORC_F_HEAT_NL];
                out[ORC_RO_VSR][g] = on ? vr * s : 0.0;
                out[ORC_RO_VST][g] = on ? or2 * vt * s : 0.0;
                out[ORC_RO_VSP][g] = on ? or2 * vp * s : 0.0;
            }
            if (flags[ORC_F_CHEM]) { /* :327-331 */
                const double xi = G(ORC_RI_XI);
                out[ORC_RO_VXIR][g] = vr * xi;
                out[ORC_RO_VXIT][g] = or2 * vt * xi;
                out[ORC_RO_VXIP][g] = or2 * vp * xi;
            }
            if (flags[ORC_F_MAG_NL]) { /* :376-391 */
                const double br = G(ORC_RI_BR), bt = G(ORC_RI_BT), bp = G(ORC_RI_BP);
                if (!lcr) {
                    out[ORC_RO_VXBR][g] = orho1 * o2 * (vt * bp - vp * bt);
                    out[ORC_RO_VXBT][g] = orho1 * or4 * (vp * br - vr * bp);
                    out[ORC_RO_VXBP][g] = orho1 * or4 * (vr * bt - vt * br);
                } else {
                    out[ORC_RO_VXBR][g] = 0.0;
                    out[ORC_RO_VXBT][g] = or4 * orho1 * vp * br;
                    out[ORC_RO_VXBP][g] = -or4 * orho1 * vt * br;
                }
            }
            if (flags[ORC_F_ANEL]) { /* :402-436 */
                const double dvrdr = G(ORC_RI_DVRDR), dvtdr = G(ORC_RI_DVTDR);
                const double dvpdr = G(ORC_RI_DVPDR), dvrdt = G(ORC_RI_DVRDT);
                const double dvrdp = G(ORC_RI_DVRDP), dvtdp = G(ORC_RI_DVTDP);
                const double dvpdp = G(ORC_RI_DVPDP), cvr = G(ORC_RI_CVR);
                double h = par[ORC_P_VISCHEAT] * or4 * orho1 * otemp1 * visc *
                           (2.0 * pow(dvrdr - (2.0 * or1 + beta) * vr, 2) +
                            2.0 * pow(cosn2 * vt + dvpdp + dvrdr - or1 * vr, 2) +
                            2.0 * pow(dvpdp + cosn2 * vt + or1 * vr, 2) +
                            pow(2.0 * dvtdp + cvr - 2.0 * cosn2 * vp, 2) +
                            o2 * (pow(r * dvtdr - (2.0 + beta * r) * vt + or1 * dvrdt, 2) +
                                  pow(r * dvpdr - (2.0 + beta * r) * vp + or1 * dvrdp, 2)) -
                            2.0 / 3.0 * pow(beta * vr, 2));
                if (flags[ORC_F_MAG_NL] && !lcr) {
                    const double cbr = G(ORC_RI_CBR), cbt = G(ORC_RI_CBT), cbp = G(ORC_RI_CBP);
                    h = h + par[ORC_P_OHMLOSS] * or2 * otemp1 * lambda *
                                (or2 * cbr * cbr + o2 * cbt * cbt + o2 * cbp * cbp);
                }
                out[ORC_RO_HEAT][g] = h;
            }
        }
    }
#undef G
}

/* src/courant.f90:161-276 (XSH_COURANT = 0) for one level, from 1e10
   (src/rIter.f90:217-218): the magnetic branch when cour_mag (l_mag,
   l_mag_LF, not l_mag_kin), else the velocity-only one.  par: courfac,
   alffac, LFfac, opm, l_cour_alf_damp; rad: or2, or4, orho1, delxr2,
   delxh2.  in: vr vt vp br bt bp (the last three unused without cour_mag). */
void orc_rl_courant(const orc_plan *p, int cour_mag, const double *par, const double *rad,
                    const double *const *in, double *dtrkc, double *dthkc)
{
    const int nt = p->n_theta, np_ = p->n_phi;
    const double courfac = par[0], alffac = par[1], LFfac = par[2], opm = par[3];
    const int damp = par[4] != 0.0;
    const double or2 = rad[0], or4 = rad[1], orho1 = rad[2], delxr2 = rad[3], delxh2 = rad[4];
    const double valri2 = damp ? (0.5 * (1.0 + opm)) * (0.5 * (1.0 + opm)) / delxr2 : 0.0;
    const double valhi2 = damp ? (0.5 * (1.0 + opm)) * (0.5 * (1.0 + opm)) / delxh2 : 0.0;
    const double cf2 = courfac * courfac, af2 = alffac * alffac, orho2 = orho1 * orho1;
    double vr2max = 0.0, vh2max = 0.0;
    for (int ip = 0; ip < np_; ++ip)
        for (int it = 0; it < nt; ++it) {
            const size_t g = (size_t)ip * nt + it;
            const double vr = in[0][g], vt = in[1][g], vp = in[2][g];
            const double o2 = p->osin2[it];
            const double vflr2 = orho2 * vr * vr;
            const double vflh2 = (vt * vt + vp * vp) * o2 * orho2;
            if (cour_mag) {
                const double br = in[3][g], bt = in[4][g], bp = in[5][g];
                const double valr = br * br * LFfac * orho1;
                const double valr2 = valr * valr / (valr + valri2);
                vr2max = fmax(vr2max, or4 * (cf2 * vflr2 + af2 * valr2));
                const double valh2 = (bt * bt + bp * bp) * LFfac * o2 * orho1;
                const double valh2m = valh2 * valh2 / (valh2 + valhi2);
                vh2max = fmax(vh2max, or2 * (cf2 * vflh2 + af2 * valh2m));
            } else {
                vr2max = fmax(vr2max, cf2 * or4 * vflr2);
                vh2max = fmax(vh2max, cf2 * or2 * vflh2);
            }
        }
    *dtrkc = 1e10;
    *dthkc = 1e10;
    if (vr2max != 0.0) *dtrkc = fmin(*dtrkc, sqrt(delxr2 / vr2max));
    if (vh2max != 0.0) *dthkc = fmin(*dthkc, sqrt(delxh2 / vh2max));
}

/* nonlinear_lm_t at one interior level (nBc = 0), src/get_td.f90: get_dwdt
   (:133-203) or get_dwdt_double_curl (:210-309), get_dzdt (:379-444),
   get_dpdt (:311-377), get_dsdt (:459-521), get_dxidt (:523-556),
   get_dbdt (:558-589); the dTheta operators of src/horizontal.f90:201-229
   and the neighbours lm2lmS / lm2lmA of src/blocking.f90:322-334.
   nl[ORC_NL_*] the forward outputs (NULL when switched off), fld = w dw ddw
   z dz, out[ORC_TO_*] (NULL where not wanted). */
void orc_rl_td(const orc_plan *p, const int *flags, const double *par, const double *rad, int lR,
               const double *const *nl_, const double *const *fld, double *const *out_)
{
    const cplx *nl[ORC_NL_N];
    for (int k = 0; k < ORC_NL_N; ++k) nl[k] = (const cplx *)nl_[k];
    const cplx *w = (const cplx *)fld[0], *dw = (const cplx *)fld[1], *ddw = (const cplx *)fld[2];
    const cplx *z = (const cplx *)fld[3], *dz = (const cplx *)fld[4];
    cplx *out[ORC_TO_N];
    for (int k = 0; k < ORC_TO_N; ++k) out[k] = (cplx *)out_[k];
    const double or1 = rad[ORC_R_OR1], or2 = rad[ORC_R_OR2], or4 = rad[ORC_R_OR4];
    const double orho1 = rad[ORC_R_ORHO1], beta = rad[ORC_R_BETA], r = rad[ORC_R_R];
    const double CorFac = par[ORC_P_CORFAC];
    const int conv_nl = flags[ORC_F_CONV_NL], corr = flags[ORC_F_CORR];
#define NLV(k, i) (nl[k] ? nl[k][i] : 0.0)
    for (int lm = 0; lm < p->lm_max; ++lm) {
        const int l = p->lm2l[lm], m = p->lm2m[lm];
        const double dl = (double)l, dLh = dl * (dl + 1.0);
        const cplx dPhi = I * (double)m;
        const double cS = td_clm(l, m), cA = td_clm(l + 1, m);
        const double t2S = (dl - 1.0) * cS, t2A = (dl + 2.0) * cA;
        const double t3S = (dl - 1.0) * (dl + 1.0) * cS, t3A = dl * (dl + 2.0) * cA;
        const double t4S = (dl + 1.0) * cS * (dl - 1.0) * dl, t4A = dl * cA * (dl + 1.0) * (dl + 2.0);
        const int lmS = (l > m) ? lm - 1 : lm;
        const int lmA = (l < p->l_max) ? lm + 1 : -1;
        const cplx zA = lmA >= 0 ? z[lmA] : 0.0, dwA = lmA >= 0 ? dw[lmA] : 0.0;
        const cplx wA = lmA >= 0 ? w[lmA] : 0.0, dzA = lmA >= 0 ? dz[lmA] : 0.0;
        /* dwdt */
        cplx dwdt, dvxvh = 0.0;
        if (lm == 0) {
            dwdt = conv_nl ? or2 * NLV(ORC_NL_ADVR, 0) : 0.0;
            if (corr && !flags[ORC_F_SINGLE]) dwdt += 2.0 * CorFac * or1 * t2A * zA;
        } else if (!flags[ORC_F_DCURL]) {
            dwdt = conv_nl ? or2 * NLV(ORC_NL_ADVR, lm) : 0.0;
            if (corr) {
                cplx c = 0.0;
                if (l < lR) c = 2.0 * CorFac * or1 * (dPhi * dw[lm] + t2A * zA - t2S * z[lmS]);
                else if (l == lR) c = 2.0 * CorFac * or1 * (dPhi * dw[lm] - t2S * z[lmS]);
                dwdt += c;
            }
        } else {
            dwdt = conv_nl ? dLh * or4 * orho1 * NLV(ORC_NL_ADVR, lm) : 0.0;
            dvxvh = conv_nl ? -orho1 * r * r * dLh * NLV(ORC_NL_ADVT, lm) : 0.0;
            if (corr) {
                cplx c = 0.0;
                const cplx ph = dPhi * (-ddw[lm] + beta * dw[lm] + (beta * or1 + or2) * dLh * w[lm]);
                if (l < lR)
                    c = 2.0 * CorFac * or2 * orho1 *
                        (ph + t3A * (dzA - beta * zA) + t3S * (dz[lmS] - beta * z[lmS]) +
                         or1 * (t4A * zA - t4S * z[lmS]));
                else if (l == lR)
                    c = 2.0 * CorFac * or2 * orho1 *
                        (ph + t3S * (dz[lmS] - beta * z[lmS]) - or1 * t4S * z[lmS]);
                dwdt += c;
            }
        }
        if (out[ORC_TO_DWDT]) out[ORC_TO_DWDT][lm] = dwdt;
        if (out[ORC_TO_DVXVH]) out[ORC_TO_DVXVH][lm] = dvxvh;
        /* dzdt */
        cplx dzdt;
        if (lm == 0) {
            dzdt = corr ? 2.0 * CorFac * or2 * (t3A * dwA + or1 * t4A * wA) : 0.0;
        } else {
            dzdt = conv_nl ? dLh * NLV(ORC_NL_ADVP, lm) : 0.0;
            if (corr) {
                if (l < lR)
                    dzdt += 2.0 * CorFac * or2 * (dPhi * z[lm] + t3A * dwA + or1 * t4A * wA +
                                                  t3S * dw[lmS] - or1 * t4S * w[lmS]);
                else if (l == lR)
                    dzdt += 2.0 * CorFac * or2 * (dPhi * z[lm] + t3S * dw[lmS] - or1 * t4S * w[lmS]);
            }
        }
        if (out[ORC_TO_DZDT]) out[ORC_TO_DZDT][lm] = dzdt;
        /* dpdt (l = 0 not set by the reference: zero) */
        if (out[ORC_TO_DPDT]) {
            cplx dpdt = 0.0;
            if (lm > 0) {
                dpdt = conv_nl ? -dLh * NLV(ORC_NL_ADVT, lm) : 0.0;
                if (corr) {
                    if (l < lR)
                        dpdt += 2.0 * CorFac * or2 * (-dPhi * (dw[lm] + or1 * dLh * w[lm]) + t3A * zA +
                                                      t3S * z[lmS]);
                    else if (l == lR)
                        dpdt += 2.0 * CorFac * or2 * (-dPhi * (dw[lm] + or1 * dLh * w[lm]) + t3S * z[lmS]);
                }
            }
            out[ORC_TO_DPDT][lm] = dpdt;
        }
        /* dsdt */
        if (out[ORC_TO_DSDT]) {
            cplx dsdt;
            const double tf = flags[ORC_F_ANEL_LIQ] ? rad[ORC_R_TEMP0] : 1.0;
            if (lm == 0) dsdt = par[ORC_P_EPSC] * rad[ORC_R_EPSCPROF];
            else dsdt = dLh * NLV(ORC_NL_VST, lm);
            if (flags[ORC_F_ANEL]) dsdt += (flags[ORC_F_ANEL_LIQ] ? tf : 1.0) * NLV(ORC_NL_HEAT, lm);
            out[ORC_TO_DSDT][lm] = dsdt;
        }
        /* dxidt */
        if (out[ORC_TO_DXIDT])
            out[ORC_TO_DXIDT][lm] = lm == 0 ? par[ORC_P_EPSCXI] : dLh * NLV(ORC_NL_VXIT, lm);
        /* dbdt, djdt, dVxBhLM */
        if (out[ORC_TO_DBDT]) {
            if (flags[ORC_F_MAG_DBDT]) {
                out[ORC_TO_DBDT][lm] = dLh * NLV(ORC_NL_VXBP, lm);
                out[ORC_TO_DVXBH][lm] = -dLh * NLV(ORC_NL_VXBT, lm) * r * r;
                out[ORC_TO_DJDT][lm] = dLh * or4 * NLV(ORC_NL_VXBR, lm);
            } else {
                out[ORC_TO_DBDT][lm] = 0.0;
                out[ORC_TO_DVXBH][lm] = 0.0;
                out[ORC_TO_DJDT][lm] = 0.0;
            }
        }
    }
#undef NLV
}
