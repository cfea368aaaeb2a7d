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
    cplx *x = (cplx *)malloc(sizeof(cplx) * n), *X = (cplx *)malloc(sizeof(cplx) * n);
    cplx *F = (cplx *)Fout;
    for (int t = 0; t < nt; ++t) {
        for (int j = 0; j < n; ++j) x[j] = grid[(size_t)j * nt + t];
        cfft(n, x, X, -1);
        for (int k = 0; k < nk; ++k) F[(size_t)k * nt + t] = X[k] / (double)n;
        F[t] = creal(F[t]);
        F[(size_t)(nk - 1) * nt + t] = creal(F[(size_t)(nk - 1) * nt + t]);
    }
    free(x); free(X);
}

/* src/fft.f90:211-252 with fft991(isign=+1): x_j = sum_k c_k e^{2 pi i jk/n},
   c_{n-k} = conj(c_k); only the real parts of c_0 and c_{n/2} enter */
void orc_fft_inverse(const orc_plan *p, const double *Fin, double *grid)
{
    int n = p->n_phi, nt = p->n_theta, nk = n / 2 + 1;
    const cplx *F = (const cplx *)Fin;
    cplx *X = (cplx *)malloc(sizeof(cplx) * n), *x = (cplx *)malloc(sizeof(cplx) * n);
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
