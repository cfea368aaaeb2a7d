// P_lm tables on the GPU: FP64 three-term recurrence per (m, theta), written
// straight into the MFMA-tiled layouts the Legendre kernels read.
//
// Recurrence and normalisation follow src/plms.f90:14-189 (norm=2: fully
// normalised with 1/sqrt(4 pi), no Condon-Shortley phase, derivative stored
// as sin(theta) dP/dtheta); weights 2 pi w_theta as src/shtransforms.f90:75-76.
//
// Layouts (doubles), for order index mc and parity p = (l - m) mod 2,
// kk = (l - m) / 2:
//   synthesis  (MFMA A operand 16 theta x 4 l):
//     syn[syn_off[2mc+p] + ((kk/4) * nt16 + th/16) * 64 + (th%16) + 16*(kk%4)]
//   analysis   (MFMA A operand 16 l x 4 theta):
//     ana[ana_off[2mc+p] + ((kk/16) * nq4 + th/4) * 64 + (kk%16) + 16*(th%4)]
// where th is the northern-hemisphere theta index.  Padding entries are 0.
#include "mhx_internal.h"

#include <cmath>

namespace mhx {

__global__ void k_plm_tables(int l_max, int minc, int n_m, int nh, int nt16, int nq4,
                             const double *__restrict__ theta, const double *__restrict__ gauss,
                             const size_t *__restrict__ syn_off, const size_t *__restrict__ ana_off,
                             double *__restrict__ sP, double *__restrict__ sdP,
                             double *__restrict__ aP, double *__restrict__ adP)
{
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n_m * nh) return;
    int mc = idx / nh, th = idx - mc * nh;
    int m = mc * minc;
    const double dnorm = 0.28209479177387814;  // 1/sqrt(4 pi)
    const double twopi = 6.283185307179586;
    double t = theta[th];
    double st = sin(t), ct = cos(t);
    double w = twopi * gauss[th];

    double fac = 1.0;
    for (int j = 3; j <= 2 * m + 1; j += 2) fac *= (double)j / (double)(j - 1);
    double p = sqrt(fac);
    if (st != 0.0) p *= pow(st, (double)m);
    else if (m != 0) p = 0.0;

    // sliding window over l: pm1 = P_{l-1}, pc = P_l (unnormalised)
    double pm1 = 0.0, pc = p;
    for (int l = m; l <= l_max; ++l) {
        // P_{l+1}
        int L1 = l + 1;
        double pn;
        if (L1 == m + 1) {
            pn = ct * sqrt((double)((2 * L1 - 1) * (2 * L1 + 1)) / (double)((L1 - m) * (L1 + m))) * pc;
        } else {
            pn = ct * sqrt((double)((2 * L1 - 1) * (2 * L1 + 1)) / (double)((L1 - m) * (L1 + m))) * pc
               - sqrt(((double)(2 * L1 + 1) * (double)(L1 + m - 1) * (double)(L1 - m - 1))
                      / ((double)(2 * L1 - 3) * (double)(L1 - m) * (double)(L1 + m))) * pm1;
        }
        double Pl = dnorm * pc;
        double dPl;
        if (l == m) {
            dPl = l / sqrt((double)(2 * l + 3)) * (dnorm * pn);
        } else {
            dPl = l * sqrt((double)((l + m + 1) * (l - m + 1)) / (double)((2 * l + 1) * (2 * l + 3))) * (dnorm * pn)
                - (l + 1) * sqrt((double)((l + m) * (l - m)) / (double)((2 * l - 1) * (2 * l + 1))) * (dnorm * pm1);
        }
        int par = (l - m) & 1, kk = (l - m) >> 1;
        size_t so = syn_off[2 * mc + par] + ((size_t)(kk >> 2) * nt16 + (th >> 4)) * 64 + (th & 15) + 16 * (kk & 3);
        size_t ao = ana_off[2 * mc + par] + ((size_t)(kk >> 4) * nq4 + (th >> 2)) * 64 + (kk & 15) + 16 * (th & 3);
        sP[so] = Pl;
        sdP[so] = dPl;
        aP[ao] = w * Pl;
        adP[ao] = w * dPl;
        pm1 = pc;
        pc = pn;
    }
}

int build_tables(Plan *p)
{
    const int nh = p->nh;
    p->syn_off.assign(2 * p->n_m, 0);
    p->ana_off.assign(2 * p->n_m, 0);
    size_t so = 0, ao = 0;
    for (int mc = 0; mc < p->n_m; ++mc) {
        int m = mc * p->minc;
        int K[2] = {(p->l_max - m) / 2 + 1, (p->l_max - m + 1) / 2};
        int Kt0 = (K[0] + 15) / 16;  // pair tiles: both parities use K[0]'s count
        for (int par = 0; par < 2; ++par) {
            p->syn_off[2 * mc + par] = so;
            so += (size_t)((K[par] + 3) / 4) * p->nt16 * 64;
            p->ana_off[2 * mc + par] = ao;
            ao += (size_t)Kt0 * p->nq4 * 64;
        }
    }
    p->syn_size = so;
    p->ana_size = ao;
    MHX_HIP_CHECK(hipMalloc(&p->d_syn_P, so * sizeof(double)));
    MHX_HIP_CHECK(hipMalloc(&p->d_syn_dP, so * sizeof(double)));
    MHX_HIP_CHECK(hipMalloc(&p->d_ana_wP, ao * sizeof(double)));
    MHX_HIP_CHECK(hipMalloc(&p->d_ana_wdP, ao * sizeof(double)));
    MHX_HIP_CHECK(hipMemset(p->d_syn_P, 0, so * sizeof(double)));
    MHX_HIP_CHECK(hipMemset(p->d_syn_dP, 0, so * sizeof(double)));
    MHX_HIP_CHECK(hipMemset(p->d_ana_wP, 0, ao * sizeof(double)));
    MHX_HIP_CHECK(hipMemset(p->d_ana_wdP, 0, ao * sizeof(double)));
    MHX_HIP_CHECK(hipMalloc(&p->d_syn_off, 2 * p->n_m * sizeof(size_t)));
    MHX_HIP_CHECK(hipMalloc(&p->d_ana_off, 2 * p->n_m * sizeof(size_t)));
    MHX_HIP_CHECK(hipMemcpy(p->d_syn_off, p->syn_off.data(), 2 * p->n_m * sizeof(size_t),
                            hipMemcpyHostToDevice));
    MHX_HIP_CHECK(hipMemcpy(p->d_ana_off, p->ana_off.data(), 2 * p->n_m * sizeof(size_t),
                            hipMemcpyHostToDevice));
    int n = p->n_m * nh;
    k_plm_tables<<<(n + 127) / 128, 128>>>(p->l_max, p->minc, p->n_m, nh, p->nt16, p->nq4,
                                           p->d_theta, p->d_gauss, p->d_syn_off, p->d_ana_off,
                                           p->d_syn_P, p->d_syn_dP, p->d_ana_wP, p->d_ana_wdP);
    MHX_HIP_CHECK(hipGetLastError());
    MHX_HIP_CHECK(hipDeviceSynchronize());
    return MHX_OK;
}

}  // namespace mhx
