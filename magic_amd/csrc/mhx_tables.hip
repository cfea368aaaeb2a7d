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

// The recurrence of src/plms.f90:46-187 at one (m, theta), one degree per
// step: after init, each step() gives P_l and sin(theta) dP_l/dtheta
// (normalised) of the next l = m, m + 1, ...
struct PlmStep {
    int m, l;
    double ct, pm1, pc;  // P_{l-1}, P_l (unnormalised)
    __device__ __forceinline__ void init(int m_, double st, double ct_)
    {
        m = m_;
        l = m_;
        ct = ct_;
        double fac = 1.0;
        for (int j = 3; j <= 2 * m + 1; j += 2) fac *= (double)j / (double)(j - 1);
        double p = sqrt(fac);
        if (st != 0.0) p *= pow(st, (double)m);
        else if (m != 0) p = 0.0;
        pm1 = 0.0;
        pc = p;
    }
    __device__ __forceinline__ void step(double &Pl, double &dPl)
    {
        const double dnorm = 0.28209479177387814;  // 1/sqrt(4 pi)
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
        Pl = dnorm * pc;
        if (l == m) {
            dPl = l / sqrt((double)(2 * l + 3)) * (dnorm * pn);
        } else {
            dPl = l * sqrt((double)((l + m + 1) * (l - m + 1)) / (double)((2 * l + 1) * (2 * l + 3))) * (dnorm * pn)
                - (l + 1) * sqrt((double)((l + m) * (l - m)) / (double)((2 * l - 1) * (2 * l + 1))) * (dnorm * pm1);
        }
        pm1 = pc;
        pc = pn;
        ++l;
    }
};

// f(l, P_l, sin(theta) dP_l/dtheta) for l = m .. l_max
template <typename F>
__device__ __forceinline__ void plm_sweep(int l_max, int m, double st, double ct, F &&f)
{
    PlmStep r;
    r.init(m, st, ct);
    for (int l = m; l <= l_max; ++l) {
        double Pl, dPl;
        r.step(Pl, dPl);
        f(l, Pl, dPl);
    }
}

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
    const double twopi = 6.283185307179586;
    double t = theta[th];
    double w = twopi * gauss[th];
    plm_sweep(l_max, m, sin(t), cos(t), [&](int l, double Pl, double dPl) {
        int par = (l - m) & 1, kk = (l - m) >> 1;
        size_t so = syn_off[2 * mc + par] + ((size_t)(kk >> 2) * nt16 + (th >> 4)) * 64 + (th & 15) + 16 * (kk & 3);
        size_t ao = ana_off[2 * mc + par] + ((size_t)(kk >> 4) * nq4 + (th >> 2)) * 64 + (kk & 15) + 16 * (th & 3);
        sP[so] = Pl;
        sdP[so] = dPl;
        aP[ao] = w * Pl;
        adP[ao] = w * dPl;
    });
}

// Row factors of the split tables straight from the recurrence (plans whose
// split kernels generate their tables in-kernel keep no FP64 tables):
// synthesis rows (mc, theta): max over l <= l_max of |P|, |dP|
__global__ void k_rowscale_syn_rec(int l_max, int minc, int n_m, int nh, const double *__restrict__ theta,
                                   float *__restrict__ rs)
{
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n_m * nh) return;
    int mc = idx / nh, th = idx - mc * nh;
    double t = theta[th], mx = 0.0;
    plm_sweep(l_max, mc * minc, sin(t), cos(t),
              [&](int, double Pl, double dPl) { mx = fmax(mx, fmax(fabs(Pl), fabs(dPl))); });
    rs[idx] = pow2_scale((float)mx);
}

// analysis rows (mc, parity, kk) of l = m + par + 2 kk: max over theta of
// |2 pi w P|, |2 pi w dP|, and 1 for the padding rows past l_max; one
// workgroup (256 threads) per mc, each thread stepping the recurrence of its
// theta rows th = tid + 256 r together, one workgroup reduction per l
__global__ __launch_bounds__(256) void k_rowscale_ana_rec(int l_max, int minc, int nh,
                                                          const double *__restrict__ theta,
                                                          const double *__restrict__ gauss,
                                                          const size_t *__restrict__ soff,
                                                          float *__restrict__ rs)
{
    constexpr int MAXT = 8;  // theta rows per thread: nh <= 2048
    const int mc = blockIdx.x, m = mc * minc, tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    __shared__ double red[4];
    const double twopi = 6.283185307179586;
    // padding rows first (every row of the order), valid rows overwritten below
    for (int par = 0; par < 2; ++par) {
        const size_t b0 = soff[2 * mc + par], b1 = soff[2 * mc + par + 1];
        for (size_t i = b0 + tid; i < b1; i += 256) rs[i] = 1.f;
    }
    PlmStep st[MAXT];
    double w[MAXT];
#pragma unroll
    for (int r = 0; r < MAXT; ++r) {
        const int th = tid + 256 * r;
        w[r] = 0.0;
        if (th < nh) {
            const double t = theta[th];
            w[r] = twopi * gauss[th];
            st[r].init(m, sin(t), cos(t));
        }
    }
    __syncthreads();
    for (int l = m; l <= l_max; ++l) {
        double mx = 0.0;
#pragma unroll
        for (int r = 0; r < MAXT; ++r) {
            if (tid + 256 * r >= nh) continue;
            double Pl, dPl;
            st[r].step(Pl, dPl);
            mx = fmax(mx, fmax(fabs(w[r] * Pl), fabs(w[r] * dPl)));
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
        if (lane == 0) red[wave] = mx;
        __syncthreads();
        if (tid == 0) {
            const double v = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
            const int par = (l - m) & 1, kk = (l - m) >> 1;
            rs[soff[2 * mc + par] + kk] = pow2_scale((float)v);
        }
        __syncthreads();
    }
}

int table_layout(Plan *p)
{
    if (p->d_syn_off) return MHX_OK;
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
    MHX_HIP_CHECK(hipMalloc(&p->d_syn_off, 2 * p->n_m * sizeof(size_t)));
    MHX_HIP_CHECK(hipMalloc(&p->d_ana_off, 2 * p->n_m * sizeof(size_t)));
    MHX_HIP_CHECK(hipMemcpy(p->d_syn_off, p->syn_off.data(), 2 * p->n_m * sizeof(size_t),
                            hipMemcpyHostToDevice));
    MHX_HIP_CHECK(hipMemcpy(p->d_ana_off, p->ana_off.data(), 2 * p->n_m * sizeof(size_t),
                            hipMemcpyHostToDevice));
    return MHX_OK;
}

int build_tables(Plan *p)
{
    int rc;
    if ((rc = table_layout(p))) return rc;
    if (p->d_syn_P) return MHX_OK;
    const int nh = p->nh;
    const size_t so = p->syn_size, ao = p->ana_size;
    MHX_HIP_CHECK(hipMalloc(&p->d_syn_P, so * sizeof(double)));
    MHX_HIP_CHECK(hipMalloc(&p->d_syn_dP, so * sizeof(double)));
    MHX_HIP_CHECK(hipMalloc(&p->d_ana_wP, ao * sizeof(double)));
    MHX_HIP_CHECK(hipMalloc(&p->d_ana_wdP, ao * sizeof(double)));
    MHX_HIP_CHECK(hipMemset(p->d_syn_P, 0, so * sizeof(double)));
    MHX_HIP_CHECK(hipMemset(p->d_syn_dP, 0, so * sizeof(double)));
    MHX_HIP_CHECK(hipMemset(p->d_ana_wP, 0, ao * sizeof(double)));
    MHX_HIP_CHECK(hipMemset(p->d_ana_wdP, 0, ao * sizeof(double)));
    int n = p->n_m * nh;
    k_plm_tables<<<(n + 127) / 128, 128>>>(p->l_max, p->minc, p->n_m, nh, p->nt16, p->nq4,
                                           p->d_theta, p->d_gauss, p->d_syn_off, p->d_ana_off,
                                           p->d_syn_P, p->d_syn_dP, p->d_ana_wP, p->d_ana_wdP);
    MHX_HIP_CHECK(hipGetLastError());
    MHX_HIP_CHECK(hipDeviceSynchronize());
    return MHX_OK;
}

int ensure_fp64_tables(Plan *p)
{
    return p->d_syn_P ? MHX_OK : build_tables(p);
}

int split_rowscales_rec(Plan *p, float *syn_rs, float *ana_rs, const size_t *ana_soff)
{
    if (p->nh > 2048) return fail(MHX_E_UNSUPPORTED, "row factors from the recurrence need n_theta <= 4096");
    const int n = p->n_m * p->nh;
    k_rowscale_syn_rec<<<(n + 127) / 128, 128>>>(p->l_max, p->minc, p->n_m, p->nh, p->d_theta, syn_rs);
    k_rowscale_ana_rec<<<p->n_m, 256>>>(p->l_max, p->minc, p->nh, p->d_theta, p->d_gauss, ana_soff,
                                        ana_rs);
    MHX_HIP_CHECK(hipGetLastError());
    return MHX_OK;
}

}  // namespace mhx
