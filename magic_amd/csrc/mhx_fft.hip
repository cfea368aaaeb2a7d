// phi FFTs, grid-space products and the axisymmetric transforms.
//
// Real FFTs of length N = n_phi over blocks of theta rows, computed as a
// complex FFT of length M = N/2 in LDS (Stockham autosort, radices 4/2/3/5,
// the radix set of src/fft.f90:67-73) plus the half-length pre/post step.
// Normalisation as the reference: forward F_k = (1/N) sum_j x_j e^{-2 pi i jk/N}
// (src/fft.f90:262-268,464), inverse unnormalised Hermitian synthesis using
// only the real parts of modes 0 and N/2 (src/fft.f90:386-398).
//
// Grids are theta-fastest (Fortran f(n_theta, n_phi)); a block of TB
// consecutive theta rows is read/written as TB-element runs per phi.
#include "mhx_internal.h"

namespace mhx {

namespace {

constexpr int NTH = 256;
constexpr int MAXB = 8;  // butterflies per thread per pass (rows*M/R <= 2048)

struct FFTArgs {
    int M, N, nfac, fac[16];
    const double2 *tw;   // exp(-2 pi i k / M), k < M
    const double2 *twr;  // exp(-2 pi i k / N), k <= M
};

__device__ inline double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
// multiply by sign*i
__device__ inline double2 cmuli(double2 a, int sign)
{
    return sign > 0 ? make_double2(-a.y, a.x) : make_double2(a.y, -a.x);
}

// in-register DFT of size R with kernel exp(sign 2 pi i / R)
template <int R> __device__ inline void dft(double2 *x, int sign);

template <> __device__ inline void dft<2>(double2 *x, int)
{
    double2 a = x[0], b = x[1];
    x[0] = cadd(a, b);
    x[1] = csub(a, b);
}
template <> __device__ inline void dft<4>(double2 *x, int sign)
{
    double2 a0 = cadd(x[0], x[2]), a1 = csub(x[0], x[2]);
    double2 b0 = cadd(x[1], x[3]), b1 = cmuli(csub(x[1], x[3]), sign);
    x[0] = cadd(a0, b0);
    x[2] = csub(a0, b0);
    x[1] = cadd(a1, b1);
    x[3] = csub(a1, b1);
}
template <> __device__ inline void dft<3>(double2 *x, int sign)
{
    const double c = -0.5, s = 0.86602540378443864676 * sign;
    double2 t1 = cadd(x[1], x[2]), t2 = csub(x[1], x[2]);
    double2 y0 = cadd(x[0], t1);
    double2 m1 = make_double2(x[0].x + c * t1.x, x[0].y + c * t1.y);
    double2 m2 = make_double2(-s * t2.y, s * t2.x);  // i s t2
    x[0] = y0;
    x[1] = cadd(m1, m2);
    x[2] = csub(m1, m2);
}
template <> __device__ inline void dft<5>(double2 *x, int sign)
{
    const double c1 = 0.30901699437494742410, c2 = -0.80901699437494742410;
    const double s1 = 0.95105651629515357212 * sign, s2 = 0.58778525229247312917 * sign;
    double2 a1 = cadd(x[1], x[4]), b1 = csub(x[1], x[4]);
    double2 a2 = cadd(x[2], x[3]), b2 = csub(x[2], x[3]);
    double2 y0 = cadd(x[0], cadd(a1, a2));
    double2 r1 = make_double2(x[0].x + c1 * a1.x + c2 * a2.x, x[0].y + c1 * a1.y + c2 * a2.y);
    double2 r2 = make_double2(x[0].x + c2 * a1.x + c1 * a2.x, x[0].y + c2 * a1.y + c1 * a2.y);
    // i (s1 b1 + s2 b2), i (s2 b1 - s1 b2)
    double2 i1 = make_double2(-(s1 * b1.y + s2 * b2.y), s1 * b1.x + s2 * b2.x);
    double2 i2 = make_double2(-(s2 * b1.y - s1 * b2.y), s2 * b1.x - s1 * b2.x);
    x[0] = y0;
    x[1] = cadd(r1, i1);
    x[4] = csub(r1, i1);
    x[2] = cadd(r2, i2);
    x[3] = csub(r2, i2);
}

// one Stockham pass over nrows rows of M complex values at buf[row*ldr + k]
template <int R>
__device__ void pass(double2 *buf, int ldr, int nrows, int M, int L, const double2 *tw, int sign)
{
    const int T = M / R;
    const int total = nrows * T;
    const int tstride = M / (L * R);
    double2 y[MAXB][R];
#pragma unroll
    for (int b = 0; b < MAXB; ++b) {
        int idx = threadIdx.x + b * NTH;
        if (idx < total) {
            int row = idx / T, j = idx - row * T;
            int k = j % L;
            double2 x[R];
#pragma unroll
            for (int s = 0; s < R; ++s) {
                double2 v = buf[row * ldr + j + s * T];
                if (s > 0) {
                    double2 w = tw[s * k * tstride];
                    if (sign > 0) w = cconj(w);
                    v = cmul(v, w);
                }
                x[s] = v;
            }
            dft<R>(x, sign);
#pragma unroll
            for (int s = 0; s < R; ++s) y[b][s] = x[s];
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < MAXB; ++b) {
        int idx = threadIdx.x + b * NTH;
        if (idx < total) {
            int row = idx / T, j = idx - row * T;
            int k = j % L;
            int base = (j - k) * R + k;
#pragma unroll
            for (int s = 0; s < R; ++s) buf[row * ldr + base + s * L] = y[b][s];
        }
    }
    __syncthreads();
}

__device__ void fft_rows(double2 *buf, int ldr, int nrows, const FFTArgs &f, int sign)
{
    int L = 1;
    for (int s = 0; s < f.nfac; ++s) {
        int R = f.fac[s];
        switch (R) {
        case 4: pass<4>(buf, ldr, nrows, f.M, L, f.tw, sign); break;
        case 2: pass<2>(buf, ldr, nrows, f.M, L, f.tw, sign); break;
        case 3: pass<3>(buf, ldr, nrows, f.M, L, f.tw, sign); break;
        default: pass<5>(buf, ldr, nrows, f.M, L, f.tw, sign); break;
        }
        L *= R;
    }
}

// inverse: F[lev][comp][k][theta] (k < n_m) -> grid[lev][comp][phi][theta]
__global__ __launch_bounds__(NTH) void k_ifft(FFTArgs f, const double *__restrict__ F, size_t F_ld,
                                              int n_m, int n_theta, int ncomp, int TB, int ntb,
                                              double *__restrict__ g, size_t g_ld)
{
    extern __shared__ double2 sbuf[];
    const int M = f.M, N = f.N, ldr = M + 1;
    int b = blockIdx.x;
    int tb = b % ntb;
    int comp = (b / ntb) % ncomp;
    int lev = b / (ntb * ncomp);
    int th0 = tb * TB;
    int nrows = min(TB, n_theta - th0);
    const double2 *Fp = reinterpret_cast<const double2 *>(F) + (size_t)lev * F_ld + (size_t)comp * n_m * n_theta;
    // load Hermitian half spectrum, zero beyond n_m
    for (int idx = threadIdx.x; idx < (M + 1) * TB; idx += NTH) {
        int r = idx % TB, k = idx / TB;
        double2 v = make_double2(0.0, 0.0);
        if (r < nrows && k < n_m) v = Fp[(size_t)k * n_theta + th0 + r];
        sbuf[r * ldr + k] = v;
    }
    __syncthreads();
    // Z_k = (X_k + conj X_{M-k}) + i e^{+2 pi i k/N} (X_k - conj X_{M-k}),
    // real parts only for X_0 and X_M
    const int npair = M / 2 + 1;
    double2 za[MAXB], zb[MAXB];
#pragma unroll
    for (int bb = 0; bb < MAXB; ++bb) {
        int idx = threadIdx.x + bb * NTH;
        if (idx < npair * TB) {
            int r = idx % TB, k = idx / TB, k2 = M - k;
            double2 a = sbuf[r * ldr + k], c = sbuf[r * ldr + k2];
            if (k == 0) { a.y = 0.0; c.y = 0.0; }
            double2 w1 = cconj(f.twr[k]), w2 = cconj(f.twr[k2]);
            za[bb] = cadd(cadd(a, cconj(c)), cmuli(cmul(w1, csub(a, cconj(c))), 1));
            zb[bb] = cadd(cadd(c, cconj(a)), cmuli(cmul(w2, csub(c, cconj(a))), 1));
        }
    }
    __syncthreads();
#pragma unroll
    for (int bb = 0; bb < MAXB; ++bb) {
        int idx = threadIdx.x + bb * NTH;
        if (idx < npair * TB) {
            int r = idx % TB, k = idx / TB, k2 = M - k;
            sbuf[r * ldr + k] = za[bb];
            if (k2 < M && k2 != k) sbuf[r * ldr + k2] = zb[bb];
        }
    }
    __syncthreads();
    fft_rows(sbuf, ldr, TB, f, +1);
    double *gp = g + (size_t)lev * g_ld + (size_t)comp * N * n_theta;
    for (int idx = threadIdx.x; idx < N * TB; idx += NTH) {
        int r = idx % TB, j = idx / TB;
        if (r < nrows) {
            double2 z = sbuf[r * ldr + (j >> 1)];
            gp[(size_t)j * n_theta + th0 + r] = (j & 1) ? z.y : z.x;
        }
    }
}

// forward: grid -> F (k < n_m), scale 1/N, times 1/sin^2(theta) when the
// comp's bit is set in osin2_mask, modes m > m_lim[lev] zeroed (m_lim
// nullptr: none)
__global__ __launch_bounds__(NTH) void k_fft(FFTArgs f, const double *__restrict__ g, size_t g_ld,
                                             int n_m, int n_theta, int ncomp, int TB, int ntb,
                                             unsigned osin2_mask, const double *__restrict__ osin2,
                                             double *__restrict__ F, size_t F_ld)
{
    extern __shared__ double2 sbuf[];
    const int M = f.M, N = f.N, ldr = M + 1;
    int b = blockIdx.x;
    int tb = b % ntb;
    int comp = (b / ntb) % ncomp;
    int lev = b / (ntb * ncomp);
    int th0 = tb * TB;
    int nrows = min(TB, n_theta - th0);
    const double *gp = g + (size_t)lev * g_ld + (size_t)comp * N * n_theta;
    for (int idx = threadIdx.x; idx < N * TB; idx += NTH) {
        int r = idx % TB, j = idx / TB;
        double v = (r < nrows) ? gp[(size_t)j * n_theta + th0 + r] : 0.0;
        double *z = reinterpret_cast<double *>(&sbuf[r * ldr + (j >> 1)]);
        z[j & 1] = v;
    }
    __syncthreads();
    fft_rows(sbuf, ldr, TB, f, -1);
    const bool os = (osin2_mask >> comp) & 1u;
    const double invN = 1.0 / (double)N;
    double2 *Fp = reinterpret_cast<double2 *>(F) + (size_t)lev * F_ld + (size_t)comp * n_m * n_theta;
    for (int idx = threadIdx.x; idx < n_m * TB; idx += NTH) {
        int r = idx % TB, k = idx / TB;
        if (r >= nrows) continue;
        double2 zk = sbuf[r * ldr + (k == M ? 0 : k)];
        double2 zc = cconj(sbuf[r * ldr + (k == 0 ? 0 : M - k)]);
        double2 E = make_double2(0.5 * (zk.x + zc.x), 0.5 * (zk.y + zc.y));
        double2 D = csub(zk, zc);                       // 2 i O
        double2 O = make_double2(0.5 * D.y, -0.5 * D.x); // D / (2i)
        double2 X = cadd(E, cmul(f.twr[k], O));
        double sc = invN;
        if (os) sc *= osin2[th0 + r];
        X.x *= sc;
        X.y *= sc;
        if (k == 0) X.y = 0.0;
        Fp[(size_t)k * n_theta + th0 + r] = X;
    }
}

// Boussinesq bulk-level products, one point per thread.
// gin comps: vr vt vp cvr cvt cvp dvr dvt dvp br bt bp cbr cbt cbp s dsdt dsdp
// gout comps: Advr Advt Advp VSr VSt VSp VxBr VxBt VxBp uGs
__global__ __launch_bounds__(NTH) void k_nl_products(int nlev, int n_theta, int n_phi,
                                                     const double *__restrict__ gin,
                                                     double *__restrict__ gout,
                                                     const double *__restrict__ osin2,
                                                     const double *__restrict__ rad, double LFfac)
{
    const size_t npts = (size_t)n_theta * n_phi;
    size_t i = (size_t)blockIdx.x * NTH + threadIdx.x;
    int lev = blockIdx.y;
    if (i >= npts || lev >= nlev) return;
    const double *a = gin + (size_t)lev * 18 * npts + i;
    double *o = gout + (size_t)lev * 10 * npts + i;
    const double or2 = rad[lev], or4 = rad[nlev + lev], orho1 = rad[2 * nlev + lev];
    const double o2 = osin2[i % n_theta];
    double vr = a[0], vt = a[npts], vp = a[2 * npts];
    double cvr = a[3 * npts], cvt = a[4 * npts], cvp = a[5 * npts];
    double br = a[9 * npts], bt = a[10 * npts], bp = a[11 * npts];
    double cbr = a[12 * npts], cbt = a[13 * npts], cbp = a[14 * npts];
    double s = a[15 * npts], dsdt = a[16 * npts], dsdp = a[17 * npts];
    // src/get_nl.f90:262-273 (curl form) + :242-258 (Lorentz force),
    // merged as src/rIter.f90:650-656
    o[0] = -o2 * (cvt * vp - cvp * vt) + LFfac * o2 * (cbt * bp - cbp * bt);
    o[npts] = -or4 * (cvp * vr - cvr * vp) + LFfac * or4 * (cbp * br - cbr * bp);
    o[2 * npts] = -or4 * (cvr * vt - cvt * vr) + LFfac * or4 * (cbr * bt - cbt * br);
    // :320-325
    o[3 * npts] = vr * s;
    o[4 * npts] = or2 * vt * s;
    o[5 * npts] = or2 * vp * s;
    // :376-388
    o[6 * npts] = orho1 * o2 * (vt * bp - vp * bt);
    o[7 * npts] = orho1 * or4 * (vp * br - vr * bp);
    o[8 * npts] = orho1 * or4 * (vr * bt - vt * br);
    // horizontal advection of s
    o[9 * npts] = or2 * o2 * (vt * dsdt + vp * dsdp);
}

// axisymmetric transforms: src/shtransforms.f90:374-403 (tor = 0) and
// :405-491 (tor = 1), one thread per northern theta
__global__ void k_axi(int l_max, int nh, const double *__restrict__ plm_syn,
                      const double *__restrict__ dplm_syn, const size_t *__restrict__ off, int nt16,
                      const double *__restrict__ fl, int lcut, int tor, double *__restrict__ f,
                      double *__restrict__ f2)
{
    int th = blockIdx.x * blockDim.x + threadIdx.x;
    if (th >= nh) return;
    double es = 0.0, ea = 0.0;  // real parts only (outputs are real)
    double n1 = 0.0, s1 = 0.0;
    for (int l = 0; l <= l_max; ++l) {
        int par = l & 1, kk = l >> 1;
        size_t o = off[par] + ((size_t)(kk >> 2) * nt16 + (th >> 4)) * 64 + (th & 15) + 16 * (kk & 3);
        if (!tor) {
            double v = fl[2 * l] * plm_syn[o];
            if (par == 0) es += v; else ea += v;
        } else if (l <= lcut) {
            // bhG = -i T, bhC = i T, PlmG = PlmC = dP: vt = 0, vp = -T dP
            double v = -fl[2 * l] * dplm_syn[o];
            n1 += v;
            if (par == 0) s1 -= v; else s1 += v;
        }
    }
    if (!tor) {
        f[2 * th] = es + ea;
        f[2 * th + 1] = es - ea;
    } else {
        f[2 * th] = 0.0;
        f[2 * th + 1] = 0.0;
        f2[2 * th] = n1;
        f2[2 * th + 1] = s1;
    }
}

FFTArgs make_args(const Plan *p)
{
    FFTArgs f;
    f.M = p->M;
    f.N = p->n_phi;
    f.nfac = (int)p->fac.size();
    for (int i = 0; i < f.nfac && i < 16; ++i) f.fac[i] = p->fac[i];
    f.tw = reinterpret_cast<const double2 *>(p->d_tw);
    f.twr = reinterpret_cast<const double2 *>(p->d_twr);
    return f;
}

int rows_per_block(const Plan *p)
{
    // LDS: TB (M+1) complex; registers: butterflies per pass TB M / 2 and
    // pre-processing pairs TB (M/2 + 1) must fit NTH * MAXB
    int TB = 16;
    while (TB > 1 && ((size_t)TB * (p->M + 1) * sizeof(double2) > 64 * 1024 ||
                      TB * p->M > 2 * NTH * MAXB || TB * (p->M / 2 + 1) > NTH * MAXB))
        TB >>= 1;
    return TB;
}

}  // namespace

int launch_ifft(const Plan *p, const double *F, size_t F_ld_lev, int ncomp, int nlev, double *grid,
                size_t grid_ld_lev, hipStream_t s)
{
    if (nlev <= 0) return MHX_OK;
    FFTArgs f = make_args(p);
    int TB = rows_per_block(p);
    int ntb = (p->n_theta + TB - 1) / TB;
    size_t shm = (size_t)TB * (p->M + 1) * sizeof(double2);
    dim3 grid_dim((unsigned)(ntb * ncomp * nlev));
    k_ifft<<<grid_dim, NTH, shm, s>>>(f, F, F_ld_lev, p->n_m, p->n_theta, ncomp, TB, ntb, grid,
                                      grid_ld_lev);
    MHX_HIP_CHECK(hipGetLastError());
    return MHX_OK;
}

int launch_fft(const Plan *p, const double *grid, size_t grid_ld_lev, int ncomp, int nlev,
               unsigned osin2_mask, double *F, size_t F_ld_lev, hipStream_t s)
{
    if (nlev <= 0) return MHX_OK;
    FFTArgs f = make_args(p);
    int TB = rows_per_block(p);
    int ntb = (p->n_theta + TB - 1) / TB;
    size_t shm = (size_t)TB * (p->M + 1) * sizeof(double2);
    dim3 grid_dim((unsigned)(ntb * ncomp * nlev));
    k_fft<<<grid_dim, NTH, shm, s>>>(f, grid, grid_ld_lev, p->n_m, p->n_theta, ncomp, TB, ntb,
                                     osin2_mask, p->d_osin2, F, F_ld_lev);
    MHX_HIP_CHECK(hipGetLastError());
    return MHX_OK;
}

int launch_nl_products(const Plan *p, int nlev, const double *gin, double *gout, const double *rad,
                       double LFfac, hipStream_t s)
{
    if (nlev <= 0) return MHX_OK;
    size_t npts = (size_t)p->n_theta * p->n_phi;
    dim3 grid_dim((unsigned)((npts + NTH - 1) / NTH), (unsigned)nlev);
    k_nl_products<<<grid_dim, NTH, 0, s>>>(nlev, p->n_theta, p->n_phi, gin, gout, p->d_osin2, rad,
                                           LFfac);
    MHX_HIP_CHECK(hipGetLastError());
    return MHX_OK;
}

int launch_axi(const Plan *p, const double *fl, double *f, int lcut, int tor, double *f2,
               hipStream_t s)
{
    k_axi<<<(p->nh + 127) / 128, 128, 0, s>>>(p->l_max, p->nh, p->d_syn_P, p->d_syn_dP,
                                               p->d_syn_off, p->nt16, fl, lcut, tor, f, f2);
    MHX_HIP_CHECK(hipGetLastError());
    return MHX_OK;
}

}  // namespace mhx
