// phi FFTs, the fused grid stage of the nonlinear step, and the
// axisymmetric transforms.
//
// Real FFTs of length N = n_phi are computed as complex FFTs of length
// M = N/2 in LDS (Stockham autosort, ping-pong between two buffers, radices
// 4/2/3/5 = the radix set of src/fft.f90:67-73) plus the half-length
// pre/post step.  Normalisation as the reference: forward
// F_k = (1/N) sum_j x_j e^{-2 pi i jk/N} (src/fft.f90:262-268,464), inverse
// the unnormalised Hermitian synthesis using only the real parts of modes 0
// and N/2 (src/fft.f90:386-398).
//
// Fourier layouts: the inverse side F[lev][comp][m][theta] (written in
// contiguous theta runs by the synthesis), the forward side
// F2[lev][comp][theta][m] (written in contiguous rows by the FFT kernels);
// m < n_m, modes up to N/2 beyond n_m are zero (src/shtransforms.f90:224-232).
#include "mhx_fftcore.h"

namespace mhx {

namespace {

// ---------------------------------------------------------------------------
// API path: Fourier rows <-> theta-fastest grids (Fortran f(n_theta, n_phi))
// WG: one (lev, comp) and TB consecutive theta rows.
template <typename T>
__global__ void k_ifft_rows(FFTArgs f, const T *__restrict__ F, size_t F_ld, int F_ncomp, int n_m,
                            int n_theta, int ncomp, int TB, int ntb, T *__restrict__ g,
                            size_t g_ld_lev, size_t g_ld_comp)
{
    typedef cpx<T> C;
    extern __shared__ char smem_raw[];
    C *s_tw = reinterpret_cast<C *>(smem_raw);
    C *s_twr = s_tw + f.M;
    C *A = s_twr + f.M + 1;
    C *B = A + (size_t)TB * f.M;
    const int M = f.M, N = f.N;
    int b = blockIdx.x;
    int tb = b % ntb;
    int comp = (b / ntb) % ncomp;
    int lev = b / (ntb * ncomp);
    int th0 = tb * TB;
    int nrows = min(TB, n_theta - th0);
    load_twiddles(s_tw, s_twr, f);
    __syncthreads();
    // inverse-side Fourier layout F[lev][comp][m][theta]; one thread team
    // per theta row for the pre-step
    const C *F0 = reinterpret_cast<const C *>(F) + (size_t)lev * F_ld + (size_t)comp * n_m * n_theta + th0;
    {
        int per = blockDim.x / TB;  // threads per row
        int r = threadIdx.x / per, t = threadIdx.x % per;
        if (r < TB) {
            if (r < nrows) irfft_pre(F0 + r, (size_t)n_theta, n_m, A + r * M, M, s_twr, t, per);
            else
                for (int k = t; k < M; k += per) A[r * M + k] = mk<C>(T(0), T(0));
        }
    }
    __syncthreads();
    C *res = fft_pp(A, B, TB, f, s_tw, +1);
    T *gp = g + (size_t)lev * g_ld_lev + (size_t)comp * g_ld_comp;
    for (int idx = threadIdx.x; idx < N * TB; idx += blockDim.x) {
        int r = idx % TB, j = idx / TB;
        if (r < nrows) {
            C z = res[r * M + (j >> 1)];
            gp[(size_t)j * n_theta + th0 + r] = (j & 1) ? z.y : z.x;
        }
    }
}

template <typename T>
__global__ void k_fft_rows(FFTArgs f, const T *__restrict__ g, size_t g_ld_lev, size_t g_ld_comp,
                           int n_m, int n_theta, int ncomp, int TB, int ntb, unsigned osin2_mask,
                           const double *__restrict__ osin2, T *__restrict__ F, size_t F_ld,
                           int blk)
{
    typedef cpx<T> C;
    extern __shared__ char smem_raw[];
    C *s_tw = reinterpret_cast<C *>(smem_raw);
    C *s_twr = s_tw + f.M;
    C *A = s_twr + f.M + 1;
    C *B = A + (size_t)TB * f.M;
    const int M = f.M, N = f.N;
    int b = blockIdx.x;
    int tb = b % ntb;
    int comp = (b / ntb) % ncomp;
    int lev = b / (ntb * ncomp);
    int th0 = tb * TB;
    int nrows = min(TB, n_theta - th0);
    load_twiddles(s_tw, s_twr, f);
    const T *gp = g + (size_t)lev * g_ld_lev + (size_t)comp * g_ld_comp;
    for (int idx = threadIdx.x; idx < N * TB; idx += blockDim.x) {
        int r = idx % TB, j = idx / TB;
        T v = (r < nrows) ? gp[(size_t)j * n_theta + th0 + r] : T(0);
        T *z = reinterpret_cast<T *>(&A[r * M + (j >> 1)]);
        z[j & 1] = v;
    }
    __syncthreads();
    C *res = fft_pp(A, B, TB, f, s_tw, -1);
    const bool os = (osin2_mask >> comp) & 1u;
    // F[lev][comp][theta][m], or theta-blocked F[lev][comp][theta/16][m][theta%16]
    // (blk; the split analysis reads 16 rows of one m as one line)
    C *Fl = reinterpret_cast<C *>(F) + (size_t)lev * F_ld;
    const int nt16 = (n_theta + 15) / 16;
    for (int idx = threadIdx.x; idx < n_m * TB; idx += blockDim.x) {
        int k = idx % n_m, r = idx / n_m;
        if (r >= nrows) continue;
        T sc = T(1.0 / (double)N);
        if (os) sc *= T(osin2[th0 + r]);
        const int th = th0 + r;
        const size_t o = blk ? (((size_t)comp * nt16 + (th >> 4)) * n_m + k) * 16 + (th & 15)
                             : ((size_t)comp * n_theta + th) * n_m + k;
        Fl[o] = rfft_post(res + r * M, k, M, s_twr, sc);
    }
}

// ---------------------------------------------------------------------------
// Fused grid stage of the nonlinear step: for one radial level and one theta
// row, inverse FFTs of the 18 synthesised components, the grid products
// (src/get_nl.f90:242-273 curl advection + Lorentz force merged as
// src/rIter.f90:650-656, :320-325 V s, :376-388 V x B, and u_h . grad_h s),
// and forward FFTs of the 10 products.  Grid values live in registers (P
// points per thread); only Fourier rows are read and written.
//
// Fin comps (layout [lev][comp][m][theta]):
//             vr vt vp | cvr cvt cvp | dvr dvt dvp | br bt bp | cbr cbt cbp |
//             s dsdt dsdp
// Fout (layout [lev][comp][theta][m]):
//             Advr Advt Advp | VSr VSt VSp | VxBr VxBt VxBp | uGs
// du/dr grids (not consumed by the products) are written to dgrid
// [lev][3][theta][phi]; with full_grids != nullptr all 18 inverse grids are
// also written in the reference layout [lev][comp][phi][theta].
constexpr int PMAX = 3;

template <typename T, int G>
__global__ void k_nl_fused(FFTArgs f, const T *__restrict__ Fin, T *__restrict__ Fout, int n_m,
                           int n_theta, int nlev, const double *__restrict__ osin2,
                           const double *__restrict__ rad, double LFfac_d,
                           T *__restrict__ dgrid, T *__restrict__ full_grids, CourArgs cour)
{
    // split form (T = float): theta-blocked forward side
    constexpr bool BLK = sizeof(T) == 4;
    typedef cpx<T> C;
    extern __shared__ char smem_raw[];
    const int M = f.M, N = f.N;
    C *s_tw = reinterpret_cast<C *>(smem_raw);
    C *s_twr = s_tw + M;
    C *A = s_twr + M + 1;
    C *B = A + (size_t)G * M;
    // XCD-aware order: consecutive theta rows (which share the cache lines
    // of F[lev][comp][m][theta]) run on the same XCD at the same time
    int lb;
    {
        const int nx = 8, nb = gridDim.x, b = blockIdx.x;
        int per = nb / nx, rem = nb % nx, x = b % nx, i = b / nx;
        lb = (x < rem ? x * (per + 1) : rem * (per + 1) + (x - rem) * per) + i;
    }
    const int th = lb % n_theta, lev = lb / n_theta;
    const int tid = threadIdx.x, nthr = blockDim.x;
    const T o2 = T(osin2[th]);
    const T or2 = T(rad[lev]), or4 = T(rad[nlev + lev]), orho1 = T(rad[2 * nlev + lev]);
    const T LFfac = T(LFfac_d);
    load_twiddles(s_tw, s_twr, f);
    __syncthreads();

    const size_t rowsz = (size_t)n_m;                   // complex per forward row
    const int ntb = (n_theta + 15) / 16;
    const C *Fi = reinterpret_cast<const C *>(Fin) + (size_t)lev * 18 * n_theta * rowsz + th;
    C *Fo = reinterpret_cast<C *>(Fout) + (size_t)lev * 10 * (BLK ? ntb * 16 : n_theta) * rowsz;

    T u[3][PMAX], adv[3][PMAX], bb[3][PMAX], vxb[3][PMAX], vs[3][PMAX], ugs[PMAX];
    T val[3][PMAX];

    // inverse FFT of comps c0..c0+2 (in groups of G rows) into val
    auto inverse3 = [&](int c0) {
#pragma unroll
        for (int g0 = 0; g0 < 3; g0 += G) {
            constexpr int gn = G;
            int per = nthr / gn;
            int r = tid / per, t = tid % per;
            if (r < gn)
                irfft_pre(Fi + (size_t)(c0 + g0 + r) * n_m * n_theta, (size_t)n_theta, n_m,
                          A + r * M, M, s_twr, t, per);
            __syncthreads();
            C *res = fft_pp(A, B, gn, f, s_tw, +1);
#pragma unroll
            for (int p = 0; p < PMAX; ++p) {
                int j = tid + p * nthr;
#pragma unroll
                for (int r2 = 0; r2 < gn; ++r2) {
                    T v = T(0);
                    if (j < N) {
                        C z = res[r2 * M + (j >> 1)];
                        v = (j & 1) ? z.y : z.x;
                        int comp = c0 + g0 + r2;
                        if (full_grids)
                            full_grids[(((size_t)lev * 18 + comp) * N + j) * n_theta + th] = v;
                        if (comp >= 6 && comp < 9 && dgrid)
                            dgrid[(((size_t)lev * 3 + comp - 6) * n_theta + th) * N + j] = v;
                    }
                    val[g0 + r2][p] = v;
                }
            }
            __syncthreads();
        }
    };
    // forward FFT of 3 (or 1) product rows given per point, rows scaled by sc[]
    auto forward = [&](T (*src)[PMAX], int ncomp, int c0, const T *sc) {
#pragma unroll
        for (int g0 = 0; g0 < 3; g0 += G) {
            if (g0 >= ncomp) break;
            const int gn = min(G, ncomp - g0);
#pragma unroll
            for (int p = 0; p < PMAX; ++p) {
                int j = tid + p * nthr;
                if (j < N)
#pragma unroll
                    for (int r2 = 0; r2 < G; ++r2) {
                        if (r2 >= gn) break;
                        T *z = reinterpret_cast<T *>(&A[r2 * M + (j >> 1)]);
                        z[j & 1] = src[g0 + r2][p] * sc[g0 + r2];
                    }
            }
            __syncthreads();
            C *res = fft_pp(A, B, gn, f, s_tw, -1);
            const T invN = T(1.0 / (double)N);
            for (int idx = tid; idx < gn * n_m; idx += nthr) {
                int r2 = idx / n_m, k = idx - r2 * n_m;
                C v = rfft_post(res + r2 * M, k, M, s_twr, invN);
                if constexpr (BLK)
                    Fo[(((size_t)(c0 + g0 + r2) * ntb + (th >> 4)) * rowsz + k) * 16 + (th & 15)] = v;
                else
                    Fo[((size_t)(c0 + g0 + r2) * n_theta + th) * rowsz + k] = v;
            }
            __syncthreads();
        }
    };

    // u
    inverse3(0);
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int p = 0; p < PMAX; ++p) u[c][p] = val[c][p];
    // curl u: Adv = -curl(u) x u terms (src/get_nl.f90:262-273)
    inverse3(3);
#pragma unroll
    for (int p = 0; p < PMAX; ++p) {
        T cvr = val[0][p], cvt = val[1][p], cvp = val[2][p];
        adv[0][p] = -o2 * (cvt * u[2][p] - cvp * u[1][p]);
        adv[1][p] = -or4 * (cvp * u[0][p] - cvr * u[2][p]);
        adv[2][p] = -or4 * (cvr * u[1][p] - cvt * u[0][p]);
    }
    // du/dr: written to dgrid only
    inverse3(6);
    // B: V x B (src/get_nl.f90:376-388)
    inverse3(9);
#pragma unroll
    for (int p = 0; p < PMAX; ++p) {
        bb[0][p] = val[0][p];
        bb[1][p] = val[1][p];
        bb[2][p] = val[2][p];
        vxb[0][p] = orho1 * o2 * (u[1][p] * bb[2][p] - u[2][p] * bb[1][p]);
        vxb[1][p] = orho1 * or4 * (u[2][p] * bb[0][p] - u[0][p] * bb[2][p]);
        vxb[2][p] = orho1 * or4 * (u[0][p] * bb[1][p] - u[1][p] * bb[0][p]);
    }
    if (cour.max) courant_max<T, PMAX>(cour, lev, nlev, o2, or2, or4, orho1, LFfac, u, bb);
    // curl B: Lorentz force added to Adv (src/get_nl.f90:242-258,
    // src/rIter.f90:650-656)
    inverse3(12);
#pragma unroll
    for (int p = 0; p < PMAX; ++p) {
        T cbr = val[0][p], cbt = val[1][p], cbp = val[2][p];
        adv[0][p] += LFfac * o2 * (cbt * bb[2][p] - cbp * bb[1][p]);
        adv[1][p] += LFfac * or4 * (cbp * bb[0][p] - cbr * bb[2][p]);
        adv[2][p] += LFfac * or4 * (cbr * bb[1][p] - cbt * bb[0][p]);
    }
    {
        const T sc[3] = {T(1), o2, o2};  // 1/sin^2 on the horizontal comps
        forward(adv, 3, 0, sc);
        forward(vxb, 3, 6, sc);
    }
    // s, ds/dtheta, ds/dphi: V s (src/get_nl.f90:320-325), u_h . grad_h s
    inverse3(15);
#pragma unroll
    for (int p = 0; p < PMAX; ++p) {
        T s = val[0][p];
        vs[0][p] = u[0][p] * s;
        vs[1][p] = or2 * u[1][p] * s;
        vs[2][p] = or2 * u[2][p] * s;
        ugs[p] = or2 * o2 * (u[1][p] * val[1][p] + u[2][p] * val[2][p]);
    }
    {
        const T sc[3] = {T(1), o2, o2};
        forward(vs, 3, 3, sc);
        T tmp[1][PMAX];
#pragma unroll
        for (int p = 0; p < PMAX; ++p) tmp[0][p] = ugs[p];
        const T one[1] = {T(1)};
        forward(tmp, 1, 9, one);
    }
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

// ---------------------------------------------------------------------------
// Fused grid stage, slot form.  Each complex FFT of length M = P Q is done
// by one "slot" of Q lanes inside a single wave (Q divides 64), every lane
// holding P points (P = 12, or 24 for M = 1536 where 12 points per lane
// would need 128 lanes); the Stockham passes (radices 12, 4 and 2, each
// dividing P) run in place in the slot's LDS row with no block barrier:
// within a wave the LDS operations execute in program order, and each pass
// loads all of its inputs before it stores any output.  Six slots work on one
// theta row: the 18 inverse FFTs in three rounds of six, the 10 forward FFTs
// in two.  Between rounds the block redistributes the row so that each thread
// holds P / 3 grid points of every component for the products.


template <typename T, int A, int B>
__global__ __launch_bounds__(384) void k_nl_slots(FFTArgs f, const T *__restrict__ Fin,
                                                  T *__restrict__ Fout, int n_m, int n_theta,
                                                  int nlev, const double *__restrict__ osin2,
                                                  const double *__restrict__ rad, double LFfac_d,
                                                  T *__restrict__ dgrid, T *__restrict__ full_grids,
                                                  CourArgs cour)
{
    typedef cpx<T> C;
    constexpr bool BLK = sizeof(T) == 4;  // split form: theta-blocked F2
    extern __shared__ char smem_raw[];
    typedef SlotFFT<A, B> FFT;
    constexpr int M = FFT::M, N = 2 * M, Q = FFT::Q;
    constexpr int NPT = FFT::P / 3;  // grid points per thread: N = NPT * 6 Q
    C *s_tw = reinterpret_cast<C *>(smem_raw);  // M + M / 8 (padded)
    C *s_twr = s_tw + M + M / 8;
    C *rows = s_twr + M + 1;  // 6 slot rows of MP (padded)
    constexpr int MP = padded_row<M, FFT::P>();
    int lb;
    {
        const int nx = 8, nb = gridDim.x, b = blockIdx.x;
        int per = nb / nx, rem = nb % nx, x = b % nx, i = b / nx;
        lb = (x < rem ? x * (per + 1) : rem * (per + 1) + (x - rem) * per) + i;
    }
    const int th = lb % n_theta, lev = lb / n_theta;
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int slot = tid / Q, t = tid - slot * Q;
    const T o2 = T(osin2[th]);
    const T or2 = T(rad[lev]), or4 = T(rad[nlev + lev]), orho1 = T(rad[2 * nlev + lev]);
    const T LFfac = T(LFfac_d);
    {
        const C *tw = reinterpret_cast<const C *>(f.tw);
        const C *twr = reinterpret_cast<const C *>(f.twr);
        for (int e = threadIdx.x; e < M; e += blockDim.x) s_tw[e + (e >> 3)] = tw[e];
        for (int e = threadIdx.x; e <= M; e += blockDim.x) s_twr[e] = twr[e];
    }
    __syncthreads();

    const int ntb = (n_theta + 15) / 16;
    const C *Fi = reinterpret_cast<const C *>(Fin) + (size_t)lev * 18 * n_theta * n_m + th;
    C *Fo = reinterpret_cast<C *>(Fout) + (size_t)lev * 10 * (BLK ? ntb * 16 : n_theta) * n_m;

    T u[3][NPT], adv[3][NPT], vxb[3][NPT];
    T val[6][NPT];

    // inverse FFTs of six comps into val: round 0 = u, curl u (comps 0-5),
    // round 1 = B, curl B (9-14), round 2 = du/dr (6-8), s, grad s (15-17)
    // without du/dr grids (dgrid null) round 2 transforms s and grad s only
    // (the products read no du/dr)
    auto inverse6 = [&](int round) {
        if (slot < 6 && (round != 2 || slot >= 3 || dgrid)) {
            C *z = rows + slot * MP;
            const int comp = round == 0 ? slot : (round == 1 ? 9 + slot : (slot < 3 ? 6 + slot : 12 + slot));
            const C *X = Fi + (size_t)comp * n_m * n_theta;
            // Z_k = (X_k + conj X_{M-k}) + i e^{+2 pi i k/N} (X_k - conj X_{M-k})
            if constexpr (FFT::P == 24) {
                // M = 1536: the lane's k = t + i Q <= M/2 take NK = 13 steps;
                // all their loads are issued first, at clamped addresses
                // (measured faster there; at 12 points per lane the guarded
                // loads one step at a time are faster: they skip the
                // coefficients past n_m)
                constexpr int NK = M / (2 * Q) + 1;
                C xa[NK], xc[NK];
#pragma unroll
                for (int i = 0; i < NK; ++i) {
                    const int k = t + i * Q;
                    xa[i] = X[(size_t)min(k, n_m - 1) * n_theta];
                    xc[i] = X[(size_t)min(M - k, n_m - 1) * n_theta];
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int i = 0; i < NK; ++i) {
                    const int k = t + i * Q, k2 = M - k;
                    if (k > M / 2) continue;
                    C a = (k < n_m) ? xa[i] : mk<C>(T(0), T(0));
                    C c = (k2 < n_m) ? xc[i] : mk<C>(T(0), T(0));
                    if (k == 0) {
                        a.y = T(0);
                        c.y = T(0);
                    }
                    z[pidx<FFT::P>(k)] = cadd(cadd(a, cconj(c)), cmuli(cmul(cconj(s_twr[k]), csub(a, cconj(c))), 1));
                    if (k2 < M && k2 != k)
                        z[pidx<FFT::P>(k2)] = cadd(cadd(c, cconj(a)), cmuli(cmul(cconj(s_twr[k2]), csub(c, cconj(a))), 1));
                }
            } else {
                // pointers stepped by Q rows (no index products per step)
                const size_t step = (size_t)Q * n_theta;
                const C *xa = X + (size_t)t * n_theta, *xc = X + (size_t)(M - t) * n_theta;
                for (int k = t; k <= M / 2; k += Q, xa += step, xc -= step) {
                    const int k2 = M - k;
                    C a = (k < n_m) ? *xa : mk<C>(T(0), T(0));
                    C c = (k2 < n_m) ? *xc : mk<C>(T(0), T(0));
                    if (k == 0) {
                        a.y = T(0);
                        c.y = T(0);
                    }
                    z[pidx<FFT::P>(k)] = cadd(cadd(a, cconj(c)), cmuli(cmul(cconj(s_twr[k]), csub(a, cconj(c))), 1));
                    if (k2 < M && k2 != k)
                        z[pidx<FFT::P>(k2)] = cadd(cadd(c, cconj(a)), cmuli(cmul(cconj(s_twr[k2]), csub(c, cconj(a))), 1));
                }
            }
            wave_order();
            FFT::run(z, t, s_tw, +1);
        }
        __syncthreads();
        // thread owns the point pairs (2q, 2q+1), q = tid + p2 nthr
#pragma unroll
        for (int p2 = 0; p2 < NPT / 2; ++p2) {
            const int q = tid + p2 * nthr;
#pragma unroll
            for (int r = 0; r < 6; ++r) {
                C zz = mk<C>(T(0), T(0));
                if (q < M) {
                    zz = rows[r * MP + pidx<FFT::P>(q)];
                    const int comp = round == 0 ? r : (round == 1 ? 9 + r : (r < 3 ? 6 + r : 12 + r));
                    if (full_grids) {
                        T *fg = full_grids + (((size_t)lev * 18 + comp) * N + 2 * q) * n_theta + th;
                        fg[0] = zz.x;
                        fg[n_theta] = zz.y;
                    }
                    if (round == 2 && r < 3 && dgrid)
                        *reinterpret_cast<C *>(dgrid + (((size_t)lev * 3 + r) * n_theta + th) * N + 2 * q) = zz;
                }
                val[r][2 * p2] = zz.x;
                val[r][2 * p2 + 1] = zz.y;
            }
        }
        __syncthreads();
    };

    // forward FFTs of the product rows held in val: round 0 = Adv (F2 comps
    // 0-2) and V x B (6-8), round 1 = V s (3-5) and u_h . grad_h s (9);
    // 1/sin^2 on the horizontal comps of the vector products
    auto forward6 = [&](int round) {
        const int nc = round == 0 ? 6 : 4;
#pragma unroll
        for (int p2 = 0; p2 < NPT / 2; ++p2) {
            const int q = tid + p2 * nthr;
            if (q < M) {
#pragma unroll
                for (int r = 0; r < 6; ++r) {
                    if (r >= nc) break;
                    const bool horiz = (r % 3 != 0) && !(round == 1 && r == 3);
                    const T sc = horiz ? o2 : T(1);
                    rows[r * MP + pidx<FFT::P>(q)] = mk<C>(val[r][2 * p2] * sc, val[r][2 * p2 + 1] * sc);
                }
            }
        }
        __syncthreads();
        if (slot < nc) {
            C *z = rows + slot * MP;
            FFT::run(z, t, s_tw, -1);
            const int comp = round == 0 ? (slot < 3 ? slot : slot + 3) : (slot < 3 ? slot + 3 : 9);
            const T invN = T(1.0 / (double)N);
            // F2 theta-blocked (split form): consecutive k are 16 apart
            constexpr int FS = BLK ? 16 : 1;
            C *fo = BLK ? Fo + (((size_t)comp * ntb + (th >> 4)) * n_m + t) * 16 + (th & 15)
                        : Fo + ((size_t)comp * n_theta + th) * n_m + t;
            for (int k = t; k < n_m; k += Q, fo += Q * FS) *fo = rfft_post_padded<FFT::P>(z, k, M, s_twr, invN);
        }
        __syncthreads();
    };

    // u, curl u: Adv = -curl(u) x u terms (src/get_nl.f90:262-273)
    inverse6(0);
#pragma unroll
    for (int p = 0; p < NPT; ++p) {
#pragma unroll
        for (int c = 0; c < 3; ++c) u[c][p] = val[c][p];
        const T cvr = val[3][p], cvt = val[4][p], cvp = val[5][p];
        adv[0][p] = -o2 * (cvt * u[2][p] - cvp * u[1][p]);
        adv[1][p] = -or4 * (cvp * u[0][p] - cvr * u[2][p]);
        adv[2][p] = -or4 * (cvr * u[1][p] - cvt * u[0][p]);
    }
    // B: V x B (src/get_nl.f90:376-388); curl B: Lorentz force added to Adv
    // (src/get_nl.f90:242-258, src/rIter.f90:650-656)
    inverse6(1);
    if (cour.max) courant_max<T, NPT>(cour, lev, nlev, o2, or2, or4, orho1, LFfac, u, val);
#pragma unroll
    for (int p = 0; p < NPT; ++p) {
        const T br = val[0][p], bt = val[1][p], bp = val[2][p];
        const T cbr = val[3][p], cbt = val[4][p], cbp = val[5][p];
        vxb[0][p] = orho1 * o2 * (u[1][p] * bp - u[2][p] * bt);
        vxb[1][p] = orho1 * or4 * (u[2][p] * br - u[0][p] * bp);
        vxb[2][p] = orho1 * or4 * (u[0][p] * bt - u[1][p] * br);
        adv[0][p] += LFfac * o2 * (cbt * bp - cbp * bt);
        adv[1][p] += LFfac * or4 * (cbp * br - cbr * bp);
        adv[2][p] += LFfac * or4 * (cbr * bt - cbt * br);
    }
#pragma unroll
    for (int p = 0; p < NPT; ++p)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            val[c][p] = adv[c][p];
            val[3 + c][p] = vxb[c][p];
        }
    forward6(0);
    // du/dr (to dgrid only); s, ds/dtheta, ds/dphi: V s (src/get_nl.f90:320-325)
    // and u_h . grad_h s
    inverse6(2);
#pragma unroll
    for (int p = 0; p < NPT; ++p) {
        const T sv = val[3][p], dst = val[4][p], dsp = val[5][p];
        val[0][p] = u[0][p] * sv;
        val[1][p] = or2 * u[1][p] * sv;
        val[2][p] = or2 * u[2][p] * sv;
        val[3][p] = or2 * o2 * (u[1][p] * dst + u[2][p] * dsp);
    }
    forward6(1);
}

// ---------------------------------------------------------------------------
// Grid stage of the radial loop for any physics switches (mhx_rloop), FP64:
// for one radial level and one theta row, the inverse FFTs of the level's
// synthesised components into LDS rows, the reference's grid products
// (src/get_nl.f90:240-436 for a bulk level, merged as src/rIter.f90:644-680)
// and the Courant maxima (src/courant.f90:209-275) per grid point, and the
// forward FFTs of the products.  Each LDS row holds one component's N = n_phi
// grid values (M complex); the products of a point overwrite its inputs in
// place (one thread owns a point), so the rows hold max(nin, nout) components.
// FFTs run G rows at a time through the block-wide Stockham passes.

// the grid products of one point: v[] the grid components (zero where
// absent), o[] the products
__device__ inline void rl_products(const RlPhys &ph, const double *v, double *o, double o2,
                                   double cosn2, double sinT, double cosT, double osinT,
                                   double phi, double r, double or1, double or2, double or4,
                                   double orho1, double beta, double otemp1, double visc,
                                   double lambda, bool lcr)
{
    const double vr = v[RI_VR], vt = v[RI_VT], vp = v[RI_VP];
#pragma unroll
    for (int k = 0; k < RO_N; ++k) o[k] = 0.0;
    // Lorentz force (:242-258), not below the conducting region
    double LFr = 0.0, LFt = 0.0, LFp = 0.0;
    const bool lf = ph.mag_LF && !lcr;
    if (lf) {
        const double br = v[RI_BR], bt = v[RI_BT], bp = v[RI_BP];
        const double cbr = v[RI_CBR], cbt = v[RI_CBT], cbp = v[RI_CBP];
        LFr = ph.LFfac * o2 * (cbt * bp - cbp * bt);
        LFt = ph.LFfac * or4 * (cbp * br - cbr * bp);
        LFp = ph.LFfac * or4 * (cbr * bt - cbt * br);
    }
    double Ar = 0.0, At = 0.0, Ap = 0.0;
    if (ph.conv_nl) {
        if (ph.adv_curl) {
            // curl(u) x u (:262-273)
            const double cvr = v[RI_CVR], cvt = v[RI_CVT], cvp = v[RI_CVP];
            Ar = -o2 * (cvt * vp - cvp * vt);
            At = -or4 * (cvp * vr - cvr * vp);
            Ap = -or4 * (cvr * vt - cvt * vr);
        } else {
            // u . grad u (:274-308)
            const double dvrdr = v[RI_DVRDR], dvtdr = v[RI_DVTDR], dvpdr = v[RI_DVPDR];
            const double dvrdt = v[RI_DVRDT], dvrdp = v[RI_DVRDP];
            const double dvtdp = v[RI_DVTDP], dvpdp = v[RI_DVPDP], cvr = v[RI_CVR];
            Ar = -or2 * orho1 *
                 (vr * (dvrdr - (2.0 * or1 + beta) * vr) +
                  o2 * (vt * (dvrdt - r * vt) + vp * (dvrdp - r * vp)));
            At = or4 * orho1 *
                 (-vr * (dvtdr - beta * vt) + vt * (cosn2 * vt + dvpdp + dvrdr) +
                  vp * (cosn2 * vp - dvtdp));
            Ap = or4 * orho1 * (-vr * (dvpdr - beta * vp) - vt * (dvtdp + cvr) - vp * dvpdp);
        }
    }
    const double s = v[RI_S];
    if (ph.phase) {
        // phase field (:333-344)
        const double ph2 = 1.0 / (ph.epsPhase * ph.epsPhase) / (ph.penaltyFac * ph.penaltyFac);
        const double pc = v[RI_PHI];
        Ar = Ar - pc * vr * ph2;
        At = At - or2 * pc * vt * ph2;
        Ap = Ap - or2 * pc * vp * ph2;
        o[RO_PHI] = -1.0 / (ph.epsPhase * ph.epsPhase) * pc * (1.0 - pc) *
                    (ph.phaseDiffFac * (1.0 - 2.0 * pc) + s - ph.tmelt);
    }
    // Adv + LF (src/rIter.f90:648-660): LF alone without nonlinear advection
    if (ph.conv_nl) {
        if (lf) {
            Ar += LFr;
            At += LFt;
            Ap += LFp;
        }
    } else if (ph.mag_LF) {
        Ar = LFr;
        At = LFt;
        Ap = LFp;
    }
    if (ph.prec) {
        // precession (:346-356), added as src/rIter.f90:662-666
        const double c = cos(ph.oek * ph.time + phi), sn = sin(ph.oek * ph.time + phi);
        Ar += ph.posnalp * osinT * r * (c * vp * cosT + sn * vt);
        At += -ph.posnalp * sinT * or2 * (c * vp + sn * or1 * vr);
        Ap += ph.posnalp * sinT * c * or2 * (vt - or1 * vr * cosT);
    }
    if (ph.centr) {
        // centrifugal acceleration (:358-374), added as src/rIter.f90:668-671
        const double s2 = sinT * sinT;
        Ar += -ph.dilution_fac * r * (s2 * s2) * ph.ra * ph.opr * s;
        At += -ph.dilution_fac * r * (s2 * sinT) * cosT * ph.ra * ph.opr * s;
    }
    o[RO_ADVR] = Ar;
    o[RO_ADVT] = At;
    o[RO_ADVP] = Ap;
    if (ph.heat_nl) {
        // V s (:320-325)
        o[RO_VSR] = vr * s;
        o[RO_VST] = or2 * vt * s;
        o[RO_VSP] = or2 * vp * s;
    }
    if (ph.chem) {
        // V xi (:327-331)
        const double xi = v[RI_XI];
        o[RO_VXIR] = vr * xi;
        o[RO_VXIT] = or2 * vt * xi;
        o[RO_VXIP] = or2 * vp * xi;
    }
    if (ph.mag_nl) {
        const double br = v[RI_BR], bt = v[RI_BT], bp = v[RI_BP];
        if (!lcr) {
            // V x B (:376-388)
            o[RO_VXBR] = orho1 * o2 * (vt * bp - vp * bt);
            o[RO_VXBT] = orho1 * or4 * (vp * br - vr * bp);
            o[RO_VXBP] = orho1 * or4 * (vr * bt - vt * br);
        } else {
            // below the conducting region (:389-391): the horizontal part only
            o[RO_VXBT] = or4 * orho1 * vp * br;
            o[RO_VXBP] = -or4 * orho1 * vt * br;
        }
    }
    if (ph.anel) {
        // viscous heating and Ohmic losses (:402-436)
        const double dvrdr = v[RI_DVRDR], dvtdr = v[RI_DVTDR], dvpdr = v[RI_DVPDR];
        const double dvrdt = v[RI_DVRDT], dvrdp = v[RI_DVRDP];
        const double dvtdp = v[RI_DVTDP], dvpdp = v[RI_DVPDP], cvr = v[RI_CVR];
        const double t1 = dvrdr - (2.0 * or1 + beta) * vr;
        const double t2 = cosn2 * vt + dvpdp + dvrdr - or1 * vr;
        const double t3 = dvpdp + cosn2 * vt + or1 * vr;
        const double t6 = 2.0 * dvtdp + cvr - 2.0 * cosn2 * vp;
        const double t4 = r * dvtdr - (2.0 + beta * r) * vt + or1 * dvrdt;
        const double t5 = r * dvpdr - (2.0 + beta * r) * vp + or1 * dvrdp;
        const double bv = beta * vr;
        double h = ph.ViscHeatFac * or4 * orho1 * otemp1 * visc *
                   (2.0 * (t1 * t1) + 2.0 * (t2 * t2) + 2.0 * (t3 * t3) + t6 * t6 +
                    o2 * (t4 * t4 + t5 * t5) - 2.0 * (1.0 / 3.0) * (bv * bv));
        if (ph.mag_nl && !lcr) {
            const double cbr = v[RI_CBR], cbt = v[RI_CBT], cbp = v[RI_CBP];
            h = h + ph.OhmLossFac * or2 * otemp1 * lambda *
                        (or2 * (cbr * cbr) + o2 * (cbt * cbt) + o2 * (cbp * cbp));
        }
        o[RO_HEAT] = h;
    }
}

// the get_nl products and Courant maxima of one (level, theta) block, point
// by point in place over the N points of its component rows in LDS; R[idx(c,
// j)] is point j of row c
template <class Idx>
__device__ __forceinline__ void rl_points(const RlGridArgs &a, double *R, Idx idx, int N, int lev, int th)
{
    const int n_theta = a.n_theta, nlev = a.nlev;
    const int tid = threadIdx.x, nthr = blockDim.x;
    const double *rad = a.rad;
    const double r = rad[RR_R * nlev + lev], or1 = rad[RR_OR1 * nlev + lev];
    const double or2 = rad[RR_OR2 * nlev + lev], or4 = rad[RR_OR4 * nlev + lev];
    const double orho1 = rad[RR_ORHO1 * nlev + lev], beta = rad[RR_BETA * nlev + lev];
    const double otemp1 = rad[RR_OTEMP1 * nlev + lev], visc = rad[RR_VISC * nlev + lev];
    const double lambda = rad[RR_LAMBDA * nlev + lev];
    const bool lcr = rad[RR_LCR * nlev + lev] != 0.0;
    const double o2 = a.th[RT_OSIN2 * n_theta + th], cosn2 = a.th[RT_COSN2 * n_theta + th];
    const double sinT = a.th[RT_SIN * n_theta + th], cosT = a.th[RT_COS * n_theta + th];
    const double osinT = a.th[RT_OSIN * n_theta + th];
    double mr = 0.0, mh = 0.0;
    const double vri2 = rad[RR_VALRI2 * nlev + lev], vhi2 = rad[RR_VALHI2 * nlev + lev];
    const double orho2 = orho1 * orho1;
    for (int j = tid; j < N; j += nthr) {
        double v[RI_N], o[RO_N];
#pragma unroll
        for (int k = 0; k < RI_N; ++k) v[k] = a.in_of[k] >= 0 ? R[idx(a.in_of[k], j)] : 0.0;
        rl_products(a.ph, v, o, o2, cosn2, sinT, cosT, osinT, j * a.dphi, r, or1, or2, or4, orho1,
                    beta, otemp1, visc, lambda, lcr);
        if (a.cmax) {
            // src/courant.f90:209-275 (XSH_COURANT = 0)
            const double vr = v[RI_VR], vt = v[RI_VT], vp = v[RI_VP];
            const double vflr2 = orho2 * vr * vr;
            const double vflh2 = (vt * vt + vp * vp) * o2 * orho2;
            if (a.ph.cour_mag) {
                const double br = v[RI_BR], bt = v[RI_BT], bp = v[RI_BP];
                const double valr = br * br * a.ph.LFfac * orho1;
                const double valr2 = valr * valr / (valr + vri2);
                mr = fmax(mr, or4 * (a.ph.cf2 * vflr2 + a.ph.af2 * valr2));
                const double valh2 = (bt * bt + bp * bp) * a.ph.LFfac * o2 * orho1;
                const double valh2m = valh2 * valh2 / (valh2 + vhi2);
                mh = fmax(mh, or2 * (a.ph.cf2 * vflh2 + a.ph.af2 * valh2m));
            } else {
                mr = fmax(mr, a.ph.cf2 * or4 * vflr2);
                mh = fmax(mh, a.ph.cf2 * or2 * vflh2);
            }
        }
#pragma unroll
        for (int k = 0; k < RO_N; ++k) {
            const int row = a.out_of[k];
            if (row >= 0) R[idx(row, j)] = ((a.out_osin2 >> row) & 1u) ? o[k] * o2 : o[k];
        }
    }
    if (a.cmax) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            mr = fmax(mr, __shfl_xor(mr, off, 64));
            mh = fmax(mh, __shfl_xor(mh, off, 64));
        }
        if ((tid & 63) == 0) {
            atomicMax(a.cmax + 2 * lev, (unsigned long long)__double_as_longlong(mr));
            atomicMax(a.cmax + 2 * lev + 1, (unsigned long long)__double_as_longlong(mh));
        }
    }
}

__global__ __launch_bounds__(512) void k_rl_grid(FFTArgs f, RlGridArgs a)
{
    typedef double2 C;
    extern __shared__ char smem_raw[];
    const int M = f.M, N = f.N;
    C *s_tw = reinterpret_cast<C *>(smem_raw);
    C *s_twr = s_tw + M;
    C *B = s_twr + M + 1;             // G scratch rows
    C *rows = B + (size_t)a.G * M;    // max(nin, nout) component rows
    int lb;
    {
        const int nx = 8, nb = gridDim.x, b = blockIdx.x;
        int per = nb / nx, rem = nb % nx, x = b % nx, i = b / nx;
        lb = (x < rem ? x * (per + 1) : rem * (per + 1) + (x - rem) * per) + i;
    }
    const int n_theta = a.n_theta, n_m = a.n_m, nlev = a.nlev;
    const int th = lb % n_theta, lev = lb / n_theta;
    const int tid = threadIdx.x, nthr = blockDim.x;
    load_twiddles(s_tw, s_twr, f);
    __syncthreads();

    // inverse FFTs, G rows at a time, into rows[0, nin)
    const C *Fi = reinterpret_cast<const C *>(a.Fin) + (size_t)lev * a.nin * n_theta * n_m + th;
    for (int g0 = 0; g0 < a.nin; g0 += a.G) {
        const int gn = min(a.G, a.nin - g0);
        const int per = nthr / gn;
        const int r = tid / per, t = tid - r * per;
        if (r < gn)
            irfft_pre(Fi + (size_t)(g0 + r) * n_m * n_theta, (size_t)n_theta, n_m, rows + (size_t)(g0 + r) * M,
                      M, s_twr, t, per);
        __syncthreads();
        C *res = fft_pp(rows + (size_t)g0 * M, B, gn, f, s_tw, +1);
        if (res != rows + (size_t)g0 * M) {
            for (int e = tid; e < gn * M; e += nthr) rows[(size_t)g0 * M + e] = res[e];
            __syncthreads();
        }
    }

    // the products, point by point, in place
    rl_points(a, reinterpret_cast<double *>(rows), [&](int c, int j) { return (size_t)c * N + j; }, N, lev, th);
    __syncthreads();

    // forward FFTs of rows[0, nout) into F2[lev][comp][theta][m]
    C *Fo = reinterpret_cast<C *>(a.Fout) + (size_t)lev * a.nout * n_theta * n_m;
    const double invN = 1.0 / (double)N;
    for (int g0 = 0; g0 < a.nout; g0 += a.G) {
        const int gn = min(a.G, a.nout - g0);
        C *res = fft_pp(rows + (size_t)g0 * M, B, gn, f, s_tw, -1);
        for (int idx = tid; idx < gn * n_m; idx += nthr) {
            const int r2 = idx / n_m, k = idx - r2 * n_m;
            Fo[((size_t)(g0 + r2) * n_theta + th) * n_m + k] = rfft_post(res + (size_t)r2 * M, k, M, s_twr, invN);
        }
        __syncthreads();
    }
}

// k_rl_grid with wave-slot FFTs (the row length M = 3 4^A 2^B): the rows
// are padded (index i at pidx<P>(i)) and every row is transformed in place
// by one slot of Q lanes of a wave, all rows at once; no scratch rows and no
// block barrier inside the FFTs
template <int A, int B>
__global__ __launch_bounds__(512) void k_rl_grid_s(FFTArgs f, RlGridArgs a)
{
    typedef double2 C;
    typedef SlotFFT<A, B> FFT;
    constexpr int M = FFT::M, N = 2 * M, Q = FFT::Q, P = FFT::P;
    constexpr int MP = padded_row<M, P>(), SPW = 64 / Q;
    extern __shared__ char smem_raw[];
    C *s_tw = reinterpret_cast<C *>(smem_raw);  // M + M / 8 (padded)
    C *s_twr = s_tw + M + M / 8;                // M + 1
    C *rows = s_twr + M + 1;                    // max(nin, nout) rows of MP
    int lb;
    {
        const int nx = 8, nb = gridDim.x, b = blockIdx.x;
        int per = nb / nx, rem = nb % nx, x = b % nx, i = b / nx;
        lb = (x < rem ? x * (per + 1) : rem * (per + 1) + (x - rem) * per) + i;
    }
    const int n_theta = a.n_theta, n_m = a.n_m;
    const int th = lb % n_theta, lev = lb / n_theta;
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int lane = tid & 63, sw = lane / Q, t = lane - sw * Q;
    const int slot = (tid >> 6) * SPW + sw, nslot = (nthr >> 6) * SPW;
    const bool act = sw < SPW;
    {
        const C *tw = reinterpret_cast<const C *>(f.tw);
        const C *twr = reinterpret_cast<const C *>(f.twr);
        for (int e = tid; e < M; e += nthr) s_tw[e + (e >> 3)] = tw[e];
        for (int e = tid; e <= M; e += nthr) s_twr[e] = twr[e];
    }
    __syncthreads();

    // inverse FFTs of rows [0, nin), one slot per row
    const C *Fi = reinterpret_cast<const C *>(a.Fin) + (size_t)lev * a.nin * n_theta * n_m + th;
    if (act)
        for (int c = slot; c < a.nin; c += nslot) {
            C *z = rows + (size_t)c * MP;
            const C *X = Fi + (size_t)c * n_m * n_theta;
            // Z_k = (X_k + conj X_{M-k}) + i e^{+2 pi i k/N} (X_k - conj X_{M-k})
            for (int k = t; k <= M / 2; k += Q) {
                const int k2 = M - k;
                C xa = k < n_m ? X[(size_t)k * n_theta] : mk<C>(0.0, 0.0);
                C xc = k2 < n_m ? X[(size_t)k2 * n_theta] : mk<C>(0.0, 0.0);
                if (k == 0) {
                    xa.y = 0.0;
                    xc.y = 0.0;
                }
                z[pidx<P>(k)] = cadd(cadd(xa, cconj(xc)), cmuli(cmul(cconj(s_twr[k]), csub(xa, cconj(xc))), 1));
                if (k2 < M && k2 != k)
                    z[pidx<P>(k2)] = cadd(cadd(xc, cconj(xa)), cmuli(cmul(cconj(s_twr[k2]), csub(xc, cconj(xa))), 1));
            }
            wave_order();
            FFT::run(z, t, s_tw, +1);
        }
    __syncthreads();

    // the products: point j of row c is the (j mod 2) part of complex j / 2
    rl_points(a, reinterpret_cast<double *>(rows),
              [&](int c, int j) { return (size_t)c * 2 * MP + 2 * pidx<P>(j >> 1) + (j & 1); }, N, lev, th);
    __syncthreads();

    // forward FFTs of rows [0, nout) in place, then F2[lev][comp][theta][m]
    if (act)
        for (int c = slot; c < a.nout; c += nslot) FFT::run(rows + (size_t)c * MP, t, s_tw, -1);
    __syncthreads();
    C *Fo = reinterpret_cast<C *>(a.Fout) + (size_t)lev * a.nout * n_theta * n_m;
    const double invN = 1.0 / (double)N;
    for (int idx = tid; idx < a.nout * n_m; idx += nthr) {
        const int r2 = idx / n_m, k = idx - r2 * n_m;
        Fo[((size_t)r2 * n_theta + th) * n_m + k] = rfft_post_padded<P>(rows + (size_t)r2 * MP, k, M, s_twr, invN);
    }
}

FFTArgs make_args(const Plan *p)
{
    FFTArgs f;
    f.M = p->M;
    f.N = p->n_phi;
    f.nfac = (int)p->fac.size();
    for (int i = 0; i < f.nfac && i < 16; ++i) f.fac[i] = p->fac[i];
    f.tw = p->d_tw;
    f.twr = p->d_twr;
    return f;
}

// rows per block of the API kernels: two LDS buffers of TB rows + twiddles
int rows_per_block(const Plan *p)
{
    int TB = 16;
    while (TB > 1 && ((size_t)(2 * TB + 2) * (p->M + 1) * sizeof(double2) > 64 * 1024)) TB >>= 1;
    return TB;
}

size_t api_shm(const Plan *p, int TB)
{
    return ((size_t)2 * TB * p->M + 2 * p->M + 1) * sizeof(double2);
}

}  // namespace

int launch_ifft(const Plan *p, const double *F, size_t F_ld_lev, int F_ncomp, int ncomp, int nlev,
                double *grid, size_t grid_ld_lev, size_t grid_ld_comp, hipStream_t s)
{
    if (nlev <= 0) return MHX_OK;
    FFTArgs f = make_args(p);
    int TB = rows_per_block(p);
    int ntb = (p->n_theta + TB - 1) / TB;
    size_t shm = api_shm(p, TB);
    int nthr = 256;
    if (nthr % TB) return fail(MHX_E_INVALID, "row block does not divide the block size");
    k_ifft_rows<double><<<dim3((unsigned)(ntb * ncomp * nlev)), nthr, shm, s>>>(
        f, F, F_ld_lev, F_ncomp, p->n_m, p->n_theta, ncomp, TB, ntb, grid, grid_ld_lev,
        grid_ld_comp);
    MHX_HIP_CHECK(hipGetLastError());
    return MHX_OK;
}

int launch_fft(const Plan *p, const double *grid, size_t grid_ld_lev, size_t grid_ld_comp,
               int ncomp, int nlev, unsigned osin2_mask, double *F, size_t F_ld_lev, hipStream_t s)
{
    if (nlev <= 0) return MHX_OK;
    FFTArgs f = make_args(p);
    int TB = rows_per_block(p);
    int ntb = (p->n_theta + TB - 1) / TB;
    size_t shm = api_shm(p, TB);
    k_fft_rows<double><<<dim3((unsigned)(ntb * ncomp * nlev)), 256, shm, s>>>(
        f, grid, grid_ld_lev, grid_ld_comp, p->n_m, p->n_theta, ncomp, TB, ntb, osin2_mask,
        p->d_osin2, F, F_ld_lev, 0);
    MHX_HIP_CHECK(hipGetLastError());
    return MHX_OK;
}

// FP32 row FFTs of the batched transform jobs, TB rows per block (a tuned
// choice); the forward side writes the theta-blocked layout of the split
// analysis
int launch_ifft_f32(const Plan *p, const float *F, size_t F_ld, int ncomp, int nrow, float *grid,
                    size_t g_ld_row, size_t g_ld_comp, int TB, hipStream_t s)
{
    if (nrow <= 0) return MHX_OK;
    FFTArgs f = make_args(p);
    f.tw = p->d_tw32;
    f.twr = p->d_twr32;
    const int ntb = (p->n_theta + TB - 1) / TB;
    const size_t shm = ((size_t)2 * TB * p->M + 2 * p->M + 1) * sizeof(float2);
    if (shm > 64 * 1024 || 256 % TB) return fail(MHX_E_UNSUPPORTED, "rows per block too large");
    k_ifft_rows<float><<<dim3((unsigned)(ntb * ncomp * nrow)), 256, shm, s>>>(
        f, F, F_ld, ncomp, p->n_m, p->n_theta, ncomp, TB, ntb, grid, g_ld_row, g_ld_comp);
    MHX_HIP_CHECK(hipGetLastError());
    return MHX_OK;
}

int launch_fft_f32(const Plan *p, const float *grid, size_t g_ld_row, size_t g_ld_comp, int ncomp,
                   int nrow, unsigned osin2_mask, float *F2, size_t F2_ld, int TB, hipStream_t s)
{
    if (nrow <= 0) return MHX_OK;
    FFTArgs f = make_args(p);
    f.tw = p->d_tw32;
    f.twr = p->d_twr32;
    const int ntb = (p->n_theta + TB - 1) / TB;
    const size_t shm = ((size_t)2 * TB * p->M + 2 * p->M + 1) * sizeof(float2);
    if (shm > 64 * 1024 || 256 % TB) return fail(MHX_E_UNSUPPORTED, "rows per block too large");
    k_fft_rows<float><<<dim3((unsigned)(ntb * ncomp * nrow)), 256, shm, s>>>(
        f, grid, g_ld_row, g_ld_comp, p->n_m, p->n_theta, ncomp, TB, ntb, osin2_mask, p->d_osin2,
        F2, F2_ld, 1);
    MHX_HIP_CHECK(hipGetLastError());
    return MHX_OK;
}

int fused_threads(int n_phi)
{
    int nt = 64 * ((n_phi + 3 * 64 - 1) / (3 * 64));
    return nt > 1024 ? 1024 : nt;
}

// the slot form applies when M = 3 4^A 2^B with 1 <= A <= 4, B <= 1 and
// min(M / 12, 64) lanes per slot dividing 64: M = 12 .. 1536 (n_phi = 24 ..
// 3072)
bool slot_form(const Plan *p, int &A, int &B)
{
    int m = p->M;
    if (m % 3) return false;
    m /= 3;
    A = 0;
    while (m % 4 == 0) { m /= 4; ++A; }
    B = (m == 2);
    if (m != 1 && m != 2) return false;
    return A >= 1 && A <= 4;
}

template <typename T, int A, int B>
int launch_slots(const Plan *p, const FFTArgs &f, int nlev, const void *Fin, void *Fout,
                 const double *rad, double LFfac, void *dgrid, void *full_grids, hipStream_t s)
{
    constexpr int M = SlotFFT<A, B>::M, NPT = SlotFFT<A, B>::P / 3, nthr = SlotFFT<A, B>::NTHR;
    if (NPT * nthr < p->n_phi) return fail(MHX_E_UNSUPPORTED, "n_phi too large for the fused stage");
    const size_t shm = ((size_t)2 * M + M / 8 + 1 + 6 * (size_t)padded_row<M, SlotFFT<A, B>::P>()) * 2 * sizeof(T);
    // set once per kernel (the attribute persists), so that a call under
    // graph capture makes no attribute call
    static int shm_set = 0;
    if (shm > 64 * 1024 && (int)shm > shm_set) {
        MHX_HIP_CHECK(hipFuncSetAttribute((const void *)k_nl_slots<T, A, B>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
        shm_set = (int)shm;
    }
    dim3 grid((unsigned)(p->n_theta * nlev));
    k_nl_slots<T, A, B><<<grid, nthr, shm, s>>>(f, static_cast<const T *>(Fin), static_cast<T *>(Fout),
                                                p->n_m, p->n_theta, nlev, p->d_osin2, rad, LFfac,
                                                static_cast<T *>(dgrid), static_cast<T *>(full_grids),
                                                p->cour);
    MHX_HIP_CHECK(hipGetLastError());
    return MHX_OK;
}

template <typename T>
int launch_fused_t(const Plan *p, int nlev, const void *Fin, void *Fout, const double *rad,
                   double LFfac, void *dgrid, void *full_grids, hipStream_t s)
{
    FFTArgs f = make_args(p);
    if (sizeof(T) == 4) {
        f.tw = p->d_tw32;
        f.twr = p->d_twr32;
    }
    int A, B;
    if (slot_form(p, A, B)) {
        const int key = 2 * A + B;
        switch (key) {
        case 2: return launch_slots<T, 1, 0>(p, f, nlev, Fin, Fout, rad, LFfac, dgrid, full_grids, s);
        case 3: return launch_slots<T, 1, 1>(p, f, nlev, Fin, Fout, rad, LFfac, dgrid, full_grids, s);
        case 4: return launch_slots<T, 2, 0>(p, f, nlev, Fin, Fout, rad, LFfac, dgrid, full_grids, s);
        case 5: return launch_slots<T, 2, 1>(p, f, nlev, Fin, Fout, rad, LFfac, dgrid, full_grids, s);
        case 6: return launch_slots<T, 3, 0>(p, f, nlev, Fin, Fout, rad, LFfac, dgrid, full_grids, s);
        case 7: return launch_slots<T, 3, 1>(p, f, nlev, Fin, Fout, rad, LFfac, dgrid, full_grids, s);
        case 8: return launch_slots<T, 4, 0>(p, f, nlev, Fin, Fout, rad, LFfac, dgrid, full_grids, s);
        case 9:
            // M = 1536 (the l_max=1023 grid): six rows of 1536 take 98 KB of
            // LDS in FP32; in FP64 they do not fit and the generic form runs
            if constexpr (sizeof(T) == 4)
                return launch_slots<T, 4, 1>(p, f, nlev, Fin, Fout, rad, LFfac, dgrid, full_grids, s);
            break;
        default: break;
        }
    }
    int nthr = fused_threads(p->n_phi);
    if (nthr * PMAX < p->n_phi) return fail(MHX_E_UNSUPPORTED, "n_phi too large for the fused stage");
    const size_t cs = 2 * sizeof(T);
    // three rows per FFT pass while they fit the default 64 KB, else one row
    // in up to the 160 KB a workgroup may hold on gfx950 (n_phi = 3072 in FP64)
    int G = ((size_t)(6 * p->M + 2 * p->M + 1) * cs <= 64 * 1024) ? 3 : 1;
    size_t shm = ((size_t)2 * G * p->M + 2 * p->M + 1) * cs;
    if (shm > 160 * 1024) return fail(MHX_E_UNSUPPORTED, "n_phi too large for the fused stage (LDS)");
    dim3 grid((unsigned)(p->n_theta * nlev));
    const T *fi = static_cast<const T *>(Fin);
    T *fo = static_cast<T *>(Fout), *dg = static_cast<T *>(dgrid), *fg = static_cast<T *>(full_grids);
    if (G == 3) {
        k_nl_fused<T, 3><<<grid, nthr, shm, s>>>(f, fi, fo, p->n_m, p->n_theta, nlev, p->d_osin2,
                                                  rad, LFfac, dg, fg, p->cour);
    } else {
        // set once per instance and raised only (the attribute persists), so
        // that a call under graph capture makes no attribute call
        static int shm_set = 0;
        if (shm > 64 * 1024 && (int)shm > shm_set) {
            MHX_HIP_CHECK(hipFuncSetAttribute((const void *)k_nl_fused<T, 1>,
                                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
            shm_set = (int)shm;
        }
        k_nl_fused<T, 1><<<grid, nthr, shm, s>>>(f, fi, fo, p->n_m, p->n_theta, nlev, p->d_osin2,
                                                  rad, LFfac, dg, fg, p->cour);
    }
    MHX_HIP_CHECK(hipGetLastError());
    return MHX_OK;
}

int launch_nl_fused(const Plan *p, int nlev, const void *Fin, void *Fout, const double *rad,
                    double LFfac, void *dgrid, void *full_grids, hipStream_t s)
{
    if (nlev <= 0) return MHX_OK;
    if (p->precision == MHX_PREC_SPLIT16) {
        if (p->grid_form == 0 && grid_tile_supported(p))
            return launch_grid_tile(p, nlev, Fin, Fout, rad, LFfac, dgrid, full_grids, s);
        return launch_fused_t<float>(p, nlev, Fin, Fout, rad, LFfac, dgrid, full_grids, s);
    }
    return launch_fused_t<double>(p, nlev, Fin, Fout, rad, LFfac, dgrid, full_grids, s);
}

template <int A, int B>
int launch_rl_slots(const Plan *p, const FFTArgs &f, const RlGridArgs &a, size_t shm, hipStream_t s)
{
    static int shm_set = 0;
    if (shm > 64 * 1024 && (int)shm > shm_set) {
        MHX_HIP_CHECK(hipFuncSetAttribute((const void *)k_rl_grid_s<A, B>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
        shm_set = (int)shm;
    }
    dim3 grid((unsigned)(p->n_theta * a.nlev));
    k_rl_grid_s<A, B><<<grid, 512, shm, s>>>(f, a);
    MHX_HIP_CHECK(hipGetLastError());
    return MHX_OK;
}

int launch_rl_grid(const Plan *p, const RlGridArgs &args, hipStream_t s)
{
    if (args.nlev <= 0) return MHX_OK;
    RlGridArgs a = args;
    const FFTArgs f = make_args(p);
    const size_t cs = sizeof(double2);
    const int nrow = std::max(a.nin, a.nout);
    if (nrow <= 0 || nrow > 32) return fail(MHX_E_INVALID, "radial loop: bad component count");
    const size_t lds_max = 160 * 1024;
    // rows of M = 3 4^A 2^B: wave-slot FFTs on padded rows, all rows at once,
    // when they fit the LDS (measured at l_max = 255 x 97, Boussinesq MHD:
    // 10.2 -> 8.7 ms per mhx_rloop call against the block-wide FFTs below)
    {
        int A, B;
        if (slot_form(p, A, B)) {
            const int M = p->M, P = M / 12 <= 64 ? 12 : M / 64;
            const size_t MP = P == 12 ? (size_t)M + (M >> 5) + 1 : (size_t)M;
            const size_t shm = ((size_t)2 * M + M / 8 + 1 + (size_t)nrow * MP) * cs;
            if (shm <= lds_max) {
                switch (2 * A + B) {
                case 2: return launch_rl_slots<1, 0>(p, f, a, shm, s);
                case 3: return launch_rl_slots<1, 1>(p, f, a, shm, s);
                case 4: return launch_rl_slots<2, 0>(p, f, a, shm, s);
                case 5: return launch_rl_slots<2, 1>(p, f, a, shm, s);
                case 6: return launch_rl_slots<3, 0>(p, f, a, shm, s);
                case 7: return launch_rl_slots<3, 1>(p, f, a, shm, s);
                case 8: return launch_rl_slots<4, 0>(p, f, a, shm, s);
                case 9: return launch_rl_slots<4, 1>(p, f, a, shm, s);
                default: break;
                }
            }
        }
    }
    const size_t fixed = ((size_t)2 * p->M + 1 + (size_t)nrow * p->M) * cs;
    if (fixed + (size_t)p->M * cs > lds_max)
        return fail(MHX_E_UNSUPPORTED, "radial loop: the grid rows of this n_phi and these physics "
                                       "switches exceed the 160 KB of LDS of a workgroup");
    // other lengths: block-wide Stockham passes over as many FFT rows per pass
    // as the LDS holds, at most 8 (64 threads per row); 512 threads per block:
    // the rows of one block use most of the LDS, so one block per CU, whose
    // eight waves hide each other's latency (measured at l_max = 255 x 97:
    // 256 threads 11.3 ms per mhx_rloop call, 512 threads 10.3 ms, 768
    // threads 10.4 ms)
    int G = (int)std::min<size_t>(8, (lds_max - fixed) / ((size_t)p->M * cs));
    G = std::max(1, std::min(G, nrow));
    a.G = G;
    const size_t shm = fixed + (size_t)G * p->M * cs;
    static int shm_set = 0;
    if (shm > 64 * 1024 && (int)shm > shm_set) {
        MHX_HIP_CHECK(hipFuncSetAttribute((const void *)k_rl_grid,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
        shm_set = (int)shm;
    }
    dim3 grid((unsigned)(p->n_theta * a.nlev));
    k_rl_grid<<<grid, 512, shm, s>>>(f, a);
    MHX_HIP_CHECK(hipGetLastError());
    return MHX_OK;
}

int launch_axi(const Plan *p, const double *fl, double *f, int lcut, int tor, double *f2,
               hipStream_t s)
{
    if (int rc = ensure_fp64_tables(const_cast<Plan *>(p))) return rc;
    k_axi<<<(p->nh + 127) / 128, 128, 0, s>>>(p->l_max, p->nh, p->d_syn_P, p->d_syn_dP,
                                               p->d_syn_off, p->nt16, fl, lcut, tor, f, f2);
    MHX_HIP_CHECK(hipGetLastError());
    return MHX_OK;
}

}  // namespace mhx
