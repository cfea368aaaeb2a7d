// Grid stage of the split-form nonlinear step on theta-row tiles, for the
// n_phi = 768 grid (l_max = 255): per radial level and TR (4 or 8)
// consecutive theta rows, the 18 inverse FFTs of the synthesised components, the grid products
// of src/get_nl.f90 (curl advection :262-273, Lorentz force :242-258 merged as
// src/rIter.f90:650-656, V s :320-325, V x B :376-388, u_h . grad_h s), the
// Courant maxima (src/courant.f90:209-275) and the 10 forward FFTs, in FP32.
//
// One workgroup = one level x TR theta rows, one wave per row.  The
// Fourier coefficients are moved between HBM and the waves through LDS tiles
// [comp][row][m], so that global accesses are runs of TR theta rows (64 bytes
// at TR = 8)
// (F[lev][comp][m][theta] on the inverse side, the theta-blocked
// F2[lev][comp][theta/16][m][theta%16] on the forward side) instead of one
// 8-byte access per lane.
//
// Two real components share one complex FFT of length N = 768 (y = x1 + i x2,
// Y_k = X1_k + i X2_k with the Hermitian halves, and back X1_k = (Y_k + conj
// Y_{N-k})/2, X2_k = (Y_k - conj Y_{N-k})/(2i)).  Each wave runs its FFTs
// alone, Stockham passes in its own LDS buffer with no block barrier:
//   inverse radices 16 (L = 1), 4 (L = 16), 12 (L = 64), decimation in time
//   forward radices 12, 4, 16, decimation in frequency (the same twiddles,
//   conjugated, held in registers)
// so that after the inverse (and before the forward) FFT lane j holds the
// grid points n = j + 64 q, q < 12, of both components: the products of a
// point are formed in the registers of the lane that owns it.  The buffer
// index paddings make every LDS address a per-lane base plus a constant.
//
// Normalisation as the reference: inverse unnormalised using the real parts
// of mode 0 (src/fft.f90:386-398), forward scaled by 1/N (src/fft.f90:464).
#include "mhx_fftcore.h"

namespace mhx {

namespace {

constexpr int GT_N = 768;
constexpr int GT_BUF = 960;              // complex per wave FFT buffer (padded)
constexpr int GT_MAXNM = 256;            // n_m <= N / 3
constexpr int GT_ROWP = GT_MAXNM + 4;    // tile row pitch (complex)
constexpr int GT_MT = GT_MAXNM / 64;     // tile loads per thread and comp
// LDS twiddle tables (complex, lane-contiguous), read once into registers:
// inverse pass 2 [3][16], inverse pass 3 [11][64]
constexpr int TW_I2 = 0, TW_I3 = TW_I2 + 3 * 16;
constexpr int TW_N = TW_I3 + 11 * 64;
// TR theta rows (waves) per workgroup: LDS of the twiddles, two tile slots
// [2 comps][TR rows][ROWP] and TR FFT buffers
template <int TR> constexpr size_t gt_lds()
{
    return (size_t)(TW_N + 2 * (2 * TR * GT_ROWP) + TR * GT_BUF) * sizeof(float2);
}

typedef float2 C;

struct GtArgs {
    const float *Fin;     // [lev][18][m][theta] complex
    float *Fout;          // [lev][10][theta/16][m][theta%16] complex
    float *dgrid;         // [lev][3][theta][phi] du/dr grids
    float *full_grids;    // optional [lev][18][phi][theta]
    const float *twn;     // exp(-2 pi i k / N), k < N
    const double *osin2;  // per theta row
    const double *rad;    // or2[nlev], or4[nlev], orho1[nlev]
    double LFfac;
    int n_m, n_theta, nlev;
    CourArgs cour;
};

// inverse FFT of the components in tile rows r1, r2 (row pointers into the
// tile, n_m coefficients each) into y[q] = x1[n] + i x2[n], n = j + 64 q
__device__ __forceinline__ void gt_inverse(const C *__restrict__ r1, const C *__restrict__ r2, int n_m,
                                           C *buf, const C (&tw2)[3], const C (&tw3)[11], int j,
                                           C (&y)[12])
{
    // pass 1 (radix 16, L = 1, 48 butterflies): Y_k, k = j + 48 s
    if (j < 48) {
        C x[16];
#pragma unroll
        for (int s = 0; s < 16; ++s) {
            x[s] = mk<C>(0.f, 0.f);
            if (s <= 5) {
                const int k = j + 48 * s;
                if (k < n_m) {
                    C a = r1[k], b = r2[k];
                    if (k == 0) a.y = b.y = 0.f;
                    x[s] = mk<C>(a.x - b.y, a.y + b.x);          // a + i b
                }
            } else if (s >= 10) {
                const int kk = GT_N - j - 48 * s;               // N - k
                if (kk < n_m) {
                    const C a = r1[kk], b = r2[kk];
                    x[s] = mk<C>(a.x + b.y, b.x - a.y);          // conj a + i conj b
                }
            }
        }
        Dft<16, C>::run(x, +1);
        // out index 16 j + s, padded i + (i >> 4) = 17 j + s
        C *o = buf + 17 * j;
#pragma unroll
        for (int s = 0; s < 16; ++s) o[s] = x[s];
    }
    wave_order();
    // pass 2 (radix 4, L = 16): butterflies b = j + 64 i, k = j mod 16
    {
        const int k = j & 15;
        C x[3][4];
        const C *ib = buf + (j + (j >> 4));                   // pad_ib(b) = pad_ib(j) + 68 i
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                C v = ib[68 * i + 204 * s];
                if (s > 0) v = cmul(v, tw2[s - 1]);
                x[i][s] = v;
            }
        wave_order();
        // out (b - k) 4 + k + 16 s, padded i + ((i >> 5) << 3):
        // 80 (j >> 4) + k + 320 i + 16 s + 8 (s >> 1)
        C *ob = buf + (80 * (j >> 4) + k);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            Dft<4, C>::run(x[i], +1);
#pragma unroll
            for (int s = 0; s < 4; ++s) ob[320 * i + 16 * s + 8 * (s >> 1)] = x[i][s];
        }
    }
    wave_order();
    // pass 3 (radix 12, L = 64): in j + 64 s, padded (j + 8 (j >> 5)) + 80 s
    {
        const C *ib = buf + (j + 8 * (j >> 5));
#pragma unroll
        for (int s = 0; s < 12; ++s) {
            C v = ib[80 * s];
            if (s > 0) v = cmul(v, tw3[s - 1]);
            y[s] = v;
        }
        Dft<12, C>::run(y, +1);
    }
    wave_order();
}

// forward FFT of y[q] = p1[n] + i p2[n], n = j + 64 q, into tile rows o1, o2
// (n_m coefficients each, scaled by 1/N).  Decimation in frequency, so that
// every twiddle is the conjugate of an inverse twiddle the lane holds
// (tw3[r - 1] = exp(2 pi i j r / N), tw2[s - 1] = exp(2 pi i (j mod 16) s / 64)):
//   radix 12 over q, times exp(-2 pi i j r / N): z_r[j] at 64 r + j;
//   radix 4 over j2 of j = j1 + 16 j2 (j1 = lane mod 16, r = 3 (lane / 16) + i),
//     times exp(-2 pi i j1 k1 / 64): at 16 (r + 12 k1) + j1;
//   radix 16 over j1 (lane b = r + 12 k1 < 48): X_{b + 48 k2} at b + 48 k2;
// the first two buffers padded i + i / 16.
__device__ __forceinline__ void gt_forward(C (&y)[12], C *__restrict__ o1, C *__restrict__ o2, int n_m,
                                           C *buf, const C (&tw2)[3], const C (&tw3)[11], int j)
{
    // pass 1 (radix 12 over q): out padded 68 r + j + j / 16
    Dft<12, C>::run(y, -1);
    {
        C *o = buf + (j + (j >> 4));
#pragma unroll
        for (int r = 0; r < 12; ++r) {
            C v = y[r];
            if (r > 0) v = cmul(v, mk<C>(tw3[r - 1].x, -tw3[r - 1].y));
            o[68 * r] = v;
        }
    }
    wave_order();
    // pass 2 (radix 4 over j2, three r per lane): in padded 68 r + j1 + 17 j2,
    // out padded 17 (r + 12 k1) + j1
    {
        const int j1 = j & 15, rg = 3 * (j >> 4);
        C x[3][4];
        const C *ib = buf + (68 * rg + j1);
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int s = 0; s < 4; ++s) x[i][s] = ib[68 * i + 17 * s];
        wave_order();
        C *ob = buf + (17 * rg + j1);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            Dft<4, C>::run(x[i], -1);
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                C v = x[i][s];
                if (s > 0) v = cmul(v, mk<C>(tw2[s - 1].x, -tw2[s - 1].y));
                ob[17 * i + 204 * s] = v;
            }
        }
    }
    wave_order();
    // pass 3 (radix 16 over j1, 48 butterflies): in padded 17 b + j1, out b + 48 s
    if (j < 48) {
        C x[16];
        const C *ib = buf + 17 * j;
#pragma unroll
        for (int s = 0; s < 16; ++s) x[s] = ib[s];
        wave_order();
        Dft<16, C>::run(x, -1);
        // outputs k = j + 48 s in [288, 480) (s = 6..9) are never read: the
        // unpacking below reads k < n_m <= N / 3 and N - k > 2 N / 3
#pragma unroll
        for (int s = 0; s < 16; ++s)
            if (s < 6 || s > 9) buf[j + 48 * s] = x[s];
    } else {
        wave_order();
    }
    wave_order();
    // X1_k = (Y_k + conj Y_{N-k}) / 2N, X2_k = (Y_k - conj Y_{N-k}) / 2iN
    const float h = 0.5f / (float)GT_N;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int k = j + 64 * q;
        if (k < n_m) {
            const C a = buf[k];
            const C c = buf[k == 0 ? 0 : GT_N - k];
            o1[k] = mk<C>(h * (a.x + c.x), h * (a.y - c.y));
            o2[k] = mk<C>(h * (a.y + c.y), h * (c.x - a.x));
        }
    }
    wave_order();
}

template <int TR>
__global__ __launch_bounds__(64 * TR) void k_grid_tile(GtArgs a)
{
    constexpr int NTHR = 64 * TR, SLOT = 2 * TR * GT_ROWP;
    extern __shared__ char smem_raw[];
    C *s_tw = reinterpret_cast<C *>(smem_raw);
    C *s_slot = s_tw + TW_N;                           // 2 slots of [2][TR][ROWP]
    C *s_buf = s_slot + 2 * SLOT;                      // per-wave FFT buffers
    const int tid = threadIdx.x, wave = tid >> 6, j = tid & 63;
    const int n_m = a.n_m, n_theta = a.n_theta, nlev = a.nlev;
    const int ntile = n_theta / TR;
    // XCD-aware order: consecutive tiles (which share the 128-byte lines of
    // the Fourier arrays) run on the same XCD at the same time
    int lb;
    {
        const int nx = 8, nb = gridDim.x, b = blockIdx.x;
        const int per = nb / nx, rem = nb % nx, x = b % nx, i = b / nx;
        lb = (x < rem ? x * (per + 1) : rem * (per + 1) + (x - rem) * per) + i;
    }
    const int tile = lb % ntile, lev = lb / ntile;
    const int th0 = tile * TR, th = th0 + wave;
    C *buf = s_buf + wave * GT_BUF;

    // twiddle tables from exp(-2 pi i k / N)
    {
        const C *tn = reinterpret_cast<const C *>(a.twn);
        for (int e = tid; e < TW_N; e += NTHR) {
            int idx;
            if (e < TW_I3) {
                const int s = e / 16 + 1, k = e % 16;
                idx = 12 * s * k;
            } else {
                const int s = (e - TW_I3) / 64 + 1, k = (e - TW_I3) % 64;
                idx = s * k;
            }
            C v = tn[idx];
            v.y = -v.y;  // exp(+2 pi i idx / N)
            s_tw[e] = v;
        }
    }

    const float o2 = (float)a.osin2[th];
    const float or2 = (float)a.rad[lev], or4 = (float)a.rad[nlev + lev];
    const float orho1 = (float)a.rad[2 * nlev + lev], LFfac = (float)a.LFfac;

    // cooperative tile moves: thread -> (row r, m = mm + 64 t)
    const int tr = tid & (TR - 1), tmm = tid / TR;
    const C *Fi = reinterpret_cast<const C *>(a.Fin) + (size_t)lev * 18 * n_m * n_theta + th0 + tr;
    C pre[2][GT_MT];
    // without du/dr grids (dgrid null) their components 6..8 are not read:
    // zeros go into the FFT halves they share
    const bool dudr = a.dgrid != nullptr;
    auto load_pair = [&](int c0, int c1) {
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const int comp = c ? c1 : c0;
            const bool skip = !dudr && comp >= 6 && comp <= 8;
            const C *src = Fi + (size_t)comp * n_m * n_theta;
#pragma unroll
            for (int t = 0; t < GT_MT; ++t) {
                const int m = tmm + 64 * t;
                pre[c][t] = (m < n_m && !skip) ? src[(size_t)m * n_theta] : mk<C>(0.f, 0.f);
            }
        }
    };
    auto store_pair = [&](C *slot) {
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int t = 0; t < GT_MT; ++t) {
                const int m = tmm + 64 * t;
                if (m < n_m) slot[(c * TR + tr) * GT_ROWP + m] = pre[c][t];
            }
    };
    const int ntb = n_theta / 16;
    C *Fo = reinterpret_cast<C *>(a.Fout) + (size_t)lev * 10 * ntb * 16 * n_m;
    auto flush_pair = [&](const C *slot, int c0, int c1) {
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            C *dst = Fo + (((size_t)(c ? c1 : c0) * ntb + (th0 >> 4)) * n_m) * 16 + (th0 & 15) + tr;
#pragma unroll
            for (int t = 0; t < GT_MT; ++t) {
                const int m = tmm + 64 * t;
                if (m < n_m) dst[(size_t)m * 16] = slot[(c * TR + tr) * GT_ROWP + m];
            }
        }
    };
    auto row = [&](C *slot, int c) { return slot + (c * TR + wave) * GT_ROWP; };
    float *fg = a.full_grids;
    auto grids_out = [&](const C (&y)[12], int c0, int c1) {
        if (!fg) return;
#pragma unroll
        for (int q = 0; q < 12; ++q) {
            const int n = j + 64 * q;
            fg[(((size_t)lev * 18 + c0) * GT_N + n) * n_theta + th] = y[q].x;
            fg[(((size_t)lev * 18 + c1) * GT_N + n) * n_theta + th] = y[q].y;
        }
    };
    float *dg = dudr ? a.dgrid + (size_t)lev * 3 * n_theta * GT_N : nullptr;
    auto dgrid_out = [&](int c, const C (&y)[12], bool imag) {
        if (!dudr) return;
        float *d = dg + ((size_t)c * n_theta + th) * GT_N + j;
#pragma unroll
        for (int q = 0; q < 12; ++q) d[64 * q] = imag ? y[q].y : y[q].x;
    };

    float u[3][12], acc[3][12], bb[3][12], pend[12];
    C y[12], tw2[3], tw3[11];
    int slot_i = 0;
    auto slot = [&]() { return s_slot + (slot_i & 1) * SLOT; };

    // one inverse pair: comps (c0, c1) already in pre; the next pair's loads
    // are issued before the FFT
    auto inverse = [&](int c0, int c1, int n0, int n1) {
        C *sl = slot();
        store_pair(sl);
        __syncthreads();
        if (n0 >= 0) load_pair(n0, n1);
        gt_inverse(row(sl, 0), row(sl, 1), n_m, buf, tw2, tw3, j, y);
        grids_out(y, c0, c1);
        ++slot_i;
    };
    // one forward pair: y already holds the products
    auto forward = [&](int c0, int c1) {
        C *sl = slot();
        gt_forward(y, row(sl, 0), row(sl, 1), n_m, buf, tw2, tw3, j);
        __syncthreads();
        flush_pair(sl, c0, c1);
        ++slot_i;
    };

    load_pair(0, 1);
    __syncthreads();  // twiddle tables
    // the inverse passes' twiddles depend on the lane only: held in registers
    // for the row's nine inverse FFTs
#pragma unroll
    for (int s = 0; s < 3; ++s) tw2[s] = s_tw[TW_I2 + s * 16 + (j & 15)];
#pragma unroll
    for (int s = 0; s < 11; ++s) tw3[s] = s_tw[TW_I3 + s * 64 + j];

    // u: (vr, vt), (vp, s); V s, forward (VSr, VSt)
    inverse(0, 1, 2, 15);
#pragma unroll
    for (int q = 0; q < 12; ++q) u[0][q] = y[q].x, u[1][q] = y[q].y;
    inverse(2, 15, 16, 17);
#pragma unroll
    for (int q = 0; q < 12; ++q) {
        u[2][q] = y[q].x;
        const float sv = y[q].y;
        pend[q] = or2 * u[2][q] * sv * o2;                               // VSp / sin^2
        y[q] = mk<C>(u[0][q] * sv, or2 * u[1][q] * sv * o2);             // VSr, VSt / sin^2
    }
    forward(3, 4);
    // grad s: u_h . grad_h s; forward (VSp, uGs)
    inverse(16, 17, 3, 4);
#pragma unroll
    for (int q = 0; q < 12; ++q)
        y[q] = mk<C>(pend[q], or2 * o2 * (u[1][q] * y[q].x + u[2][q] * y[q].y));
    forward(5, 9);
    // curl u: Adv = -curl(u) x u (src/get_nl.f90:262-273)
    inverse(3, 4, 5, 6);
#pragma unroll
    for (int q = 0; q < 12; ++q) {
        const float cvr = y[q].x, cvt = y[q].y;
        acc[0][q] = -o2 * (cvt * u[2][q]);
        acc[1][q] = or4 * (cvr * u[2][q]);
        acc[2][q] = -or4 * (cvr * u[1][q] - cvt * u[0][q]);
    }
    inverse(5, 6, 9, 10);
#pragma unroll
    for (int q = 0; q < 12; ++q) {
        const float cvp = y[q].x;
        acc[0][q] += o2 * (cvp * u[1][q]);
        acc[1][q] -= or4 * (cvp * u[0][q]);
    }
    dgrid_out(0, y, true);
    // B; V x B (src/get_nl.f90:376-388) and the Courant maxima
    inverse(9, 10, 11, 7);
#pragma unroll
    for (int q = 0; q < 12; ++q) bb[0][q] = y[q].x, bb[1][q] = y[q].y;
    inverse(11, 7, 12, 13);
#pragma unroll
    for (int q = 0; q < 12; ++q) bb[2][q] = y[q].x;
    dgrid_out(1, y, true);
    if (a.cour.max) courant_max<float, 12>(a.cour, lev, nlev, o2, or2, or4, orho1, LFfac, u, bb);
#pragma unroll
    for (int q = 0; q < 12; ++q) {
        const float vxr = orho1 * o2 * (u[1][q] * bb[2][q] - u[2][q] * bb[1][q]);
        const float vxt = orho1 * or4 * (u[2][q] * bb[0][q] - u[0][q] * bb[2][q]);
        pend[q] = orho1 * or4 * (u[0][q] * bb[1][q] - u[1][q] * bb[0][q]) * o2;   // VxBp / sin^2
        y[q] = mk<C>(vxr, vxt * o2);
    }
    forward(6, 7);
    // curl B: Lorentz force added to Adv (src/get_nl.f90:242-258)
    inverse(12, 13, 14, 8);
#pragma unroll
    for (int q = 0; q < 12; ++q) {
        const float cbr = y[q].x, cbt = y[q].y;
        acc[0][q] += LFfac * o2 * (cbt * bb[2][q]);
        acc[1][q] -= LFfac * or4 * (cbr * bb[2][q]);
        acc[2][q] += LFfac * or4 * (cbr * bb[1][q] - cbt * bb[0][q]);
    }
    inverse(14, 8, -1, -1);
#pragma unroll
    for (int q = 0; q < 12; ++q) {
        const float cbp = y[q].x;
        acc[0][q] -= LFfac * o2 * (cbp * bb[1][q]);
        acc[1][q] += LFfac * or4 * (cbp * bb[0][q]);
    }
    dgrid_out(2, y, true);
    // forward (VxBp, Advr), (Advt, Advp), the horizontal comps / sin^2
#pragma unroll
    for (int q = 0; q < 12; ++q) y[q] = mk<C>(pend[q], acc[0][q]);
    forward(8, 0);
#pragma unroll
    for (int q = 0; q < 12; ++q) y[q] = mk<C>(acc[1][q] * o2, acc[2][q] * o2);
    forward(1, 2);
}

}  // namespace

bool grid_tile_supported(const Plan *p)
{
    return p->n_phi == GT_N && p->n_m <= GT_MAXNM && 3 * p->n_m <= GT_N && p->n_theta % 16 == 0 &&
           p->d_twn32 != nullptr;
}

template <int TR>
int launch_tr(const Plan *p, const GtArgs &a, int nlev, hipStream_t s)
{
    constexpr size_t lds = gt_lds<TR>();
    static bool shm_set = false;
    if (!shm_set) {
        MHX_HIP_CHECK(hipFuncSetAttribute((const void *)k_grid_tile<TR>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        shm_set = true;
    }
    dim3 grid((unsigned)(p->n_theta / TR * nlev));
    k_grid_tile<TR><<<grid, 64 * TR, lds, s>>>(a);
    MHX_HIP_CHECK(hipGetLastError());
    return MHX_OK;
}

int launch_grid_tile(const Plan *p, int nlev, const void *Fin, void *Fout, const double *rad,
                     double LFfac, void *dgrid, void *full_grids, hipStream_t s)
{
    if (nlev <= 0) return MHX_OK;
    if (!grid_tile_supported(p)) return fail(MHX_E_UNSUPPORTED, "grid tile kernel: unsupported grid");
    GtArgs a;
    a.Fin = static_cast<const float *>(Fin);
    a.Fout = static_cast<float *>(Fout);
    a.dgrid = static_cast<float *>(dgrid);
    a.full_grids = static_cast<float *>(full_grids);
    a.twn = p->d_twn32;
    a.osin2 = p->d_osin2;
    a.rad = rad;
    a.LFfac = LFfac;
    a.n_m = p->n_m;
    a.n_theta = p->n_theta;
    a.nlev = nlev;
    a.cour = p->cour;
    // 8 rows per workgroup: measured 0.952 ms per launch at l_max=255 x 97
    // against 0.993 ms with 4 rows (two workgroups per CU, 32-byte runs)
    return launch_tr<8>(p, a, nlev, s);
}

}  // namespace mhx
