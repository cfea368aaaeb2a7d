// Split-precision Legendre stage: scaled FP16 hi/lo operands on
// v_mfma_f32_16x16x32_f16 with FP32 accumulation (three products per term:
// hi*hi + hi*lo + lo*hi), batched over radial levels.
//
// The contraction structure is that of the FP64 kernels (mhx_legendre.hip):
// per order m, GEMM rows are theta (synthesis) or l (analysis), K is l or
// theta of one parity, columns are 8 radial levels x re/im, with the
// equatorial-symmetry accumulators Sym/Anti of src/shtransforms.f90:172-219.
//
// Scaling keeps every operand inside FP16's normal range:
//   - tables: one power-of-2 factor per output row (per (m, theta) for the
//     synthesis, per (m, l) for the analysis), common to P and dP, chosen so
//     that the row maximum lies in [2^14, 2^15); generated in FP64 then split;
//   - B operand: one power-of-2 factor per (chunk of K, radial level), common
//     to all columns of that level, found while staging; each chunk is
//     accumulated in fresh FP32 registers and added to the running FP32 sum
//     after multiplying by that factor.
// Row factors are applied in the epilogue; results are stored in FP32
// (Fourier values) or FP64 (spectral coefficients).
//
// f16 16x16x32 lane layouts (checked on gfx950): A lane l holds row l%16,
// k = 8(l/16)..8(l/16)+7; B lane l holds column l%16, the same k; C element
// j of lane l is row 4(l/16)+j, column l%16.
#include "mhx_internal.h"

#include <cmath>

namespace mhx {

namespace {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int WG = 256;
constexpr int KS = 32;    // K per MFMA step
constexpr int NKS = 2;    // MFMA steps per staged chunk
constexpr int CK = KS * NKS;  // K per chunk (per parity for the synthesis)

__device__ inline f4 mfma16(h8 a, h8 b, f4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

__device__ inline int xcd_logical(int b, int nb)
{
    const int nx = 8;
    int per = nb / nx, rem = nb % nx;
    int x = b % nx, i = b / nx;
    return (x < rem ? x * (per + 1) : rem * (per + 1) + (x - rem) * per) + i;
}

// power-of-2 factor s with max*s in [2^14, 2^15); 1 for max == 0; capped
// at 2^115 so that rows of tiny values (P_lm near the poles) stay finite
__device__ inline float pow2_scale(float mx)
{
    if (!(mx > 0.f)) return 1.f;
    int e;
    frexpf(mx, &e);          // mx = f 2^e, f in [0.5, 1)
    if (e < -100) e = -100;
    return ldexpf(1.f, 15 - e);
}

__device__ inline void split_h(float v, _Float16 &hi, _Float16 &lo)
{
    hi = (_Float16)v;
    lo = (_Float16)(v - (float)hi);
}

__device__ inline void apply_factor(double &re, double &im, int ipow, int mpow, double m, double scale)
{
    double f = scale;
    for (int k = 0; k < mpow; ++k) f *= m;
    re *= f;
    im *= f;
    for (int k = 0; k < (ipow & 3); ++k) {
        double t = re;
        re = -im;
        im = t;
    }
}

struct Prod {
    int acc, tab, col;
};

// same product tables as the FP64 kernels
template <int KIND> struct Syn16;
template <> struct Syn16<SYN_SCAL> { static constexpr int NOUT = 1, NCOL = 1, NPROD = 1; static constexpr Prod prod[1] = {{0, 0, 0}}; };
template <> struct Syn16<SYN_GRAD> { static constexpr int NOUT = 2, NCOL = 2, NPROD = 2; static constexpr Prod prod[2] = {{0, 1, 0}, {1, 0, 1}}; };
template <> struct Syn16<SYN_SG> { static constexpr int NOUT = 3, NCOL = 2, NPROD = 3; static constexpr Prod prod[3] = {{0, 0, 0}, {1, 1, 0}, {2, 0, 1}}; };
template <> struct Syn16<SYN_QST> { static constexpr int NOUT = 3, NCOL = 5, NPROD = 5; static constexpr Prod prod[5] = {{0, 0, 0}, {1, 1, 1}, {1, 0, 2}, {2, 0, 3}, {2, 1, 4}}; };
template <> struct Syn16<SYN_SPHTOR> { static constexpr int NOUT = 2, NCOL = 4, NPROD = 4; static constexpr Prod prod[4] = {{0, 1, 0}, {0, 0, 1}, {1, 0, 2}, {1, 1, 3}}; };

template <int KIND> struct Ana16;
template <> struct Ana16<ANA_SH> { static constexpr int NOUT = 1, NCOL = 1, NPROD = 1; static constexpr Prod prod[1] = {{0, 0, 0}}; static constexpr bool divll[1] = {false}; };
template <> struct Ana16<ANA_SPHTOR> { static constexpr int NOUT = 2, NCOL = 4, NPROD = 4; static constexpr Prod prod[4] = {{0, 1, 0}, {0, 0, 1}, {1, 1, 2}, {1, 0, 3}}; static constexpr bool divll[2] = {true, true}; };
template <> struct Ana16<ANA_QST> { static constexpr int NOUT = 3, NCOL = 5, NPROD = 5; static constexpr Prod prod[5] = {{0, 0, 0}, {1, 1, 1}, {1, 0, 2}, {2, 1, 3}, {2, 0, 4}}; static constexpr bool divll[3] = {false, true, true}; };

struct Syn16Args {
    SynJob job;
    const _Float16 *tP, *tdP;   // split tables (hi/lo interleaved per lane)
    const size_t *off;          // [2 mc + par] offsets in halves
    const float *rowscale;      // [mc][nh]
    const int *lstart;
    const double *osin2;
    int l_max, minc, n_m, nh, n_theta, nt16;
};

struct Ana16Args {
    AnaJob job;
    const _Float16 *tP, *tdP;
    const size_t *off;
    const float *rowscale;      // [2 mc + par] blocks of K0 entries: aoff_s
    const size_t *soff;
    const int *lstart;
    int l_max, m_max, minc, n_m, nh, n_theta, nq32;
};

// ------------------------------------------------------------- synthesis
template <int KIND>
__global__ __launch_bounds__(WG) void k_syn16(Syn16Args a)
{
    using S = Syn16<KIND>;
    constexpr int NOUT = S::NOUT, NCOL = S::NCOL, NPROD = S::NPROD;
    // B chunk: [col][par][16 cols][CK k] halves, hi and lo planes
    __shared__ _Float16 bhi[NCOL * 2 * 16 * CK];
    __shared__ _Float16 blo[NCOL * 2 * 16 * CK];
    __shared__ unsigned lvmax[LEV_GROUP];
    __shared__ float lvscale[LEV_GROUP];

    const SynJob &job = a.job;
    const int lb = xcd_logical(blockIdx.x, gridDim.x);
    const int mc = lb % a.n_m;
    const int ng = (job.nlev + LEV_GROUP - 1) / LEV_GROUP;
    const int gy = (lb / a.n_m) % ng, gz = lb / (a.n_m * ng);
    const int m = mc * a.minc;
    const double dm = (double)m;
    const int lev0 = gy * LEV_GROUP;
    const int nlev_here = min(LEV_GROUP, job.nlev - lev0);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int tt = gz * 4 + wave;
    const bool active = tt < a.nt16;
    const int K[2] = {(a.l_max - m) / 2 + 1, (a.l_max - m + 1) / 2};
    const int nks[2] = {(K[0] + KS - 1) / KS, (K[1] + KS - 1) / KS};
    const int l0 = a.lstart[mc];
    const int nchunk = (K[0] + CK - 1) / CK;

    f4 acc[NOUT][2], tmp[NOUT][2];
#pragma unroll
    for (int o = 0; o < NOUT; ++o)
#pragma unroll
        for (int s = 0; s < 2; ++s) acc[o][s] = f4{0.f, 0.f, 0.f, 0.f};

    // staging entries: (par, kk in chunk, level): 2 * CK * 8 = 1024 -> 4/thread
    constexpr int NE = (2 * CK * LEV_GROUP) / WG;
    for (int c = 0; c < nchunk; ++c) {
        if (threadIdx.x < LEV_GROUP) lvmax[threadIdx.x] = 0u;
        __syncthreads();
        double cre[NE][NCOL], cim[NE][NCOL];
#pragma unroll
        for (int e = 0; e < NE; ++e) {
            int idx = threadIdx.x + e * WG;
            int levl = idx % LEV_GROUP;
            int kkl = (idx / LEV_GROUP) % CK;
            int par = idx / (LEV_GROUP * CK);
            int l = m + par + 2 * (c * CK + kkl);
            int lev = lev0 + levl;
            bool valid = (levl < nlev_here) && (l <= a.l_max);
            if (valid) valid = l <= job.lcut[lev];
            double xr[MAX_SRC], xi[MAX_SRC];
            size_t o = ((size_t)lev * job.src_ld + (size_t)(l0 + l - m)) * 2;
#pragma unroll
            for (int sidx = 0; sidx < MAX_SRC; ++sidx) {
                double2 v = make_double2(0.0, 0.0);
                if (valid && job.src[sidx]) v = *reinterpret_cast<const double2 *>(job.src[sidx] + o);
                xr[sidx] = v.x;
                xi[sidx] = v.y;
            }
            double ll = (double)l * (double)(l + 1);
            float mx = 0.f;
#pragma unroll
            for (int col = 0; col < NCOL; ++col) {
                double re = 0.0, im = 0.0;
                if (valid) {
                    const ColOp &op = job.col[col];
                    double g = op.gkind == G_ONE ? 1.0 : (op.gkind == G_LL ? ll : job.or2[lev] * ll);
                    double x0r = 0, x0i = 0, x1r = 0, x1i = 0;
#pragma unroll
                    for (int sidx = 0; sidx < MAX_SRC; ++sidx) {
                        if (op.src0 == sidx) { x0r = xr[sidx]; x0i = xi[sidx]; }
                        if (op.src1 == sidx) { x1r = xr[sidx]; x1i = xi[sidx]; }
                    }
                    re = g * x0r + op.hcoef * x1r;
                    im = g * x0i + op.hcoef * x1i;
                    apply_factor(re, im, op.ipow, op.mpow, dm, op.scale);
                }
                cre[e][col] = re;
                cim[e][col] = im;
                mx = fmaxf(mx, fmaxf(fabsf((float)re), fabsf((float)im)));
            }
            atomicMax(&lvmax[levl], __float_as_uint(mx));
        }
        __syncthreads();
        if (threadIdx.x < LEV_GROUP) lvscale[threadIdx.x] = pow2_scale(__uint_as_float(lvmax[threadIdx.x]));
        __syncthreads();
#pragma unroll
        for (int e = 0; e < NE; ++e) {
            int idx = threadIdx.x + e * WG;
            int levl = idx % LEV_GROUP;
            int kkl = (idx / LEV_GROUP) % CK;
            int par = idx / (LEV_GROUP * CK);
            double sc = (double)lvscale[levl];
#pragma unroll
            for (int col = 0; col < NCOL; ++col) {
                _Float16 h, l;
                int base = ((col * 2 + par) * 16 + levl * 2) * CK + kkl;
                split_h((float)(cre[e][col] * sc), h, l);
                bhi[base] = h;
                blo[base] = l;
                split_h((float)(cim[e][col] * sc), h, l);
                bhi[base + CK] = h;
                blo[base + CK] = l;
            }
        }
        __syncthreads();
        if (active) {
#pragma unroll
            for (int o = 0; o < NOUT; ++o)
#pragma unroll
                for (int s = 0; s < 2; ++s) tmp[o][s] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int par = 0; par < 2; ++par) {
#pragma unroll
                for (int q = 0; q < NKS; ++q) {
                    int ks = c * NKS + q;
                    if (ks >= nks[par]) break;
                    size_t to = a.off[2 * mc + par] + (((size_t)ks * a.nt16 + tt) * 64 + lane) * 16;
                    h8 aPh = *reinterpret_cast<const h8 *>(a.tP + to);
                    h8 aPl = *reinterpret_cast<const h8 *>(a.tP + to + 8);
                    h8 adPh = *reinterpret_cast<const h8 *>(a.tdP + to);
                    h8 adPl = *reinterpret_cast<const h8 *>(a.tdP + to + 8);
#pragma unroll
                    for (int pr = 0; pr < NPROD; ++pr) {
                        const Prod P = S::prod[pr];
                        int bo = ((P.col * 2 + par) * 16 + (lane & 15)) * CK + q * KS + 8 * (lane >> 4);
                        h8 bh = *reinterpret_cast<const h8 *>(bhi + bo);
                        h8 bl = *reinterpret_cast<const h8 *>(blo + bo);
                        int tgt = (P.tab == 0) ? par : (1 - par);
                        h8 ah = P.tab == 0 ? aPh : adPh;
                        h8 al = P.tab == 0 ? aPl : adPl;
                        f4 t = tmp[P.acc][tgt];
                        t = mfma16(ah, bl, t);
                        t = mfma16(al, bh, t);
                        t = mfma16(ah, bh, t);
                        tmp[P.acc][tgt] = t;
                    }
                }
            }
            float inv = 1.f / lvscale[(lane & 15) >> 1];
#pragma unroll
            for (int o = 0; o < NOUT; ++o)
#pragma unroll
                for (int s = 0; s < 2; ++s)
#pragma unroll
                    for (int j = 0; j < 4; ++j) acc[o][s][j] += tmp[o][s][j] * inv;
        }
        __syncthreads();
    }

    // epilogue: north = Sym + Anti, south = Sym - Anti, times the row factor;
    // float2 stores into F[lev][comp][m][theta]
    if (active) {
        const int col = lane & 15, levl = col >> 1, ri = col & 1;
        if (levl < nlev_here) {
            float *fo = reinterpret_cast<float *>(job.fout) + (size_t)(lev0 + levl) * job.fout_ld * 2;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                int th = tt * 16 + 4 * (lane >> 4) + j;
                if (th >= a.nh) continue;
                float rs = 1.f / a.rowscale[(size_t)mc * a.nh + th];
#pragma unroll
                for (int o = 0; o < NOUT; ++o) {
                    float sc = rs;
                    if (job.row_scale[o]) sc *= (float)a.osin2[2 * th];
                    size_t base = (((size_t)job.out_comp[o] * a.n_m + mc) * a.n_theta + 2 * th) * 2;
                    fo[base + ri] = (acc[o][0][j] + acc[o][1][j]) * sc;
                    fo[base + 2 + ri] = (acc[o][0][j] - acc[o][1][j]) * sc;
                }
            }
        }
    }
}

// -------------------------------------------------------------- analysis
template <int KIND>
__global__ __launch_bounds__(WG) void k_ana16(Ana16Args a)
{
    using S = Ana16<KIND>;
    constexpr int NOUT = S::NOUT, NCOL = S::NCOL, NPROD = S::NPROD;
    // B chunk: [col][sym/anti][16 cols][CK theta] halves
    __shared__ _Float16 bhi[NCOL * 2 * 16 * CK];
    __shared__ _Float16 blo[NCOL * 2 * 16 * CK];
    __shared__ unsigned lvmax[LEV_GROUP];
    __shared__ float lvscale[LEV_GROUP];

    const AnaJob &job = a.job;
    const int lb = xcd_logical(blockIdx.x, gridDim.x);
    const int mc = lb % a.n_m;
    const int ng = (job.nlev + LEV_GROUP - 1) / LEV_GROUP;
    const int gy = (lb / a.n_m) % ng, gz = lb / (a.n_m * ng);
    const int m = mc * a.minc;
    const double dm = (double)m;
    const int lev0 = gy * LEV_GROUP;
    const int nlev_here = min(LEV_GROUP, job.nlev - lev0);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int t = gz * 4 + wave;  // pair tile: l = m + 32 t + 0..31
    const int K0 = (a.l_max - m) / 2 + 1;
    const int ntile = (K0 + 15) / 16;
    const bool active = t < ntile;
    const int l0 = a.lstart[mc];
    const int nchunk = (a.nh + CK - 1) / CK;

    int csrc[NCOL], cip[NCOL], cmp[NCOL];
    double csc[NCOL];
    if constexpr (KIND == ANA_SH) {
        csrc[0] = 0; cip[0] = 0; cmp[0] = 0; csc[0] = 1.0;
    } else {
        const int b = (KIND == ANA_QST) ? 1 : 0;
        if constexpr (KIND == ANA_QST) { csrc[0] = 0; cip[0] = 0; cmp[0] = 0; csc[0] = 1.0; }
        csrc[b + 0] = 1; cip[b + 0] = 0; cmp[b + 0] = 0; csc[b + 0] = 1.0;
        csrc[b + 1] = 2; cip[b + 1] = 1; cmp[b + 1] = 1; csc[b + 1] = -1.0;
        csrc[b + 2] = 2; cip[b + 2] = 0; cmp[b + 2] = 0; csc[b + 2] = -1.0;
        csrc[b + 3] = 1; cip[b + 3] = 1; cmp[b + 3] = 1; csc[b + 3] = -1.0;
    }
    bool mok[LEV_GROUP];
#pragma unroll
    for (int i = 0; i < LEV_GROUP; ++i) mok[i] = false;

    f4 acc[NOUT][2], tmp[NOUT][2];
#pragma unroll
    for (int o = 0; o < NOUT; ++o)
#pragma unroll
        for (int s = 0; s < 2; ++s) acc[o][s] = f4{0.f, 0.f, 0.f, 0.f};

    constexpr int NE = (CK * LEV_GROUP) / WG;  // 2
    const float *fin = reinterpret_cast<const float *>(job.fin);
    for (int c = 0; c < nchunk; ++c) {
        if (threadIdx.x < LEV_GROUP) lvmax[threadIdx.x] = 0u;
        __syncthreads();
        float sy[NE][NCOL][2], an[NE][NCOL][2];
#pragma unroll
        for (int e = 0; e < NE; ++e) {
            int idx = threadIdx.x + e * WG;
            int levl = idx % LEV_GROUP;
            int thl = idx / LEV_GROUP;
            int th = c * CK + thl;
            int lev = lev0 + levl;
            bool valid = (levl < nlev_here) && (th < a.nh);
            if (valid) {
                int lc = job.lcut[lev];
                int mlim = lc < a.m_max ? lc : a.m_max;
                valid = m <= mlim;
            }
            float nr[3], ni[3], sr[3], si[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) { nr[k] = ni[k] = sr[k] = si[k] = 0.f; }
            if (valid) {
                int nsrc = (KIND == ANA_SH) ? 1 : 3;
                for (int k = 0; k < nsrc; ++k) {
                    int comp = job.in_comp[k];
                    if (comp < 0) continue;
                    size_t o = ((size_t)lev * job.fin_ld + ((size_t)comp * a.n_theta + 2 * th) * a.n_m + mc) * 2;
                    size_t o2 = o + (size_t)a.n_m * 2;
                    nr[k] = fin[o];
                    ni[k] = fin[o + 1];
                    sr[k] = fin[o2];
                    si[k] = fin[o2 + 1];
                }
            }
            float mx = 0.f;
#pragma unroll
            for (int col = 0; col < NCOL; ++col) {
                int k = csrc[col];
                double syr = (double)nr[k] + sr[k], syi = (double)ni[k] + si[k];
                double anr = (double)nr[k] - sr[k], ani = (double)ni[k] - si[k];
                apply_factor(syr, syi, cip[col], cmp[col], dm, csc[col]);
                apply_factor(anr, ani, cip[col], cmp[col], dm, csc[col]);
                sy[e][col][0] = (float)syr; sy[e][col][1] = (float)syi;
                an[e][col][0] = (float)anr; an[e][col][1] = (float)ani;
                mx = fmaxf(mx, fmaxf(fmaxf(fabsf((float)syr), fabsf((float)syi)),
                                     fmaxf(fabsf((float)anr), fabsf((float)ani))));
            }
            atomicMax(&lvmax[levl], __float_as_uint(mx));
        }
        __syncthreads();
        if (threadIdx.x < LEV_GROUP) lvscale[threadIdx.x] = pow2_scale(__uint_as_float(lvmax[threadIdx.x]));
        __syncthreads();
#pragma unroll
        for (int e = 0; e < NE; ++e) {
            int idx = threadIdx.x + e * WG;
            int levl = idx % LEV_GROUP;
            int thl = idx / LEV_GROUP;
            float sc = lvscale[levl];
#pragma unroll
            for (int col = 0; col < NCOL; ++col) {
#pragma unroll
                for (int ri = 0; ri < 2; ++ri) {
                    _Float16 h, l;
                    int b0 = ((col * 2 + 0) * 16 + levl * 2 + ri) * CK + thl;
                    int b1 = ((col * 2 + 1) * 16 + levl * 2 + ri) * CK + thl;
                    split_h(sy[e][col][ri] * sc, h, l);
                    bhi[b0] = h;
                    blo[b0] = l;
                    split_h(an[e][col][ri] * sc, h, l);
                    bhi[b1] = h;
                    blo[b1] = l;
                }
            }
        }
        __syncthreads();
        if (active) {
#pragma unroll
            for (int o = 0; o < NOUT; ++o)
#pragma unroll
                for (int s = 0; s < 2; ++s) tmp[o][s] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int q = 0; q < NKS; ++q) {
                int kc = c * NKS + q;
                if (kc >= a.nq32) break;
#pragma unroll
                for (int par = 0; par < 2; ++par) {
                    size_t to = a.off[2 * mc + par] + (((size_t)t * a.nq32 + kc) * 64 + lane) * 16;
                    h8 aPh = *reinterpret_cast<const h8 *>(a.tP + to);
                    h8 aPl = *reinterpret_cast<const h8 *>(a.tP + to + 8);
                    h8 adPh = *reinterpret_cast<const h8 *>(a.tdP + to);
                    h8 adPl = *reinterpret_cast<const h8 *>(a.tdP + to + 8);
#pragma unroll
                    for (int pr = 0; pr < NPROD; ++pr) {
                        const Prod P = S::prod[pr];
                        int use = (P.tab == 0) ? par : (1 - par);
                        int bo = ((P.col * 2 + use) * 16 + (lane & 15)) * CK + q * KS + 8 * (lane >> 4);
                        h8 bh = *reinterpret_cast<const h8 *>(bhi + bo);
                        h8 bl = *reinterpret_cast<const h8 *>(blo + bo);
                        h8 ah = P.tab == 0 ? aPh : adPh;
                        h8 al = P.tab == 0 ? aPl : adPl;
                        f4 tt2 = tmp[P.acc][par];
                        tt2 = mfma16(ah, bl, tt2);
                        tt2 = mfma16(al, bh, tt2);
                        tt2 = mfma16(ah, bh, tt2);
                        tmp[P.acc][par] = tt2;
                    }
                }
            }
            float inv = 1.f / lvscale[(lane & 15) >> 1];
#pragma unroll
            for (int o = 0; o < NOUT; ++o)
#pragma unroll
                for (int s = 0; s < 2; ++s)
#pragma unroll
                    for (int j = 0; j < 4; ++j) acc[o][s][j] += tmp[o][s][j] * inv;
        }
        __syncthreads();
    }

    // epilogue: row r -> l = m + par + 2 (16 t + r), times 1/rowscale,
    // divided by l(l+1) for the spheroidal/toroidal outputs
    if (active) {
        const int col = lane & 15, levl = col >> 1, ri = col & 1;
        if (levl < nlev_here) {
            int lev = lev0 + levl;
            int lc = job.lcut[lev];
            bool mkeep = m <= (lc < a.m_max ? lc : a.m_max);
#pragma unroll
            for (int par = 0; par < 2; ++par)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    int r = 4 * (lane >> 4) + j;
                    int kk = 16 * t + r;
                    int l = m + par + 2 * kk;
                    if (l > a.l_max) continue;
                    float rs = a.rowscale[a.soff[2 * mc + par] + kk];
                    bool keep = mkeep && l <= lc;
#pragma unroll
                    for (int o = 0; o < NOUT; ++o) {
                        double v = keep ? (double)acc[o][par][j] / (double)rs : 0.0;
                        if (S::divll[o] && l > 0) v /= (double)l * (double)(l + 1);
                        job.out[o][((size_t)lev * job.out_ld + (size_t)(l0 + l - m)) * 2 + ri] = v;
                    }
                }
        }
    }
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
        size_t o16 = offx[2 * mc + par] + (((size_t)kc * nt16 + (th >> 4)) * 64 + lane) * 16 + (kin & 7);
        _Float16 h, lo;
        split_h((float)(syn_p(sP, off, mc, m, l, l_max, th, nt16, ct) * sc), h, lo);
        hP[o16] = h;
        hP[o16 + 8] = lo;
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
        size_t o16 = offx[2 * mc + par] + (((size_t)(kk >> 4) * nq32 + kc) * 64 + lane) * 16 + (kin & 7);
        _Float16 h, lo;
        split_h((float)(ana_p(aP, off, mc, m, l, l_max, th, nq4, ct) * s), h, lo);
        hP[o16] = h;
        hP[o16 + 8] = lo;
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
    k_rowscale_syn<<<(n + 127) / 128, 128>>>(p->l_max, p->minc, p->n_m, nh, p->nt16, p->d_syn_P,
                                             p->d_syn_dP, p->d_syn_off, p->d_syn16_rs);
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
    {
        const size_t ncr = (size_t)p->lstart[p->n_m - 1] + 2 * (p->n_m - 1) +
                           (p->l_max + 3 - (p->n_m - 1) * p->minc);
        MHX_HIP_CHECK(hipMalloc(&p->d_crec, ncr * sizeof(float)));
        dim3 g((p->l_max + 3 + 127) / 128, p->n_m);
        k_crec<<<g, 128>>>(p->l_max, p->minc, p->n_m, p->d_lstart, p->d_crec);
    }
    MHX_HIP_CHECK(hipGetLastError());
    MHX_HIP_CHECK(hipDeviceSynchronize());
    return MHX_OK;
}

int launch_syn16(const Plan *p, const SynJob &job, hipStream_t s)
{
    if (job.nlev <= 0) return MHX_OK;
    Syn16Args a;
    a.job = job;
    a.tP = p->d_syn16_P;
    a.tdP = p->d_syn16_dP;
    a.off = p->d_syn16_off;
    a.rowscale = p->d_syn16_rs;
    a.lstart = p->d_lstart;
    a.osin2 = p->d_osin2;
    a.l_max = p->l_max;
    a.minc = p->minc;
    a.n_m = p->n_m;
    a.nh = p->nh;
    a.n_theta = p->n_theta;
    a.nt16 = p->nt16;
    dim3 grid((unsigned)(p->n_m * ((job.nlev + LEV_GROUP - 1) / LEV_GROUP) * ((p->nt16 + 3) / 4)));
    switch (job.kind) {
    case SYN_SCAL: k_syn16<SYN_SCAL><<<grid, WG, 0, s>>>(a); break;
    case SYN_GRAD: k_syn16<SYN_GRAD><<<grid, WG, 0, s>>>(a); break;
    case SYN_SG: k_syn16<SYN_SG><<<grid, WG, 0, s>>>(a); break;
    case SYN_QST: k_syn16<SYN_QST><<<grid, WG, 0, s>>>(a); break;
    case SYN_SPHTOR: k_syn16<SYN_SPHTOR><<<grid, WG, 0, s>>>(a); break;
    default: return fail(MHX_E_INVALID, "unknown synthesis kind");
    }
    MHX_HIP_CHECK(hipGetLastError());
    return MHX_OK;
}

int launch_ana16(const Plan *p, const AnaJob &job, hipStream_t s)
{
    if (job.nlev <= 0) return MHX_OK;
    Ana16Args a;
    a.job = job;
    a.tP = p->d_ana16_P;
    a.tdP = p->d_ana16_dP;
    a.off = p->d_ana16_off;
    a.rowscale = p->d_ana16_rs;
    a.soff = p->d_ana16_soff;
    a.lstart = p->d_lstart;
    a.l_max = p->l_max;
    a.m_max = p->m_max;
    a.minc = p->minc;
    a.n_m = p->n_m;
    a.nh = p->nh;
    a.n_theta = p->n_theta;
    a.nq32 = p->nq32;
    int ntile0 = ((p->l_max) / 2 + 1 + 15) / 16;
    dim3 grid((unsigned)(p->n_m * ((job.nlev + LEV_GROUP - 1) / LEV_GROUP) * ((ntile0 + 3) / 4)));
    switch (job.kind) {
    case ANA_SH: k_ana16<ANA_SH><<<grid, WG, 0, s>>>(a); break;
    case ANA_SPHTOR: k_ana16<ANA_SPHTOR><<<grid, WG, 0, s>>>(a); break;
    case ANA_QST: k_ana16<ANA_QST><<<grid, WG, 0, s>>>(a); break;
    default: return fail(MHX_E_INVALID, "unknown analysis kind");
    }
    MHX_HIP_CHECK(hipGetLastError());
    return MHX_OK;
}

}  // namespace mhx
