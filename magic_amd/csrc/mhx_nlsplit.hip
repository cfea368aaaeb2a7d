// Split-precision Legendre kernels of the nonlinear step (MHX_PREC_SPLIT16).
//
// Same contraction as the generic split kernels (mhx_split.hip): per order
// m a dense product over l (synthesis) or theta (analysis) on
// v_mfma_f32_16x16x32_f16 with scaled FP16 hi/lo operands and FP32
// accumulation, three products per term (hi*lo + lo*hi + hi*hi), the
// equatorial-symmetry accumulators of src/shtransforms.f90:172-219.  These
// kernels serve only the fixed transforms of the nonlinear step
// (src/rIter.f90:483-707), so their column operations are compile-time:
//   synthesis  SG    s                     -> s, ds/dtheta_c, ds/dphi
//              QST   Q = l(l+1) a, S = b, T = c               (torpol_to_spat)
//              QSTC  Q = l(l+1) a, S = b, T = or2 l(l+1) c - d (torpol_to_curl_spat)
//   analysis   QST   f, vt, vp -> Q, S, T   (spat_to_qst, crossed vt/vp)
//              SH    f -> Q                 (scal_to_SH)
// One block owns one order m and 16 radial levels (two 16-column MFMA tiles:
// 8 levels x re/im each) and covers every theta tile (synthesis) or l tile
// (analysis), so each B value is built once per block.  The B operand is
// scaled by one power of two per (m, level, column group), chosen before the
// contraction: group 0 feeds the radial (Q) output, group 1 the tangential
// (S, T) outputs, which carry the i m factors.  The synthesis bounds the
// column values over all l of the block; the analysis reads this order's
// column of F2 once (2 max|N +- S|, times max(1, m) for group 1).  The
// accumulators thus carry a single scale and need no per-chunk rescaling.
// Blocks are dealt to the 8 XCDs by m modulo 8 (balanced work per XCD; the
// level groups of one m run back to back on one XCD and share its L2).
#include "mhx_internal.h"

#include <cmath>

namespace mhx {

namespace {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef _Float16 h2 __attribute__((ext_vector_type(2)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int LGS = 16;        // radial levels per block
constexpr int KS = 32;         // K per MFMA step (= per staged chunk)
constexpr int ROWH = KS + 8;   // halves per LDS row (padded)
constexpr int PLANE = 2 * 32 * ROWH;  // halves per B plane: [Sym/Anti or parity][32 cols][ROWH]

__device__ inline f4 mfma16(h8 a, h8 b, f4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// block -> (m slot, rest): XCD x = b % 8 takes orders mc = x, x+8, ...
__device__ inline void xcd_deal(int b, int per_m, int &mc, int &rest)
{
    const int x = b & 7, i = b >> 3;
    mc = x + 8 * (i / per_m);
    rest = i % per_m;
}

// power of two s with mx * s in [2^14, 2^15); 1 for mx == 0
__device__ inline float bscale(double mx)
{
    if (!(mx > 0.0)) return 1.f;
    int e;
    frexp(mx, &e);
    e = e < -100 ? -100 : (e > 125 ? 125 : e);
    return ldexpf(1.f, 15 - e);
}

__device__ inline void split_h(float v, _Float16 &hi, _Float16 &lo)
{
    hi = (_Float16)v;
    lo = (_Float16)(v - (float)hi);
}

// ------------------------------------------------------------- synthesis
template <int V> struct NSyn;
// planes: 0 s, 1 i m s; products (out, table 0 = P / 1 = dP, plane)
template <> struct NSyn<NS_SG> {
    static constexpr int NSRC = 1, NPL = 2, NOUT = 3, NPROD = 3;
    static constexpr int prod[3][3] = {{0, 0, 0}, {1, 1, 0}, {2, 0, 1}};
    static constexpr int pgrp[2] = {0, 1}, ogrp[3] = {0, 0, 1};
};
// planes: 0 Q, 1 S, 2 i m T, 3 i m S, 4 -T
// vr = P Q, vt = dP S + P (i m T), vp = P (i m S) + dP (-T)
template <> struct NSyn<NS_QST> {
    static constexpr int NSRC = 3, NPL = 5, NOUT = 3, NPROD = 5;
    static constexpr int prod[5][3] = {{0, 0, 0}, {1, 1, 1}, {1, 0, 2}, {2, 0, 3}, {2, 1, 4}};
    static constexpr int pgrp[5] = {0, 1, 1, 1, 1}, ogrp[3] = {0, 1, 1};
};
template <> struct NSyn<NS_QSTC> {
    static constexpr int NSRC = 4, NPL = 5, NOUT = 3, NPROD = 5;
    static constexpr int prod[5][3] = {{0, 0, 0}, {1, 1, 1}, {1, 0, 2}, {2, 0, 3}, {2, 1, 4}};
    static constexpr int pgrp[5] = {0, 1, 1, 1, 1}, ogrp[3] = {0, 1, 1};
};

template <int V>
__global__ __launch_bounds__(384, 2) void k_nsyn16(NSynArgs a)
{
    using S = NSyn<V>;
    constexpr int NPL = S::NPL, NOUT = S::NOUT, NPROD = S::NPROD, NSRC = S::NSRC;
    __shared__ __align__(16) _Float16 bhi[NPL * PLANE];
    __shared__ __align__(16) _Float16 blo[NPL * PLANE];
    __shared__ unsigned lvmax[2 * LGS];  // [column group][level]
    __shared__ float lvsc[2 * LGS];

    const int tid = threadIdx.x, nthr = blockDim.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int ng = (a.nlev + LGS - 1) / LGS;
    int mc, rest;
    xcd_deal(blockIdx.x, ng * a.ngz, mc, rest);
    if (mc >= a.n_m) return;
    const int gy = rest % ng, gz = rest / ng;
    const int m = mc * a.minc;
    const double dm = (double)m;
    const int lev0 = gy * LGS, nlh = min(LGS, a.nlev - lev0);
    const int ntn = nlh > 8 ? 2 : 1;
    const int l0 = a.lstart[mc];
    const int K0 = (a.l_max - m) / 2 + 1, K1 = (a.l_max - m + 1) / 2;
    const int nks0 = (K0 + KS - 1) / KS, nks1 = (K1 + KS - 1) / KS;

    // ---- B scales per (level, column group): bound of the column values over l
    if (tid < 2 * LGS) lvmax[tid] = 0u;
    __syncthreads();
    {
        const int tpl = nthr / LGS;
        const int levl = tid / tpl, t = tid - levl * tpl;
        double mx0 = 0.0, mx1 = 0.0;
        if (levl < nlh) {
            const int lev = lev0 + levl;
            const int lhi = min(a.l_max, a.lcut[lev]);
            const double o2 = (V == NS_QSTC) ? a.or2[lev] : 0.0;
            const double mf = dm > 1.0 ? dm : 1.0;
            for (int l = m + t; l <= lhi; l += tpl) {
                const size_t o = ((size_t)lev * a.src_ld + (size_t)(l0 + l - m)) * 2;
                double x[NSRC];
#pragma unroll
                for (int s = 0; s < NSRC; ++s) {
                    double2 v = *reinterpret_cast<const double2 *>(a.src[s] + o);
                    x[s] = fmax(fabs(v.x), fabs(v.y));
                }
                const double ll = (double)l * (double)(l + 1);
                if constexpr (V == NS_SG) {
                    mx0 = fmax(mx0, x[0]);
                    mx1 = fmax(mx1, mf * x[0]);
                } else {
                    double tb = (V == NS_QSTC) ? o2 * ll * x[2] + x[3] : x[2];
                    mx0 = fmax(mx0, ll * x[0]);
                    mx1 = fmax(mx1, mf * fmax(x[1], tb));
                }
            }
            atomicMax(&lvmax[levl], __float_as_uint((float)mx0));
            atomicMax(&lvmax[LGS + levl], __float_as_uint((float)mx1));
        }
    }
    __syncthreads();
    if (tid < 2 * LGS) lvsc[tid] = bscale((double)__uint_as_float(lvmax[tid]));

    f4 acc[NOUT][2][2][2];  // [out][Sym/Anti][theta tile][column tile]
#pragma unroll
    for (int o = 0; o < NOUT; ++o)
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int t = 0; t < 2; ++t)
#pragma unroll
                for (int n = 0; n < 2; ++n) acc[o][s][t][n] = f4{0.f, 0.f, 0.f, 0.f};
    const int tile0 = gz * a.tpb + wave * 2;
    bool tok[2];
    tok[0] = tile0 < a.nt16;
    tok[1] = tile0 + 1 < a.nt16;

    for (int c = 0; c < nks0; ++c) {
        __syncthreads();  // lvsc ready / previous chunk consumed
        // ---- stage l = m + 64 c + 4 g + j (j < 4) for 16 levels: 256 threads;
        // parity par = j & 1 holds k = 32 c + 2 g + (j >> 1)
        if (tid < 256) {
            const int g = tid & 15, levl = tid >> 4;
            const int lev = lev0 + levl;
            const bool lok = levl < nlh;
            const int lhi = lok ? min(a.l_max, a.lcut[lev]) : -1;
            const double o2 = (V == NS_QSTC && lok) ? a.or2[lev] : 0.0;
            const float sc0 = lvsc[levl], sc1 = lvsc[LGS + levl];
            const float fm = (float)dm;
#pragma unroll
            for (int par = 0; par < 2; ++par) {
                // planes of the two l of this parity, FP32, scaled
                float pr[NPL][2], pi[NPL][2];
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    const int l = m + 64 * c + 4 * g + 2 * q + par;
                    const bool valid = l <= lhi;
                    const size_t o = ((size_t)(lok ? lev : 0) * a.src_ld + (size_t)(l0 + l - m)) * 2;
                    double xr[NSRC], xi[NSRC];
#pragma unroll
                    for (int s2 = 0; s2 < NSRC; ++s2) {
                        double2 v = make_double2(0.0, 0.0);
                        if (valid) v = *reinterpret_cast<const double2 *>(a.src[s2] + o);
                        xr[s2] = v.x;
                        xi[s2] = v.y;
                    }
                    const double ll = (double)l * (double)(l + 1);
                    if constexpr (V == NS_SG) {
                        const float sr = (float)xr[0], si = (float)xi[0];
                        pr[0][q] = sr;        pi[0][q] = si;
                        pr[1][q] = -fm * si;  pi[1][q] = fm * sr;
                    } else {
                        float tr, ti;
                        if constexpr (V == NS_QSTC) {
                            tr = (float)(o2 * ll * xr[2] - xr[3]);
                            ti = (float)(o2 * ll * xi[2] - xi[3]);
                        } else {
                            tr = (float)xr[2];
                            ti = (float)xi[2];
                        }
                        const float br = (float)xr[1], bi = (float)xi[1];
                        pr[0][q] = (float)(ll * xr[0]); pi[0][q] = (float)(ll * xi[0]);
                        pr[1][q] = br;        pi[1][q] = bi;
                        pr[2][q] = -fm * ti;  pi[2][q] = fm * tr;
                        pr[3][q] = -fm * bi;  pi[3][q] = fm * br;
                        pr[4][q] = -tr;       pi[4][q] = -ti;
                    }
                }
#pragma unroll
                for (int pl = 0; pl < NPL; ++pl)
#pragma unroll
                    for (int ri = 0; ri < 2; ++ri) {
                        const float sc = S::pgrp[pl] ? sc1 : sc0;
                        h2 vh, vl;
                        _Float16 h, lo;
                        split_h((ri ? pi[pl][0] : pr[pl][0]) * sc, h, lo);
                        vh[0] = h;
                        vl[0] = lo;
                        split_h((ri ? pi[pl][1] : pr[pl][1]) * sc, h, lo);
                        vh[1] = h;
                        vl[1] = lo;
                        const int idx = pl * PLANE + (par * 32 + levl * 2 + ri) * ROWH + 2 * g;
                        *reinterpret_cast<h2 *>(bhi + idx) = vh;
                        *reinterpret_cast<h2 *>(blo + idx) = vl;
                    }
            }
        }
        __syncthreads();
        // ---- contraction: this wave's two theta tiles x ntn column tiles
#pragma unroll
        for (int par = 0; par < 2; ++par) {
            if (c >= (par ? nks1 : nks0)) continue;
            h8 aPh[2], aPl[2], adPh[2], adPl[2];
#pragma unroll
            for (int tt = 0; tt < 2; ++tt) {
                if (tok[tt]) {
                    const size_t to = a.off[2 * mc + par] +
                                      (((size_t)c * a.nt16 + tile0 + tt) * 64 + lane) * 16;
                    aPh[tt] = *reinterpret_cast<const h8 *>(a.tP + to);
                    aPl[tt] = *reinterpret_cast<const h8 *>(a.tP + to + 8);
                    adPh[tt] = *reinterpret_cast<const h8 *>(a.tdP + to);
                    adPl[tt] = *reinterpret_cast<const h8 *>(a.tdP + to + 8);
                }
            }
#pragma unroll
            for (int tn = 0; tn < 2; ++tn) {
                if (tn >= ntn) break;
#pragma unroll
                for (int pr = 0; pr < NPROD; ++pr) {
                    const int o = S::prod[pr][0], tab = S::prod[pr][1], pl = S::prod[pr][2];
                    const int tgt = tab == 0 ? par : 1 - par;
                    const int bo = pl * PLANE + (par * 32 + tn * 16 + (lane & 15)) * ROWH + 8 * (lane >> 4);
                    const h8 bh = *reinterpret_cast<const h8 *>(bhi + bo);
                    const h8 bl = *reinterpret_cast<const h8 *>(blo + bo);
#pragma unroll
                    for (int tt = 0; tt < 2; ++tt) {
                        if (!tok[tt]) continue;
                        const h8 ah = tab == 0 ? aPh[tt] : adPh[tt];
                        const h8 al = tab == 0 ? aPl[tt] : adPl[tt];
                        f4 t = acc[o][tgt][tt][tn];
                        t = mfma16(ah, bl, t);
                        t = mfma16(al, bh, t);
                        t = mfma16(ah, bh, t);
                        acc[o][tgt][tt][tn] = t;
                    }
                }
            }
        }
    }

    // ---- epilogue: north = Sym + Anti, south = Sym - Anti; remove the row
    // and level factors; F[lev][comp][m][theta] FP32 complex
#pragma unroll
    for (int tt = 0; tt < 2; ++tt) {
        if (!tok[tt]) continue;
#pragma unroll
        for (int tn = 0; tn < 2; ++tn) {
            if (tn >= ntn) break;
            const int n = tn * 16 + (lane & 15), levl = n >> 1, ri = n & 1;
            if (levl >= nlh) continue;
            const float inv0 = 1.f / lvsc[levl], inv1 = 1.f / lvsc[LGS + levl];
            float *fo = a.fout + (size_t)(lev0 + levl) * a.fout_ld * 2;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int th = (tile0 + tt) * 16 + 4 * (lane >> 4) + j;
                if (th >= a.nh) continue;
                const float rrs = 1.f / a.rowscale[(size_t)mc * a.nh + th];
#pragma unroll
                for (int o = 0; o < NOUT; ++o) {
                    const float rs = rrs * (S::ogrp[o] ? inv1 : inv0);
                    const size_t base =
                        (((size_t)(a.comp0 + o) * a.n_m + mc) * a.n_theta + 2 * th) * 2;
                    const float sy = acc[o][0][tt][tn][j], an = acc[o][1][tt][tn][j];
                    fo[base + ri] = (sy + an) * rs;
                    fo[base + 2 + ri] = (sy - an) * rs;
                }
            }
        }
    }
}

// -------------------------------------------------------------- analysis
template <int V> struct NAna;
// planes: 0 f, 1 vt, 2 -i m vp, 3 -vp, 4 -i m vt
// Q = wP f, S = wdP vt + wP (-i m vp), T = wdP (-vp) + wP (-i m vt)
template <> struct NAna<NA_QST> {
    static constexpr int NIN = 3, NPL = 5, NOUT = 3, NPROD = 5;
    static constexpr int prod[5][3] = {{0, 0, 0}, {1, 1, 1}, {1, 0, 2}, {2, 1, 3}, {2, 0, 4}};
    static constexpr bool divll[3] = {false, true, true};
    static constexpr int pgrp[5] = {0, 1, 1, 1, 1}, ogrp[3] = {0, 1, 1};
};
template <> struct NAna<NA_SH> {
    static constexpr int NIN = 1, NPL = 1, NOUT = 1, NPROD = 1;
    static constexpr int prod[1][3] = {{0, 0, 0}};
    static constexpr bool divll[1] = {false};
    static constexpr int pgrp[1] = {0}, ogrp[1] = {0};
};

template <int V>
__global__ __launch_bounds__(256) void k_nana16(NAnaArgs a)
{
    using S = NAna<V>;
    constexpr int NPL = S::NPL, NOUT = S::NOUT, NPROD = S::NPROD, NIN = S::NIN;
    __shared__ __align__(16) _Float16 bhi[NPL * PLANE];
    __shared__ __align__(16) _Float16 blo[NPL * PLANE];
    __shared__ unsigned lvmax[2 * LGS];  // [column group][level]
    __shared__ float lvsc[2 * LGS];

    const int tid = threadIdx.x, nthr = blockDim.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int ng = (a.nlev + LGS - 1) / LGS;
    int mc, rest;
    xcd_deal(blockIdx.x, ng * a.ngz, mc, rest);
    if (mc >= a.n_m) return;
    const int gy = rest % ng, gz = rest / ng;
    const int m = mc * a.minc;
    const double dm = (double)m;
    const int lev0 = gy * LGS, nlh = min(LGS, a.nlev - lev0);
    const int ntn = nlh > 8 ? 2 : 1;
    const int l0 = a.lstart[mc];
    const int K0 = (a.l_max - m) / 2 + 1;
    const int ntl = (K0 + 15) / 16;           // l tiles per parity
    const int t0 = gz * 8 + wave * 2;         // this wave's l tiles t0, t0+1
    bool tok[2];
    tok[0] = t0 < ntl;
    tok[1] = t0 + 1 < ntl;

    // ---- B scales per (level, column group) from this order's F2 column:
    // 2 max|f| (N +- S) and 2 max(1, m) max(|vt|, |vp|)
    const float *fin = a.fin;
    if (tid < 2 * LGS) lvmax[tid] = 0u;
    __syncthreads();
    {
        // (level, theta block) pairs; 128 bytes = 16 rows per comp
        for (int idx = tid; idx < LGS * a.ntb; idx += nthr) {
            const int levl = idx % LGS, tb = idx / LGS;
            if (levl >= nlh) continue;
            const int lev = lev0 + levl;
            float mx[NIN];
#pragma unroll
            for (int k = 0; k < NIN; ++k) {
                const f4 *src = reinterpret_cast<const f4 *>(
                    fin + ((size_t)lev * a.fin_ld + (((size_t)a.in_comp[k] * a.ntb + tb) * a.n_m + mc) * 16) * 2);
                const int nrow = min(16, a.n_theta - 16 * tb);  // rows past n_theta are padding
                float v = 0.f;
                for (int q = 0; q < nrow / 2; ++q) {
                    const f4 w = src[q];
                    v = fmaxf(v, fmaxf(fmaxf(fabsf(w[0]), fabsf(w[1])), fmaxf(fabsf(w[2]), fabsf(w[3]))));
                }
                mx[k] = v;
            }
            atomicMax(&lvmax[levl], __float_as_uint(mx[0]));
            if constexpr (NIN == 3) atomicMax(&lvmax[LGS + levl], __float_as_uint(fmaxf(mx[1], mx[2])));
        }
    }
    __syncthreads();
    if (tid < 2 * LGS) {
        const double f = tid < LGS ? 2.0 : 2.0 * (dm > 1.0 ? dm : 1.0);
        lvsc[tid] = bscale(f * (double)__uint_as_float(lvmax[tid]));
    }

    f4 acc[NOUT][2][2][2];  // [out][parity][l tile][column tile]
#pragma unroll
    for (int o = 0; o < NOUT; ++o)
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int t = 0; t < 2; ++t)
#pragma unroll
                for (int n = 0; n < 2; ++n) acc[o][s][t][n] = f4{0.f, 0.f, 0.f, 0.f};

    // F2 layout: [lev][comp][theta block of 16 rows][m][16] complex
    for (int kc = 0; kc < a.nq32; ++kc) {
        __syncthreads();
        // ---- stage theta_h = 32 kc + 8 tb + 2 q + (0, 1): rows 64 kc + 16 tb + 4 q .. +3
        {
            const int q = tid & 3, tb = (tid >> 2) & 3, levl = tid >> 4;
            const int lev = lev0 + levl;
            const int thb = kc * 4 + tb;        // theta block (16 rows)
            const int th0 = 32 * kc + 8 * tb + 2 * q;
            bool valid = levl < nlh && th0 < a.nh;
            const bool ok1 = th0 + 1 < a.nh;  // rows past n_theta are padding
            if (valid) {
                const int lc = a.lcut[lev];
                valid = m <= (lc < a.m_max ? lc : a.m_max);
            }
            const float sc0 = lvsc[levl], sc1 = lvsc[LGS + levl];
            // rows 2 th, 2 th + 1 for th = th0, th0 + 1: 4 complex, 32 bytes
            float x[NIN][8];
#pragma unroll
            for (int k = 0; k < NIN; ++k) {
                f4 v0 = f4{0.f, 0.f, 0.f, 0.f}, v1 = v0;
                if (valid) {
                    const size_t o = ((size_t)lev * a.fin_ld +
                                      (((size_t)a.in_comp[k] * a.ntb + thb) * a.n_m + mc) * 16 + 4 * q) * 2;
                    v0 = *reinterpret_cast<const f4 *>(fin + o);
                    if (ok1) v1 = *reinterpret_cast<const f4 *>(fin + o + 4);
                }
                // x = N0re N0im S0re S0im N1re N1im S1re S1im
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    x[k][e] = v0[e];
                    x[k][4 + e] = v1[e];
                }
            }
            // planes per (Sym/Anti, re/im), two theta_h per half2
            const float fm = (float)dm;
            const int kk = 8 * tb + 2 * q;  // theta_h within the chunk
#pragma unroll
            for (int sa = 0; sa < 2; ++sa) {
                float pr[NPL][2], pi[NPL][2];
#pragma unroll
                for (int p = 0; p < 2; ++p) {
                    float vr[NIN], vi[NIN];
#pragma unroll
                    for (int k = 0; k < NIN; ++k) {
                        const float nr = x[k][4 * p], ni = x[k][4 * p + 1];
                        const float sr = x[k][4 * p + 2], si = x[k][4 * p + 3];
                        vr[k] = sa ? nr - sr : nr + sr;
                        vi[k] = sa ? ni - si : ni + si;
                    }
                    pr[0][p] = vr[0]; pi[0][p] = vi[0];
                    if constexpr (V == NA_QST) {
                        pr[1][p] = vr[1];       pi[1][p] = vi[1];
                        pr[2][p] = fm * vi[2];  pi[2][p] = -fm * vr[2];   // -i m vp
                        pr[3][p] = -vr[2];      pi[3][p] = -vi[2];
                        pr[4][p] = fm * vi[1];  pi[4][p] = -fm * vr[1];   // -i m vt
                    }
                }
#pragma unroll
                for (int pl = 0; pl < NPL; ++pl)
#pragma unroll
                    for (int ri = 0; ri < 2; ++ri) {
                        const float sc = S::pgrp[pl] ? sc1 : sc0;
                        h2 vh, vl;
                        _Float16 h, lo;
                        split_h((ri ? pi[pl][0] : pr[pl][0]) * sc, h, lo);
                        vh[0] = h;
                        vl[0] = lo;
                        split_h((ri ? pi[pl][1] : pr[pl][1]) * sc, h, lo);
                        vh[1] = h;
                        vl[1] = lo;
                        const int idx = pl * PLANE + (sa * 32 + levl * 2 + ri) * ROWH + kk;
                        *reinterpret_cast<h2 *>(bhi + idx) = vh;
                        *reinterpret_cast<h2 *>(blo + idx) = vl;
                    }
            }
        }
        __syncthreads();
        // ---- contraction
#pragma unroll
        for (int par = 0; par < 2; ++par) {
            h8 aPh[2], aPl[2], adPh[2], adPl[2];
#pragma unroll
            for (int tt = 0; tt < 2; ++tt) {
                if (tok[tt]) {
                    const size_t to = a.off[2 * mc + par] +
                                      (((size_t)(t0 + tt) * a.nq32 + kc) * 64 + lane) * 16;
                    aPh[tt] = *reinterpret_cast<const h8 *>(a.tP + to);
                    aPl[tt] = *reinterpret_cast<const h8 *>(a.tP + to + 8);
                    if constexpr (V == NA_QST) {
                        adPh[tt] = *reinterpret_cast<const h8 *>(a.tdP + to);
                        adPl[tt] = *reinterpret_cast<const h8 *>(a.tdP + to + 8);
                    }
                }
            }
#pragma unroll
            for (int tn = 0; tn < 2; ++tn) {
                if (tn >= ntn) break;
#pragma unroll
                for (int pr = 0; pr < NPROD; ++pr) {
                    const int o = S::prod[pr][0], tab = S::prod[pr][1], pl = S::prod[pr][2];
                    const int use = tab == 0 ? par : 1 - par;
                    const int bo = pl * PLANE + (use * 32 + tn * 16 + (lane & 15)) * ROWH + 8 * (lane >> 4);
                    const h8 bh = *reinterpret_cast<const h8 *>(bhi + bo);
                    const h8 bl = *reinterpret_cast<const h8 *>(blo + bo);
#pragma unroll
                    for (int tt = 0; tt < 2; ++tt) {
                        if (!tok[tt]) continue;
                        const h8 ah = tab == 0 ? aPh[tt] : adPh[tt];
                        const h8 al = tab == 0 ? aPl[tt] : adPl[tt];
                        f4 t = acc[o][par][tt][tn];
                        t = mfma16(ah, bl, t);
                        t = mfma16(al, bh, t);
                        t = mfma16(ah, bh, t);
                        acc[o][par][tt][tn] = t;
                    }
                }
            }
        }
    }

    // ---- epilogue: row r of tile t -> l = m + par + 2 (16 t + r); remove the
    // row and level factors, / l(l+1) for S and T, lcut masks
#pragma unroll
    for (int tn = 0; tn < 2; ++tn) {
        if (tn >= ntn) break;
        const int n = tn * 16 + (lane & 15), levl = n >> 1, ri = n & 1;
        if (levl >= nlh) continue;
        const int lev = lev0 + levl;
        const int lc = a.lcut[lev];
        const bool mkeep = m <= (lc < a.m_max ? lc : a.m_max);
        const double inv0 = 1.0 / (double)lvsc[levl], inv1 = 1.0 / (double)lvsc[LGS + levl];
#pragma unroll
        for (int tt = 0; tt < 2; ++tt) {
            if (!tok[tt]) continue;
#pragma unroll
            for (int par = 0; par < 2; ++par)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int kk = 16 * (t0 + tt) + 4 * (lane >> 4) + j;
                    const int l = m + par + 2 * kk;
                    if (l > a.l_max) continue;
                    const double rrs = 1.0 / (double)a.rowscale[a.soff[2 * mc + par] + kk];
                    const bool keep = mkeep && l <= lc;
                    const double dll = (double)l * (double)(l + 1);
#pragma unroll
                    for (int o = 0; o < NOUT; ++o) {
                        const double rs = rrs * (S::ogrp[o] ? inv1 : inv0);
                        double v = keep ? (double)acc[o][par][tt][tn][j] * rs : 0.0;
                        if (S::divll[o] && l > 0) v /= dll;
                        a.out[o][((size_t)lev * a.out_ld + (size_t)(l0 + l - m)) * 2 + ri] = v;
                    }
                }
        }
    }
}

}  // namespace

int launch_nsyn16(const Plan *p, int kind, const NSynArgs &args, hipStream_t s)
{
    if (args.nlev <= 0) return MHX_OK;
    NSynArgs a = args;
    a.tP = p->d_syn16_P;
    a.tdP = p->d_syn16_dP;
    a.off = p->d_syn16_off;
    a.rowscale = p->d_syn16_rs;
    a.lstart = p->d_lstart;
    a.l_max = p->l_max;
    a.minc = p->minc;
    a.n_m = p->n_m;
    a.nh = p->nh;
    a.n_theta = p->n_theta;
    a.nt16 = p->nt16;
    const int nw = 6;  // 12 theta tiles per block (all of l_max = 255); 256 threads stage B
    a.tpb = 2 * nw;
    a.ngz = (p->nt16 + a.tpb - 1) / a.tpb;
    const int ng = (a.nlev + LGS - 1) / LGS;
    dim3 grid((unsigned)(8 * ((p->n_m + 7) / 8) * ng * a.ngz));
    const int nthr = 64 * nw;
    switch (kind) {
    case NS_SG: k_nsyn16<NS_SG><<<grid, nthr, 0, s>>>(a); break;
    case NS_QST: k_nsyn16<NS_QST><<<grid, nthr, 0, s>>>(a); break;
    case NS_QSTC: k_nsyn16<NS_QSTC><<<grid, nthr, 0, s>>>(a); break;
    default: return fail(MHX_E_INVALID, "unknown nonlinear-step synthesis kind");
    }
    MHX_HIP_CHECK(hipGetLastError());
    return MHX_OK;
}

int launch_nana16(const Plan *p, int kind, const NAnaArgs &args, hipStream_t s)
{
    if (args.nlev <= 0) return MHX_OK;
    NAnaArgs a = args;
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
    a.ntb = (p->n_theta + 15) / 16;
    a.nq32 = p->nq32;
    const int ntl0 = (p->l_max / 2 + 1 + 15) / 16;
    a.ngz = (ntl0 + 7) / 8;
    const int ng = (a.nlev + LGS - 1) / LGS;
    dim3 grid((unsigned)(8 * ((p->n_m + 7) / 8) * ng * a.ngz));
    switch (kind) {
    case NA_QST: k_nana16<NA_QST><<<grid, 256, 0, s>>>(a); break;
    case NA_SH: k_nana16<NA_SH><<<grid, 256, 0, s>>>(a); break;
    default: return fail(MHX_E_INVALID, "unknown nonlinear-step analysis kind");
    }
    MHX_HIP_CHECK(hipGetLastError());
    return MHX_OK;
}

}  // namespace mhx
