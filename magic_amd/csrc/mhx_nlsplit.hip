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
// One block owns one order m and 8 radial levels (one 16-column MFMA tile:
// 8 levels x re/im) and covers every theta tile (synthesis) or l tile
// (analysis), so each B value is built once per block.  The B operand is
// scaled by one power of two per (m, level, column group), chosen before the
// contraction: group 0 feeds the radial (Q) output, group 1 the tangential
// (S, T) outputs, which carry the i m factors.  The synthesis bounds the
// column values over all l of the block; the analysis reads this order's
// column of F2 once (2 max|N +- S|, times max(1, m) for group 1).  The
// accumulators thus carry a single scale and need no per-chunk rescaling.
// Blocks are dealt to the 8 XCDs by m modulo 8 (balanced work per XCD; the
// level groups of one m run back to back on one XCD and share its L2).
#include "mhx_fftcore.h"
#include <utility>

#include <cmath>
#include <type_traits>

namespace mhx {

namespace {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef _Float16 h2 __attribute__((ext_vector_type(2)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int TNC = 2;         // 16-column MFMA tiles per block (8 levels each)
constexpr int LGS = 8 * TNC;   // radial levels per block
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

// REC: the table fragments are not read but generated per chunk by the FP64
// three-term recurrence in l (src/plms.f90:81-97 in the form
// P_{l+1} = A_{l+1} (cos(theta) P_l - eps_l P_{l-1}), A = 1/eps,
// eps_l = sqrt((l^2 - m^2) / (4 l^2 - 1))) and the identity
// sin(theta) dP_l/dtheta = l eps_{l+1} P_{l+1} - (l + 1) eps_l P_{l-1}
// (src/plms.f90:142-162).  Lane (row r, group g) of a wave owns theta rows
// r of its two tiles and, in chunk c, l = m + 64 c + 16 g + j (j < 16): it
// walks 64 steps per chunk, keeping the first 16, from seeds P_mm(theta)
// times the row scale held as FP64 mantissa and power-of-two exponent (the
// exponent is renormalised once per chunk, so no value leaves the double
// range however small P_mm is).  The coefficients of the block's order sit
// in LDS.
constexpr int RLMAX = 1152;  // coefficient entries l - m (l_max <= 1023)

template <int V, bool REC>
__global__ __launch_bounds__(REC ? 256 : 384, 2) void k_nsyn16(NSynArgs a)
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
    const int ntn = nlh > 8 ? 2 : 1;  // column tiles in use (<= TNC)
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
            // four l per trip, their loads issued before use (clamped to lhi;
            // a clamped repeat does not change the maxima)
            for (int lb = m + t; lb <= lhi; lb += 4 * tpl) {
                double2 v[4][NSRC];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int l = min(lb + u * tpl, lhi);
                    const size_t o = ((size_t)lev * a.src_ld + (size_t)(l0 + l - m)) * 2;
#pragma unroll
                    for (int s = 0; s < NSRC; ++s) v[u][s] = *reinterpret_cast<const double2 *>(a.src[s] + o);
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int l = min(lb + u * tpl, lhi);
                    double x[NSRC];
#pragma unroll
                    for (int s = 0; s < NSRC; ++s) x[s] = fmax(fabs(v[u][s].x), fabs(v[u][s].y));
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
            }
            atomicMax(&lvmax[levl], __float_as_uint((float)mx0));
            atomicMax(&lvmax[LGS + levl], __float_as_uint((float)mx1));
        }
    }
    __syncthreads();
    if (tid < 2 * LGS) lvsc[tid] = bscale((double)__uint_as_float(lvmax[tid]));

    f4 acc[NOUT][2][2][TNC];  // [out][Sym/Anti][theta tile][column tile]
#pragma unroll
    for (int o = 0; o < NOUT; ++o)
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int t = 0; t < 2; ++t)
#pragma unroll
                for (int n = 0; n < TNC; ++n) acc[o][s][t][n] = f4{0.f, 0.f, 0.f, 0.f};
    const int tile0 = gz * a.tpb + wave * 2;
    bool tok[2];
    tok[0] = tile0 < a.nt16;
    tok[1] = tile0 + 1 < a.nt16;

    // REC: recurrence coefficients (A_l, eps_l), l = m + i, zero past l_max + 1
    __shared__ double2 rco[REC ? RLMAX : 1];
    const int rg = lane >> 4;  // lane group: l offset 16 rg within a chunk
    double rp[2], rq[2], rct[2], reps = 0.0;  // P_{l-1}, P_l (x 2^-E), cos(theta), eps_l
    int rE[2];
    if constexpr (REC) {
        for (int i = tid; i < RLMAX; i += nthr) {
            const int l = m + i;
            double A = 0.0, e = 0.0;
            if (i > 0 && l <= a.l_max + 1) {
                const double dl = (double)l;
                const double num = (dl - dm) * (dl + dm), den = (2.0 * dl - 1.0) * (2.0 * dl + 1.0);
                A = sqrt(den / num);
                e = sqrt(num / den);
            }
            rco[i] = make_double2(A, e);
        }
#pragma unroll
        for (int tt = 0; tt < 2; ++tt) {
            const int th = (tile0 + tt) * 16 + (lane & 15);
            const bool ok = tok[tt] && th < a.nh;
            const double2 sd = ok ? a.seed[(size_t)mc * a.nh + th] : make_double2(0.0, 0.0);
            rq[tt] = sd.x;
            rp[tt] = 0.0;
            rE[tt] = (int)sd.y;
            rct[tt] = ok ? a.cost[th] : 0.0;
        }
        __syncthreads();
        // group rg starts at l = m + 16 rg
        for (int j = 0; j < 48; ++j) {
            if (j < 16 * rg) {
                const double2 ce = rco[j + 1];
#pragma unroll
                for (int tt = 0; tt < 2; ++tt) {
                    const double pn = ce.x * (rct[tt] * rq[tt] - reps * rp[tt]);
                    rp[tt] = rq[tt];
                    rq[tt] = pn;
                }
                reps = ce.y;
            }
        }
    }

    for (int c = 0; c < nks0; ++c) {
        __syncthreads();  // lvsc ready / previous chunk consumed
        // ---- stage l = m + 64 c + 4 g + j (j < 4) for LGS levels: 16 LGS
        // threads; parity par = j & 1 holds k = 32 c + 2 g + (j >> 1)
        if (tid < 16 * LGS) {
            const int g = tid & 15, levl = tid >> 4;
            const int lev = lev0 + levl;
            const bool lok = levl < nlh;
            const int lhi = lok ? min(a.l_max, a.lcut[lev]) : -1;
            const double o2 = (V == NS_QSTC && lok) ? a.or2[lev] : 0.0;
            const float sc0 = lvsc[levl], sc1 = lvsc[LGS + levl];
            const float fm = (float)dm;
            // the four l of this thread, all loads issued first at addresses
            // clamped into the order's range, then masked
            double2 xv[2][2][NSRC];  // [par][q][source]
#pragma unroll
            for (int par = 0; par < 2; ++par)
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    const int l = min(m + 64 * c + 4 * g + 2 * q + par, a.l_max);
                    const size_t o = ((size_t)(lok ? lev : 0) * a.src_ld + (size_t)(l0 + l - m)) * 2;
#pragma unroll
                    for (int s2 = 0; s2 < NSRC; ++s2)
                        xv[par][q][s2] = *reinterpret_cast<const double2 *>(a.src[s2] + o);
                }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int par = 0; par < 2; ++par) {
                // planes of the two l of this parity, FP32, scaled
                float pr[NPL][2], pi[NPL][2];
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    const int l = m + 64 * c + 4 * g + 2 * q + par;
                    const bool valid = l <= lhi;
                    double xr[NSRC], xi[NSRC];
#pragma unroll
                    for (int s2 = 0; s2 < NSRC; ++s2) {
                        xr[s2] = valid ? xv[par][q][s2].x : 0.0;
                        xi[s2] = valid ? xv[par][q][s2].y : 0.0;
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
        if constexpr (REC) {
            // ---- per theta tile: generate the chunk's fragments of both
            // parities (element e of parity par is l = l0 + 2 e + par,
            // l0 = m + 64 c + 16 rg), then contract them
            const int i0 = 64 * c + 16 * rg;
            const double dl0 = (double)(m + i0), reps0 = reps, dlmax = (double)a.l_max;
#pragma unroll
            for (int tt = 0; tt < 2; ++tt) {
                if (!tok[tt]) continue;
                // renormalise: |P| within [2^-256, 2^256] x 2^E at the chunk start
                {
                    const double v = fabs(rq[tt]) + fabs(rp[tt]);
                    const bool lo = v != 0.0 && v < 0x1p-256, hi = v > 0x1p256;
                    const double f = lo ? 0x1p256 : (hi ? 0x1p-256 : 1.0);
                    rq[tt] *= f;
                    rp[tt] *= f;
                    rE[tt] += lo ? -256 : (hi ? 256 : 0);
                }
                const double sc = rE[tt] >= -1000 ? ldexp(1.0, rE[tt]) : 0.0;
                h8 gPh[2], gPl[2], gdPh[2], gdPl[2];  // [parity]
                double ep = reps0, q = rq[tt], pv = rp[tt];
                const double x = rct[tt];
#pragma unroll
                for (int j = 0; j < 16; ++j) {
                    const double2 ce = rco[i0 + j + 1];  // A_{l+1}, eps_{l+1}, l = l0 + j
                    const double dl = dl0 + (double)j;
                    const double pn = ce.x * (x * q - ep * pv);
                    const double dp = dl * ce.y * pn - (dl + 1.0) * ep * pv;
                    // entries past l_max are zero, as in the stored tables
                    // (P_{l_max+1} is outside the row scale's range)
                    const bool in = dl <= dlmax;
                    _Float16 h, l;
                    split_h(in ? (float)(q * sc) : 0.f, h, l);
                    gPh[j & 1][j >> 1] = h;
                    gPl[j & 1][j >> 1] = l;
                    split_h(in ? (float)(dp * sc) : 0.f, h, l);
                    gdPh[j & 1][j >> 1] = h;
                    gdPl[j & 1][j >> 1] = l;
                    pv = q;
                    q = pn;
                    ep = ce.y;
                    if (j & 1) __builtin_amdgcn_sched_barrier(0);
                }
#pragma unroll 4
                for (int j = 16; j < 64; ++j) {
                    const double2 ce = rco[i0 + j + 1];
                    const double pn = ce.x * (x * q - ep * pv);
                    pv = q;
                    q = pn;
                    ep = ce.y;
                }
                rq[tt] = q;
                rp[tt] = pv;
                reps = ep;
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int par = 0; par < 2; ++par) {
                    if (c >= (par ? nks1 : nks0)) continue;
#pragma unroll
                    for (int tn = 0; tn < TNC; ++tn) {
                        if (tn >= ntn) break;
#pragma unroll
                        for (int pr = 0; pr < NPROD; ++pr) {
                            const int o = S::prod[pr][0], tab = S::prod[pr][1], pl = S::prod[pr][2];
                            const int tgt = tab == 0 ? par : 1 - par;
                            const int bo = pl * PLANE + (par * 32 + tn * 16 + (lane & 15)) * ROWH + 8 * (lane >> 4);
                            const h8 bh = *reinterpret_cast<const h8 *>(bhi + bo);
                            const h8 bl = *reinterpret_cast<const h8 *>(blo + bo);
                            const h8 ah = tab == 0 ? gPh[par] : gdPh[par];
                            const h8 al = tab == 0 ? gPl[par] : gdPl[par];
                            f4 t = acc[o][tgt][tt][tn];
                            t = mfma16(ah, bl, t);
                            t = mfma16(al, bh, t);
                            t = mfma16(ah, bh, t);
                            acc[o][tgt][tt][tn] = t;
                        }
                    }
                }
            }
            continue;
        }
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
            for (int tn = 0; tn < TNC; ++tn) {
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
        for (int tn = 0; tn < TNC; ++tn) {
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

// ------------------------- whole-K / whole-theta forms (l_max <= 255)
// One block covers one order m, 8 radial levels and all theta rows (4 waves x
// 3 theta tiles in synthesis); its B operand for the whole K range sits in LDS.
constexpr int WLG = 8;                   // levels per block
constexpr int WTT = 3;                   // theta tiles per wave
constexpr int WNT = 384;                 // n_theta at most (one block covers all rows)

// compile-time loop: f(integral_constant<int, S>) for S = 0 .. N - 1
template <class F, int... S>
__device__ __forceinline__ void static_for_impl(F &&f, std::integer_sequence<int, S...>)
{
    (f(std::integral_constant<int, S>()), ...);
}
template <int N, class F> __device__ __forceinline__ void static_for(F &&f)
{
    static_for_impl(f, std::make_integer_sequence<int, N>());
}

// ----------------------------------- synthesis, whole-K form on P only
// The dP products are folded onto P by
//   sin(theta) dP_l = l c_{l+1} P_{l+1} - (l+1) c_l P_{l-1},
//   c_l = sqrt((l^2 - m^2) / (4 l^2 - 1))   (src/plms.f90:81-97, :142-162),
// so sum_l dP_l x_l = sum_l P_l x'_l with
//   x'_l = (l-1) c_l x_{l-1} - (l+2) c_{l+1} x_{l+1},   l = m .. l_max + 1.
// Columns against P (table to l_max + 1):
//   QST: C0 = Q, C1 = S' + i m T (-> vt), C2 = i m S - T' (-> vp)
//   SG : C0 = s, C1 = s' (-> ds/dtheta_c); ds/dphi = i m (P C0) in the epilogue
// Staging: 1. Q, S, T (SG: s) in FP32 into the LDS rows of the columns;
// 2. the neighbour combinations in place (all reads before all writes);
// 3. one power of two per (column, level); rows split in place to hi/lo;
// 4. contraction over the k steps of both parities, the table fragments
//    loaded PFD steps ahead (the step's MFMA work is shorter than an L2 hit);
// 5. epilogue: outputs staged per level in LDS, written as theta runs.
// table fragments in flight: a ring of SRING (synthesis) / ARING (analysis)
// steps, loaded SRING - 1 / ARING - 1 steps ahead of their MFMAs (deeper
// rings measured no faster: 5 / 6 gave 0.1351 / 0.1612 ms per QST launch
// against 0.1362 / 0.1599, and the scalar analysis lost a wave per SIMD)
constexpr int SRING = 4, ARING = 4;
constexpr int PKP = 160;                 // k per parity held (l_max + 1 - m)/2 + 1 <= 160
constexpr int PROW = 4 * PKP + 16;       // bytes per LDS row (padded)

// LDS of one whole-K synthesis item: B rows [column][parity][16 columns][hi |
// lo], per-(column, level) maxima and scales, the order's c_l (l - m <= l_max
// + 2 - m) and row factors
template <int V> struct NSynPL {
    static constexpr int NCOL = V == NS_S ? 1 : (V == NS_SG ? 2 : 3);
    static constexpr size_t SB = (size_t)NCOL * 2 * 16 * PROW;
    static constexpr size_t O_LVMAX = SB, O_LVSC = O_LVMAX + 4 * NCOL * WLG;
    static constexpr size_t O_SCR = O_LVSC + 4 * NCOL * WLG, O_SRSV = O_SCR + 4 * (2 * PKP + 8);
    static constexpr size_t BYTES = O_SRSV + 4 * (WNT / 2);
};

// One whole-K synthesis item: order mc, level group gy, theta group gz, on
// the LDS area `lds` (NSynPL<V>::BYTES), 256 threads; tid = threadIdx.x
// (a parameter, so that a caller looping over items can keep the values
// derived from it inside the loop).
// NRL: 128-wide rounds of l that can hold l <= l_max (1 for l_max < 128, else 2)
template <int V, int NRL>
__device__ __forceinline__ void nsyn16p_item(const NSynArgs &a, int mc, int gy, int gz, unsigned char *lds,
                                             const int tid)
{
    constexpr bool SCAL = V == NS_SG || V == NS_S;
    constexpr int NSRC = SCAL ? 1 : (V == NS_QSTC ? 4 : 3);
    constexpr int NCOL = NSynPL<V>::NCOL, NROW = NCOL * 2 * 16;
    constexpr int NOUT = V == NS_S ? 1 : 3;
    constexpr int SB = NROW * PROW;
    unsigned char *sb = lds;
    unsigned *lvmax = reinterpret_cast<unsigned *>(lds + NSynPL<V>::O_LVMAX);
    float *lvsc = reinterpret_cast<float *>(lds + NSynPL<V>::O_LVSC);
    float *scr = reinterpret_cast<float *>(lds + NSynPL<V>::O_SCR);
    float *srsv = reinterpret_cast<float *>(lds + NSynPL<V>::O_SRSV);

    const int wave = tid >> 6, lane = tid & 63;
    const int m = mc * a.minc;
    const float fm = (float)m;
    const int lev0 = gy * WLG, nlh = min(WLG, a.nlev - lev0);
    const int l0 = a.lstart[mc];
    const int K0 = (a.l_max + 1 - m) / 2 + 1, K1 = (a.l_max + 2 - m) / 2;
    const int nks0 = (K0 + KS - 1) / KS, nks1 = (K1 + KS - 1) / KS;
    if (tid < NCOL * WLG) lvmax[tid] = 0u;
    {
        const float *cr = a.crec + a.lstart[mc] + 2 * mc;
        const int dmax = a.l_max + 2 - m;
        for (int d = tid; d < 2 * PKP + 8; d += 256) scr[d] = d <= dmax ? cr[d] : 0.f;
        for (int t = tid; t < a.nh; t += 256) srsv[t] = a.rowscale[(size_t)mc * a.nh + t];
    }
    const int levl = tid >> 5, g0 = tid & 31;   // staging: level, group of 4 l
    const int lev = lev0 + levl;
    const bool lok = levl < nlh;
    const int lhi = lok ? min(a.l_max, a.lcut[lev]) : -1;
    __syncthreads();

    // c_l for the thread's l = lb .. lb + 4 of each round (0 past l_max + 2)
    float cv[3][5];
#pragma unroll
    for (int it = 0; it < 3; ++it)
#pragma unroll
        for (int q = 0; q < 5; ++q) {
            const int d = 4 * (g0 + 32 * it) + q;
            cv[it][q] = d < 2 * PKP + 8 ? scr[d] : 0.f;
        }
    // ---- 1. Q, S, T (SG: s) in FP32 registers, l = m + 4 (g0 + 32 it) + j.
    // The sources are read lane-consecutive (thread g0 of a level reads
    // l = m + 32 r + g0, r < 4 NRL: each four-lane group one 64-byte run),
    // the pointwise FP64 column values rounded to FP32 and moved through the
    // B-row area of LDS (free until step 3) to the layout of steps 2 and 3.
    // Rounds it >= NRL hold only l > l_max (zero sources): they are not read.
    float2 cq[3][4], cs[3][4], ct[3][4];  // column 0 (Q or s), S (or s), T
    {
        constexpr int NSC = SCAL ? 1 : 3;      // staged columns
        constexpr int SRP = 128 * NRL + 4;     // staged row pitch (float2)
        static_assert((size_t)WLG * NSC * SRP * sizeof(float2) <= (size_t)SB,
                      "staged columns exceed the B-row area");
        float2 *srow = reinterpret_cast<float2 *>(sb) + (size_t)levl * NSC * SRP;
        const double o2 = (V == NS_QSTC && lok) ? a.or2[lev] : 0.0;
        // all loads first, unconditional at a clamped l (a load under a branch
        // waits for itself before the next one); values past lhi are zeroed
        double2 xa[4 * NRL][NSRC];
#pragma unroll
        for (int r = 0; r < 4 * NRL; ++r) {
            const int l = m + 32 * r + g0;
            const int lc = min(l, max(lhi, m));
            const size_t o = ((size_t)(lok ? lev : lev0) * a.src_ld + (size_t)(l0 + lc - m)) * 2;
#pragma unroll
            for (int s2 = 0; s2 < NSRC; ++s2) xa[r][s2] = *reinterpret_cast<const double2 *>(a.src[s2] + o);
        }
        __builtin_amdgcn_sched_barrier(0);
        const float2 z = make_float2(0.f, 0.f);
#pragma unroll
        for (int r = 0; r < 4 * NRL; ++r) {
            const int l = m + 32 * r + g0, li = 32 * r + g0;
            const bool in = l <= lhi;
            const double2 *x = xa[r];
            const double ll = (double)l * (double)(l + 1);
            if constexpr (SCAL) {
                srow[li] = in ? make_float2((float)x[0].x, (float)x[0].y) : z;
            } else {
                srow[li] = in ? make_float2((float)(ll * x[0].x), (float)(ll * x[0].y)) : z;
                srow[SRP + li] = in ? make_float2((float)x[1].x, (float)x[1].y) : z;
                if constexpr (V == NS_QSTC)
                    srow[2 * SRP + li] = in ? make_float2((float)(o2 * ll * x[2].x - x[3].x),
                                                          (float)(o2 * ll * x[2].y - x[3].y))
                                            : z;
                else
                    srow[2 * SRP + li] = in ? make_float2((float)x[2].x, (float)x[2].y) : z;
            }
        }
        __syncthreads();
#pragma unroll
        for (int it = 0; it < 3; ++it)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int li = 4 * (g0 + 32 * it) + j;
                if (it < NRL) {
                    cq[it][j] = srow[li];
                    cs[it][j] = SCAL ? cq[it][j] : srow[SRP + li];
                    if constexpr (!SCAL) ct[it][j] = srow[2 * SRP + li];
                } else {
                    cq[it][j] = z;
                    cs[it][j] = z;
                    ct[it][j] = z;
                }
            }
    }
    // ---- 2. C1 (and C2) from the neighbours l - 1, l + 1: within a round
    // from the thread itself or the lanes g0 -+ 1 of its level (32-wide
    // shuffles); across rounds from lane 31 of round it - 1 / lane 0 of it + 1
    float2 c1[3][4], c2[3][4];
    (void)c1; (void)c2;
    float mx0 = 0.f, mx1 = 0.f, mx2 = 0.f;
    if constexpr (NCOL == 1) {
#pragma unroll
        for (int it = 0; it < 3; ++it)
#pragma unroll
            for (int j = 0; j < 4; ++j) mx0 = fmaxf(mx0, fmaxf(fabsf(cq[it][j].x), fabsf(cq[it][j].y)));
    } else {
        auto sh = [&](float2 v, int src) {
            return make_float2(__shfl(v.x, src, 32), __shfl(v.y, src, 32));
        };
        float2 sl[3], sr[3], tl[3], tr[3];  // S, T at l = lb - 1 and lb + 4
#pragma unroll
        for (int it = 0; it < 3; ++it) {
            const float2 z = make_float2(0.f, 0.f);
            const float2 a1 = sh(cs[it][3], (g0 + 31) & 31), b1 = it > 0 ? sh(cs[it - 1][3], 31) : z;
            const float2 a2 = sh(cs[it][0], (g0 + 1) & 31), b2 = it < 2 ? sh(cs[it + 1][0], 0) : z;
            sl[it] = g0 > 0 ? a1 : b1;
            sr[it] = g0 < 31 ? a2 : b2;
            if constexpr (NCOL == 3) {
                const float2 a3 = sh(ct[it][3], (g0 + 31) & 31), b3 = it > 0 ? sh(ct[it - 1][3], 31) : z;
                const float2 a4 = sh(ct[it][0], (g0 + 1) & 31), b4 = it < 2 ? sh(ct[it + 1][0], 0) : z;
                tl[it] = g0 > 0 ? a3 : b3;
                tr[it] = g0 < 31 ? a4 : b4;
            }
        }
#pragma unroll
        for (int it = 0; it < 3; ++it)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int l = m + 4 * (g0 + 32 * it) + j;
                const float am = (float)(l - 1) * cv[it][j], ap = (float)(l + 2) * cv[it][j + 1];
                const float2 sm = j > 0 ? cs[it][j - 1] : sl[it], sp = j < 3 ? cs[it][j + 1] : sr[it];
                // x' = am x_{l-1} - ap x_{l+1}
                const float spr = am * sm.x - ap * sp.x, spi = am * sm.y - ap * sp.y;
                if constexpr (NCOL == 2) {
                    c1[it][j] = make_float2(spr, spi);
                } else {
                    const float2 tm = j > 0 ? ct[it][j - 1] : tl[it], tp = j < 3 ? ct[it][j + 1] : tr[it];
                    const float tpr = am * tm.x - ap * tp.x, tpi = am * tm.y - ap * tp.y;
                    // C1 = S' + i m T, C2 = i m S - T'
                    c1[it][j] = make_float2(spr - fm * ct[it][j].y, spi + fm * ct[it][j].x);
                    c2[it][j] = make_float2(-fm * cs[it][j].y - tpr, fm * cs[it][j].x - tpi);
                    mx2 = fmaxf(mx2, fmaxf(fabsf(c2[it][j].x), fabsf(c2[it][j].y)));
                }
                mx0 = fmaxf(mx0, fmaxf(fabsf(cq[it][j].x), fabsf(cq[it][j].y)));
                mx1 = fmaxf(mx1, fmaxf(fabsf(c1[it][j].x), fabsf(c1[it][j].y)));
            }
    }
    // ---- 3. one power of two per (column, level) from the exact maxima,
    // then hi/lo pairs (k = 2 g, 2 g + 1 of one parity) written once
#pragma unroll
    for (int off = 16; off >= 1; off >>= 1) {
        mx0 = fmaxf(mx0, __shfl_xor(mx0, off));
        if constexpr (NCOL >= 2) mx1 = fmaxf(mx1, __shfl_xor(mx1, off));
        if constexpr (NCOL == 3) mx2 = fmaxf(mx2, __shfl_xor(mx2, off));
    }
    if (lok && g0 == 0) {
        atomicMax(&lvmax[levl], __float_as_uint(mx0));
        if constexpr (NCOL >= 2) atomicMax(&lvmax[WLG + levl], __float_as_uint(mx1));
        if constexpr (NCOL == 3) atomicMax(&lvmax[2 * WLG + levl], __float_as_uint(mx2));
    }
    __syncthreads();
    if (tid < NCOL * WLG) lvsc[tid] = bscale((double)__uint_as_float(lvmax[tid]));
    __syncthreads();
    {
        float scl[NCOL];
#pragma unroll
        for (int c = 0; c < NCOL; ++c) scl[c] = lvsc[c * WLG + levl];
#pragma unroll
        for (int it = 0; it < 3; ++it) {
            const int g = g0 + 32 * it;
            if (2 * g >= PKP) continue;
#pragma unroll
            for (int c = 0; c < NCOL; ++c) {
                const float2 *v = c == 0 ? cq[it] : (c == 1 ? c1[it] : c2[it]);
#pragma unroll
                for (int par = 0; par < 2; ++par)
#pragma unroll
                    for (int ri = 0; ri < 2; ++ri) {
                        const float v0 = (ri ? v[par].y : v[par].x) * scl[c];
                        const float v1 = (ri ? v[par + 2].y : v[par + 2].x) * scl[c];
                        _Float16 h0, l0h, h1, l1h;
                        split_h(v0, h0, l0h);
                        split_h(v1, h1, l1h);
                        unsigned char *row = sb + ((c * 2 + par) * 16 + levl * 2 + ri) * PROW;
                        typedef _Float16 h2 __attribute__((ext_vector_type(2)));
                        h2 hv, lv;
                        hv[0] = h0; hv[1] = h1;
                        lv[0] = l0h; lv[1] = l1h;
                        *reinterpret_cast<h2 *>(row + 4 * g) = hv;
                        *reinterpret_cast<h2 *>(row + 2 * PKP + 4 * g) = lv;
                    }
            }
        }
    }
    __syncthreads();

    // ---- 4. contraction: NCOL products against P per k step; steps
    // j < nks0 are parity 0 (ks = j), the rest parity 1 (ks = j - nks0)
    f4 acc[NCOL][2][WTT];  // [column][Sym/Anti][theta tile]
#pragma unroll
    for (int o = 0; o < NCOL; ++o)
#pragma unroll
        for (int s2 = 0; s2 < 2; ++s2)
#pragma unroll
            for (int t = 0; t < WTT; ++t) acc[o][s2][t] = f4{0.f, 0.f, 0.f, 0.f};
    const int tile0 = gz * 4 * WTT + wave * WTT;
    bool tok[WTT];
#pragma unroll
    for (int t = 0; t < WTT; ++t) tok[t] = tile0 + t < a.nt16;
    // The contraction is straight-line code: a fixed 2 NCH steps (parity
    // 0 chunks, then parity 1 chunks), every global load unconditional, so
    // that the fragments of the next three steps stay in flight (a load under
    // a branch makes the wait before the next MFMA cover every outstanding
    // load).  Chunks past a parity's K hold zero B rows and read a clamped,
    // finite table chunk; tiles past nt16 compute on a copy of the last tile
    // and are dropped in the epilogue.
    constexpr int NCH = PKP / KS;
    float rsv[WTT][4];  // row factors of the wave's theta rows, for the epilogue
#pragma unroll
    for (int tt = 0; tt < WTT; ++tt)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int th = min((tile0 + tt) * 16 + 4 * (lane >> 4) + j, a.nh - 1);
            rsv[tt][j] = srsv[th];
        }
    struct Frag { h8 h[WTT], l[WTT]; };
    const size_t toff0 = a.off[2 * mc], toff1 = a.off[2 * mc + 1];
    // NC chunks per parity (nks0 >= nks1); one straight-line body per NC
    auto contract = [&](auto ncc) {
        constexpr int NC = decltype(ncc)::value, NS = 2 * NC;
        Frag fr[SRING];
        auto load_s = [&](auto sc) {
            constexpr int S = decltype(sc)::value;
            if constexpr (S < NS) {
                constexpr int PAR = S / NC;
                const int ks = min(S % NC, (PAR ? nks1 : nks0) - 1);
#pragma unroll
                for (int tt = 0; tt < WTT; ++tt) {
                    const int tile = min(tile0 + tt, a.nt16 - 1);
                    // the fragment's 64 hi parts, then its 64 lo parts
                    const size_t to = (PAR ? toff1 : toff0) + (((size_t)ks * a.nt16 + tile) * 128 + lane) * 8;
                    fr[S % SRING].h[tt] = *reinterpret_cast<const h8 *>(a.tP + to);
                    fr[S % SRING].l[tt] = *reinterpret_cast<const h8 *>(a.tP + to + 512);
                }
            }
        };
        auto comp_s = [&](auto sc) {
            constexpr int S = decltype(sc)::value;
            constexpr int PAR = S / NC, KC = S % NC;
            const Frag &f = fr[S % SRING];
            const int kb = 2 * (KC * KS + 8 * (lane >> 4));
#pragma unroll
            for (int c = 0; c < NCOL; ++c) {
                const int row = (c * 2 + PAR) * 16 + (lane & 15);
                const h8 bh = *reinterpret_cast<const h8 *>(sb + row * PROW + kb);
                const h8 bl = *reinterpret_cast<const h8 *>(sb + row * PROW + 2 * PKP + kb);
#pragma unroll
                for (int tt = 0; tt < WTT; ++tt) {
                    f4 t = acc[c][PAR][tt];
                    t = mfma16(f.h[tt], bl, t);
                    t = mfma16(f.l[tt], bh, t);
                    t = mfma16(f.h[tt], bh, t);
                    acc[c][PAR][tt] = t;
                }
            }
        };
        static_for<SRING - 1>([&](auto sc) { load_s(sc); });
        __builtin_amdgcn_sched_barrier(0);
        static_for<NS>([&](auto sc) {
            constexpr int S = decltype(sc)::value;
            load_s(std::integral_constant<int, S + SRING - 1>());
            __builtin_amdgcn_sched_barrier(0);  // keep the loads ahead of this step
            comp_s(sc);
            __builtin_amdgcn_sched_barrier(0);
        });
    };
    switch (nks0) {
    case 1: contract(std::integral_constant<int, 1>()); break;
    case 2: contract(std::integral_constant<int, 2>()); break;
    case 3: contract(std::integral_constant<int, 3>()); break;
    case 4: contract(std::integral_constant<int, 4>()); break;
    default: contract(std::integral_constant<int, NCH>()); break;
    }

    // ---- 5. epilogue straight from the accumulators: lane (level, re/im)
    // holds theta_h = 4 (lane / 16) + j of a tile; after one exchange with its
    // re/im partner each lane owns two theta_h (j = p + 2 ri) and stores the
    // complex (North, South) pair of each as 16 contiguous bytes
    const int n = lane & 15, ol = n >> 1, ri = n & 1;
    const bool lvok = ol < nlh;
#pragma unroll
    for (int o = 0; o < NOUT; ++o) {
        const int oc = (NCOL == 2 && o == 2) ? 0 : o;  // SG: i m (P C0)
        const float inv = 1.f / lvsc[oc * WLG + ol];
        float *fo = a.fout + ((size_t)(lev0 + (lvok ? ol : 0)) * a.fout_ld +
                              ((size_t)(a.comp0 + o) * a.n_m + mc) * a.n_theta) * 2;
#pragma unroll
        for (int tt = 0; tt < WTT; ++tt) {
            if (!tok[tt]) continue;
            float vn[4], vs[4];  // North, South of this lane's re or im part
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                float sy = acc[oc][0][tt][j], an = acc[oc][1][tt][j];
                if (NCOL == 2 && o == 2) {
                    // (i m) X: re = -m X.im, im = m X.re, from lane n ^ 1
                    const float osy = __shfl_xor(sy, 1), oan = __shfl_xor(an, 1);
                    sy = ri ? fm * osy : -fm * osy;
                    an = ri ? fm * oan : -fm * oan;
                }
                const float rs = inv / rsv[tt][j];
                vn[j] = (sy + an) * rs;
                vs[j] = (sy - an) * rs;
            }
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                // send the partner the part of the theta_h it stores (selects
                // with fixed indices keep vn, vs in registers)
                const float on = ri ? vn[p + 2] : vn[p], os = ri ? vs[p + 2] : vs[p];
                const float gn = __shfl_xor(ri ? vn[p] : vn[p + 2], 1);
                const float gs = __shfl_xor(ri ? vs[p] : vs[p + 2], 1);
                const int th = (tile0 + tt) * 16 + 4 * (lane >> 4) + p + 2 * ri;
                f4 w;
                w[0] = ri ? gn : on;
                w[1] = ri ? on : gn;
                w[2] = ri ? gs : os;
                w[3] = ri ? os : gs;
                if (lvok && th < a.nh) *reinterpret_cast<f4 *>(fo + 4 * th) = w;
            }
        }
    }
}

template <int V, int NRL>
__global__ __launch_bounds__(256, 2) void k_nsyn16p(NSynArgs a)
{
    __shared__ __align__(16) unsigned char lds[NSynPL<V>::BYTES];
    const int ng = (a.nlev + WLG - 1) / WLG;
    int mc, rest;
    xcd_deal(blockIdx.x, ng * a.ngz, mc, rest);
    if (mc >= a.n_m) return;
    nsyn16p_item<V, NRL>(a, mc, rest % ng, rest / ng, lds, threadIdx.x);
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

// REC: the table fragments are generated per theta chunk instead of read:
// thread (theta row thl of the chunk, segment sg) runs the FP64 recurrence
// over the 32 degrees l = m + 256 gz + 32 sg + j of its segment from the
// stored FP64 seeds 2 pi w (P_{l-1}, P_l) (common power-of-two exponent)
// and writes w P and w sin(theta) dP/dtheta (the identity of the synthesis),
// times the row factor of l, as FP16 hi/lo to LDS [parity][k][theta]; the
// contraction reads its fragments there.
constexpr int ROWA = 32 + 8;   // halves per LDS row of generated fragments (padded)
constexpr int RNG = 1;         // level groups per block with generated fragments
constexpr int RLB = 256 + 8;   // recurrence coefficients held per block

template <int V, bool REC>
__global__ __launch_bounds__(256) void k_nana16(NAnaArgs a)
{
    using S = NAna<V>;
    constexpr int NPL = S::NPL, NOUT = S::NOUT, NPROD = S::NPROD, NIN = S::NIN;
    constexpr bool DP = V == NA_QST;  // dP fragments needed
    // level groups of LGS per block: REC blocks run RNG groups on each
    // generated chunk (the generation is then shared by RNG x 16 levels)
    constexpr int NG = REC ? RNG : 1;
    __shared__ __align__(16) _Float16 bhi[NPL * PLANE];
    __shared__ __align__(16) _Float16 blo[NPL * PLANE];
    __shared__ unsigned lvmax[2 * LGS * NG];  // [level group][column group][level]
    __shared__ float lvsc[2 * LGS * NG];
    // REC: generated fragments [P hi, P lo, dP hi, dP lo][parity][k of the block][theta]
    __shared__ __align__(16) _Float16 gA[REC ? (DP ? 4 : 2) * 2 * 128 * ROWA : 8];
    __shared__ double2 rcb[REC ? RLB : 1];  // (A_l, eps_l), l = m + 256 gz + i
    __shared__ float grs[REC ? 2 * 128 : 1];  // row factors [parity][k of the block]

    const int tid = threadIdx.x, nthr = blockDim.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int ng = (a.nlev + LGS * NG - 1) / (LGS * NG);
    int mc, rest;
    xcd_deal(blockIdx.x, ng * a.ngz, mc, rest);
    if (mc >= a.n_m) return;
    const int gy = rest % ng, gz = rest / ng;
    const int m = mc * a.minc;
    const double dm = (double)m;
    const int levb = gy * LGS * NG;  // first level of the block
    auto nlh_of = [&](int s) { return max(0, min(LGS, a.nlev - levb - LGS * s)); };
    const int l0 = a.lstart[mc];
    const int K0 = (a.l_max - m) / 2 + 1;
    const int ntl = (K0 + 15) / 16;           // l tiles per parity
    const int t0 = gz * 8 + wave * 2;         // this wave's l tiles t0, t0+1
    if (gz * 8 >= ntl) return;                // no l tile of this order here
    bool tok[2];
    tok[0] = t0 < ntl;
    tok[1] = t0 + 1 < ntl;
    if constexpr (REC) {
        for (int i = tid; i < RLB; i += nthr) {
            const int l = m + 256 * gz + i;
            double A = 0.0, e = 0.0;
            if (l > m && l <= a.l_max + 1) {
                const double dl = (double)l;
                const double num = (dl - dm) * (dl + dm), den = (2.0 * dl - 1.0) * (2.0 * dl + 1.0);
                A = sqrt(den / num);
                e = sqrt(num / den);
            }
            rcb[i] = make_double2(A, e);
        }
        for (int i = tid; i < 256; i += nthr) {
            const int par = i >> 7, kk = 128 * gz + (i & 127);
            grs[i] = kk < ntl * 16 ? a.rowscale[a.soff[2 * mc + par] + kk] : 0.f;
        }
    }

    // ---- B scales per (level, column group) from this order's F2 column:
    // 2 max|f| (N +- S) and 2 max(1, m) max(|vt|, |vp|)
    const float *fin = a.fin;
    if (tid < 2 * LGS * NG) lvmax[tid] = 0u;
    __syncthreads();
    {
        // (level, theta block) pairs; 128 bytes = 16 rows per comp
        for (int idx = tid; idx < LGS * NG * a.ntb; idx += nthr) {
            const int levg = idx % (LGS * NG), tb = idx / (LGS * NG);
            const int lev = levb + levg;
            if (lev >= a.nlev) continue;
            const int sgrp = levg / LGS, levl = levg % LGS;
            float mx[NIN];
#pragma unroll
            for (int k = 0; k < NIN; ++k) {
                const f4 *src = reinterpret_cast<const f4 *>(
                    fin + ((size_t)lev * a.fin_ld + (((size_t)a.in_comp[k] * a.ntb + tb) * a.n_m + mc) * 16) * 2);
                const int nrow = min(16, a.n_theta - 16 * tb);  // rows past n_theta are padding
                // all eight row pairs loaded first (clamped to the block's
                // rows; a repeat does not change the maximum)
                f4 w[8];
#pragma unroll
                for (int q = 0; q < 8; ++q) w[q] = src[min(q, nrow / 2 - 1)];
                float v = 0.f;
#pragma unroll
                for (int q = 0; q < 8; ++q)
                    v = fmaxf(v, fmaxf(fmaxf(fabsf(w[q][0]), fabsf(w[q][1])), fmaxf(fabsf(w[q][2]), fabsf(w[q][3]))));
                mx[k] = v;
            }
            atomicMax(&lvmax[sgrp * 2 * LGS + levl], __float_as_uint(mx[0]));
            if constexpr (NIN == 3)
                atomicMax(&lvmax[sgrp * 2 * LGS + LGS + levl], __float_as_uint(fmaxf(mx[1], mx[2])));
        }
    }
    __syncthreads();
    if (tid < 2 * LGS * NG) {
        const double f = (tid % (2 * LGS)) < LGS ? 2.0 : 2.0 * (dm > 1.0 ? dm : 1.0);
        lvsc[tid] = bscale(f * (double)__uint_as_float(lvmax[tid]));
    }

    f4 acc[NG][NOUT][2][2][2];  // [level group][out][parity][l tile][column tile]
#pragma unroll
    for (int g = 0; g < NG; ++g)
#pragma unroll
        for (int o = 0; o < NOUT; ++o)
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int t = 0; t < 2; ++t)
#pragma unroll
                    for (int n = 0; n < 2; ++n) acc[g][o][s][t][n] = f4{0.f, 0.f, 0.f, 0.f};

    // this thread's staging level in each group (fixed over the chunks): in
    // range and m kept
    bool stage_lev_ok[NG];
#pragma unroll
    for (int g = 0; g < NG; ++g) {
        const int levl = tid >> 4;
        stage_lev_ok[g] = false;
        if (levl < nlh_of(g)) {
            const int lc = a.lcut[levb + LGS * g + levl];
            stage_lev_ok[g] = m <= (lc < a.m_max ? lc : a.m_max);
        }
    }
    // REC: seeds of the first chunk (segment of this thread, theta = thl)
    double2 nsd = make_double2(0.0, 0.0);
    int nE = 0;
    double nx = 0.0;
    size_t seg_base = 0;
    if constexpr (REC) {
        const int segi = 8 * gz + (tid >> 5);
        seg_base = (a.segoff[mc] + (size_t)min(segi, (a.l_max - m) >> 5)) * a.nh;
        const int th0 = min(tid & 31, a.nh - 1);
        nsd = a.aseed[seg_base + th0];
        nE = a.aexp[seg_base + th0];
        nx = a.cost[th0];
    }
    // F2 layout: [lev][comp][theta block of 16 rows][m][16] complex
    for (int kc = 0; kc < a.nq32; ++kc) {
        __syncthreads();
        if constexpr (REC) {
            // ---- generate this chunk's fragments (see above); the seeds of
            // the next chunk are loaded now, used one chunk later
            const int thl = tid & 31, seg = tid >> 5;
            const int th = 32 * kc + thl;
            const int segi = 8 * gz + seg;
            const int ls = m + 32 * segi;
            const bool gen = th < a.nh && ls <= a.l_max;
            const double2 sd = nsd;
            const int E = nE;
            const double x = nx;
            if (kc + 1 < a.nq32) {
                const int th1 = min(th + 32, a.nh - 1);
                const size_t so = seg_base + th1;
                nsd = a.aseed[so];
                nE = a.aexp[so];
                nx = a.cost[th1];
            }
            const double sc = (gen && E >= -1000) ? ldexp(1.0, E) : 0.0;
            double pv = sd.x, q = sd.y;
            double eps = rcb[32 * seg].y;  // eps_{ls}
            const double dls = (double)ls, dlmax = (double)a.l_max;
#pragma unroll 4
            for (int j = 0; j < 32; ++j) {
                const double2 ce = rcb[32 * seg + j + 1];  // A_{l+1}, eps_{l+1}, l = ls + j
                const double dl = dls + (double)j;
                const double pn = ce.x * (x * q - eps * pv);
                const int par = j & 1, kl = 16 * seg + (j >> 1);
                const double r = dl <= dlmax ? sc * (double)grs[par * 128 + kl] : 0.0;
                _Float16 h, lo;
                split_h((float)(q * r), h, lo);
                const int ai = (par * 128 + kl) * ROWA + thl;
                gA[ai] = h;
                gA[2 * 128 * ROWA + ai] = lo;
                if constexpr (DP) {
                    const double dp = dl * ce.y * pn - (dl + 1.0) * eps * pv;
                    split_h((float)(dp * r), h, lo);
                    gA[4 * 128 * ROWA + ai] = h;
                    gA[6 * 128 * ROWA + ai] = lo;
                }
                pv = q;
                q = pn;
                eps = ce.y;
            }
        }
#pragma unroll
      for (int sg = 0; sg < NG; ++sg) {
        if (sg > 0) __syncthreads();  // the previous group's contraction read B
        const int lev0 = levb + LGS * sg, nlh = nlh_of(sg);
        const int ntn = nlh > 8 ? 2 : 1;
        // ---- stage theta_h = 32 kc + 8 tb + 2 q + (0, 1): rows 64 kc + 16 tb + 4 q .. +3
        {
            const int q = tid & 3, tb = (tid >> 2) & 3, levl = tid >> 4;
            const int lev = lev0 + levl;
            const int thb = kc * 4 + tb;        // theta block (16 rows)
            const int th0 = 32 * kc + 8 * tb + 2 * q;
            const bool valid = stage_lev_ok[sg] && th0 < a.nh;
            const bool ok1 = th0 + 1 < a.nh;  // rows past n_theta are padding
            const float sc0 = lvsc[sg * 2 * LGS + levl], sc1 = lvsc[sg * 2 * LGS + LGS + levl];
            // rows 2 th, 2 th + 1 for th = th0, th0 + 1: 4 complex, 32 bytes
            float x[NIN][8];
            // loads issued unconditionally at clamped addresses (level and
            // theta block in range; the second row pair of a block is padding
            // past n_theta), then masked
            f4 xv0[NIN], xv1[NIN];
#pragma unroll
            for (int k = 0; k < NIN; ++k) {
                const size_t o = ((size_t)(levl < nlh ? lev : 0) * a.fin_ld +
                                  (((size_t)a.in_comp[k] * a.ntb + min(thb, a.ntb - 1)) * a.n_m + mc) * 16 +
                                  4 * q) * 2;
                xv0[k] = *reinterpret_cast<const f4 *>(fin + o);
                xv1[k] = *reinterpret_cast<const f4 *>(fin + o + 4);
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int k = 0; k < NIN; ++k) {
                const f4 z4 = f4{0.f, 0.f, 0.f, 0.f};
                const f4 v0 = valid ? xv0[k] : z4, v1 = (valid && ok1) ? xv1[k] : z4;
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
                if constexpr (REC) {
                    const int ai = (par * 128 + 16 * (wave * 2 + tt) + (lane & 15)) * ROWA + 8 * (lane >> 4);
                    aPh[tt] = *reinterpret_cast<const h8 *>(gA + ai);
                    aPl[tt] = *reinterpret_cast<const h8 *>(gA + 2 * 128 * ROWA + ai);
                    if constexpr (DP) {
                        adPh[tt] = *reinterpret_cast<const h8 *>(gA + 4 * 128 * ROWA + ai);
                        adPl[tt] = *reinterpret_cast<const h8 *>(gA + 6 * 128 * ROWA + ai);
                    }
                } else if (tok[tt]) {
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
                        f4 t = acc[sg][o][par][tt][tn];
                        t = mfma16(ah, bl, t);
                        t = mfma16(al, bh, t);
                        t = mfma16(ah, bh, t);
                        acc[sg][o][par][tt][tn] = t;
                    }
                }
            }
        }
      }
    }

    // ---- epilogue: row r of tile t -> l = m + par + 2 (16 t + r); remove the
    // row and level factors, / l(l+1) for S and T, lcut masks
#pragma unroll
    for (int sg = 0; sg < NG; ++sg) {
      const int lev0 = levb + LGS * sg, nlh = nlh_of(sg);
      const int ntn = nlh > 8 ? 2 : 1;
#pragma unroll
      for (int tn = 0; tn < 2; ++tn) {
        if (tn >= ntn) break;
        const int n = tn * 16 + (lane & 15), levl = n >> 1, ri = n & 1;
        if (levl >= nlh) continue;
        const int lev = lev0 + levl;
        const int lc = a.lcut[lev];
        const bool mkeep = m <= (lc < a.m_max ? lc : a.m_max);
        const double inv0 = 1.0 / (double)lvsc[sg * 2 * LGS + levl];
        const double inv1 = 1.0 / (double)lvsc[sg * 2 * LGS + LGS + levl];
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
                        double v = keep ? (double)acc[sg][o][par][tt][tn][j] * rs : 0.0;
                        if (S::divll[o] && l > 0) v /= dll;
                        a.out[o][((size_t)lev * a.out_ld + (size_t)(l0 + l - m)) * 2 + ri] = v;
                    }
                }
        }
      }
    }
}

constexpr int AKP = 192;                 // theta_h held by the whole-theta analysis
constexpr int AROW = 4 * AKP + 16;       // bytes per LDS row (padded)

// ------------------------- analysis, whole-theta form on wP only (QST, SH)
// The wdP products are folded onto wP by the same identity as in synthesis:
// with G_l = sum_theta wP_l v,
//   sum_theta wdP_l v = l c_{l+1} G_{l+1} - (l+1) c_l G_{l-1},
// so three products G^f, G^vt, G^vp (l = m .. l_max + 1) replace five, and
//   Q = G^f, S = (l c_{l+1} G^vt_{l+1} - (l+1) c_l G^vt_{l-1}) - i m G^vp_l,
//   T = -(l c_{l+1} G^vp_{l+1} - (l+1) c_l G^vp_{l-1}) - i m G^vt_l,
// S, T / l(l+1) (src/shtransforms.f90:731-872 with the crossed vt/vp); SH
// keeps only G^f.  The block stages Sym/Anti of all theta pairs of its 8
// levels (hi/lo rows [hi 192][lo 192] in LDS, one power of two per (level,
// group) from the exact maxima), then contracts over theta with the wave's l
// tiles wave, wave + 4, wave + 8.  The epilogue keeps G (FP32, row factors
// removed) in LDS as [group][level][l - m] and forms the outputs per (level, l).
constexpr int PTT = 3;  // l tiles per wave

// LDS of one whole-theta analysis item: B rows, per-(group, level) maxima and
// scales, and per item the lcut of the levels, c_l (l - m = 0 .. l_max + 2 -
// m) and row factors
template <int V> struct NAnaPL {
    static constexpr int NCOL = V == NA_SH ? 1 : 3;
    static constexpr size_t SB = (size_t)NCOL * 2 * 16 * AROW;
    static constexpr size_t O_LVMAX = SB, O_LVSC = O_LVMAX + 4 * 2 * WLG, O_SLC = O_LVSC + 4 * 2 * WLG;
    static constexpr size_t O_SCR = O_SLC + 4 * WLG, O_SRS = O_SCR + 4 * (4 * 16 * PTT * 2);
    static constexpr size_t BYTES = O_SRS + 4 * (2 * 16 * 4 * PTT);
};

// One whole-theta analysis item: order mc, level group gy, on the LDS area
// `lds` (NAnaPL<V>::BYTES), 256 threads; tid = threadIdx.x (as above)
template <int V>
__device__ __forceinline__ void nana16p_item(const NAnaArgs &a, int mc, int gy, unsigned char *lds,
                                             const int tid)
{
    constexpr int NCOL = NAnaPL<V>::NCOL, NROW = NCOL * 2 * 16;
    constexpr int SB = NROW * AROW;
    unsigned char *sb = lds;
    unsigned *lvmax = reinterpret_cast<unsigned *>(lds + NAnaPL<V>::O_LVMAX);
    float *lvsc = reinterpret_cast<float *>(lds + NAnaPL<V>::O_LVSC);
    constexpr int cgrp[3] = {0, 1, 1};
    int *slc = reinterpret_cast<int *>(lds + NAnaPL<V>::O_SLC);
    float *scr = reinterpret_cast<float *>(lds + NAnaPL<V>::O_SCR);
    float (*srs)[16 * 4 * PTT] = reinterpret_cast<float (*)[16 * 4 * PTT]>(lds + NAnaPL<V>::O_SRS);

    const int wave = tid >> 6, lane = tid & 63;
    const int m = mc * a.minc;
    const int lev0 = gy * WLG, nlh = min(WLG, a.nlev - lev0);
    const int l0 = a.lstart[mc];
    const int K0 = (a.l_max + 1 - m) / 2 + 1;
    const int ntl = (K0 + 15) / 16;
    bool tok[PTT];
#pragma unroll
    for (int i = 0; i < PTT; ++i) tok[i] = wave + 4 * i < ntl;
    // ---- 1. Sym/Anti of the theta pairs in registers; exact maxima per
    // (level, group); after the scales, hi/lo pairs (two theta_h) written once
    // to the LDS rows [hi 192][lo 192] (theta_h >= nh zero).
    // The block's Fourier loads go first, before the per-block staging and
    // the barrier that waits for it: unconditional (a load under a branch
    // waits for itself before the next one) at in-bounds addresses (entries
    // past the group's levels read its first level; rows past nh lie in the
    // padding of their 16-row block), zeroed after the barrier where the
    // level, the row or the level's lcut leaves them out.
    const int nent = WLG * a.ntb * 4;
    constexpr int NE = (WLG * (AKP / 8) * 4 + 255) / 256;  // 3
    f4 xv[NE][NCOL][2];
#pragma unroll
    for (int i = 0; i < NE; ++i) {
        const int e = tid + 256 * i;
        const int q = e & 3, tb = (e >> 2) % a.ntb, levl = (e >> 2) / a.ntb;
        const int lev = lev0 + ((e < nent && levl < nlh) ? levl : 0);
#pragma unroll
        for (int k = 0; k < NCOL; ++k) {
            const size_t o = ((size_t)lev * a.fin_ld +
                              (((size_t)a.in_comp[k] * a.ntb + tb) * a.n_m + mc) * 16 + 4 * q) * 2;
            xv[i][k][0] = *reinterpret_cast<const f4 *>(a.fin + o);
            xv[i][k][1] = *reinterpret_cast<const f4 *>(a.fin + o + 4);
        }
    }
    __builtin_amdgcn_sched_barrier(0);
    if (tid < 2 * WLG) lvmax[tid] = 0u;
    if (tid < WLG) slc[tid] = tid < nlh ? a.lcut[lev0 + tid] : -1;
    {
        const float *cr = a.crec + a.lstart[mc] + 2 * mc;
        for (int d = tid; d <= a.l_max + 2 - m; d += 256) scr[d] = cr[d];
        for (int e = tid; e < 2 * 16 * ntl; e += 256) {
            const int par = e / (16 * ntl), kk = e - par * 16 * ntl;
            srs[par][kk] = a.rowscale[a.soff[2 * mc + par] + kk];
        }
    }
    __syncthreads();
    {
#pragma unroll
        for (int i = 0; i < NE; ++i) {
            const int e = tid + 256 * i;
            const int q = e & 3, tb = (e >> 2) % a.ntb, levl = (e >> 2) / a.ntb;
            const int th0 = 8 * tb + 2 * q;
            bool valid = e < nent && levl < nlh && th0 < a.nh;
            if (valid) {
                const int lc = slc[levl];
                valid = m <= (lc < a.m_max ? lc : a.m_max);
            }
            const bool ok1 = valid && th0 + 1 < a.nh;
#pragma unroll
            for (int k = 0; k < NCOL; ++k) {
                if (!valid) xv[i][k][0] = f4{0.f, 0.f, 0.f, 0.f};
                if (!ok1) xv[i][k][1] = f4{0.f, 0.f, 0.f, 0.f};
            }
        }
#pragma unroll
        for (int i = 0; i < NE; ++i) {
            const int e = tid + 256 * i;
            const int q = e & 3, tb = (e >> 2) % a.ntb, levl = (e >> 2) / a.ntb;
            const int th0 = 8 * tb + 2 * q;
            if (e >= nent || th0 >= a.nh) continue;
            float mx0 = 0.f, mx1 = 0.f;
#pragma unroll
            for (int k = 0; k < NCOL; ++k)
#pragma unroll
                for (int p = 0; p < 2; ++p) {
                    const f4 w = xv[i][k][p];  // N re, N im, S re, S im; |N +- S| <= |N| + |S|
                    float &mx = cgrp[k] ? mx1 : mx0;
                    mx = fmaxf(mx, fmaxf(fmaxf(fabsf(w[0] + w[2]), fabsf(w[0] - w[2])),
                                         fmaxf(fabsf(w[1] + w[3]), fabsf(w[1] - w[3]))));
                }
            if (levl < nlh) {
                atomicMax(&lvmax[levl], __float_as_uint(mx0));
                atomicMax(&lvmax[WLG + levl], __float_as_uint(mx1));
            }
        }
        typedef _Float16 h2 __attribute__((ext_vector_type(2)));
        const h2 hz = {(_Float16)0.f, (_Float16)0.f};
        for (int e = tid; e < NROW * (AKP / 2); e += 256) {
            const int row = e / (AKP / 2), k2 = 2 * (e % (AKP / 2));
            if (k2 >= a.nh) {
                *reinterpret_cast<h2 *>(sb + row * AROW + 2 * k2) = hz;
                *reinterpret_cast<h2 *>(sb + row * AROW + 2 * AKP + 2 * k2) = hz;
            }
        }
    }
    __syncthreads();
    if (tid < 2 * WLG) lvsc[tid] = bscale((double)__uint_as_float(lvmax[tid]));
    __syncthreads();
    {
        typedef _Float16 h2 __attribute__((ext_vector_type(2)));
#pragma unroll
        for (int i = 0; i < NE; ++i) {
            const int e = tid + 256 * i;
            const int q = e & 3, tb = (e >> 2) % a.ntb, levl = (e >> 2) / a.ntb;
            const int th0 = 8 * tb + 2 * q;
            if (e >= nent || th0 >= a.nh) continue;
#pragma unroll
            for (int k = 0; k < NCOL; ++k) {
                const float sc = lvsc[(cgrp[k] ? WLG : 0) + levl];
#pragma unroll
                for (int sa = 0; sa < 2; ++sa) {
                    float cr[2], ci[2];
#pragma unroll
                    for (int p = 0; p < 2; ++p) {
                        const f4 w = xv[i][k][p];
                        cr[p] = (sa ? w[0] - w[2] : w[0] + w[2]) * sc;
                        ci[p] = (sa ? w[1] - w[3] : w[1] + w[3]) * sc;
                    }
#pragma unroll
                    for (int ri = 0; ri < 2; ++ri) {
                        _Float16 h0, l0h, h1, l1h;
                        split_h(ri ? ci[0] : cr[0], h0, l0h);
                        split_h(ri ? ci[1] : cr[1], h1, l1h);
                        h2 hv, lv;
                        hv[0] = h0; hv[1] = h1;
                        lv[0] = l0h; lv[1] = l1h;
                        unsigned char *rowp = sb + ((k * 2 + sa) * 16 + levl * 2 + ri) * AROW;
                        *reinterpret_cast<h2 *>(rowp + 2 * th0) = hv;
                        *reinterpret_cast<h2 *>(rowp + 2 * AKP + 2 * th0) = lv;
                    }
                }
            }
        }
    }
    __syncthreads();

    // ---- 3. contraction over theta_h: G^c (parity par) += wP(par) x col(c, Sym/Anti = par)
    f4 acc[NCOL][2][PTT];
#pragma unroll
    for (int o = 0; o < NCOL; ++o)
#pragma unroll
        for (int s2 = 0; s2 < 2; ++s2)
#pragma unroll
            for (int t = 0; t < PTT; ++t) acc[o][s2][t] = f4{0.f, 0.f, 0.f, 0.f};
    // Straight-line contraction over a fixed 2 AKP / KS half steps (theta
    // chunk H / 2, parity H % 2), all loads unconditional, the fragments of
    // half step H loaded at H - 3 (ring of four).  Chunks past nq32 read a
    // clamped table chunk against zero B rows (theta_h >= nh is zero in
    // LDS); tiles past ntl are dropped in the epilogue.
    constexpr int NHS = 2 * (AKP / KS);
    struct AF { h8 h[PTT], l[PTT]; };
    AF fr[ARING];
    const size_t toff0 = a.off[2 * mc], toff1 = a.off[2 * mc + 1];
    auto load_h = [&](auto hc) {
        constexpr int H = decltype(hc)::value;
        if constexpr (H < NHS) {
            constexpr int PAR = H % 2;
            const int kc = min(H / 2, a.nq32 - 1);
#pragma unroll
            for (int tt = 0; tt < PTT; ++tt) {
                const int tile = min(wave + 4 * tt, ntl - 1);
                // the fragment's 64 hi parts, then its 64 lo parts
                const size_t to = (PAR ? toff1 : toff0) + (((size_t)tile * a.nq32 + kc) * 128 + lane) * 8;
                fr[H % ARING].h[tt] = *reinterpret_cast<const h8 *>(a.tP + to);
                fr[H % ARING].l[tt] = *reinterpret_cast<const h8 *>(a.tP + to + 512);
            }
        }
    };
    auto comp_h = [&](auto hc) {
        constexpr int H = decltype(hc)::value;
        constexpr int PAR = H % 2, KC = H / 2;
        const AF &f = fr[H % ARING];
        const int kb = 2 * (KC * KS + 8 * (lane >> 4));
#pragma unroll
        for (int c = 0; c < NCOL; ++c) {
            const int row = (c * 2 + PAR) * 16 + (lane & 15);
            const h8 bh = *reinterpret_cast<const h8 *>(sb + row * AROW + kb);
            const h8 bl = *reinterpret_cast<const h8 *>(sb + row * AROW + 2 * AKP + kb);
#pragma unroll
            for (int tt = 0; tt < PTT; ++tt) {
                if (!tok[tt]) continue;  // wave-uniform, no loads inside
                f4 t = acc[c][PAR][tt];
                t = mfma16(f.h[tt], bl, t);
                t = mfma16(f.l[tt], bh, t);
                t = mfma16(f.h[tt], bh, t);
                acc[c][PAR][tt] = t;
            }
        }
    };
    static_for<ARING - 1>([&](auto hc) { load_h(hc); });
    __builtin_amdgcn_sched_barrier(0);
    static_for<NHS>([&](auto hc) {
        constexpr int H = decltype(hc)::value;
        load_h(std::integral_constant<int, H + ARING - 1>());
        __builtin_amdgcn_sched_barrier(0);  // keep the loads ahead of this step
        comp_h(hc);
        __builtin_amdgcn_sched_barrier(0);
    });

    // ---- 4. G (row factors removed) into LDS: [c][level][l - m] complex float
    float *sg = reinterpret_cast<float *>(sb);
    const int rowg = 2 * (a.l_max + 2) + 2;
    const int kmax = a.l_max + 1 - m;  // l - m at most
    __syncthreads();  // B rows consumed
    {
        const int n = lane & 15, levl = n >> 1, ri = n & 1;
#pragma unroll
        for (int tt = 0; tt < PTT; ++tt) {
            if (!tok[tt]) continue;
#pragma unroll
            for (int par = 0; par < 2; ++par)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int kk = 16 * (wave + 4 * tt) + 4 * (lane >> 4) + j;
                    const int d = par + 2 * kk;
                    if (d > kmax) continue;
                    const float rr = 1.f / srs[par][kk];
#pragma unroll
                    for (int c = 0; c < NCOL; ++c)
                        sg[(c * WLG + levl) * rowg + 2 * d + ri] = acc[c][par][tt][j] * rr;
                }
        }
    }
    __syncthreads();
    // ---- 5. Q, S, T per (level, l), coalesced double2 stores
    const int nl = a.l_max - m + 1;
    for (int i = tid; i < WLG * nl; i += 256) {
        const int d = i % nl, lv = i / nl;
        if (lv >= nlh) continue;
        const int lev = lev0 + lv;
        const int l = m + d;
        const int lc = slc[lv];
        const bool keep = m <= (lc < a.m_max ? lc : a.m_max) && l <= lc;
        double2 q = make_double2(0.0, 0.0), s = q, t = q;
        if (keep) {
            const double inv0 = 1.0 / (double)lvsc[lv];
            const float *gf = sg + (0 * WLG + lv) * rowg + 2 * d;
            q = make_double2((double)gf[0] * inv0, (double)gf[1] * inv0);
            if constexpr (V == NA_QST) {
                const double inv1 = 1.0 / (double)lvsc[WLG + lv];
                const double dm = (double)m;
                const double cp = (double)l * (double)scr[d + 1];   // l c_{l+1}
                const double cm = (double)(l + 1) * (double)scr[d]; // (l+1) c_l
                const float *gt = sg + (1 * WLG + lv) * rowg + 2 * d;
                const float *gp = sg + (2 * WLG + lv) * rowg + 2 * d;
                const double tmr = d > 0 ? (double)gt[-2] : 0.0, tmi = d > 0 ? (double)gt[-1] : 0.0;
                const double pmr = d > 0 ? (double)gp[-2] : 0.0, pmi = d > 0 ? (double)gp[-1] : 0.0;
                const double dtr = cp * gt[2] - cm * tmr, dti = cp * gt[3] - cm * tmi;
                const double dpr = cp * gp[2] - cm * pmr, dpi = cp * gp[3] - cm * pmi;
                // -i m X: re = m X.im, im = -m X.re
                const double oll = l > 0 ? 1.0 / ((double)l * (double)(l + 1)) : 1.0;
                s = make_double2((dtr + dm * gp[1]) * inv1 * oll, (dti - dm * gp[0]) * inv1 * oll);
                t = make_double2((-dpr + dm * gt[1]) * inv1 * oll, (-dpi - dm * gt[0]) * inv1 * oll);
            }
        }
        const size_t o = ((size_t)lev * a.out_ld + (size_t)(l0 + d)) * 2;
        *reinterpret_cast<double2 *>(a.out[0] + o) = q;
        if constexpr (V == NA_QST) {
            *reinterpret_cast<double2 *>(a.out[1] + o) = s;
            *reinterpret_cast<double2 *>(a.out[2] + o) = t;
        }
    }
}

template <int V>
__global__ __launch_bounds__(256, 2) void k_nana16p(NAnaArgs a)
{
    __shared__ __align__(16) unsigned char lds[NAnaPL<V>::BYTES];
    const int ng = (a.nlev + WLG - 1) / WLG;
    int mc, rest;
    xcd_deal(blockIdx.x, ng, mc, rest);
    if (mc >= a.n_m) return;
    nana16p_item<V>(a, mc, rest, lds, threadIdx.x);
}

// ------------------------------------------------ persistent batch launch
// LDS of a row item: FP32 twiddles (M, then M + 1 entries, padded to an even
// count) and two ping-pong buffers of tbr rows of M points.  The tile of the
// synthesis output (tbr theta rows x n_m orders) is read into the second
// buffer, which the pre-step has consumed before the first FFT pass writes it.
__host__ __device__ constexpr size_t pb_row_bytes(int M, int tbr)
{
    return ((size_t)2 * M + 2 + (size_t)2 * tbr * M) * sizeof(float2);
}
constexpr size_t cmax(size_t x, size_t y) { return x > y ? x : y; }
constexpr size_t PB_LDS = cmax(cmax(NSynPL<NS_QST>::BYTES, NSynPL<NS_S>::BYTES),
                               cmax(NAnaPL<NA_QST>::BYTES, NAnaPL<NA_SH>::BYTES));
// a wait gives up after this many polls of its counter (each a device-scope
// load and a short sleep: about a second in all), far beyond any real wait
constexpr int PB_SPIN = 1 << 20;

// one row item: the theta rows th0 .. th0 + tbr of the rows (field x level)
// of level group g, every comp: the synthesis output F[row][comp][m][theta]
// through the inverse real FFT (src/fft.f90:386-398) into a grid row held in
// LDS, and that row through the forward real FFT (scale 1/n_phi, and
// 1/sin^2(theta) for the comps of osin2_mask) into the theta-blocked F2
// that the analysis items read.  Row for row the arithmetic of the per-call
// row FFT kernels (the same passes and twiddles).
__device__ __forceinline__ void pb_row_item(const PBJob &J, int g, int tb, unsigned char *lds)
{
    typedef float2 C;
    const int M = J.M, TBR = J.tbr, n_m = J.n_m, nth = J.n_theta;
    C *s_tw = reinterpret_cast<C *>(lds);
    C *s_twr = s_tw + M;
    C *A = s_tw + 2 * M + 2;
    C *B = A + (size_t)TBR * M;
    const int tid = threadIdx.x;
    for (int k = tid; k < M; k += 256) s_tw[k] = reinterpret_cast<const C *>(J.tw)[k];
    for (int k = tid; k <= M; k += 256) s_twr[k] = reinterpret_cast<const C *>(J.twr)[k];
    // complex FFT of the TBR rows in buffer a (b scratch), radices read per
    // pass from the job; returns the buffer that holds the result
    auto fft = [&](C *a, C *b, int sign) {
        int L = 1;
        for (int s = 0; s < J.nfac; ++s) {
            const int R = J.fac[s];
            switch (R) {
            case 4: pass_pp<4>(a, b, TBR, M, L, s_tw, sign); break;
            case 2: pass_pp<2>(a, b, TBR, M, L, s_tw, sign); break;
            case 3: pass_pp<3>(a, b, TBR, M, L, s_tw, sign); break;
            default: pass_pp<5>(a, b, TBR, M, L, s_tw, sign); break;
            }
            __syncthreads();
            C *t = a;
            a = b;
            b = t;
            L *= R;
        }
        return a;
    };
    const int th0 = tb * TBR, nrows = min(TBR, nth - th0);
    const int r0 = g * WLG, r1 = min(r0 + WLG, J.rows);
    const float invN = (float)(1.0 / (double)J.N);
    const int per = 256 / TBR, pr = tid / per, pq = tid - pr * per;
    for (int r = r0; r < r1; ++r)
        for (int c = 0; c < J.nc; ++c) {
            __syncthreads();  // twiddles in; the previous row's buffers read
            const C *Fp = reinterpret_cast<const C *>(J.F) + (size_t)r * J.f_ld + (size_t)c * n_m * nth + th0;
            for (int idx = tid; idx < n_m * TBR; idx += 256) {
                const int m = idx / TBR, t = idx - m * TBR;
                B[t * n_m + m] = t < nrows ? Fp[(size_t)m * nth + t] : make_float2(0.f, 0.f);
            }
            __syncthreads();
            irfft_pre(B + pr * n_m, (size_t)1, n_m, A + pr * M, M, s_twr, pq, per);
            __syncthreads();
            C *res = fft(A, B, +1);
            res = fft(res, res == A ? B : A, -1);
            const bool os = (J.osin2_mask >> c) & 1u;
            C *Fo = reinterpret_cast<C *>(J.F2) + (size_t)r * J.f2_ld;
            for (int idx = tid; idx < n_m * TBR; idx += 256) {
                const int k = idx / TBR, t = idx - k * TBR;
                if (t >= nrows) continue;
                const int th = th0 + t;
                float sc = invN;
                if (os) sc *= (float)J.osin2[th];
                Fo[(((size_t)c * J.nt16 + (th >> 4)) * n_m + k) * 16 + (th & 15)] =
                    rfft_post(res + t * M, k, M, s_twr, sc);
            }
        }
}

// Row item on wave slots, for M = 12 R1 R2 ... (the batch's n_phi = 192,
// 384, 576, 768): one slot of Q = M / 12 lanes (64 / Q slots per wave) per
// theta row holds the row in LDS (padded by one point per 32) and runs its
// inverse and its forward FFT in place with no block barrier between the
// passes.  The block's NS = 4 (64 / Q) theta rows of one row and comp: the
// Fourier tile (orders x theta rows) read in theta runs into LDS, the
// pre-step per slot, both FFTs, the post-step back into the tile, the tile
// written in theta runs of the theta-blocked F2.
template <int M> struct PBSlot;
template <> struct PBSlot<96> {
    template <typename C> __device__ static void run(C *z, int t, const C *tw, int sign)
    {
        slot_chain<96, 1, 12, 4, 2>(z, t, tw, sign);
    }
};
template <> struct PBSlot<192> {
    template <typename C> __device__ static void run(C *z, int t, const C *tw, int sign)
    {
        slot_chain<192, 1, 12, 4, 4>(z, t, tw, sign);
    }
};
template <> struct PBSlot<288> {
    template <typename C> __device__ static void run(C *z, int t, const C *tw, int sign)
    {
        slot_chain<288, 1, 12, 4, 3, 2>(z, t, tw, sign);
    }
};
template <> struct PBSlot<384> {
    template <typename C> __device__ static void run(C *z, int t, const C *tw, int sign)
    {
        slot_chain<384, 1, 12, 4, 4, 2>(z, t, tw, sign);
    }
};

// theta rows per slot row item (0: M has no slot form here)
__host__ __device__ constexpr int pb_slot_rows(int M)
{
    return (M == 96 || M == 192 || M == 288 || M == 384) ? 4 * (64 / (M / 12)) : 0;
}
// LDS of a slot row item: padded twiddles, twr, NS padded rows, the tile
__host__ __device__ constexpr size_t pb_slot_bytes(int M)
{
    return ((size_t)M + M / 8 + M + 2 + (size_t)pb_slot_rows(M) * (M + M / 32 + 1) +
            (size_t)(2 * M / 3) * (pb_slot_rows(M) + 1)) * sizeof(float2);
}
static_assert(pb_slot_bytes(384) <= PB_LDS && pb_slot_bytes(288) <= PB_LDS && pb_slot_bytes(192) <= PB_LDS &&
              pb_slot_bytes(96) <= PB_LDS, "slot row item exceeds the launch's LDS");

template <int M>
__device__ __forceinline__ void pb_row_slots(const PBJob &J, int g, int tb, unsigned char *lds, const int tid)
{
    typedef float2 C;
    constexpr int P = 12, Q = M / P, SPW = 64 / Q, NS = 4 * SPW, MP = padded_row<M, P>();
    constexpr int TS = NS + 1;  // tile stride per order (theta fastest)
    static_assert(NS == pb_slot_rows(M), "slot rows");
    C *s_tw = reinterpret_cast<C *>(lds);
    C *s_twr = s_tw + M + M / 8;
    C *rows = s_twr + M + 2;
    C *tile = rows + NS * MP;
    const int wave = tid >> 6, lane = tid & 63;
    const int sw = lane / Q, t = lane - sw * Q, slot = wave * SPW + sw;
    const bool act = sw < SPW;
    {
        const C *tw = reinterpret_cast<const C *>(J.tw);
        const C *twr = reinterpret_cast<const C *>(J.twr);
        for (int e = tid; e < M; e += 256) s_tw[e + (e >> 3)] = tw[e];
        for (int e = tid; e <= M; e += 256) s_twr[e] = twr[e];
    }
    const int n_m = J.n_m, nth = J.n_theta, nc = J.nc;
    const int th0 = tb * NS, nrows = min(NS, nth - th0);
    const int r0 = g * WLG, r1 = min(r0 + WLG, J.rows);
    const float invN = (float)(1.0 / (double)J.N);
    C *z = rows + (act ? slot : 0) * MP;
    // the tile of step (r, c) in registers: element idx = tid + 256 i is
    // (order idx / NS, theta row idx % NS); the next step's tile is loaded
    // while the slots work on the current one
    constexpr int NV = ((2 * M / 3) * NS + 255) / 256;
    C v[NV];
    auto load = [&](int r, int c) {
        const C *Fp = reinterpret_cast<const C *>(J.F) + (size_t)r * J.f_ld + (size_t)c * n_m * nth + th0;
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            const int idx = tid + 256 * i, m = idx / NS, tt = idx - m * NS;
            v[i] = (m < n_m && tt < nrows) ? Fp[(size_t)m * nth + tt] : make_float2(0.f, 0.f);
        }
    };
    load(r0, 0);
    for (int r = r0; r < r1; ++r)
        for (int c = 0; c < nc; ++c) {
            __syncthreads();  // twiddles in; the previous tile written out
#pragma unroll
            for (int i = 0; i < NV; ++i) {
                const int idx = tid + 256 * i, m = idx / NS, tt = idx - m * NS;
                if (m < n_m) tile[m * TS + tt] = v[i];
            }
            __syncthreads();
            if (c + 1 < nc) load(r, c + 1);
            else if (r + 1 < r1) load(r + 1, 0);
            if (act) {
                // Z_k = (X_k + conj X_{M-k}) + i e^{+2 pi i k/N} (X_k - conj X_{M-k})
                for (int k = t; k <= M / 2; k += Q) {
                    const int k2 = M - k;
                    C a = k < n_m ? tile[k * TS + slot] : make_float2(0.f, 0.f);
                    C b = k2 < n_m ? tile[k2 * TS + slot] : make_float2(0.f, 0.f);
                    if (k == 0) {
                        a.y = 0.f;
                        b.y = 0.f;
                    }
                    z[pidx<P>(k)] = cadd(cadd(a, cconj(b)), cmuli(cmul(cconj(s_twr[k]), csub(a, cconj(b))), 1));
                    if (k2 < M && k2 != k)
                        z[pidx<P>(k2)] = cadd(cadd(b, cconj(a)), cmuli(cmul(cconj(s_twr[k2]), csub(b, cconj(a))), 1));
                }
                wave_order();
                PBSlot<M>::run(z, t, s_tw, +1);
                PBSlot<M>::run(z, t, s_tw, -1);
            }
            __syncthreads();  // every slot has read its tile column
            if (act) {
                float sc = invN;
                if (((J.osin2_mask >> c) & 1u) && slot < nrows) sc *= (float)J.osin2[th0 + slot];
                for (int k = t; k < n_m; k += Q) tile[k * TS + slot] = rfft_post_padded<P>(z, k, M, s_twr, sc);
            }
            __syncthreads();
            C *Fo = reinterpret_cast<C *>(J.F2) + (size_t)r * J.f2_ld;
            for (int idx = tid; idx < n_m * NS; idx += 256) {
                const int k = idx / NS, tt = idx - k * NS;
                if (tt >= nrows) continue;
                const int th = th0 + tt;
                Fo[(((size_t)c * J.nt16 + (th >> 4)) * n_m + k) * 16 + (th & 15)] = tile[k * TS + tt];
            }
        }
}

__device__ __forceinline__ void pb_syn_unit(const PBJob &J, int mc, int g, unsigned char *lds, int tid)
{
    if (J.syn_kind == NS_S) {
        if (J.nrl == 1) nsyn16p_item<NS_S, 1>(J.syn, mc, g, 0, lds, tid);
        else nsyn16p_item<NS_S, 2>(J.syn, mc, g, 0, lds, tid);
    } else {
        if (J.nrl == 1) nsyn16p_item<NS_QST, 1>(J.syn, mc, g, 0, lds, tid);
        else nsyn16p_item<NS_QST, 2>(J.syn, mc, g, 0, lds, tid);
    }
}

__device__ __forceinline__ void pb_ana_unit(const PBJob &J, int mc, int g, unsigned char *lds, int tid)
{
    if (J.ana_kind == NA_SH) nana16p_item<NA_SH>(J.ana, mc, g, lds, tid);
    else nana16p_item<NA_QST>(J.ana, mc, g, lds, tid);
}

// polls counter c until it reaches need, or until the wait limit or another
// workgroup's failure: false.  No cache maintenance follows: the producers
// ran on this XCD and completed their stores into its L2 before counting,
// and no CU holds an earlier copy of the lines in its L1 (a level group's
// Fourier rows are read only by its consumers, only after all its producers
// have counted; the L1 starts each launch empty); the block barrier after
// the wait orders the consumer's loads after it.
__device__ bool pb_wait(int *c, int need, int *fail_flag)
{
    for (int i = 0; i < PB_SPIN; ++i) {
        if (__hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= need) return true;
        if (__hip_atomic_load(fail_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return false;
        __builtin_amdgcn_s_sleep(8);
    }
    return false;
}

// XCC_ID hardware register (bits 3:0): the XCD the wave runs on
__device__ __forceinline__ int xcc_id() { return __builtin_amdgcn_s_getreg(20 | (0 << 6) | (3 << 11)); }

__global__ void k_xcc_ids(int *ids)
{
    if (threadIdx.x == 0) ids[blockIdx.x] = xcc_id();
}

// The work list of a batch, claimed by `gridDim.x` workgroups of 256
// threads, each from the queue of its XCD in ticket order (a workgroup claims
// its next ticket while it works on the current item).  A producer item's
// stores are complete (every thread waits for its own) before the barrier
// after which it counts; its consumers, on the same XCD, read them through
// the shared L2.
__global__ __launch_bounds__(256, 2) void k_pbatch(const PBJob *__restrict__ jobs,
                                                   const PBItem *__restrict__ items, PBQueues qs,
                                                   int *ctr, int nctr)
{
    __shared__ __align__(16) unsigned char lds[PB_LDS];
    __shared__ int s_tk[2];
    __shared__ int s_flag;
    const int tid = threadIdx.x;
    const int q = qs.xmap[xcc_id() & (PB_MAXQ - 1)];
    const int qb = q >= 0 ? qs.off[q] : 0, qe = q >= 0 ? qs.off[q + 1] : 0;
    if (tid == 0) s_tk[0] = qe > qb ? qb + atomicAdd(&ctr[q], 1) : qe;
    __syncthreads();
    int t = __builtin_amdgcn_readfirstlane(s_tk[0]);
    for (int it = 0; t < qe; ++it) {
        int nxt = 0;
        if (tid == 0) nxt = qb + atomicAdd(&ctr[q], 1);
        const PBItem w = items[t];
        const PBJob &J = jobs[w.job];
        if (w.kind != PB_SYN) {
            if (tid == 0) {
                // a row item: every synthesis unit of its level group; an
                // analysis item: every row item of each of its level groups.
                // No acquire fence: the arrays waited for are read only after
                // their counts are complete (no item reads them earlier in
                // the launch, and stores do not fill the CU's L1), so the
                // CU's L1 holds no stale line of them
                bool ok = true;
                if (w.kind == PB_ROW) ok = pb_wait(&ctr[J.c_syn + w.g0], J.n_m, &ctr[PB_ERR]);
                else
                    for (int g = w.g0; ok && g < w.g0 + w.ng; ++g) ok = pb_wait(&ctr[J.c_row + g], J.ntbr, &ctr[PB_ERR]);
                if (!ok) atomicOr(&ctr[PB_ERR], 1);
                s_flag = ok ? 0 : 1;
            }
            __syncthreads();
            if (__builtin_amdgcn_readfirstlane(s_flag)) break;
        }
        if (w.kind == PB_ROW) {
            const PBJob *jp = &J;
            int ltid = tid;
            asm volatile("" : "+s"(jp), "+v"(ltid));
            switch (jp->M) {
            case 96: pb_row_slots<96>(*jp, w.g0, w.tb, lds, ltid); break;
            case 192: pb_row_slots<192>(*jp, w.g0, w.tb, lds, ltid); break;
            case 288: pb_row_slots<288>(*jp, w.g0, w.tb, lds, ltid); break;
            case 384: pb_row_slots<384>(*jp, w.g0, w.tb, lds, ltid); break;
            default: pb_row_item(*jp, w.g0, w.tb, lds); break;
            }
        } else {
            for (int u = 0; u < 2; ++u) {
                const int mc = u ? w.m1 : w.m0;
                if (mc < 0) continue;
                for (int g = w.g0; g < w.g0 + w.ng; ++g) {
                    // the job's fields and everything derived from the
                    // thread index are computed inside each unit, not
                    // hoisted out of these loops and held across them
                    const PBJob *jp = &J;
                    int ltid = tid;
                    asm volatile("" : "+s"(jp), "+v"(ltid));
                    if (w.kind == PB_SYN) pb_syn_unit(*jp, mc, g, lds, ltid);
                    else pb_ana_unit(*jp, mc, g, lds, ltid);
                    __syncthreads();  // the LDS is the next unit's
                }
            }
        }
        // every thread's stores complete (in the XCD's L2) before the item counts
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (tid == 0) {
            if (w.kind == PB_SYN) {
                const int nord = w.m1 >= 0 ? 2 : 1;
                for (int g = w.g0; g < w.g0 + w.ng; ++g) atomicAdd(&ctr[J.c_syn + g], nord);
            } else if (w.kind == PB_ROW) {
                atomicAdd(&ctr[J.c_row + w.g0], 1);
            }
            s_tk[(it + 1) & 1] = nxt;
        }
        __syncthreads();
        t = __builtin_amdgcn_readfirstlane(s_tk[(it + 1) & 1]);
    }
    // the last workgroup out checks that every queue was claimed to its end
    // (else its items never ran: error) and clears the counters for the next
    // launch (the error flag stays for the host)
    __syncthreads();
    if (tid == 0) s_flag = atomicAdd(&ctr[PB_DONE], 1) == (int)gridDim.x - 1;
    __syncthreads();
    if (s_flag) {
        __threadfence();
        // thread i checks ticket i (i < nq) before it clears counters i, i + 256, ...
        if (tid < qs.nq && __hip_atomic_load(&ctr[tid], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <
                               qs.off[tid + 1] - qs.off[tid])
            atomicOr(&ctr[PB_ERR], 2);
        for (int i = tid; i < nctr; i += 256)
            if (i != PB_ERR) ctr[i] = 0;
    }
}

}  // namespace

int launch_nsyn16(const Plan *p, int kind, const NSynArgs &args, hipStream_t s)
{
    if (args.nlev <= 0) return MHX_OK;
    NSynArgs a = args;
    a.crec = p->d_crec;
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
    if (p->syn_form != 2 && p->l_max <= 255 && p->n_theta <= WNT && p->d_syn16x) {
        // whole-K form on P only: 12 theta tiles per block
        a.tP = p->d_syn16x;
        a.off = p->d_syn16x_off;
        a.rowscale = p->d_syn16x_rs;
        a.tpb = 4 * WTT;
        a.ngz = (p->nt16 + a.tpb - 1) / a.tpb;
        const int ng = (a.nlev + WLG - 1) / WLG;
        dim3 grid((unsigned)(8 * ((p->n_m + 7) / 8) * ng * a.ngz));
        const bool one = p->l_max + 1 <= 128;
        switch (kind) {
        case NS_SG:
            if (one) k_nsyn16p<NS_SG, 1><<<grid, 256, 0, s>>>(a);
            else k_nsyn16p<NS_SG, 2><<<grid, 256, 0, s>>>(a);
            break;
        case NS_QST:
            if (one) k_nsyn16p<NS_QST, 1><<<grid, 256, 0, s>>>(a);
            else k_nsyn16p<NS_QST, 2><<<grid, 256, 0, s>>>(a);
            break;
        case NS_QSTC:
            if (one) k_nsyn16p<NS_QSTC, 1><<<grid, 256, 0, s>>>(a);
            else k_nsyn16p<NS_QSTC, 2><<<grid, 256, 0, s>>>(a);
            break;
        case NS_S:
            if (one) k_nsyn16p<NS_S, 1><<<grid, 256, 0, s>>>(a);
            else k_nsyn16p<NS_S, 2><<<grid, 256, 0, s>>>(a);
            break;
        default: return fail(MHX_E_INVALID, "unknown nonlinear-step synthesis kind");
        }
        MHX_HIP_CHECK(hipGetLastError());
        return MHX_OK;
    }
    // 12 theta tiles per block from stored tables (all of l_max = 255); 8 with
    // the recurrence (its fragments of both parities live in registers);
    // 256 threads stage B
    const int nw = p->rec ? 4 : 6;
    a.tpb = 2 * nw;
    a.ngz = (p->nt16 + a.tpb - 1) / a.tpb;
    const int ng = (a.nlev + LGS - 1) / LGS;
    dim3 grid((unsigned)(8 * ((p->n_m + 7) / 8) * ng * a.ngz));
    const int nthr = 64 * nw;
    if (p->rec) {
        if (p->l_max > 1023 || !p->d_rec_seed)
            return fail(MHX_E_UNSUPPORTED, "recurrence tables need l_max <= 1023");
        a.seed = p->d_rec_seed;
        a.cost = p->d_cost;
        switch (kind) {
        case NS_SG: k_nsyn16<NS_SG, true><<<grid, nthr, 0, s>>>(a); break;
        case NS_QST: k_nsyn16<NS_QST, true><<<grid, nthr, 0, s>>>(a); break;
        case NS_QSTC: k_nsyn16<NS_QSTC, true><<<grid, nthr, 0, s>>>(a); break;
        default: return fail(MHX_E_INVALID, "unknown nonlinear-step synthesis kind");
        }
    } else {
        switch (kind) {
        case NS_SG: k_nsyn16<NS_SG, false><<<grid, nthr, 0, s>>>(a); break;
        case NS_QST: k_nsyn16<NS_QST, false><<<grid, nthr, 0, s>>>(a); break;
        case NS_QSTC: k_nsyn16<NS_QSTC, false><<<grid, nthr, 0, s>>>(a); break;
        default: return fail(MHX_E_INVALID, "unknown nonlinear-step synthesis kind");
        }
    }
    MHX_HIP_CHECK(hipGetLastError());
    return MHX_OK;
}

int launch_nana16(const Plan *p, int kind, const NAnaArgs &args, hipStream_t s)
{
    if (args.nlev <= 0) return MHX_OK;
    NAnaArgs a = args;
    a.crec = p->d_crec;
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
    const int ntlx = ((p->l_max + 1) / 2 + 1 + 15) / 16;  // l tiles with l <= l_max + 1
    const bool gfit = 3 * WLG * (2 * (p->l_max + 2) + 2) * 4 <= 6 * 16 * AROW;
    if (p->ana_form != 2 && (kind == NA_QST || kind == NA_SH) && p->nh <= AKP && ntlx <= 4 * PTT && gfit &&
        p->d_ana16x) {
        // whole-theta form on wP only
        a.tP = p->d_ana16x;
        a.off = p->d_ana16x_off;
        a.rowscale = p->d_ana16x_rs;
        a.soff = p->d_ana16x_soff;
        a.ngz = 1;
        const int ng = (a.nlev + WLG - 1) / WLG;
        dim3 grid((unsigned)(8 * ((p->n_m + 7) / 8) * ng));
        if (kind == NA_QST) k_nana16p<NA_QST><<<grid, 256, 0, s>>>(a);
        else k_nana16p<NA_SH><<<grid, 256, 0, s>>>(a);
        MHX_HIP_CHECK(hipGetLastError());
        return MHX_OK;
    }
    const int ng = (a.nlev + LGS * (p->rec ? RNG : 1) - 1) / (LGS * (p->rec ? RNG : 1));
    dim3 grid((unsigned)(8 * ((p->n_m + 7) / 8) * ng * a.ngz));
    if (p->rec) {
        if (p->l_max > 1023 || !p->d_rec_aseed)
            return fail(MHX_E_UNSUPPORTED, "recurrence tables need l_max <= 1023");
        a.aseed = p->d_rec_aseed;
        a.aexp = p->d_rec_aexp;
        a.segoff = p->d_rec_segoff;
        a.cost = p->d_cost;
        switch (kind) {
        case NA_QST: k_nana16<NA_QST, true><<<grid, 256, 0, s>>>(a); break;
        case NA_SH: k_nana16<NA_SH, true><<<grid, 256, 0, s>>>(a); break;
        default: return fail(MHX_E_INVALID, "unknown nonlinear-step analysis kind");
        }
    } else {
        switch (kind) {
        case NA_QST: k_nana16<NA_QST, false><<<grid, 256, 0, s>>>(a); break;
        case NA_SH: k_nana16<NA_SH, false><<<grid, 256, 0, s>>>(a); break;
        default: return fail(MHX_E_INVALID, "unknown nonlinear-step analysis kind");
        }
    }
    MHX_HIP_CHECK(hipGetLastError());
    return MHX_OK;
}

// the plan's tables and sizes admit the whole-K synthesis, the whole-theta
// analysis and row items of 4 theta rows within the persistent launch's LDS
bool pbatch_supported(const Plan *p)
{
    const int ntlx = ((p->l_max + 1) / 2 + 1 + 15) / 16;
    const bool gfit = 3 * WLG * (2 * (p->l_max + 2) + 2) * 4 <= 6 * 16 * AROW;
    return p->l_max <= 255 && p->n_theta <= WNT && p->d_syn16x && p->d_ana16x && p->nh <= AKP &&
           ntlx <= 4 * PTT && gfit && (int)p->fac.size() <= PB_MAXFAC && pb_row_bytes(p->M, 4) <= PB_LDS &&
           (!pb_slot_rows(p->M) || 3 * p->n_m <= 2 * p->M);  // slot row items: tiles of 2 M / 3 orders
}

size_t pbatch_row_lds(int M, int tbr) { return pb_row_bytes(M, tbr); }
int pbatch_slot_rows(int M) { return pb_slot_rows(M); }
size_t pbatch_lds() { return PB_LDS; }

// the plan's part of a job: tables and sizes of the whole-K synthesis and
// the whole-theta analysis (as their launchers set them), the FP32 FFT
// factors and twiddles; the job's arrays, counts and kinds are the caller's
int fill_pbatch_job(const Plan *p, PBJob &j)
{
    if (!pbatch_supported(p)) return fail(MHX_E_UNSUPPORTED, "plan does not admit the persistent batch form");
    NSynArgs &a = j.syn;
    a.crec = p->d_crec;
    a.tP = p->d_syn16x;
    a.tdP = nullptr;
    a.off = p->d_syn16x_off;
    a.rowscale = p->d_syn16x_rs;
    a.lstart = p->d_lstart;
    a.l_max = p->l_max;
    a.minc = p->minc;
    a.n_m = p->n_m;
    a.nh = p->nh;
    a.n_theta = p->n_theta;
    a.nt16 = p->nt16;
    a.tpb = 4 * WTT;
    a.ngz = 1;
    NAnaArgs &c = j.ana;
    c.crec = p->d_crec;
    c.tP = p->d_ana16x;
    c.tdP = nullptr;
    c.off = p->d_ana16x_off;
    c.rowscale = p->d_ana16x_rs;
    c.soff = p->d_ana16x_soff;
    c.lstart = p->d_lstart;
    c.l_max = p->l_max;
    c.m_max = p->m_max;
    c.minc = p->minc;
    c.n_m = p->n_m;
    c.nh = p->nh;
    c.n_theta = p->n_theta;
    c.ntb = (p->n_theta + 15) / 16;
    c.nq32 = p->nq32;
    c.ngz = 1;
    j.nrl = p->l_max + 1 <= 128 ? 1 : 2;
    j.M = p->M;
    j.N = p->n_phi;
    j.nfac = (int)p->fac.size();
    for (int i = 0; i < PB_MAXFAC; ++i) j.fac[i] = i < j.nfac ? p->fac[i] : 1;
    j.tw = p->d_tw32;
    j.twr = p->d_twr32;
    j.osin2 = p->d_osin2;
    j.n_m = p->n_m;
    j.n_theta = p->n_theta;
    j.nt16 = (p->n_theta + 15) / 16;
    return MHX_OK;
}

int probe_xcc_ids(int *ids, int n, hipStream_t s)
{
    k_xcc_ids<<<dim3((unsigned)n), 64, 0, s>>>(ids);
    MHX_HIP_CHECK(hipGetLastError());
    return MHX_OK;
}

int launch_pbatch(const PBJob *jobs, const PBItem *items, const PBQueues &q, int *ctr, int nctr,
                  int grid, hipStream_t s)
{
    if (q.nq <= 0 || q.off[q.nq] <= 0) return MHX_OK;
    k_pbatch<<<dim3((unsigned)grid), 256, 0, s>>>(jobs, items, q, ctr, nctr);
    MHX_HIP_CHECK(hipGetLastError());
    return MHX_OK;
}

}  // namespace mhx
