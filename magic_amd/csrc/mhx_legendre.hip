// Legendre stage on FP64 MFMA (v_mfma_f64_16x16x4f64), batched over radial
// levels.
//
// Synthesis (src/shtransforms.f90:103-372,493-646): for one order m, the
// Fourier coefficient of every output component at every theta is a sum over
// l of P_lm(theta) or dP_lm(theta) times a complex column built from the input
// coefficients.  With theta rows as GEMM rows, l as K and (8 radial levels x
// re/im) as 16 GEMM columns this is a dense contraction per m.  The
// equatorial symmetry of the reference (ES/EA sums, :172-189) is kept: K runs
// over l of one parity at a time, and two accumulators per output gather
//     Sym  = sum_{l-m even} P-terms + sum_{l-m odd} dP-terms
//     Anti = sum_{l-m odd}  P-terms + sum_{l-m even} dP-terms
// so that north = Sym + Anti and south = Sym - Anti (P is even and sin dP/dth
// odd under theta -> pi - theta for l - m even).
//
// Analysis (src/shtransforms.f90:648-872): rows are l of one parity, K is the
// northern theta, columns as above; the ES/EA combinations of north/south
// Fourier rows are formed while staging B, Gauss weights are folded into the
// tables (wPlm, wdPlm).
//
// The factors i m of the vector transforms are folded into the B columns so
// that every output is a plain sum of MFMA products held in the same lanes.
#include "mhx_internal.h"

namespace mhx {

namespace {

constexpr int WG = 256;
constexpr int KQC = 8;          // quads (of 4 l or 4 theta) per staged chunk
constexpr int CHUNK = 4 * KQC;  // 32 rows of B per parity / per theta chunk

struct Prod {
    int acc, tab, col;  // tab 0: P, 1: dP
};

template <int KIND> struct SynSpec;
template <> struct SynSpec<SYN_SCAL> {
    static constexpr int NOUT = 1, NCOL = 1, NPROD = 1;
    static constexpr Prod prod[1] = {{0, 0, 0}};
};
template <> struct SynSpec<SYN_GRAD> {  // gt = dP.x ; gp = P.(i m x)
    static constexpr int NOUT = 2, NCOL = 2, NPROD = 2;
    static constexpr Prod prod[2] = {{0, 1, 0}, {1, 0, 1}};
};
template <> struct SynSpec<SYN_SG> {    // s = P.x ; gt = dP.x ; gp = P.(i m x)
    static constexpr int NOUT = 3, NCOL = 2, NPROD = 3;
    static constexpr Prod prod[3] = {{0, 0, 0}, {1, 1, 0}, {2, 0, 1}};
};
template <> struct SynSpec<SYN_QST> {   // cols: Q, S, i m T, i m S, -T
    static constexpr int NOUT = 3, NCOL = 5, NPROD = 5;
    static constexpr Prod prod[5] = {{0, 0, 0}, {1, 1, 1}, {1, 0, 2}, {2, 0, 3}, {2, 1, 4}};
};
template <> struct SynSpec<SYN_SPHTOR> {  // cols: S, i m T, i m S, -T
    static constexpr int NOUT = 2, NCOL = 4, NPROD = 4;
    static constexpr Prod prod[4] = {{0, 1, 0}, {0, 0, 1}, {1, 0, 2}, {1, 1, 3}};
};

template <int KIND> struct AnaSpec;
template <> struct AnaSpec<ANA_SH> {      // cols: f
    static constexpr int NOUT = 1, NCOL = 1, NPROD = 1;
    static constexpr Prod prod[1] = {{0, 0, 0}};
    static constexpr bool divll[1] = {false};
};
template <> struct AnaSpec<ANA_SPHTOR> {  // cols: vt, -i m vp, -vp, -i m vt
    static constexpr int NOUT = 2, NCOL = 4, NPROD = 4;
    static constexpr Prod prod[4] = {{0, 1, 0}, {0, 0, 1}, {1, 1, 2}, {1, 0, 3}};
    static constexpr bool divll[2] = {true, true};
};
template <> struct AnaSpec<ANA_QST> {     // cols: f, vt, -i m vp, -vp, -i m vt
    static constexpr int NOUT = 3, NCOL = 5, NPROD = 5;
    static constexpr Prod prod[5] = {{0, 0, 0}, {1, 1, 1}, {1, 0, 2}, {2, 1, 3}, {2, 0, 4}};
    static constexpr bool divll[3] = {false, true, true};
};

struct SynArgs {
    SynJob job;
    const double *sP, *sdP;
    const size_t *off;
    const int *lstart;
    const double *osin2;
    int l_max, minc, n_m, nh, n_theta, nt16;
};

struct AnaArgs {
    AnaJob job;
    const double *aP, *adP;
    const size_t *off;
    const int *lstart;
    int l_max, m_max, minc, n_m, nh, n_theta, nq4;
};

// v_mfma_f64_16x16x4f64 operands: A lane l holds A[l % 16][l / 16], B lane l
// holds B[l / 16][l % 16]; D element j of lane l is D[l / 16 + 4 j][l % 16].
__device__ inline d4 mfma64(double a, double b, d4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Blocks are dispatched round-robin over the 8 XCDs (each with its own L2).
// Remap the linear block id so that each XCD works through a contiguous range
// of logical blocks; with the order index m fastest, the blocks resident on
// one XCD at a time cover neighbouring m, whose values share cache lines of
// the row-contiguous forward-side layout F2[lev][comp][theta][m].
__device__ inline int xcd_logical(int b, int nb)
{
    const int nx = 8;
    int per = nb / nx, rem = nb % nx;
    int x = b % nx, i = b / nx;
    return (x < rem ? x * (per + 1) : rem * (per + 1) + (x - rem) * per) + i;
}

// i^ipow * m^mpow * scale applied to (re, im)
__device__ inline void apply_factor(double &re, double &im, int ipow, int mpow, double m,
                                    double scale)
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

// ------------------------------------------------------------- synthesis
template <int KIND>
__global__ __launch_bounds__(WG) void k_syn(SynArgs a)
{
    using S = SynSpec<KIND>;
    constexpr int NOUT = S::NOUT, NCOL = S::NCOL, NPROD = S::NPROD;
    // B chunk: [col][parity][32 kk][16 cols]; reused by the epilogue
    constexpr int BS_DOUBLES = NCOL * 2 * CHUNK * 16;
    constexpr int EP_DOUBLES = 4 * LEV_GROUP * 32 * 2;  // per wave: 8 lev x 32 theta x re/im
    __shared__ double smem[BS_DOUBLES > EP_DOUBLES ? BS_DOUBLES : EP_DOUBLES];

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
    const int K0 = (a.l_max - m) / 2 + 1, K1 = (a.l_max - m + 1) / 2;
    const int Kq[2] = {(K0 + 3) / 4, (K1 + 3) / 4};
    const int l0 = a.lstart[mc];

    d4 acc[NOUT][2];
#pragma unroll
    for (int o = 0; o < NOUT; ++o)
        for (int s = 0; s < 2; ++s) acc[o][s] = d4{0.0, 0.0, 0.0, 0.0};

    // B staging is software-pipelined: the input coefficients of chunk c+1
    // are loaded into registers while the MFMAs of chunk c run, then built
    // into columns in LDS.  The table fragments of a chunk are all issued
    // before its MFMAs.
    constexpr int NE = (2 * CHUNK * LEV_GROUP) / WG;  // staging entries per thread
    double pxr[NE][MAX_SRC], pxi[NE][MAX_SRC];
    double pll[NE];
    bool pval[NE];
    auto load_B = [&](int c) {
#pragma unroll
        for (int e = 0; e < NE; ++e) {
            int idx = threadIdx.x + e * WG;
            int levl = idx % LEV_GROUP;
            int kkl = (idx / LEV_GROUP) % CHUNK;
            int par = idx / (LEV_GROUP * CHUNK);
            int l = m + par + 2 * (c * CHUNK + kkl);
            int lev = lev0 + levl;
            bool valid = (levl < nlev_here) && (l <= a.l_max);
            if (valid) valid = l <= job.lcut[lev];
            pval[e] = valid;
            size_t o = ((size_t)lev * job.src_ld + (size_t)(l0 + l - m)) * 2;
#pragma unroll
            for (int sidx = 0; sidx < MAX_SRC; ++sidx) {
                double2 v = make_double2(0.0, 0.0);
                if (valid && job.src[sidx]) v = *reinterpret_cast<const double2 *>(job.src[sidx] + o);
                pxr[e][sidx] = v.x;
                pxi[e][sidx] = v.y;
            }
        }
    };
    auto store_B = [&]() {
#pragma unroll
        for (int e = 0; e < NE; ++e) {
            int idx = threadIdx.x + e * WG;
            int levl = idx % LEV_GROUP;
            int kkl = (idx / LEV_GROUP) % CHUNK;
            int par = idx / (LEV_GROUP * CHUNK);
            int lev = lev0 + levl;
#pragma unroll
            for (int col = 0; col < NCOL; ++col) {
                double re = 0.0, im = 0.0;
                if (pval[e]) {
                    const ColOp &op = job.col[col];
                    double g = 1.0;
                    if (op.gkind != G_ONE) {
                        double ll = pll[e];
                        g = (op.gkind == G_LL) ? ll : job.or2[lev] * ll;
                    }
                    double xr0 = 0.0, xi0 = 0.0, xr1 = 0.0, xi1 = 0.0;
#pragma unroll
                    for (int sidx = 0; sidx < MAX_SRC; ++sidx) {
                        if (op.src0 == sidx) { xr0 = pxr[e][sidx]; xi0 = pxi[e][sidx]; }
                        if (op.src1 == sidx) { xr1 = pxr[e][sidx]; xi1 = pxi[e][sidx]; }
                    }
                    re = g * xr0 + op.hcoef * xr1;
                    im = g * xi0 + op.hcoef * xi1;
                    apply_factor(re, im, op.ipow, op.mpow, dm, op.scale);
                }
                double *b = smem + (((col * 2 + par) * CHUNK + kkl) * 16 + levl * 2);
                b[0] = re;
                b[1] = im;
            }
        }
    };

    const int nchunk = (Kq[0] + KQC - 1) / KQC;
    auto set_ll = [&](int c) {
#pragma unroll
        for (int e = 0; e < NE; ++e) {
            int idx = threadIdx.x + e * WG;
            int kkl = (idx / LEV_GROUP) % CHUNK;
            int par = idx / (LEV_GROUP * CHUNK);
            double l = (double)(m + par + 2 * (c * CHUNK + kkl));
            pll[e] = l * (l + 1.0);
        }
    };
    load_B(0);
    set_ll(0);
    store_B();
    __syncthreads();
    const double *tP[2], *tdP[2];
#pragma unroll
    for (int par = 0; par < 2; ++par) {
        tP[par] = a.sP + a.off[2 * mc + par] + (size_t)tt * 64 + lane;
        tdP[par] = a.sdP + a.off[2 * mc + par] + (size_t)tt * 64 + lane;
    }
    const size_t qstride = (size_t)a.nt16 * 64;
    for (int c = 0; c < nchunk; ++c) {
        double aP[2][KQC], adP[2][KQC];
        int nq[2];
#pragma unroll
        for (int par = 0; par < 2; ++par) {
            nq[par] = min(KQC, Kq[par] - c * KQC);
#pragma unroll
            for (int q = 0; q < KQC; ++q) {
                bool ok = active && q < nq[par];
                size_t kq = (size_t)(c * KQC + q);
                aP[par][q] = ok ? tP[par][kq * qstride] : 0.0;
                adP[par][q] = ok ? tdP[par][kq * qstride] : 0.0;
            }
        }
        if (c + 1 < nchunk) load_B(c + 1);
        if (active) {
#pragma unroll
            for (int par = 0; par < 2; ++par)
#pragma unroll
                for (int q = 0; q < KQC; ++q) {
                    if (q >= nq[par]) break;
#pragma unroll
                    for (int pr = 0; pr < NPROD; ++pr) {
                        const Prod P = S::prod[pr];
                        double b = smem[((P.col * 2 + par) * CHUNK + q * 4 + (lane >> 4)) * 16 + (lane & 15)];
                        // P-table: even parity -> Sym (0); dP-table: even -> Anti (1)
                        int tgt = (P.tab == 0) ? par : (1 - par);
                        acc[P.acc][tgt] = mfma64(P.tab == 0 ? aP[par][q] : adP[par][q], b, acc[P.acc][tgt]);
                    }
                }
        }
        __syncthreads();
        if (c + 1 < nchunk) {
            set_ll(c + 1);
            store_B();
        }
        __syncthreads();
    }

    // epilogue: north = Sym + Anti, south = Sym - Anti; transposed through
    // LDS into contiguous runs of 32 theta rows of F[lev][comp][m][theta]
    __syncthreads();
    double *ep = smem + wave * (LEV_GROUP * 32 * 2);
    const int col = lane & 15, levl_c = col >> 1, ri = col & 1;
#pragma unroll
    for (int o = 0; o < NOUT; ++o) {
        if (active) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                int row = (lane >> 4) + 4 * j;
                int th = tt * 16 + row;
                double sc = 1.0;
                if (job.row_scale[o] && th < a.nh) sc = a.osin2[2 * th];
                ep[(levl_c * 32 + 2 * row) * 2 + ri] = (acc[o][0][j] + acc[o][1][j]) * sc;
                ep[(levl_c * 32 + 2 * row + 1) * 2 + ri] = (acc[o][0][j] - acc[o][1][j]) * sc;
            }
        }
        __syncthreads();
        if (active) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                int e = r * 64 + lane;
                int levl = e >> 5, thl = e & 31;
                int theta = tt * 32 + thl;
                if (levl < nlev_here && theta < a.n_theta) {
                    size_t o2 = ((size_t)(lev0 + levl) * job.fout_ld +
                                 ((size_t)job.out_comp[o] * a.n_m + mc) * a.n_theta + theta) * 2;
                    *reinterpret_cast<double2 *>(job.fout + o2) =
                        make_double2(ep[(levl * 32 + thl) * 2], ep[(levl * 32 + thl) * 2 + 1]);
                }
            }
        }
        __syncthreads();
    }
}

// -------------------------------------------------------------- analysis
template <int KIND>
__global__ __launch_bounds__(WG) void k_ana(AnaArgs a)
{
    using S = AnaSpec<KIND>;
    constexpr int NOUT = S::NOUT, NCOL = S::NCOL, NPROD = S::NPROD;
    constexpr int BS_DOUBLES = NCOL * 2 * CHUNK * 16;     // [col][sym/anti][32 th][16]
    constexpr int EP_DOUBLES = 4 * LEV_GROUP * 32 * 2;
    __shared__ double smem[BS_DOUBLES > EP_DOUBLES ? BS_DOUBLES : EP_DOUBLES];

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

    // column definitions: source comp slot, i-power, m-power, scale
    // SH: f | SPHTOR: vt, -i m vp, -vp, -i m vt | QST: f + SPHTOR's
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

    d4 acc[NOUT][2];
#pragma unroll
    for (int o = 0; o < NOUT; ++o)
        for (int s = 0; s < 2; ++s) acc[o][s] = d4{0.0, 0.0, 0.0, 0.0};

    const int nchunk = (a.nq4 + KQC - 1) / KQC;
    for (int c = 0; c < nchunk; ++c) {
        __syncthreads();
        for (int idx = threadIdx.x; idx < CHUNK * LEV_GROUP; idx += WG) {
            int levl = idx % LEV_GROUP;
            int thl = idx / LEV_GROUP;
            int th = c * CHUNK + thl;
            int lev = lev0 + levl;
            bool valid = (levl < nlev_here) && (th < a.nh);
            if (valid) {
                int lc = job.lcut[lev];
                int mlim = lc < a.m_max ? lc : a.m_max;
                valid = m <= mlim;
            }
            double nr[3], ni[3], sr[3], si[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) { nr[k] = ni[k] = sr[k] = si[k] = 0.0; }
            if (valid) {
                int nsrc = (KIND == ANA_SH) ? 1 : 3;
                for (int k = 0; k < nsrc; ++k) {
                    int comp = job.in_comp[k];
                    if (comp < 0) continue;
                    size_t o = ((size_t)lev * job.fin_ld + ((size_t)comp * a.n_theta + 2 * th) * a.n_m + mc) * 2;
                    size_t o2 = o + (size_t)a.n_m * 2;
                    nr[k] = job.fin[o];
                    ni[k] = job.fin[o + 1];
                    sr[k] = job.fin[o2];
                    si[k] = job.fin[o2 + 1];
                }
            }
#pragma unroll
            for (int col = 0; col < NCOL; ++col) {
                int k = csrc[col];
                double syr = nr[k] + sr[k], syi = ni[k] + si[k];
                double anr = nr[k] - sr[k], ani = ni[k] - si[k];
                apply_factor(syr, syi, cip[col], cmp[col], dm, csc[col]);
                apply_factor(anr, ani, cip[col], cmp[col], dm, csc[col]);
                double *b0 = smem + (((col * 2 + 0) * CHUNK + thl) * 16 + levl * 2);
                double *b1 = smem + (((col * 2 + 1) * CHUNK + thl) * 16 + levl * 2);
                b0[0] = syr; b0[1] = syi;
                b1[0] = anr; b1[1] = ani;
            }
        }
        __syncthreads();
        if (active) {
            int nq = min(KQC, a.nq4 - c * KQC);
            for (int q = 0; q < nq; ++q) {
                size_t tq = (size_t)(c * KQC + q);
#pragma unroll
                for (int par = 0; par < 2; ++par) {
                    size_t o = a.off[2 * mc + par] + ((size_t)t * a.nq4 + tq) * 64 + lane;
                    double aP = a.aP[o];
                    double adP = a.adP[o];
#pragma unroll
                    for (int pr = 0; pr < NPROD; ++pr) {
                        const Prod P = S::prod[pr];
                        // even rows: P uses the symmetric combination, dP the antisymmetric
                        int use = (P.tab == 0) ? par : (1 - par);
                        double b = smem[((P.col * 2 + use) * CHUNK + q * 4 + (lane >> 4)) * 16 + (lane & 15)];
                        acc[P.acc][par] = mfma64(P.tab == 0 ? aP : adP, b, acc[P.acc][par]);
                    }
                }
            }
        }
    }

    // epilogue: rows r -> l = m + par + 2 (16 t + r); transpose to 32
    // consecutive l per level
    __syncthreads();
    double *ep = smem + wave * (LEV_GROUP * 32 * 2);
    const int col = lane & 15, levl_c = col >> 1, ri = col & 1;
#pragma unroll
    for (int o = 0; o < NOUT; ++o) {
        if (active) {
#pragma unroll
            for (int par = 0; par < 2; ++par)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    int r = (lane >> 4) + 4 * j;
                    int l = m + par + 2 * (16 * t + r);
                    double v = acc[o][par][j];
                    if (S::divll[o] && l > 0) v /= (double)l * (double)(l + 1);
                    ep[(levl_c * 32 + 2 * r + par) * 2 + ri] = v;
                }
        }
        __syncthreads();
        if (active) {
#pragma unroll
            for (int rr = 0; rr < 4; ++rr) {
                int e = rr * 64 + lane;
                int levl = e >> 5, ll = e & 31;
                int l = m + 32 * t + ll;
                if (levl < nlev_here && l <= a.l_max) {
                    int lev = lev0 + levl;
                    int lc = job.lcut[lev];
                    bool keep = (l <= lc) && (m <= (lc < a.m_max ? lc : a.m_max));
                    double2 v = keep ? make_double2(ep[(levl * 32 + ll) * 2], ep[(levl * 32 + ll) * 2 + 1])
                                     : make_double2(0.0, 0.0);
                    size_t o2 = ((size_t)lev * job.out_ld + (size_t)(l0 + l - m)) * 2;
                    *reinterpret_cast<double2 *>(job.out[o] + o2) = v;
                }
            }
        }
        __syncthreads();
    }
}

}  // namespace

int launch_syn(const Plan *p, const SynJob &job, hipStream_t s)
{
    if (int rc = ensure_fp64_tables(const_cast<Plan *>(p))) return rc;
    if (job.nlev <= 0) return MHX_OK;
    SynArgs a;
    a.job = job;
    a.sP = p->d_syn_P;
    a.sdP = p->d_syn_dP;
    a.off = p->d_syn_off;
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
    case SYN_SCAL: k_syn<SYN_SCAL><<<grid, WG, 0, s>>>(a); break;
    case SYN_GRAD: k_syn<SYN_GRAD><<<grid, WG, 0, s>>>(a); break;
    case SYN_SG: k_syn<SYN_SG><<<grid, WG, 0, s>>>(a); break;
    case SYN_QST: k_syn<SYN_QST><<<grid, WG, 0, s>>>(a); break;
    case SYN_SPHTOR: k_syn<SYN_SPHTOR><<<grid, WG, 0, s>>>(a); break;
    default: return fail(MHX_E_INVALID, "unknown synthesis kind");
    }
    MHX_HIP_CHECK(hipGetLastError());
    return MHX_OK;
}

int launch_ana(const Plan *p, const AnaJob &job, hipStream_t s)
{
    if (int rc = ensure_fp64_tables(const_cast<Plan *>(p))) return rc;
    if (job.nlev <= 0) return MHX_OK;
    AnaArgs a;
    a.job = job;
    a.aP = p->d_ana_wP;
    a.adP = p->d_ana_wdP;
    a.off = p->d_ana_off;
    a.lstart = p->d_lstart;
    a.l_max = p->l_max;
    a.m_max = p->m_max;
    a.minc = p->minc;
    a.n_m = p->n_m;
    a.nh = p->nh;
    a.n_theta = p->n_theta;
    a.nq4 = p->nq4;
    int ntile0 = ((p->l_max) / 2 + 1 + 15) / 16;  // pair tiles at m = 0
    dim3 grid((unsigned)(p->n_m * ((job.nlev + LEV_GROUP - 1) / LEV_GROUP) * ((ntile0 + 3) / 4)));
    switch (job.kind) {
    case ANA_SH: k_ana<ANA_SH><<<grid, WG, 0, s>>>(a); break;
    case ANA_SPHTOR: k_ana<ANA_SPHTOR><<<grid, WG, 0, s>>>(a); break;
    case ANA_QST: k_ana<ANA_QST><<<grid, WG, 0, s>>>(a); break;
    default: return fail(MHX_E_INVALID, "unknown analysis kind");
    }
    MHX_HIP_CHECK(hipGetLastError());
    return MHX_OK;
}

}  // namespace mhx
