// Internal declarations shared by the mhx HIP sources.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/mhx.h"

namespace mhx {

typedef double d4 __attribute__((ext_vector_type(4)));

// row factor of the split tables: the power of two s with max * s in
// [2^14, 2^15); 1 for max == 0; capped at 2^115 so that rows of tiny values
// (P_lm near the poles) stay finite
__device__ inline float pow2_scale(float mx)
{
    if (!(mx > 0.f)) return 1.f;
    int e;
    frexpf(mx, &e);          // mx = f 2^e, f in [0.5, 1)
    if (e < -100) e = -100;
    return ldexpf(1.f, 15 - e);
}

// ---------------------------------------------------------------- errors
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);
// true for device (or managed) memory, false for host memory
bool is_device_ptr(const void *ptr);
#define MHX_HIP_CHECK(expr)                                                   \
    do {                                                                      \
        hipError_t e_ = (expr);                                               \
        if (e_ != hipSuccess)                                                 \
            return ::mhx::fail(MHX_E_HIP, std::string(#expr) + ": " +         \
                                               hipGetErrorString(e_));        \
    } while (0)

// ----------------------------------------------------- Legendre job model
// A transform "job" runs one kind of Legendre contraction for a group of
// radial levels.  B columns are complex coefficient vectors built on the fly
// from up to two input arrays:
//     col = i^ipow * m^mpow * (g(l, lev) * x0 + hcoef * x1)
// with g = 1, l(l+1) or or2[lev] l(l+1).
enum GKind { G_ONE = 0, G_LL = 1, G_OR2LL = 2 };

struct ColOp {
    int src0;     // index into the job's source list
    int src1;     // -1 if unused
    int gkind;
    int ipow;     // power of i (0..3)
    int mpow;     // power of m (0..2)
    double hcoef; // coefficient of x1
    double scale;
};

// synthesis kinds (see mhx_legendre.hip for the product tables)
enum SynKind { SYN_SCAL = 0, SYN_GRAD = 1, SYN_SG = 2, SYN_QST = 3, SYN_SPHTOR = 4 };
// analysis kinds
enum AnaKind { ANA_SH = 0, ANA_SPHTOR = 1, ANA_QST = 2 };

constexpr int MAX_SRC = 4;
constexpr int MAX_COL = 5;
constexpr int MAX_OUT = 3;
constexpr int LEV_GROUP = 8;   // radial levels per MFMA column tile (re/im -> 16)

struct SynJob {
    int kind;
    int nlev;
    // sources: coefficient arrays (complex, interleaved), level stride ld (complex)
    const double *src[MAX_SRC];
    size_t src_ld;
    ColOp col[MAX_COL];
    // outputs: Fourier arrays F[lev][comp][m][theta] complex; comp index and
    // per-level stride (complex elements)
    double *fout;
    size_t fout_ld;      // complex elements per level
    int out_comp[MAX_OUT];
    int row_scale[MAX_OUT]; // 1: multiply by 1/sin^2(theta) (dphspat)
    const int *lcut;     // device, per level
    const double *or2;   // device, per level (may be null)
};

struct AnaJob {
    int kind;
    int nlev;
    // inputs: Fourier arrays F[lev][comp][m][theta]
    const double *fin;
    size_t fin_ld;       // complex elements per level
    int in_comp[3];      // f (SH/QST), vt, vp
    // outputs: coefficient arrays
    double *out[MAX_OUT];
    size_t out_ld;       // complex elements per level
    const int *lcut;     // device
};

// Legendre kernels of the nonlinear step in the split form (mhx_nlsplit.hip)
// NS_S: the scalar alone (no gradient), split whole-K form only
enum NSynKind { NS_SG = 0, NS_QST = 1, NS_QSTC = 2, NS_S = 3 };
enum NAnaKind { NA_QST = 0, NA_SH = 1 };

struct NSynArgs {
    const double *src[4];   // coefficient arrays, complex, level stride src_ld
    size_t src_ld;
    float *fout;            // F[lev][comp][m][theta] FP32 complex
    size_t fout_ld;         // complex elements per level
    int comp0;              // first output comp
    const int *lcut;        // device, per level
    const double *or2;      // device, per level
    int nlev;
    // filled by the launcher
    const _Float16 *tP, *tdP;
    const size_t *off;
    const float *rowscale;
    const int *lstart;
    const float *crec;      // c_l per (m, l), filled by the launcher
    const double2 *seed;    // recurrence seeds per (mc, theta_h)
    const double *cost;     // cos(theta_h)
    int l_max, minc, n_m, nh, n_theta, nt16, tpb, ngz;
};

struct NAnaArgs {
    const float *fin;       // F2[lev][comp][theta/16][m][theta%16] FP32 complex
    size_t fin_ld;          // complex elements per level
    int in_comp[3];         // f (QST, SH), vt, vp
    double *out[3];         // coefficient arrays, complex, level stride out_ld
    size_t out_ld;
    const int *lcut;        // device
    int nlev;
    // filled by the launcher
    const _Float16 *tP, *tdP;
    const size_t *off;
    const float *rowscale;
    const size_t *soff;
    const int *lstart;
    const float *crec;      // c_l per (m, l), filled by the launcher
    const double2 *aseed;   // recurrence seeds per (segment, theta_h)
    const int *aexp;
    const size_t *segoff;
    const double *cost;
    int l_max, m_max, minc, n_m, nh, n_theta, ntb, nq32, ngz;
};

// Courant maxima fused into the grid stage (src/courant.f90): par = [courfac^2,
// alffac^2, valri2[nlev], valhi2[nlev]] (device), max = [vr2max, vh2max] per
// level as the bit patterns of non-negative doubles (device, zeroed first);
// max == nullptr: not computed
struct CourArgs {
    const double *par = nullptr;
    unsigned long long *max = nullptr;
};

// lm-space assembly of the explicit terms (mhx_td.hip)
struct TdArgs {
    const double *in[10];   // AdvrLM AdvtLM AdvpLM VStLM VxBrLM VxBtLM VxBpLM w dw z
    double *out[7];         // dwdt dzdt dpdt dsdt dbdt djdt dVxBhLM
    size_t ld_lm;           // complex elements per level
    const double *rad;      // device: or2[nlev], or4[nlev], epsc[nlev]
    const int *lcut;        // device: l_R per level
    double CorFac;
    int nlev;
    // filled by the launcher
    const int *lm2l, *lm2m;
    int l_max, lm_max;
};

// ---- the radial loop of any physics switches (mhx_rloop.hip) -------------
// grid components of its inverse side (the reference's grid_space_arrays_t
// fields, src/get_nl.f90:39-67) and products of its forward side
enum RlIn {
    RI_VR = 0, RI_VT, RI_VP, RI_CVR, RI_CVT, RI_CVP, RI_DVRDR, RI_DVTDR, RI_DVPDR,
    RI_DVRDT, RI_DVRDP, RI_DVTDP, RI_DVPDP, RI_S, RI_XI, RI_PHI,
    RI_BR, RI_BT, RI_BP, RI_CBR, RI_CBT, RI_CBP, RI_N
};
enum RlOut {
    RO_ADVR = 0, RO_ADVT, RO_ADVP, RO_VSR, RO_VST, RO_VSP, RO_VXIR, RO_VXIT, RO_VXIP,
    RO_VXBR, RO_VXBT, RO_VXBP, RO_HEAT, RO_PHI, RO_N
};
// per-level radial factors of the grid stage, RL_NRAD arrays of nlev
enum RlRad {
    RR_R = 0, RR_OR1, RR_OR2, RR_OR4, RR_ORHO1, RR_BETA, RR_OTEMP1, RR_VISC, RR_LAMBDA,
    RR_LCR, RR_VALRI2, RR_VALHI2, RL_NRAD
};
// per-theta-row factors (reference's scrambled rows), RL_NTH arrays of n_theta
enum RlTheta { RT_OSIN2 = 0, RT_COSN2, RT_SIN, RT_COS, RT_OSIN, RL_NTH };

struct RlPhys {
    int conv_nl, adv_curl, heat_nl, chem, phase, prec, centr, mag_LF, mag_nl, anel;
    int cour_mag;   // magnetic branch of the Courant maxima
    double LFfac, ViscHeatFac, OhmLossFac, epsPhase, penaltyFac, phaseDiffFac, tmelt;
    double posnalp, oek, time, dilution_fac, ra, opr, cf2, af2;
};

struct RlGridArgs {
    const double *Fin;   // F[lev][nin][m][theta] complex
    double *Fout;        // F2[lev][nout][theta][m] complex
    int nin, nout;
    int in_of[RI_N];     // row (compact comp index) of each grid component, -1 absent
    int out_of[RO_N];    // row of each product, -1 absent
    unsigned out_osin2;  // rows multiplied by 1/sin^2(theta) before the forward FFT
    int n_m, n_theta, nlev, G;
    double dphi;         // 2 pi / (n_phi minc)
    const double *rad;   // RL_NRAD x nlev
    const double *th;    // RL_NTH x n_theta
    RlPhys ph;
    unsigned long long *cmax;  // [lev][2] vr2max, vh2max (bit patterns), zeroed first
};
struct RlBuffers;  // the radial loop's device buffers (mhx_rloop.hip)

constexpr int EK_CHUNKS = 16;  // lm chunks (blocks) per level of the energy sums
struct EkArgs {
    const double *w, *dw, *z;  // complex, level stride ld_lm (b, db, aj for mag)
    size_t ld_lm;
    const double *or2;         // device: or2[nlev]
    const double *orho1;       // device: orho1[nlev], null for 1 (mag: unused)
    double *part;              // device: nsum x nlev x EK_CHUNKS partial sums
    double *out;               // device: nsum sums x nlev
    int nlev;
    int mag = 0;               // magnetic energies (src/magnetic_energy.f90:264-340)
    int nsum = 8;              // 8 (kinetic) or 10 (magnetic: + dipole, axial dipole)
    // filled by the launcher
    const int *lm2l, *lm2m;
    int lm_max;
};

// ------------------------------------------------------------------ plan
struct Plan {
    int l_max, m_max, minc, n_theta, n_phi, nh, n_m, lm_max, precision;
    int device = 0;     // the device current at plan creation
    int nt16;   // theta tiles of 16 (northern half)
    int nq4;    // theta quads (northern half)
    int M;      // complex FFT length n_phi/2
    std::vector<int> lstart;        // n_m
    std::vector<int> fac;           // FFT radices of M
    std::vector<size_t> syn_off;    // [mc*2+p] offsets (doubles) into syn tables
    std::vector<size_t> ana_off;    // [mc*2+p]
    size_t syn_size = 0, ana_size = 0;
    std::vector<double> theta_ord, gauss;  // nh each (host)

    // device
    int *d_lstart = nullptr;
    double *d_syn_P = nullptr, *d_syn_dP = nullptr;   // FP64 MFMA-tiled
    double *d_ana_wP = nullptr, *d_ana_wdP = nullptr;
    size_t *d_syn_off = nullptr, *d_ana_off = nullptr;
    double *d_osin2 = nullptr;      // n_theta (scrambled)
    double *d_theta = nullptr, *d_gauss = nullptr; // nh
    double *d_tw = nullptr;         // M complex: exp(-2 pi i k/M)
    double *d_twr = nullptr;        // (M+1) complex: exp(-2 pi i k/n_phi)
    float *d_tw32 = nullptr, *d_twr32 = nullptr;  // the same in FP32
    float *d_twn32 = nullptr;       // n_phi complex: exp(-2 pi i k/n_phi), FP32
    int grid_form = 0;              // grid stage of the split step: 0 tiles where supported, 1 slots
    // the nonlinear step's du/dr grids (synthesis of dw, ddw, dz and their
    // transforms): mhx_set_dudr; off inside mhx_rloop's fused route
    bool dudr = true;

    // split-precision tables (MHX_PREC_SPLIT16): FP16 hi/lo pairs in the
    // 16x16x32 MFMA operand layout, with power-of-2 row factors
    int nq32 = 0;                   // theta chunks of 32 (northern half)
    size_t syn16_size = 0, ana16_size = 0;   // halves
    // Courant maxima of mhx_nl_step_courant (set for the call's launches)
    CourArgs cour;
    double *d_cour_par = nullptr;
    unsigned long long *d_cour_max = nullptr;
    int cour_nlev = 0;
    std::vector<double> h_cour;
    // per-level factors of mhx_get_td (or2, or4, epsc; l_R)
    double *d_ek = nullptr;  int ek_nlev = 0;  // energy sums: radial factors + outputs
    double *d_td_rad = nullptr;
    int *d_td_lcut = nullptr;
    int td_nlev = 0;
    std::vector<double> h_td;
    std::vector<int> h_td_lcut;
    // recurrence-generated synthesis tables (rec): FP64 seeds P_mm(theta) x
    // row scale as (mantissa, power-of-two exponent) per (mc, theta_h)
    bool rec = false;
    double2 *d_rec_seed = nullptr;
    double *d_cost = nullptr;  // cos(theta_h)
    // analysis: 2 pi w(theta) (P_{l-1}, P_l) at l = m + 32 s, mantissas and a
    // common power-of-two exponent, per (mc, segment s, theta_h)
    double2 *d_rec_aseed = nullptr;
    int *d_rec_aexp = nullptr;
    size_t *d_rec_segoff = nullptr;  // first segment of each order
    _Float16 *d_syn16_P = nullptr, *d_syn16_dP = nullptr;
    _Float16 *d_ana16_P = nullptr, *d_ana16_dP = nullptr;
    size_t *d_syn16_off = nullptr, *d_ana16_off = nullptr, *d_ana16_soff = nullptr;
    float *d_syn16_rs = nullptr, *d_ana16_rs = nullptr;
    // P-only split tables with l <= l_max + 1 (dP products folded onto P by
    // sin(theta) dP_l = l c_{l+1} P_{l+1} - (l+1) c_l P_{l-1}); same layouts
    // with K'[par] = #{l in [m, l_max+1] : l - m = par mod 2}
    _Float16 *d_syn16x = nullptr, *d_ana16x = nullptr;
    size_t *d_syn16x_off = nullptr, *d_ana16x_off = nullptr, *d_ana16x_soff = nullptr;
    float *d_syn16x_rs = nullptr, *d_ana16x_rs = nullptr;
    // c_l = sqrt((l^2 - m^2) / (4 l^2 - 1)) for l = m .. l_max + 2 at lstart[mc] + 2 mc
    float *d_crec = nullptr;
    // split Legendre forms (0: by shape, 1: whole-K / whole-theta, 2: chunked),
    // set by the batch autotuner
    int syn_form = 0, ana_form = 0;

    // workspaces
    int ws_levels = 0;
    double *ws_F = nullptr;     size_t ws_F_bytes = 0;    // Fourier, inverse side
    double *ws_G = nullptr;     size_t ws_G_bytes = 0;    // grids
    double *ws_F2 = nullptr;    size_t ws_F2_bytes = 0;   // Fourier, forward side
    double *ws_C = nullptr;     size_t ws_C_bytes = 0;    // staged coefficients
    double *ws_H = nullptr;     size_t ws_H_bytes = 0;    // nl step host arrays, staged
    int *d_lm2l = nullptr;      // lm_max
    int *d_lm2m = nullptr;      // lm_max
    int *d_lcut = nullptr;      int d_lcut_n = 0;
    double *d_rad = nullptr;    int d_rad_n = 0;          // or2, or4, orho1
    std::vector<double> h_rad;  // host copy of the per-level radial factors
    // pinned staging of the nonlinear step's per-level factors and lcut
    void *h_pin = nullptr;      size_t h_pin_bytes = 0;
    // the split nonlinear step's launches as a replayed graph, keyed by the
    // call's arrays and sizes (captured on the second call with one key)
    hipGraphExec_t gexec = nullptr;
    std::vector<unsigned long long> gkey, gkey_seen;
    hipStream_t stream = nullptr;
    hipEvent_t ev_user = nullptr;  // orders the plan stream after a caller's stream
    hipEvent_t ev_done = nullptr;  // orders a caller's stream after the plan stream
    // branch streams of the nonlinear step: its six synthesis launches and
    // its four analysis launches are mutually independent and are dealt
    // over nbranch streams (the plan stream and nbranch - 1 of these)
    static constexpr int MAXBR = 4;
    int nbranch = 3;
    hipStream_t br[MAXBR - 1] = {};
    hipEvent_t ev_fork = nullptr, ev_join[MAXBR - 1] = {};
    hipEvent_t ev_pin = nullptr;   // after the uploads from h_pin
    bool pin_pending = false;      // ev_pin recorded and not yet waited for
    bool last_host_out = false;    // the last nonlinear step copied outputs to host

    RlBuffers *rl = nullptr;       // buffers of mhx_rloop, made on its first call

    bool timing = false;
    static constexpr int NEV = 16;
    int n_launch = 0;                 // launches timed in the last step
    double launch_ms[NEV] = {};       // their durations
    double stage_ms[5] = {0, 0, 0, 0, 0};
    hipEvent_t ev[NEV] = {};
};

// ------------------------------------------------------------- launchers
int build_tables(Plan *p);
int build_split_tables(Plan *p);
int launch_nsyn16(const Plan *p, int kind, const NSynArgs &args, hipStream_t s);
int launch_get_td(const Plan *p, const TdArgs &args, hipStream_t s);
int launch_e_kin(const Plan *p, const EkArgs &args, hipStream_t s);
int launch_nana16(const Plan *p, int kind, const NAnaArgs &args, hipStream_t s);
int launch_syn(const Plan *p, const SynJob &job, hipStream_t s);
int launch_ana(const Plan *p, const AnaJob &job, hipStream_t s);
// Fourier data (complex, m < n_m): inverse side F[lev][comp][m][theta],
// forward side F2[lev][comp][theta][m]; grids in the reference layout (theta
// fastest).  F points at the first comp; F_ld_lev is the level stride in
// complex elements; grid comp c of level l starts at
// grid + l*grid_ld_lev + c*grid_ld_comp (doubles).
int launch_ifft(const Plan *p, const double *F, size_t F_ld_lev, int F_ncomp, int ncomp,
                int nlev, double *grid, size_t grid_ld_lev, size_t grid_ld_comp,
                hipStream_t s);
// forward: scale 1/n_phi, times 1/sin^2(theta) for comps flagged in osin2_mask
int launch_fft(const Plan *p, const double *grid, size_t grid_ld_lev, size_t grid_ld_comp,
               int ncomp, int nlev, unsigned osin2_mask, double *F, size_t F_ld_lev,
               hipStream_t s);
// fused grid stage of the nonlinear step: Fin (18 comps per level) -> Fout
// (10 comps per level); du/dr grids to dgrid [lev][3][theta][phi]; all 18
// inverse grids in the reference layout to full_grids when not null
// (FP32 Fourier data and grids for a split-precision plan; there Fout is
// theta-blocked, F2[lev][comp][theta/16][m][theta%16])
int launch_ifft_f32(const Plan *p, const float *F, size_t F_ld, int ncomp, int nrow, float *grid,
                    size_t g_ld_row, size_t g_ld_comp, int TB, hipStream_t s);
int launch_fft_f32(const Plan *p, const float *grid, size_t g_ld_row, size_t g_ld_comp, int ncomp,
                   int nrow, unsigned osin2_mask, float *F2, size_t F2_ld, int TB, hipStream_t s);
int launch_nl_fused(const Plan *p, int nlev, const void *Fin, void *Fout,
                    const double *rad, double LFfac, void *dgrid, void *full_grids,
                    hipStream_t s);

// ------------------------------------------ persistent launch of a batch
// One launch runs every job of an uneven batch of round-trip jobs as a work
// list whose items the workgroups claim in order from a ticket counter:
// synthesis items (orders x level groups of 8 rows), then row items (a
// level group's theta rows through the inverse and the forward FFT in LDS,
// no grid in HBM), then analysis items.  A row item waits until every
// synthesis unit of its level group is done, an analysis item until every
// row item of its level group is done; waits are only on items of lower
// tickets, claimed by running workgroups, so they end.
enum PBKind { PB_SYN = 0, PB_ROW = 1, PB_ANA = 2 };
constexpr int PB_MAXFAC = 12;
struct PBJob {
    NSynArgs syn;              // whole-K synthesis, filled by fill_pbatch_job
    NAnaArgs ana;              // whole-theta analysis, filled by fill_pbatch_job
    int syn_kind, ana_kind;    // NS_S / NS_QST, NA_SH / NA_QST
    int nrl;                   // 128-wide rounds of l holding l <= l_max
    int M, N, nfac, fac[PB_MAXFAC];
    const float *tw, *twr;     // FP32 exp(-2 pi i k / M), exp(-2 pi i k / N)
    const double *osin2;       // 1/sin^2(theta) of the plan's rows
    unsigned osin2_mask;       // comps scaled by it in the forward FFT
    const float *F;            // F[row][comp][m][theta] (synthesis output)
    float *F2;                 // F2[row][comp][theta/16][m][theta%16]
    size_t f_ld, f2_ld;        // complex elements per row
    int nc, n_m, n_theta, nt16, rows, tbr, ntbr;
    int c_syn, c_row;          // first done counter of the job's level groups
};
// item: kind, job; PB_SYN/PB_ANA: orders m0 (and m1 >= 0), level groups
// g0 .. g0 + ng - 1; PB_ROW: level group g0, theta block tb
struct PBItem {
    short kind, job, m0, m1, g0, ng, tb, pad;
};
// The work list is split into one queue per XCD (accelerator complex die,
// with its own L2): a level group's synthesis, row and analysis items all lie
// in one queue, claimed only by workgroups running on that XCD, so a
// producer's stores reach the consumer through the shared L2 once they are
// complete.  The workgroup's XCD is read from its hardware register (XCC_ID)
// and mapped to a queue by `xmap` (measured when the batch is made).
constexpr int PB_MAXQ = 16;
struct PBQueues {
    int nq;                    // queues
    int off[PB_MAXQ + 1];      // items of queue q: off[q] .. off[q + 1] - 1
    int xmap[PB_MAXQ];         // XCC_ID -> queue (-1: none)
};
// counters: [q] ticket of queue q, [PB_MAXQ] workgroups done, [PB_MAXQ + 1]
// error (a wait ran out, or a queue was left unclaimed; left set), then the
// jobs' done counters; the last workgroup out resets all but the error flag
constexpr int PB_DONE = PB_MAXQ, PB_ERR = PB_MAXQ + 1, PB_CTR0 = PB_MAXQ + 2;
bool pbatch_supported(const Plan *p);
size_t pbatch_row_lds(int M, int tbr);
// theta rows per row item where the row FFTs run on wave slots (0: the
// generic LDS FFT, tbr chosen by the autotuner)
int pbatch_slot_rows(int M);
size_t pbatch_lds();
int fill_pbatch_job(const Plan *p, PBJob &j);
// XCC_ID of each of n workgroups (device array ids), for the queue map
int probe_xcc_ids(int *ids, int n, hipStream_t s);
int launch_pbatch(const PBJob *jobs, const PBItem *items, const PBQueues &q, int *ctr, int nctr,
                  int grid, hipStream_t s);
int launch_axi(const Plan *p, const double *fl, double *f, int lcut, int tor,
               double *f2, hipStream_t s);
void rl_free(Plan *p);
int launch_rl_grid(const Plan *p, const RlGridArgs &a, hipStream_t s);
// FP64 tables (mhx_tables.hip): table_layout fills the offsets, build_tables
// also allocates and computes the tables; ensure_fp64_tables builds them on
// first use where plan creation left them out (recurrence-table plans);
// split_rowscales_rec computes the split tables' row factors from the
// recurrence, without FP64 tables
int table_layout(Plan *p);
int ensure_fp64_tables(Plan *p);
int split_rowscales_rec(Plan *p, float *syn_rs, float *ana_rs, const size_t *ana_soff);
// the split step's grid stage on theta-row tiles (mhx_grid.hip): n_phi = 768
bool grid_tile_supported(const Plan *p);
int launch_grid_tile(const Plan *p, int nlev, const void *Fin, void *Fout, const double *rad,
                     double LFfac, void *dgrid, void *full_grids, hipStream_t s);

}  // namespace mhx

struct mhx_plan {
    mhx::Plan p;
};
