// The radial loop of a time-step stage for any physics switches, interior
// levels: the body of rIter_single_t%radialLoop (src/rIter.f90:188-444) for
// nBc = 0, batched over all levels of the call, FP64 on the GPU:
//
//   synthesis   transform_to_grid_space (:483-620): u = torpol(w, dw, z);
//               with l_adv_curl curl u = torpol_to_curl(w, ddw, z, dz), else
//               du/dr = torpol(dw, ddw, dz), cvr = pol_to_curlr(z),
//               grad(l(l+1) w) = pol_to_grad(w), d/dphi = torpol_to_dphspat
//               (dw, z); s, xi, phi scalars; B = torpol(b, db, aj),
//               curl B = torpol_to_curl(b, ddb, aj, dj)
//   grid stage  get_nl (src/get_nl.f90:240-436) merged as
//               transform_to_lm_space (:644-680), Courant maxima
//               (src/courant.f90:209-275) - k_rl_grid, mhx_fft.hip
//   analysis    spat_to_qst of Adv, V s, V xi, V x B; scal_to_SH of the
//               heating terms and of the phase-field terms (:682-707)
//   lm space    nonlinear_lm_t (src/get_td.f90:133-589) - k_rl_td below
//
// One launch per stage for all levels; the synthesis and analysis launches
// are the plan's FP64 Legendre kernels (mhx_legendre.hip).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "mhx_internal.h"

namespace mhx {

namespace {

// the nonlinear_lm_t arrays of the analysis (complex, [lev][lm])
enum RlNl { NL_ADVR = 0, NL_ADVT, NL_ADVP, NL_VST, NL_VSP, NL_VXIT, NL_VXIP, NL_VXBR, NL_VXBT,
            NL_VXBP, NL_HEAT, NL_N };

// per-level factors of the lm-space assembly, RT_NRAD arrays of nlev
enum RlTdRad { TR_OR1 = 0, TR_OR2, TR_OR4, TR_ORHO1, TR_BETA, TR_R, TR_EPSCPROF, TR_TEMP0, TR_NRAD };

struct RlTdArgs {
    const double *nl[NL_N];       // analysis outputs (may be null when switched off)
    const double *w, *dw, *ddw, *z, *dz;
    double *dwdt, *dzdt, *dpdt, *dsdt, *dxidt, *dbdt, *djdt, *dVxBhLM, *dVxVhLM;
    size_t ld_nl, ld_f, ld_o;     // complex elements per level
    const double *rad;            // TR_NRAD x nlev
    const int *lcut;              // l_R per level
    int conv_nl, corr, single_matrix, dcurl, dpdt_on, heat, anel, anel_liq, chem, mag;
    int mag_dbdt;                 // l_mag_nl or l_mag_kin
    double CorFac, epsc, epscXi;
    int nlev, lm_max, l_max;
    const int *lm2l, *lm2m;
};

__device__ inline double2 cr(double a, double2 z) { return make_double2(a * z.x, a * z.y); }
__device__ inline double2 ca(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 cs(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
// dPhi z = i m z (src/horizontal.f90:208)
__device__ inline double2 cim(double m, double2 z) { return make_double2(-m * z.y, m * z.x); }

__device__ inline double clm_(int l, int m)
{
    if (l <= m) return 0.0;
    const double dl = (double)l, dm = (double)m;
    return sqrt((dl + dm) * (dl - dm) / ((2.0 * dl - 1.0) * (2.0 * dl + 1.0)));
}

// one thread per (lm, level); the l -/+ 1 neighbours (lm2lmS, lm2lmA of
// src/blocking.f90:322-334) are the adjacent entries of the m-major rows, and
// the operators dTheta2-4 A/S those of src/horizontal.f90:209-229
__global__ __launch_bounds__(256) void k_rl_td(RlTdArgs a)
{
    const int lm = blockIdx.x * blockDim.x + threadIdx.x, lev = blockIdx.y;
    if (lm >= a.lm_max) return;
    const int l = a.lm2l[lm], m = a.lm2m[lm], L = a.l_max, nlev = a.nlev;
    const double dm = (double)m, dl = (double)l, dLh = dl * (dl + 1.0);
    const double or1 = a.rad[TR_OR1 * nlev + lev], or2 = a.rad[TR_OR2 * nlev + lev];
    const double or4 = a.rad[TR_OR4 * nlev + lev], orho1 = a.rad[TR_ORHO1 * nlev + lev];
    const double beta = a.rad[TR_BETA * nlev + lev], r = a.rad[TR_R * nlev + lev];
    const int lR = a.lcut[lev];
    const double c2 = 2.0 * a.CorFac;
    const double cS = clm_(l, m), cA = clm_(l + 1, m);
    // dTheta2S/A, dTheta3S/A, dTheta4S/A at this lm
    const double t2S = (dl - 1.0) * cS, t2A = (dl + 2.0) * cA;
    const double t3S = (dl - 1.0) * (dl + 1.0) * cS, t3A = dl * (dl + 2.0) * cA;
    const double t4S = (dl + 1.0) * cS * ((dl - 1.0) * dl), t4A = dl * cA * ((dl + 1.0) * (dl + 2.0));
    const size_t fi = (size_t)lev * a.ld_f + lm, ni = (size_t)lev * a.ld_nl + lm;
    const size_t oi = (size_t)lev * a.ld_o + lm;
    const bool hasA = l < L;
    const size_t fS = l > m ? fi - 1 : fi, fA = hasA ? fi + 1 : fi;
    const double2 zero = make_double2(0.0, 0.0);
    auto F = [&](const double *p, size_t i) { return reinterpret_cast<const double2 *>(p)[i]; };
    auto NL = [&](int k) { return a.nl[k] ? reinterpret_cast<const double2 *>(a.nl[k])[ni] : zero; };
    auto put = [&](double *p, double2 v) {
        if (p) reinterpret_cast<double2 *>(p)[oi] = v;
    };
    const double2 wl = F(a.w, fi), dwl = F(a.dw, fi), zl = F(a.z, fi);
    const double2 wS = F(a.w, fS), dwS = F(a.dw, fS), zS = F(a.z, fS);
    const double2 wA = hasA ? F(a.w, fA) : zero, dwA = hasA ? F(a.dw, fA) : zero;
    const double2 zA = hasA ? F(a.z, fA) : zero;
    const bool below = l < lR, at = l == lR;

    // dwdt: get_dwdt (:133-203) or get_dwdt_double_curl (:210-309)
    double2 dwdt, dvxvh = zero;
    if (l == 0) {
        dwdt = a.conv_nl ? cr(or2, NL(NL_ADVR)) : zero;
        if (a.corr && !a.single_matrix) dwdt = ca(dwdt, cr(c2 * or1 * t2A, zA));
    } else if (!a.dcurl) {
        dwdt = a.conv_nl ? cr(or2, NL(NL_ADVR)) : zero;
        if (a.corr && (below || at)) {
            double2 c = cs(cim(dm, dwl), cr(t2S, zS));
            if (below) c = ca(c, cr(t2A, zA));
            dwdt = ca(dwdt, cr(c2 * or1, c));
        }
    } else {
        dwdt = a.conv_nl ? cr(dLh * or4 * orho1, NL(NL_ADVR)) : zero;
        dvxvh = a.conv_nl ? cr(-orho1 * r * r * dLh, NL(NL_ADVT)) : zero;
        if (a.corr && (below || at)) {
            const double2 ddwl = F(a.ddw, fi), dzS = F(a.dz, fS);
            // dPhi (-ddw + beta dw + (beta or1 + or2) dLh w)
            double2 c = cim(dm, ca(ca(cr(-1.0, ddwl), cr(beta, dwl)), cr((beta * or1 + or2) * dLh, wl)));
            if (below) {
                const double2 dzA = hasA ? F(a.dz, fA) : zero;
                c = ca(c, cr(t3A, cs(dzA, cr(beta, zA))));
                c = ca(c, cr(t3S, cs(dzS, cr(beta, zS))));
                c = ca(c, cr(or1, cs(cr(t4A, zA), cr(t4S, zS))));
            } else {
                c = ca(c, cr(t3S, cs(dzS, cr(beta, zS))));
                c = cs(c, cr(or1 * t4S, zS));
            }
            dwdt = ca(dwdt, cr(c2 * or2 * orho1, c));
        }
    }
    put(a.dwdt, dwdt);
    if (a.dcurl) put(a.dVxVhLM, dvxvh);

    // dzdt: get_dzdt (:379-444)
    double2 dzdt;
    if (l == 0) {
        dzdt = a.corr ? cr(c2 * or2, ca(cr(t3A, dwA), cr(or1 * t4A, wA))) : zero;
    } else {
        dzdt = a.conv_nl ? cr(dLh, NL(NL_ADVP)) : zero;
        if (a.corr && (below || at)) {
            double2 c = ca(cim(dm, zl), cr(t3S, dwS));
            c = cs(c, cr(or1 * t4S, wS));
            if (below) c = ca(c, ca(cr(t3A, dwA), cr(or1 * t4A, wA)));
            dzdt = ca(dzdt, cr(c2 * or2, c));
        }
    }
    put(a.dzdt, dzdt);

    // dpdt: get_dpdt (:311-377); l = 0 is not set by the reference (zero)
    if (a.dpdt_on) {
        double2 dpdt = zero;
        if (l > 0) {
            dpdt = a.conv_nl ? cr(-dLh, NL(NL_ADVT)) : zero;
            if (a.corr && (below || at)) {
                double2 c = cr(-1.0, cim(dm, ca(dwl, cr(or1 * dLh, wl))));
                c = ca(c, cr(t3S, zS));
                if (below) c = ca(c, cr(t3A, zA));
                dpdt = ca(dpdt, cr(c2 * or2, c));
            }
        }
        put(a.dpdt, dpdt);
    }

    // dsdt: get_dsdt (:459-521); dVSrLM is the analysis's own output
    if (a.heat) {
        double2 dsdt;
        const double tf = a.anel_liq ? a.rad[TR_TEMP0 * nlev + lev] : 1.0;
        if (l == 0) {
            dsdt = make_double2(a.epsc * a.rad[TR_EPSCPROF * nlev + lev], 0.0);
            if (a.anel) dsdt = ca(dsdt, a.anel_liq ? cr(tf, NL(NL_HEAT)) : NL(NL_HEAT));
        } else {
            dsdt = cr(dLh, NL(NL_VST));
            if (a.anel) dsdt = ca(dsdt, a.anel_liq ? cr(tf, NL(NL_HEAT)) : NL(NL_HEAT));
        }
        put(a.dsdt, dsdt);
    }

    // dxidt: get_dxidt (:523-556)
    if (a.chem) put(a.dxidt, l == 0 ? make_double2(a.epscXi, 0.0) : cr(dLh, NL(NL_VXIT)));

    // dbdt, djdt, dVxBhLM: get_dbdt (:558-589)
    if (a.mag) {
        if (a.mag_dbdt) {
            put(a.dbdt, cr(dLh, NL(NL_VXBP)));
            put(a.dVxBhLM, cr(-dLh * r * r, NL(NL_VXBT)));
            put(a.djdt, cr(dLh * or4, NL(NL_VXBR)));
        } else {
            put(a.dbdt, zero);
            put(a.djdt, zero);
            put(a.dVxBhLM, zero);
        }
    }
}

// dtrkc = min(1e10, sqrt(delxr2 / vr2max)) where vr2max != 0, likewise dthkc
// (src/rIter.f90:291 and src/courant.f90:270-274)
__global__ void k_rl_courant(const unsigned long long *__restrict__ cmax, const double *__restrict__ dx,
                             int nlev, double *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nlev) return;
    const double vr2 = __longlong_as_double((long long)cmax[2 * i]);
    const double vh2 = __longlong_as_double((long long)cmax[2 * i + 1]);
    double dtr = 1e10, dth = 1e10;
    if (vr2 != 0.0) dtr = fmin(dtr, sqrt(dx[i] / vr2));
    if (vh2 != 0.0) dth = fmin(dth, sqrt(dx[nlev + i] / vh2));
    out[i] = dtr;
    out[nlev + i] = dth;
}

__global__ void k_rl_zero(unsigned long long *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

template <typename T> void free_dev_(T *&p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

int grow(double *&buf, size_t &have, size_t need_doubles)
{
    if (have >= need_doubles) return MHX_OK;
    free_dev_(buf);
    have = 0;
    if (hipMalloc(&buf, need_doubles * sizeof(double)) != hipSuccess) {
        (void)hipGetLastError();
        return fail(MHX_E_NOMEM, "radial loop: device allocation of " +
                                     std::to_string(need_doubles * sizeof(double)) + " bytes failed");
    }
    have = need_doubles;
    return MHX_OK;
}

ColOp rop(int s0, int g = G_ONE, int ipow = 0, int mpow = 0, double scale = 1.0, int s1 = -1,
          double h = 0.0)
{
    ColOp c;
    c.src0 = s0;
    c.src1 = s1;
    c.gkind = g;
    c.ipow = ipow;
    c.mpow = mpow;
    c.hcoef = h;
    c.scale = scale;
    return c;
}

// columns Q, S, i m T, i m S, -T of a QST synthesis; S, i m T, i m S, -T of a
// sph-tor synthesis (the forms of the per-call entries)
void rl_qst_cols(SynJob &j, ColOp Q, ColOp S, ColOp T, bool with_q)
{
    ColOp miT = T;
    miT.ipow = (miT.ipow + 1) & 3;
    miT.mpow += 1;
    ColOp miS = S;
    miS.ipow = (miS.ipow + 1) & 3;
    miS.mpow += 1;
    ColOp mT = T;
    mT.scale = -mT.scale;
    int k = 0;
    if (with_q) j.col[k++] = Q;
    j.col[k++] = S;
    j.col[k++] = miT;
    j.col[k++] = miS;
    j.col[k++] = mT;
}

}  // namespace

// device buffers of the radial loop (per plan)
struct RlBuffers {
    double *F = nullptr, *F2 = nullptr, *nl = nullptr, *stage = nullptr, *rad = nullptr;
    size_t nF = 0, nF2 = 0, nnl = 0, nstage = 0, nrad = 0;
    int *lcut = nullptr;
    int nlcut = 0;
    double *th = nullptr;
    unsigned long long *cmax = nullptr;
    int ncmax = 0;
};

void rl_free(Plan *p)
{
    if (!p->rl) return;
    RlBuffers *b = p->rl;
    free_dev_(b->F);
    free_dev_(b->F2);
    free_dev_(b->nl);
    free_dev_(b->stage);
    free_dev_(b->rad);
    free_dev_(b->th);
    if (b->lcut) (void)hipFree(b->lcut);
    if (b->cmax) (void)hipFree(b->cmax);
    delete b;
    p->rl = nullptr;
}

// The switch sets the fused step covers (src/rIter.f90:483-707 with curl
// advection, heat, the full MHD terms and no double curl, composition, phase
// field, precession or centrifugal term): the fused step with its Courant
// maxima (mhx_nl_step_courant), then its lm-space assembly (mhx_get_td), in
// the plan's precision, as magic_amd/rloop.py composes them; returns when
// done.  The Courant steps of src/courant.f90:271-272 (1e10 where a level has
// no flow) from the maxima on the host.
bool rl_fused_covers(const mhx_rloop_phys *ph)
{
    return (ph->l_conv || ph->l_mag_kin) && ph->l_conv_nl && ph->l_adv_curl && !ph->l_anel &&
           !ph->l_anelastic_liquid && ph->l_heat && ph->l_heat_nl && ph->l_mag && ph->l_mag_LF &&
           ph->l_mag_nl && !ph->l_mag_kin && !ph->l_chemical_conv && !ph->l_phase_field &&
           !ph->l_precession && !ph->l_centrifuge && !ph->l_double_curl && !ph->l_single_matrix &&
           ph->l_cour_alf_damp && ph->n_r_LCR == 0;
}

int rl_fused(mhx_plan *h, const mhx_rloop_phys *ph, int nlev, const mhx_rloop_radial *rad,
             const mhx_rloop_fields *in, size_t ld_lm, const mhx_rloop_out *out, void *stream)
{
    Plan *P = &h->p;
    if (!P->rl) P->rl = new RlBuffers();
    RlBuffers *B = P->rl;
    const size_t arr = (size_t)nlev * ld_lm * 2;
    int rc;
    if ((rc = grow(B->nl, B->nnl, 9 * arr))) return rc;
    // the step's forward outputs: VSrLM is the caller's dVSrLM, the rest scratch
    double *nl[10];
    for (int k = 0, j = 0; k < 10; ++k) nl[k] = k == 3 ? out->dVSrLM : B->nl + (size_t)(j++) * arr;
    const double *fin[11] = {in->w, in->dw, in->ddw, in->z, in->dz, in->s,
                             in->b, in->db, in->ddb, in->aj, in->dj};
    // valri2, valhi2 (src/courant.f90:177-183), epsc epscProf (src/get_td.f90:480)
    const double damp = (0.5 * (1.0 + ph->opm)) * (0.5 * (1.0 + ph->opm));
    std::vector<double> vri(nlev), vhi(nlev), eps(nlev), vr2(nlev), vh2(nlev), dt(2 * (size_t)nlev);
    for (int i = 0; i < nlev; ++i) {
        vri[i] = damp / rad->delxr2[i];
        vhi[i] = damp / rad->delxh2[i];
        eps[i] = ph->epsc * rad->epscProf[i];
    }
    // the step's du/dr grids are not among the loop's outputs
    const bool keep = P->dudr;
    P->dudr = false;
    rc = mhx_nl_step_courant(h, nlev, fin, ld_lm, nl, rad->or2, rad->or4, rad->orho1, ph->LFfac, rad->l_R,
                             ph->courfac, ph->alffac, vri.data(), vhi.data(), vr2.data(), vh2.data(), stream);
    P->dudr = keep;
    if (rc) return rc;
    // AdvrLM AdvtLM AdvpLM VStLM VxBrLM VxBtLM VxBpLM, then w dw z
    const double *tin[10] = {nl[0], nl[1], nl[2], nl[4], nl[6], nl[7], nl[8], in->w, in->dw, in->z};
    double *tout[7] = {out->dwdt, out->dzdt, out->dpdt, out->dsdt, out->dbdt, out->djdt, out->dVxBhLM};
    if ((rc = mhx_get_td(h, nlev, tin, ld_lm, tout, rad->or2, rad->or4, eps.data(),
                         ph->l_corr ? ph->CorFac : 0.0, rad->l_R, stream)))
        return rc;
    for (int i = 0; i < nlev; ++i) {
        dt[i] = vr2[i] != 0.0 ? std::min(1e10, std::sqrt(rad->delxr2[i] / vr2[i])) : 1e10;
        dt[nlev + i] = vh2[i] != 0.0 ? std::min(1e10, std::sqrt(rad->delxh2[i] / vh2[i])) : 1e10;
    }
    MHX_HIP_CHECK(hipMemcpyAsync(out->dtrkc, dt.data(), (size_t)nlev * sizeof(double), hipMemcpyDefault,
                                 P->stream));
    MHX_HIP_CHECK(hipMemcpyAsync(out->dthkc, dt.data() + nlev, (size_t)nlev * sizeof(double),
                                 hipMemcpyDefault, P->stream));
    MHX_HIP_CHECK(hipStreamSynchronize(P->stream));  // dt leaves scope
    return MHX_OK;
}

}  // namespace mhx

using namespace mhx;

int mhx_rloop(mhx_plan *h, const mhx_rloop_phys *ph, int nlev, const mhx_rloop_radial *rad,
              const mhx_rloop_fields *in, size_t ld_lm, const mhx_rloop_out *out, void *stream)
{
    if (!h) return fail(MHX_E_INVALID, "null plan");
    Plan *P = &h->p;
    if (nlev <= 0) return MHX_OK;
    if (!ph || !rad || !in || !out) return fail(MHX_E_INVALID, "null argument");
    if (ld_lm < (size_t)P->lm_max) return fail(MHX_E_INVALID, "ld_lm below lm_max");
    const bool conv = ph->l_conv || ph->l_mag_kin, heat = ph->l_heat, chem = ph->l_chemical_conv;
    const bool phase = ph->l_phase_field, mag = ph->l_mag || ph->l_mag_LF;
    const bool curl = ph->l_adv_curl, anel = ph->l_anel && heat;
    const bool dcurl = ph->l_double_curl, dpdt_on = !dcurl || ph->lPressNext;
    // the grid components this stage needs (src/rIter.f90:483-620)
    const bool need_grad = !curl || anel;
    if (!rad->r || !rad->or1 || !rad->or2 || !rad->or4 || !rad->orho1 || !rad->beta || !rad->otemp1 ||
        !rad->visc || !rad->lambda || !rad->temp0 || !rad->epscProf || !rad->delxr2 || !rad->delxh2 ||
        !rad->nR || !rad->l_R)
        return fail(MHX_E_INVALID, "null radial array");
    if (!in->w || !in->dw || !in->ddw || !in->z || !in->dz) return fail(MHX_E_INVALID, "null velocity input");
    if (heat && !in->s) return fail(MHX_E_INVALID, "l_heat without s");
    if (chem && !in->xi) return fail(MHX_E_INVALID, "l_chemical_conv without xi");
    if (phase && !in->phi) return fail(MHX_E_INVALID, "l_phase_field without phi");
    if (mag && (!in->b || !in->db || !in->ddb || !in->aj || !in->dj))
        return fail(MHX_E_INVALID, "l_mag without b, db, ddb, aj, dj");
    if (!out->dwdt || !out->dzdt || (dpdt_on && !out->dpdt) || !out->dtrkc || !out->dthkc)
        return fail(MHX_E_INVALID, "null output array");
    if (heat && (!out->dsdt || !out->dVSrLM)) return fail(MHX_E_INVALID, "l_heat without dsdt, dVSrLM");
    if (chem && (!out->dxidt || !out->dVXirLM)) return fail(MHX_E_INVALID, "l_chemical_conv without dxidt, dVXirLM");
    if (phase && !out->dphidt) return fail(MHX_E_INVALID, "l_phase_field without dphidt");
    if (ph->l_mag && (!out->dbdt || !out->djdt || !out->dVxBhLM))
        return fail(MHX_E_INVALID, "l_mag without dbdt, djdt, dVxBhLM");
    if (dcurl && !out->dVxVhLM) return fail(MHX_E_INVALID, "l_double_curl without dVxVhLM");
    for (int i = 0; i < nlev; ++i)
        if (rad->l_R[i] < 0 || rad->l_R[i] > P->l_max) return fail(MHX_E_INVALID, "l_R out of range");
    if (rl_fused_covers(ph)) return rl_fused(h, ph, nlev, rad, in, ld_lm, out, stream);

    int rc;
    if (!P->rl) P->rl = new RlBuffers();
    RlBuffers *B = P->rl;
    hipStream_t s = P->stream;
    if (stream) {
        MHX_HIP_CHECK(hipEventRecord(P->ev_user, (hipStream_t)stream));
        MHX_HIP_CHECK(hipStreamWaitEvent(s, P->ev_user, 0));
    }
    // previous call's host staging and uploads must be done before refilling
    MHX_HIP_CHECK(hipStreamSynchronize(s));

    // ---- component lists (compact indices) ----
    int in_of[RI_N], out_of[RO_N];
    std::fill(in_of, in_of + RI_N, -1);
    std::fill(out_of, out_of + RO_N, -1);
    int nin = 0, nout = 0;
    auto addin = [&](int c) { in_of[c] = nin++; };
    if (conv) {
        addin(RI_VR); addin(RI_VT); addin(RI_VP);
        if (curl) {
            addin(RI_CVR); addin(RI_CVT); addin(RI_CVP);
        }
        if (need_grad) {
            if (!curl) addin(RI_CVR);
            addin(RI_DVRDR); addin(RI_DVTDR); addin(RI_DVPDR);
            addin(RI_DVRDT); addin(RI_DVRDP); addin(RI_DVTDP); addin(RI_DVPDP);
        }
    }
    if (heat) addin(RI_S);
    if (chem) addin(RI_XI);
    if (phase) addin(RI_PHI);
    if (mag) {
        addin(RI_BR); addin(RI_BT); addin(RI_BP);
        addin(RI_CBR); addin(RI_CBT); addin(RI_CBP);
    }
    unsigned osin2 = 0;
    auto addvec = [&](int c0) {
        out_of[c0] = nout++;
        out_of[c0 + 1] = nout++;
        out_of[c0 + 2] = nout++;
        osin2 |= 3u << (nout - 2);
    };
    const bool adv = ph->l_conv_nl || ph->l_mag_LF;
    if (adv) addvec(RO_ADVR);
    if (heat) addvec(RO_VSR);
    if (chem) addvec(RO_VXIR);
    if (ph->l_mag_nl) addvec(RO_VXBR);
    if (anel) out_of[RO_HEAT] = nout++;
    if (phase) out_of[RO_PHI] = nout++;
    if (nin == 0 || nout == 0) return fail(MHX_E_UNSUPPORTED, "radial loop: nothing to transform");

    // ---- buffers ----
    const size_t four = (size_t)P->n_m * P->n_theta * 2;  // doubles per Fourier comp
    const size_t lmd = (size_t)nlev * ld_lm * 2;           // doubles per nl array
    if ((rc = grow(B->F, B->nF, (size_t)nlev * nin * four)) ||
        (rc = grow(B->F2, B->nF2, (size_t)nlev * nout * four)) ||
        (rc = grow(B->nl, B->nnl, (size_t)NL_N * lmd)) ||
        (rc = grow(B->rad, B->nrad, (size_t)(RL_NRAD + TR_NRAD + 2) * nlev)))
        return rc;
    if (B->nlcut < nlev) {
        if (B->lcut) (void)hipFree(B->lcut);
        B->lcut = nullptr;
        B->nlcut = 0;
        MHX_HIP_CHECK(hipMalloc(&B->lcut, (size_t)nlev * sizeof(int)));
        B->nlcut = nlev;
    }
    if (B->ncmax < nlev) {
        if (B->cmax) (void)hipFree(B->cmax);
        B->cmax = nullptr;
        B->ncmax = 0;
        MHX_HIP_CHECK(hipMalloc(&B->cmax, 2 * (size_t)nlev * sizeof(unsigned long long)));
        B->ncmax = nlev;
    }
    if (!B->th) {
        // per scrambled theta row: 1/sin^2, cos/sin^2, sin, cos, 1/sin
        // (src/horizontal.f90:141-157)
        std::vector<double> t((size_t)RL_NTH * P->n_theta);
        for (int k = 0; k < P->nh; ++k) {
            const double c = P->theta_ord[k], sn = std::sin(c), co = std::cos(c);
            for (int hs = 0; hs < 2; ++hs) {
                const int row = 2 * k + hs;
                const double sg = hs ? -1.0 : 1.0;
                t[RT_OSIN2 * P->n_theta + row] = 1.0 / (sn * sn);
                t[RT_COSN2 * P->n_theta + row] = sg * co / sn / sn;
                t[RT_SIN * P->n_theta + row] = sn;
                t[RT_COS * P->n_theta + row] = sg * co;
                t[RT_OSIN * P->n_theta + row] = 1.0 / sn;
            }
        }
        MHX_HIP_CHECK(hipMalloc(&B->th, t.size() * sizeof(double)));
        MHX_HIP_CHECK(hipMemcpy(B->th, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    // per-level factors: grid stage, lm assembly, Courant lengths
    {
        std::vector<double> hr((size_t)(RL_NRAD + TR_NRAD + 2) * nlev);
        const double damp = ph->l_cour_alf_damp ? (0.5 * (1.0 + ph->opm)) * (0.5 * (1.0 + ph->opm)) : 0.0;
        for (int i = 0; i < nlev; ++i) {
            double *g = hr.data();
            g[RR_R * nlev + i] = rad->r[i];
            g[RR_OR1 * nlev + i] = rad->or1[i];
            g[RR_OR2 * nlev + i] = rad->or2[i];
            g[RR_OR4 * nlev + i] = rad->or4[i];
            g[RR_ORHO1 * nlev + i] = rad->orho1[i];
            g[RR_BETA * nlev + i] = rad->beta[i];
            g[RR_OTEMP1 * nlev + i] = rad->otemp1[i];
            g[RR_VISC * nlev + i] = rad->visc[i];
            g[RR_LAMBDA * nlev + i] = rad->lambda[i];
            g[RR_LCR * nlev + i] = rad->nR[i] <= ph->n_r_LCR ? 1.0 : 0.0;
            // valri2, valhi2 (src/courant.f90:177-183)
            g[RR_VALRI2 * nlev + i] = ph->l_cour_alf_damp ? damp / rad->delxr2[i] : 0.0;
            g[RR_VALHI2 * nlev + i] = ph->l_cour_alf_damp ? damp / rad->delxh2[i] : 0.0;
            double *t = g + (size_t)RL_NRAD * nlev;
            t[TR_OR1 * nlev + i] = rad->or1[i];
            t[TR_OR2 * nlev + i] = rad->or2[i];
            t[TR_OR4 * nlev + i] = rad->or4[i];
            t[TR_ORHO1 * nlev + i] = rad->orho1[i];
            t[TR_BETA * nlev + i] = rad->beta[i];
            t[TR_R * nlev + i] = rad->r[i];
            t[TR_EPSCPROF * nlev + i] = rad->epscProf[i];
            t[TR_TEMP0 * nlev + i] = rad->temp0[i];
            double *dx = t + (size_t)TR_NRAD * nlev;
            dx[i] = rad->delxr2[i];
            dx[nlev + i] = rad->delxh2[i];
        }
        MHX_HIP_CHECK(hipMemcpyAsync(B->rad, hr.data(), hr.size() * sizeof(double), hipMemcpyHostToDevice, s));
        MHX_HIP_CHECK(hipMemcpyAsync(B->lcut, rad->l_R, (size_t)nlev * sizeof(int), hipMemcpyHostToDevice, s));
        MHX_HIP_CHECK(hipStreamSynchronize(s));  // host vectors leave scope
    }
    const double *d_or2 = B->rad + (size_t)RR_OR2 * nlev;
    const double *d_trad = B->rad + (size_t)RL_NRAD * nlev;
    const double *d_dx = d_trad + (size_t)TR_NRAD * nlev;

    // ---- host arrays staged (inputs in, outputs back after the launches) ----
    const size_t arr = (size_t)nlev * ld_lm * 2;
    const double *fin[13] = {in->w, in->dw, in->ddw, in->z, in->dz, heat ? in->s : nullptr,
                             chem ? in->xi : nullptr, phase ? in->phi : nullptr,
                             mag ? in->b : nullptr, mag ? in->db : nullptr, mag ? in->ddb : nullptr,
                             mag ? in->aj : nullptr, mag ? in->dj : nullptr};
    double *fout[12] = {out->dwdt, out->dzdt, dpdt_on ? out->dpdt : nullptr, heat ? out->dsdt : nullptr,
                        heat ? out->dVSrLM : nullptr, chem ? out->dxidt : nullptr,
                        chem ? out->dVXirLM : nullptr, phase ? out->dphidt : nullptr,
                        ph->l_mag ? out->dbdt : nullptr, ph->l_mag ? out->djdt : nullptr,
                        ph->l_mag ? out->dVxBhLM : nullptr, dcurl ? out->dVxVhLM : nullptr};
    bool hin[13] = {}, hout[12] = {};
    int nh = 0;
    for (int i = 0; i < 13; ++i)
        if (fin[i] && !is_device_ptr(fin[i])) hin[i] = true, ++nh;
    for (int i = 0; i < 12; ++i)
        if (fout[i] && !is_device_ptr(fout[i])) hout[i] = true, ++nh;
    const bool host_cour = !is_device_ptr(out->dtrkc) || !is_device_ptr(out->dthkc);
    if ((rc = grow(B->stage, B->nstage, (size_t)nh * arr + 2 * (size_t)nlev))) return rc;
    const double *din[13];
    double *dout[12];
    {
        int slot = 0;
        for (int i = 0; i < 13; ++i) {
            din[i] = fin[i];
            if (!hin[i]) continue;
            double *d = B->stage + (size_t)slot++ * arr;
            MHX_HIP_CHECK(hipMemcpyAsync(d, fin[i], arr * sizeof(double), hipMemcpyHostToDevice, s));
            din[i] = d;
        }
        for (int i = 0; i < 12; ++i) dout[i] = hout[i] ? B->stage + (size_t)slot++ * arr : fout[i];
    }
    double *dcour = B->stage + (size_t)nh * arr;  // dtrkc, dthkc

    // ---- synthesis ----
    enum { W = 0, DW, DDW, Z, DZ, S_, XI, PHI, BB, DB, DDB, AJ, DJ };
    auto syn = [&](int kind, std::initializer_list<int> srcs, int c0, int c1, int c2) -> SynJob {
        SynJob j;
        std::memset(&j, 0, sizeof(j));
        j.kind = kind;
        j.nlev = nlev;
        j.src_ld = ld_lm;
        int k = 0;
        for (int sidx : srcs) j.src[k++] = din[sidx];
        j.fout = B->F;
        j.fout_ld = (size_t)nin * P->n_m * P->n_theta;
        j.out_comp[0] = c0;
        j.out_comp[1] = c1;
        j.out_comp[2] = c2;
        j.lcut = B->lcut;
        j.or2 = d_or2;
        return j;
    };
    // independent launches (the syntheses, then the analyses) dealt over
    // the plan's branch streams, forked from and joined back into s
    int nl_launch = 0;
    auto on = [&]() -> hipStream_t {
        const int i = nl_launch++ % P->nbranch;
        return i == 0 ? s : P->br[i - 1];
    };
    auto fork = [&]() -> int {
        nl_launch = 0;
        if (P->nbranch < 2) return MHX_OK;
        MHX_HIP_CHECK(hipEventRecord(P->ev_fork, s));
        for (int i = 0; i + 1 < P->nbranch; ++i) MHX_HIP_CHECK(hipStreamWaitEvent(P->br[i], P->ev_fork, 0));
        return MHX_OK;
    };
    auto join = [&]() -> int {
        for (int i = 0; i + 1 < P->nbranch; ++i) {
            MHX_HIP_CHECK(hipEventRecord(P->ev_join[i], P->br[i]));
            MHX_HIP_CHECK(hipStreamWaitEvent(s, P->ev_join[i], 0));
        }
        return MHX_OK;
    };
    if ((rc = fork())) return rc;
    if (conv) {
        // torpol_to_spat(w, dw, z): Q = l(l+1) w, S = dw, T = z
        SynJob j = syn(SYN_QST, {W, DW, Z}, in_of[RI_VR], in_of[RI_VT], in_of[RI_VP]);
        rl_qst_cols(j, rop(0, G_LL), rop(1), rop(2), true);
        if ((rc = launch_syn(P, j, on()))) return rc;
        if (curl) {
            // torpol_to_curl_spat(or2, w, ddw, z, dz): Q = l(l+1) z, S = dz,
            // T = or2 l(l+1) w - ddw
            SynJob c = syn(SYN_QST, {W, DDW, Z, DZ}, in_of[RI_CVR], in_of[RI_CVT], in_of[RI_CVP]);
            rl_qst_cols(c, rop(2, G_LL), rop(3), rop(0, G_OR2LL, 0, 0, 1.0, 1, -1.0), true);
            if ((rc = launch_syn(P, c, on()))) return rc;
        }
        if (need_grad) {
            if (!curl) {
                // pol_to_curlr_spat(z): l(l+1) z
                SynJob c = syn(SYN_SCAL, {Z}, in_of[RI_CVR], 0, 0);
                c.col[0] = rop(0, G_LL);
                if ((rc = launch_syn(P, c, on()))) return rc;
            }
            // torpol_to_spat(dw, ddw, dz)
            SynJob d = syn(SYN_QST, {DW, DDW, DZ}, in_of[RI_DVRDR], in_of[RI_DVTDR], in_of[RI_DVPDR]);
            rl_qst_cols(d, rop(0, G_LL), rop(1), rop(2), true);
            if ((rc = launch_syn(P, d, on()))) return rc;
            // pol_to_grad_spat(w): gradient of l(l+1) w
            SynJob g = syn(SYN_GRAD, {W}, in_of[RI_DVRDT], in_of[RI_DVRDP], 0);
            g.col[0] = rop(0, G_LL);
            g.col[1] = rop(0, G_LL, 1, 1);
            if ((rc = launch_syn(P, g, on()))) return rc;
            // torpol_to_dphspat(dw, z): S = i m dw, T = i m z, times 1/sin^2
            SynJob q = syn(SYN_SPHTOR, {DW, Z}, in_of[RI_DVTDP], in_of[RI_DVPDP], 0);
            rl_qst_cols(q, ColOp(), rop(0, G_ONE, 1, 1), rop(1, G_ONE, 1, 1), false);
            q.row_scale[0] = q.row_scale[1] = 1;
            if ((rc = launch_syn(P, q, on()))) return rc;
        }
    }
    auto scalar = [&](int src, int comp) -> int {
        SynJob j = syn(SYN_SCAL, {src}, comp, 0, 0);
        j.col[0] = rop(0);
        return launch_syn(P, j, on());
    };
    if (heat && (rc = scalar(S_, in_of[RI_S]))) return rc;
    if (chem && (rc = scalar(XI, in_of[RI_XI]))) return rc;
    if (phase && (rc = scalar(PHI, in_of[RI_PHI]))) return rc;
    if (mag) {
        SynJob b = syn(SYN_QST, {BB, DB, AJ}, in_of[RI_BR], in_of[RI_BT], in_of[RI_BP]);
        rl_qst_cols(b, rop(0, G_LL), rop(1), rop(2), true);
        if ((rc = launch_syn(P, b, on()))) return rc;
        SynJob c = syn(SYN_QST, {BB, DDB, AJ, DJ}, in_of[RI_CBR], in_of[RI_CBT], in_of[RI_CBP]);
        rl_qst_cols(c, rop(2, G_LL), rop(3), rop(0, G_OR2LL, 0, 0, 1.0, 1, -1.0), true);
        if ((rc = launch_syn(P, c, on()))) return rc;
    }

    if ((rc = join())) return rc;

    // ---- grid stage ----
    RlGridArgs g;
    std::memset(&g, 0, sizeof(g));
    g.Fin = B->F;
    g.Fout = B->F2;
    g.nin = nin;
    g.nout = nout;
    std::memcpy(g.in_of, in_of, sizeof(in_of));
    std::memcpy(g.out_of, out_of, sizeof(out_of));
    g.out_osin2 = osin2;
    g.n_m = P->n_m;
    g.n_theta = P->n_theta;
    g.nlev = nlev;
    g.dphi = 2.0 * 3.14159265358979323846 / ((double)P->n_phi * (double)P->minc);
    g.rad = B->rad;
    g.th = B->th;
    RlPhys &q = g.ph;
    q.conv_nl = ph->l_conv_nl;
    q.adv_curl = curl;
    q.heat_nl = heat && ph->l_heat_nl;
    q.chem = chem;
    q.phase = phase;
    q.prec = ph->l_precession;
    q.centr = ph->l_centrifuge;
    q.mag_LF = ph->l_mag_LF;
    q.mag_nl = ph->l_mag_nl;
    q.anel = anel;
    q.cour_mag = ph->l_mag && ph->l_mag_LF && !ph->l_mag_kin;
    q.LFfac = ph->LFfac;
    q.ViscHeatFac = ph->ViscHeatFac;
    q.OhmLossFac = ph->OhmLossFac;
    q.epsPhase = ph->epsPhase;
    q.penaltyFac = ph->penaltyFac;
    q.phaseDiffFac = ph->phaseDiffFac;
    q.tmelt = ph->tmelt;
    q.posnalp = -2.0 * ph->oek * ph->po * std::sin(ph->prec_angle);  // src/get_nl.f90:233
    q.oek = ph->oek;
    q.time = ph->timeStage;
    q.dilution_fac = ph->dilution_fac;
    q.ra = ph->ra;
    q.opr = ph->opr;
    q.cf2 = ph->courfac * ph->courfac;
    q.af2 = ph->alffac * ph->alffac;
    g.cmax = B->cmax;
    k_rl_zero<<<(2 * nlev + 255) / 256, 256, 0, s>>>(B->cmax, 2 * nlev);
    MHX_HIP_CHECK(hipGetLastError());
    if ((rc = launch_rl_grid(P, g, s))) return rc;
    k_rl_courant<<<(nlev + 127) / 128, 128, 0, s>>>(B->cmax, d_dx, nlev, dcour);
    MHX_HIP_CHECK(hipGetLastError());

    // ---- analysis ----
    double *nl[NL_N];
    for (int k = 0; k < NL_N; ++k) nl[k] = B->nl + (size_t)k * lmd;
    auto ana = [&](int kind, int c0, double *o0, double *o1, double *o2, size_t ld) -> int {
        AnaJob j;
        std::memset(&j, 0, sizeof(j));
        j.kind = kind;
        j.nlev = nlev;
        j.fin = B->F2;
        j.fin_ld = (size_t)nout * P->n_m * P->n_theta;
        if (kind == ANA_SH) {
            j.in_comp[0] = c0;
            j.in_comp[1] = j.in_comp[2] = -1;
        } else {
            j.in_comp[0] = c0;
            j.in_comp[1] = c0 + 1;
            j.in_comp[2] = c0 + 2;
        }
        j.out[0] = o0;
        j.out[1] = o1;
        j.out[2] = o2;
        j.out_ld = ld;
        j.lcut = B->lcut;
        return launch_ana(P, j, on());
    };
    if ((rc = fork())) return rc;
    // the nonlinear_lm_t arrays live in nl[] with the caller's level stride
    // ld_lm, so that the outputs the reference writes straight from the
    // analysis (dVSrLM, dVXirLM, dphidt) share a launch with them
    if (adv && (rc = ana(ANA_QST, out_of[RO_ADVR], nl[NL_ADVR], nl[NL_ADVT], nl[NL_ADVP], ld_lm))) return rc;
    if (heat && (rc = ana(ANA_QST, out_of[RO_VSR], dout[4], nl[NL_VST], nl[NL_VSP], ld_lm))) return rc;
    if (chem && (rc = ana(ANA_QST, out_of[RO_VXIR], dout[6], nl[NL_VXIT], nl[NL_VXIP], ld_lm))) return rc;
    if (ph->l_mag_nl && (rc = ana(ANA_QST, out_of[RO_VXBR], nl[NL_VXBR], nl[NL_VXBT], nl[NL_VXBP], ld_lm)))
        return rc;
    if (anel && (rc = ana(ANA_SH, out_of[RO_HEAT], nl[NL_HEAT], nullptr, nullptr, ld_lm))) return rc;
    if (phase && (rc = ana(ANA_SH, out_of[RO_PHI], dout[7], nullptr, nullptr, ld_lm))) return rc;
    if ((rc = join())) return rc;

    // ---- lm-space assembly ----
    RlTdArgs t;
    std::memset(&t, 0, sizeof(t));
    for (int k = 0; k < NL_N; ++k) t.nl[k] = nullptr;
    if (adv) t.nl[NL_ADVR] = nl[NL_ADVR], t.nl[NL_ADVT] = nl[NL_ADVT], t.nl[NL_ADVP] = nl[NL_ADVP];
    if (heat) t.nl[NL_VST] = nl[NL_VST], t.nl[NL_VSP] = nl[NL_VSP];
    if (chem) t.nl[NL_VXIT] = nl[NL_VXIT], t.nl[NL_VXIP] = nl[NL_VXIP];
    if (ph->l_mag_nl) t.nl[NL_VXBR] = nl[NL_VXBR], t.nl[NL_VXBT] = nl[NL_VXBT], t.nl[NL_VXBP] = nl[NL_VXBP];
    if (anel) t.nl[NL_HEAT] = nl[NL_HEAT];
    t.w = din[W];
    t.dw = din[DW];
    t.ddw = din[DDW];
    t.z = din[Z];
    t.dz = din[DZ];
    t.dwdt = dout[0];
    t.dzdt = dout[1];
    t.dpdt = dout[2];
    t.dsdt = dout[3];
    t.dxidt = dout[5];
    t.dbdt = dout[8];
    t.djdt = dout[9];
    t.dVxBhLM = dout[10];
    t.dVxVhLM = dout[11];
    t.ld_nl = ld_lm;
    t.ld_f = ld_lm;
    t.ld_o = ld_lm;
    t.rad = d_trad;
    t.lcut = B->lcut;
    t.conv_nl = ph->l_conv_nl;
    t.corr = ph->l_corr;
    t.single_matrix = ph->l_single_matrix;
    t.dcurl = dcurl;
    t.dpdt_on = dpdt_on;
    t.heat = heat;
    t.anel = anel;
    t.anel_liq = ph->l_anelastic_liquid;
    t.chem = chem;
    t.mag = ph->l_mag;
    t.mag_dbdt = ph->l_mag_nl || ph->l_mag_kin;
    t.CorFac = ph->CorFac;
    t.epsc = ph->epsc;
    t.epscXi = ph->epscXi;
    t.nlev = nlev;
    t.lm_max = P->lm_max;
    t.l_max = P->l_max;
    t.lm2l = P->d_lm2l;
    t.lm2m = P->d_lm2m;
    k_rl_td<<<dim3((unsigned)((P->lm_max + 255) / 256), (unsigned)nlev), 256, 0, s>>>(t);
    MHX_HIP_CHECK(hipGetLastError());

    // ---- results ----
    for (int i = 0; i < 12; ++i)
        if (hout[i])
            MHX_HIP_CHECK(hipMemcpyAsync(fout[i], dout[i], arr * sizeof(double), hipMemcpyDeviceToHost, s));
    MHX_HIP_CHECK(hipMemcpyAsync(out->dtrkc, dcour, (size_t)nlev * sizeof(double), hipMemcpyDefault, s));
    MHX_HIP_CHECK(hipMemcpyAsync(out->dthkc, dcour + nlev, (size_t)nlev * sizeof(double), hipMemcpyDefault, s));
    if (stream && nh == 0 && !host_cour) {
        MHX_HIP_CHECK(hipEventRecord(P->ev_done, s));
        MHX_HIP_CHECK(hipStreamWaitEvent((hipStream_t)stream, P->ev_done, 0));
        return MHX_OK;
    }
    MHX_HIP_CHECK(hipStreamSynchronize(s));
    return MHX_OK;
}
