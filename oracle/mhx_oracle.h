/*
 * ORACLE - test infrastructure only.
 *
 * Plain-C FP64 restatement of the reference's native spherical-harmonic
 * transform path (MagIC src/plms.f90, src/horizontal.f90, src/fft.f90,
 * src/shtransforms.f90, src/sht_native.f90, src/get_nl.f90).  Only tests,
 * the smoke check and the CPU-baseline leg of the benchmark may load it, and
 * only as the checker.  The product library (libmhx) never calls it.
 *
 * Conventions (all as in the reference):
 *   - coefficient arrays: complex128 interleaved (re,im), standard m-major
 *     lm order (m = 0, minc, 2 minc, ..., l = m..l_max), length lm_max;
 *   - grid arrays: float64 f[phi][theta] (theta fastest, Fortran f(nt,np)),
 *     theta rows north/south interleaved (row 2k = theta_k, 2k+1 = pi-theta_k,
 *     0-based);
 *   - Fourier arrays: complex128 F[m][theta], m = 0..n_phi/2.
 */
#ifndef MHX_ORACLE_H
#define MHX_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_plan orc_plan;

/* Gauss-Legendre colatitudes (ascending from the north pole) and weights
   (sum 2) for n points, src/horizontal.f90:279-340. */
void orc_gauleg(int n, double *theta_ord, double *gauss);

/* Fully normalised P_lm(cos theta) and sin(theta) dP_lm/dtheta, m-major,
   for m = m_min..m_max step minc, l = m..l_max (norm=2 of
   src/plms.f90:14-189). */
void orc_plm_theta(double theta, int l_max, int m_min, int m_max, int minc,
                   double *plm, double *dtheta_plm);

orc_plan *orc_plan_create(int l_max, int m_max, int minc, int n_theta,
                          int n_phi);
void orc_plan_destroy(orc_plan *p);
int orc_plan_lm_max(const orc_plan *p);
int orc_plan_n_m_max(const orc_plan *p);
/* copies of the plan's tables: theta_ord[nh], gauss[nh] (northern half),
   plm/dplm/wplm/wdplm[nh][lm_max] (lm fastest), osin2[n_theta] (scrambled) */
void orc_plan_tables(const orc_plan *p, double *theta_ord, double *gauss,
                     double *plm, double *dplm, double *wplm, double *wdplm,
                     double *osin2);

/* phi FFTs, src/fft.f90:164-252: forward F = rfft(f)/n_phi, inverse the
   unnormalised Hermitian synthesis (imaginary parts of modes 0 and n_phi/2
   ignored).  F is [n_phi/2+1][n_theta] complex. */
void orc_fft_forward(const orc_plan *p, const double *grid, double *F);
void orc_fft_inverse(const orc_plan *p, const double *F, double *grid);

/* native transforms, src/shtransforms.f90 */
void orc_sph_to_spat(const orc_plan *p, const double *Slm, double *sc, int lcut);
void orc_sph_to_grad_spat(const orc_plan *p, const double *Slm, double *gt,
                          double *gp, int lcut);
void orc_qst_to_spat(const orc_plan *p, const double *Q, const double *S,
                     const double *T, double *vr, double *vt, double *vp,
                     int lcut);
void orc_sphtor_to_spat(const orc_plan *p, const double *S, const double *T,
                        double *vt, double *vp, int lcut);
void orc_axi_to_spat(const orc_plan *p, const double *Sl, double *sc);
void orc_toraxi_to_spat(const orc_plan *p, const double *Tl, double *vt,
                        double *vp, int lcut);
void orc_spat_to_sph(const orc_plan *p, const double *f, double *flm, int lcut);
void orc_spat_to_sph_tor(const orc_plan *p, const double *vt, const double *vp,
                         double *slm, double *tlm, int lcut);

/* module sht API, src/sht_native.f90 */
void orc_scal_to_spat(const orc_plan *p, const double *S, double *f, int lcut);
void orc_scal_to_grad_spat(const orc_plan *p, const double *S, double *gt,
                           double *gp, int lcut);
void orc_pol_to_grad_spat(const orc_plan *p, const double *S, double *gt,
                          double *gp, int lcut);
void orc_torpol_to_spat(const orc_plan *p, const double *W, const double *dW,
                        const double *Z, double *vr, double *vt, double *vp,
                        int lcut);
void orc_pol_to_curlr_spat(const orc_plan *p, const double *Q, double *cvr,
                           int lcut);
void orc_torpol_to_curl_spat(const orc_plan *p, double or2, const double *B,
                             const double *ddB, const double *J,
                             const double *dJ, double *cvr, double *cvt,
                             double *cvp, int lcut);
void orc_torpol_to_dphspat(const orc_plan *p, const double *dW, const double *Z,
                           double *dvt, double *dvp, int lcut);
void orc_torpol_to_spat_IC(const orc_plan *p, double r, double r_icb,
                           const double *W, const double *dW, const double *Z,
                           double *br, double *bt, double *bp);
void orc_torpol_to_curl_spat_IC(const orc_plan *p, double r, double r_icb,
                                const double *dB, const double *ddB,
                                const double *J, const double *dJ, double *cbr,
                                double *cbt, double *cbp);
void orc_scal_to_SH(const orc_plan *p, const double *f, double *flm, int lcut);
void orc_spat_to_qst(const orc_plan *p, const double *f, const double *g,
                     const double *h, double *q, double *s, double *t, int lcut);
void orc_spat_to_sphertor(const orc_plan *p, const double *f, const double *g,
                          double *flm, double *glm, int lcut);

/* grid-space products of one bulk level (src/get_nl.f90:240-400 with
   l_adv_curl, l_mag_LF, l_heat_nl, l_mag_nl; Boussinesq):
   in[13] = vr vt vp cvr cvt cvp br bt bp cbr cbt cbp s,
   out[12] = Adv(r,t,p) LF(r,t,p) VxB(r,t,p) VS(r,t,p).  Each array is one
   n_theta x n_phi grid. */
void orc_get_nl(const orc_plan *p, double or2, double or4, double orho1,
                double LFfac, const double *const *in, double *const *out);

/* src/get_td.f90 (interior level, rotating Boussinesq MHD, no double curl):
   in[10] = AdvrLM AdvtLM AdvpLM VStLM VxBrLM VxBtLM VxBpLM w dw z, out[7] =
   dwdt dzdt dpdt dsdt dbdt djdt dVxBhLM; complex lm_max arrays. */
void orc_get_td(const orc_plan *p, int lR, double or2, double or4, double CorFac,
                double epsc, const double *const *in, double *const *out);

/* src/courant.f90 (magnetic branch): the level's radial and horizontal
   maxima; dtrkc = sqrt(delxr2 / vr2max), dthkc = sqrt(delxh2 / vh2max).
   in[6] = vr vt vp br bt bp grids (n_phi x n_theta, theta fastest). */
void orc_courant(const orc_plan *p, double or2, double or4, double orho1, double LFfac,
                 double courfac, double alffac, double valri2, double valhi2,
                 const double *const *in, double *vr2max, double *vh2max);

/* src/nonlinear_bcs.f90:24-68 (get_br_v_bcs): br_vt_lm, br_vp_lm (complex
   lm_max) of the boundary products on grids br, vt, vp. */
void orc_get_br_v_bcs(const orc_plan *p, const double *br, const double *vt, const double *vp,
                      double omega, double O_r_E_2, double O_rho, double *br_vt_lm,
                      double *br_vp_lm);
/* src/nonlinear_bcs.f90:71-118 (get_b_nl_bcs): bc 0 'CMB': b_nl_bc =
   -fac br_vt_lm, aj_nl_bc = -fac br_vp_lm; bc 1 'ICB': aj_nl_bc = -fac
   br_vp_lm (b_nl_bc untouched); fac = cond_or_sigma * prmag; lm = 1 zero.
   Complex arrays of lm_max; returns -1 for another bc. */
int orc_get_b_nl_bcs(int lm_max, int bc, const double *br_vt_lm, const double *br_vp_lm,
                     double cond_or_sigma, double prmag, double *b_nl_bc, double *aj_nl_bc);

/* ---- the radial loop of any physics switches, one interior level ---- */
/* grid components (the reference's grid_space_arrays_t fields) */
enum { ORC_RI_VR = 0, ORC_RI_VT, ORC_RI_VP, ORC_RI_CVR, ORC_RI_CVT, ORC_RI_CVP, ORC_RI_DVRDR,
       ORC_RI_DVTDR, ORC_RI_DVPDR, ORC_RI_DVRDT, ORC_RI_DVRDP, ORC_RI_DVTDP, ORC_RI_DVPDP,
       ORC_RI_S, ORC_RI_XI, ORC_RI_PHI, ORC_RI_BR, ORC_RI_BT, ORC_RI_BP, ORC_RI_CBR, ORC_RI_CBT,
       ORC_RI_CBP, ORC_RI_N };
/* grid products */
enum { ORC_RO_ADVR = 0, ORC_RO_ADVT, ORC_RO_ADVP, ORC_RO_VSR, ORC_RO_VST, ORC_RO_VSP, ORC_RO_VXIR,
       ORC_RO_VXIT, ORC_RO_VXIP, ORC_RO_VXBR, ORC_RO_VXBT, ORC_RO_VXBP, ORC_RO_HEAT, ORC_RO_PHI,
       ORC_RO_N };
/* switches (0/1) */
enum { ORC_F_CONV_NL = 0, ORC_F_ADV_CURL, ORC_F_HEAT_NL, ORC_F_CHEM, ORC_F_PHASE, ORC_F_PREC,
       ORC_F_CENTR, ORC_F_MAG_LF, ORC_F_MAG_NL, ORC_F_ANEL, ORC_F_LCR, ORC_F_CORR, ORC_F_SINGLE,
       ORC_F_DCURL, ORC_F_ANEL_LIQ, ORC_F_MAG_DBDT, ORC_F_N };
/* constants */
enum { ORC_P_LFFAC = 0, ORC_P_VISCHEAT, ORC_P_OHMLOSS, ORC_P_EPSPHASE, ORC_P_PENALTY,
       ORC_P_PHASEDIFF, ORC_P_TMELT, ORC_P_PO, ORC_P_PREC_ANGLE, ORC_P_OEK, ORC_P_TIME,
       ORC_P_DILUTION, ORC_P_RA, ORC_P_OPR, ORC_P_MINC, ORC_P_CORFAC, ORC_P_EPSC, ORC_P_EPSCXI,
       ORC_P_N };
/* per-level radial values */
enum { ORC_R_R = 0, ORC_R_OR1, ORC_R_OR2, ORC_R_OR4, ORC_R_ORHO1, ORC_R_BETA, ORC_R_OTEMP1,
       ORC_R_VISC, ORC_R_LAMBDA, ORC_R_TEMP0, ORC_R_EPSCPROF, ORC_R_N };
/* forward outputs entering the lm-space assembly */
enum { ORC_NL_ADVR = 0, ORC_NL_ADVT, ORC_NL_ADVP, ORC_NL_VST, ORC_NL_VXIT, ORC_NL_VXBR,
       ORC_NL_VXBT, ORC_NL_VXBP, ORC_NL_HEAT, ORC_NL_N };
/* lm-space assembly outputs */
enum { ORC_TO_DWDT = 0, ORC_TO_DZDT, ORC_TO_DPDT, ORC_TO_DSDT, ORC_TO_DXIDT, ORC_TO_DBDT,
       ORC_TO_DJDT, ORC_TO_DVXBH, ORC_TO_DVXVH, ORC_TO_N };

/* src/get_nl.f90:240-436 and src/rIter.f90:644-680 on one level's grids
   (in[ORC_RI_N], NULL where absent; out[ORC_RO_N], NULL where not wanted) */
void orc_rl_products(const orc_plan *p, const int *flags, const double *par, const double *rad,
                     const double *const *in, double *const *out);
/* src/courant.f90:161-276 for one level: par = courfac, alffac, LFfac,
   opm, l_cour_alf_damp; rad = or2, or4, orho1, delxr2, delxh2; in = vr vt
   vp br bt bp grids */
void orc_rl_courant(const orc_plan *p, int cour_mag, const double *par, const double *rad,
                    const double *const *in, double *dtrkc, double *dthkc);
/* src/get_td.f90 nonlinear_lm_t at one interior level */
void orc_rl_td(const orc_plan *p, const int *flags, const double *par, const double *rad, int lR,
               const double *const *nl, const double *const *fld, double *const *out);

#ifdef __cplusplus
}
#endif
#endif
