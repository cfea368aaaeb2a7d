/*
 * mhx - MI355X-native spherical-harmonic transforms and nonlinear-term path
 * for MagIC-style dynamo codes.  C ABI (plain pointers and sizes).
 *
 * The per-call entry points mirror the reference's transform API, module sht
 * (MagIC src/sht_native.f90:16-20), one function per public routine, with
 * the same argument meaning and order after the leading plan handle.  The
 * batched entry points replace the reference's per-radius loop
 * (src/rIter.f90:190-444) with one launch per stage for all radial levels.
 *
 * Data conventions (identical to the reference):
 *   coefficient vector: complex128 as interleaved (re, im) doubles, length
 *       lm_max, standard m-major order: m = 0, minc, ..., m_max and
 *       l = m..l_max inside each m (src/blocking.f90:309-317);
 *   grid: n_theta x n_phi doubles, theta fastest (Fortran f(n_theta,n_phi)),
 *       theta rows north/south interleaved: 0-based row 2k is colatitude
 *       theta_k, row 2k+1 is pi - theta_k (src/horizontal.f90:141-191),
 *       phi_j = j 2 pi / (n_phi minc);
 *   vector grids in Robert form (sin(theta) v_theta, sin(theta) v_phi);
 *   forward transforms scale by 1/n_phi (src/fft.f90:464).
 *
 * Pointers may be host or device memory; the library detects which.  Host
 * pointers are staged through the plan's device buffers (PCIe-bound; the
 * batched entry points with device pointers are the fast path).  All calls
 * are synchronous with respect to the host unless they take a stream.
 *
 * Errors: every function returns 0 on success or a negative MHX_E* code; the
 * message of the last error on the calling thread is mhx_last_error().
 * The reference aborts the run instead (src/useful.f90:271-302); the
 * Fortran shim (magic_amd/fortran/sht_mhx.f90) turns a nonzero code into
 * that abort.
 */
#ifndef MHX_H
#define MHX_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MHX_OK 0
#define MHX_E_INVALID (-1)   /* bad argument (sizes, lcut, null pointer) */
#define MHX_E_HIP (-2)       /* HIP runtime error */
#define MHX_E_NOMEM (-3)     /* device allocation failed */
#define MHX_E_UNSUPPORTED (-4)

/* arithmetic of the Legendre stage */
#define MHX_PREC_FP64 0      /* plain form: FP64 MFMA */
#define MHX_PREC_SPLIT16 1   /* scaled FP16 hi/lo split, FP32 accumulation */
/* the split form with the nonlinear step's P_lm and dP_lm, for l_max > 255,
 * recomputed inside the Legendre kernels by an FP64 three-term recurrence
 * from FP64 seeds (P_mm(theta) for the synthesis, every 32nd degree for the
 * analysis) held as mantissa and power-of-two exponent, instead of read from
 * stored FP16 tables; 255 < l_max <= 1023 */
#define MHX_PREC_SPLIT16_REC 2

typedef struct mhx_plan mhx_plan;

const char *mhx_last_error(void);

/* Plan for one truncation and grid; replaces initialize_sht +
 * initialize_transforms (src/sht_native.f90:24-32,
 * src/shtransforms.f90:38-91).  The P_lm tables are generated on the GPU in
 * FP64 (recurrence of src/plms.f90:14-189) from Gauss-Legendre nodes
 * (src/horizontal.f90:279-340).  n_theta must be a multiple of 4 and n_phi a
 * multiple of 4 whose half factors into 2, 3 and 5 (src/fft.f90:55-73).
 * max_levels sizes the batched workspaces (may be 0; grown on demand). */
int mhx_plan_create(int l_max, int m_max, int minc, int n_theta, int n_phi,
                    int precision, int max_levels, mhx_plan **plan);
/* finalize_sht (src/sht_native.f90:34-38) */
int mhx_plan_destroy(mhx_plan *plan);
int mhx_plan_info(const mhx_plan *plan, int *lm_max, int *n_m_max,
                  int *n_theta, int *n_phi);
/* copies of the plan's FP64 tables in the reference's layout:
 * theta_ord[n_theta/2], gauss[n_theta/2], plm/dplm[n_theta/2][lm_max]
 * (lm fastest); any pointer may be NULL */
int mhx_plan_tables(const mhx_plan *plan, double *theta_ord, double *gauss,
                    double *plm, double *dplm);

/* ---- module sht, one entry per public routine (src/sht_native.f90) ---- */
/* :40-52 */
int mhx_scal_to_spat(mhx_plan *p, const double *Slm, double *fieldc, int lcut);
/* :54-68 */
int mhx_scal_to_grad_spat(mhx_plan *p, const double *Slm, double *gradtc,
                          double *gradpc, int lcut);
/* :70-97 */
int mhx_pol_to_grad_spat(mhx_plan *p, const double *Slm, double *gradtc,
                         double *gradpc, int lcut);
/* :99-127 */
int mhx_torpol_to_spat(mhx_plan *p, const double *Wlm, const double *dWlm,
                       const double *Zlm, double *vrc, double *vtc, double *vpc,
                       int lcut);
/* :129-141 */
int mhx_sphtor_to_spat(mhx_plan *p, const double *dWlm, const double *Zlm,
                       double *vtc, double *vpc, int lcut);
/* :143-186 */
int mhx_torpol_to_curl_spat_IC(mhx_plan *p, double r, double r_ICB,
                               const double *dBlm, const double *ddBlm,
                               const double *Jlm, const double *dJlm,
                               double *cbr, double *cbt, double *cbp);
/* :188-229 */
int mhx_torpol_to_spat_IC(mhx_plan *p, double r, double r_ICB,
                          const double *Wlm, const double *dWlm,
                          const double *Zlm, double *Br, double *Bt, double *Bp);
/* :231-272 */
int mhx_torpol_to_dphspat(mhx_plan *p, const double *dWlm, const double *Zlm,
                          double *dvtdp, double *dvpdp, int lcut);
/* :274-300 */
int mhx_pol_to_curlr_spat(mhx_plan *p, const double *Qlm, double *cvrc,
                          int lcut);
/* :302-336 */
int mhx_torpol_to_curl_spat(mhx_plan *p, double or2, const double *Blm,
                            const double *ddBlm, const double *Jlm,
                            const double *dJlm, double *cvrc, double *cvtc,
                            double *cvpc, int lcut);
/* :338-346; f is intent(inout) in the reference and is left unchanged here */
int mhx_scal_to_SH(mhx_plan *p, const double *f, double *fLM, int lcut);
/* :348-364 */
int mhx_spat_to_qst(mhx_plan *p, const double *f, const double *g,
                    const double *h, double *qLM, double *sLM, double *tLM,
                    int lcut);
/* :366-379 */
int mhx_spat_to_sphertor(mhx_plan *p, const double *f, const double *g,
                         double *fLM, double *gLM, int lcut);
/* :381-391; fl_ax has l_max+1 complex entries (m = 0), f has n_theta */
int mhx_axi_to_spat(mhx_plan *p, const double *fl_ax, double *f);
/* :393-405 */
int mhx_toraxi_to_spat(mhx_plan *p, const double *fl_ax, double *ft,
                       double *fp, int lcut);

/* src/nonlinear_bcs.f90:24-68 (get_br_v_bcs, called at src/rIter.f90:270
 * and :275 for the outer and inner boundary): the induction products of a
 * conducting mantle or inner core, br_vt = fac br vt and br_vp = br (fac vp
 * - omega sin^2 theta) with fac = O_r_E_2 O_rho, analysed by
 * spat_to_sphertor at l_max into br_vt_lm, br_vp_lm (complex lm_max).
 * Grids n_theta x n_phi as the other per-level routines; host or device. */
int mhx_get_br_v_bcs(mhx_plan *p, const double *br, const double *vt, const double *vp,
                     double omega, double O_r_E_2, double O_rho, double *br_vt_lm,
                     double *br_vp_lm);
/* src/nonlinear_bcs.f90:71-118 (get_b_nl_bcs, called from
 * src/step_time.f90:619 and :631 on the products above): the nonlinear term
 * of the magnetic boundary condition of a conducting mantle, bc =
 * MHX_BC_CMB ('CMB'): b_nl_bc = -fac br_vt_lm, aj_nl_bc = -fac br_vp_lm with
 * fac = conductance_ma prmag; or inner core, MHX_BC_ICB ('ICB'): aj_nl_bc =
 * -fac br_vp_lm with fac = sigma_ratio prmag (b_nl_bc and br_vt_lm not
 * used, may be null).  Entry lm = 1 (l = 0) is zero.  cond_or_sigma is
 * conductance_ma or sigma_ratio.  Complex lm_max arrays, host or device;
 * another bc is MHX_E_INVALID (the reference's abortRun). */
#define MHX_BC_CMB 0
#define MHX_BC_ICB 1
int mhx_get_b_nl_bcs(mhx_plan *p, int bc, const double *br_vt_lm, const double *br_vp_lm,
                     double cond_or_sigma, double prmag, double *b_nl_bc, double *aj_nl_bc);

/* ---- batched over radial levels (device pointers, one launch per stage) ----
 * Level i of an input/output coefficient array starts at ptr + 2*i*ld_lm
 * doubles; level i of a grid at ptr + i*ld_grid doubles.  lcut[nlev] is the
 * per-level cut-off (l_R(nR) of src/radial.f90:290-307), host memory.
 * These replace nlev calls of the per-level routine of the same name. */
int mhx_torpol_to_spat_batch(mhx_plan *p, int nlev, const double *W,
                             const double *dW, const double *Z, size_t ld_lm,
                             double *vr, double *vt, double *vp, size_t ld_grid,
                             const int *lcut, void *stream);
int mhx_spat_to_qst_batch(mhx_plan *p, int nlev, const double *f,
                          const double *g, const double *h, size_t ld_grid,
                          double *q, double *s, double *t, size_t ld_lm,
                          const int *lcut, void *stream);

/* ---- the nonlinear-term path of one radial loop, all levels at once ----
 * Replaces transform_to_grid_space / get_nl / transform_to_lm_space of
 * src/rIter.f90:466-712 for bulk levels of a Boussinesq MHD run with
 * curl-form advection (l_adv_curl) and the derivative fields:
 *   inverse: scal_to_spat(s), scal_to_grad_spat(s), torpol_to_spat(w,dw,z),
 *            torpol_to_curl_spat(or2,w,ddw,z,dz), torpol_to_spat(dw,ddw,dz),
 *            torpol_to_spat(b,db,aj), torpol_to_curl_spat(or2,b,ddb,aj,dj);
 *   grid:    Adv = -curl(u) x u terms + Lorentz force (src/get_nl.f90:242-273,
 *            merged as src/rIter.f90:650-656), VS = u s (:320-325),
 *            VxB (:376-388), and the scalar u_h . grad_h s
 *            = or2/sin^2(theta) (vt ds/dtheta_c + vp ds/dphi_c);
 *   forward: spat_to_qst(Adv), spat_to_qst(VS), spat_to_qst(VxB),
 *            scal_to_SH(u_h . grad_h s).
 * in[11]  = w dw ddw z dz s b db ddb aj dj  (coefficients, ld_lm)
 * out[10] = AdvrLM AdvtLM AdvpLM VSrLM VStLM VSpLM VxBrLM VxBtLM VxBpLM
 *           uGsLM (coefficients, ld_lm)
 * radial: per-level or2, or4, orho1 (nlev each) and LFfac; lcut[nlev].
 * Coefficient arrays may be host or device memory; lcut/radial are host.
 * With a non-null stream and device outputs the call returns without
 * waiting: the work is ordered after `stream` and `stream` after the work
 * (the host arrays lcut/radial are copied before the call returns).
 * Otherwise the call returns when the outputs are written. */
int mhx_nl_step(mhx_plan *p, int nlev, const double *const *in, size_t ld_lm,
                double *const *out, const double *or2, const double *or4,
                const double *orho1, double LFfac, const int *lcut,
                void *stream);
/* the same step, also writing the 18 inverse grid components
 * (vr vt vp cvr cvt cvp dvr dvt dvp br bt bp cbr cbt cbp s dsdt dsdp), each
 * grid component at grids + ((lev*18 + comp) * n_phi * n_theta) */
int mhx_nl_step_grids(mhx_plan *p, int nlev, const double *const *in,
                      size_t ld_lm, double *const *out, const double *or2,
                      const double *or4, const double *orho1, double LFfac,
                      const int *lcut, double *grids, void *stream);

/* the same step, also returning the Courant maxima of each level, reduced
 * inside the grid stage from its velocity and magnetic grids: the
 * reference's courant (src/courant.f90:54-277, XSH_COURANT = 0, magnetic
 * branch l_mag, l_mag_LF, not l_mag_kin), called per level from
 * src/rIter.f90:292-299 with the same grids,
 *   vr2max = max or4 (courfac^2 orho2 vr^2 + alffac^2 valr^2/(valr+valri2)),
 *   vh2max = max or2 (courfac^2 orho2 (vt^2+vp^2)/sin^2 +
 *                     alffac^2 valh^2/(valh+valhi2)),
 * valr = LFfac orho1 br^2, valh = LFfac orho1 (bt^2+bp^2)/sin^2, orho2 =
 * orho1^2; valri2/valhi2 per level as the reference forms them
 * ((1+opm)^2/4 / delxr2, / delxh2 with l_cour_alf_damp, else 0).  The
 * caller then takes dtrkc = sqrt(delxr2/vr2max), dthkc = sqrt(delxh2/vh2max)
 * (vr2max, vh2max: host arrays of nlev).  In the split precision the maxima
 * come from the FP32 grids. */
int mhx_nl_step_courant(mhx_plan *p, int nlev, const double *const *in, size_t ld_lm,
                        double *const *out, const double *or2, const double *or4,
                        const double *orho1, double LFfac, const int *lcut,
                        double courfac, double alffac, const double *valri2,
                        const double *valhi2, double *vr2max, double *vh2max,
                        void *stream);

/* lm-space assembly of the explicit terms at interior radial levels, the
 * step after the forward transforms of the nonlinear step: the reference's
 * nonlinear_lm_t get_dwdt, get_dzdt, get_dpdt, get_dsdt and get_dbdt
 * (src/get_td.f90:133-589, called per level at src/rIter.f90:412-429) for a
 * rotating Boussinesq MHD run without the double-curl form (l_corr,
 * l_conv_nl, l_mag_nl; not l_anel, l_double_curl, l_single_matrix), batched
 * over nlev levels.
 * in[10] = AdvrLM AdvtLM AdvpLM VStLM VxBrLM VxBtLM VxBpLM (the outputs of
 *          mhx_nl_step) w dw z (the fields at the levels), ld_lm
 * out[7] = dwdt dzdt dpdt dsdt dbdt djdt dVxBhLM, ld_lm; dpdt of l = 0 is
 *          zero (the reference leaves it unset)
 * per level (host arrays): or2, or4, epsc = epsc * epscProf(nR) (the l = 0
 * term of dsdt), lcut = l_R(nR); CorFac as src/get_td.f90 reads it; or1 =
 * sqrt(or2) and r^2 = 1/or2.  Arrays host or device memory; stream as in
 * mhx_nl_step.  FP64 in every precision. */
int mhx_get_td(mhx_plan *p, int nlev, const double *const *in, size_t ld_lm,
               double *const *out, const double *or2, const double *or4,
               const double *epsc, double CorFac, const int *lcut, void *stream);

/* mhx_nl_step followed by mhx_get_td on its outputs in one call: the
 * radial loop's explicit terms of the interior levels (src/rIter.f90:
 * 412-429 after :483-707) with the lm-space assembly as the step's last
 * launch, inside the step's captured graph.  out[10] as mhx_nl_step (still
 * written: dVSrLM = VSrLM is the reference's own output); td_out[7] = dwdt
 * dzdt dpdt dsdt dbdt djdt dVxBhLM as mhx_get_td's out; epsc, CorFac as
 * mhx_get_td, the other arguments as mhx_nl_step.  The assembly is FP64 on
 * the step's outputs in either precision. */
int mhx_nl_step_td(mhx_plan *p, int nlev, const double *const *in, size_t ld_lm,
                   double *const *out, double *const *td_out, const double *or2,
                   const double *or4, const double *orho1, double LFfac, const double *epsc,
                   double CorFac, const int *lcut, void *stream);

/* ---- the radial loop for any physics switches, interior levels ----------
 * One call replaces the loop body of rIter_single_t%radialLoop
 * (src/rIter.f90:188-444) for nlev interior levels (nBc = 0) of a time step
 * stage, on the GPU: the inverse
 * transforms of transform_to_grid_space (:466-622), the grid products of
 * get_nl (src/get_nl.f90:213-441) merged as transform_to_lm_space
 * (:644-680), the forward transforms (:682-707), the Courant steps
 * (src/courant.f90:54-277, XSH_COURANT = 0) and the lm-space assembly of
 * nonlinear_lm_t (src/get_td.f90: get_dwdt or get_dwdt_double_curl,
 * get_dzdt, get_dpdt, get_dsdt, get_dxidt, get_dbdt).  Not covered (the
 * caller runs the reference's own loop for such stages): boundary levels,
 * the diagnostics computed inside the loop (graph, movie, TO, RMS, power,
 * helicity, ...), l_full_sphere's centre, finite-difference l_parallel_solve
 * and the temperature-diffusion single-matrix form.  The switch sets the
 * fused step covers (curl advection, heat, the full MHD terms with
 * l_cour_alf_damp; no anelastic, double-curl, composition, phase-field,
 * precession or centrifugal terms, n_r_LCR = 0) run mhx_nl_step_courant and
 * mhx_get_td in the plan's precision and return when done; the others run
 * FP64 kernels whatever the plan's precision.
 *
 * The switches and constants the loop reads from the reference's logic,
 * physical_parameters and time-scheme data; booleans as 0/1. */
typedef struct mhx_rloop_phys {
    int l_conv, l_conv_nl, l_heat, l_heat_nl, l_chemical_conv, l_phase_field;
    int l_mag, l_mag_LF, l_mag_nl, l_mag_kin, l_anel, l_anelastic_liquid;
    int l_adv_curl, l_double_curl, l_corr, l_single_matrix, l_precession;
    int l_centrifuge, l_cour_alf_damp, lPressNext;
    int n_r_LCR;  /* levels with nR <= n_r_LCR lie outside the conducting region */
    double LFfac, CorFac, epsc, epscXi, ViscHeatFac, OhmLossFac, opm;
    double courfac, alffac;  /* tscheme%courfac, tscheme%alffac */
    double epsPhase, penaltyFac, phaseDiffFac, tmelt;
    double po, prec_angle, oek, timeStage, dilution_fac, ra, opr;
} mhx_rloop_phys;

/* per-level values (host arrays of nlev entries): the reference's
 * radial_functions at each level, its radial index nR (1 = outer boundary)
 * and cut-off l_R(nR) */
typedef struct mhx_rloop_radial {
    const double *r, *or1, *or2, *or4, *orho1, *beta, *otemp1, *visc, *lambda;
    const double *temp0, *epscProf, *delxr2, *delxh2;
    const int *nR, *l_R;
} mhx_rloop_radial;

/* coefficient inputs (complex, level stride ld_lm; the fields module's
 * *_Rloc): w dw ddw z dz always, s with l_heat, xi with l_chemical_conv, phi
 * with l_phase_field, b db ddb aj dj with l_mag; others may be NULL */
typedef struct mhx_rloop_fields {
    const double *w, *dw, *ddw, *z, *dz, *s, *xi, *phi, *b, *db, *ddb, *aj, *dj;
} mhx_rloop_fields;

/* outputs (complex, level stride ld_lm; the radialLoop_if arguments):
 * dwdt dzdt dpdt always (dpdt skipped, may be NULL, with l_double_curl and
 * not lPressNext), dsdt dVSrLM with l_heat, dxidt dVXirLM with
 * l_chemical_conv, dphidt with l_phase_field, dbdt djdt dVxBhLM with l_mag,
 * dVxVhLM with l_double_curl; dtrkc dthkc (double, nlev) the Courant steps
 * (1e10 where the level imposes none) */
typedef struct mhx_rloop_out {
    double *dwdt, *dzdt, *dpdt, *dsdt, *dVSrLM, *dxidt, *dVXirLM, *dphidt;
    double *dbdt, *djdt, *dVxBhLM, *dVxVhLM;
    double *dtrkc, *dthkc;
} mhx_rloop_out;

/* Arrays host or device memory (the library detects which; host arrays are
 * staged); the stream as in mhx_nl_step.  MHX_E_UNSUPPORTED for switches
 * outside the list above (l_mag_kin, l_single_matrix with l_corr). */
int mhx_rloop(mhx_plan *p, const mhx_rloop_phys *phys, int nlev, const mhx_rloop_radial *rad,
              const mhx_rloop_fields *in, size_t ld_lm, const mhx_rloop_out *out,
              void *stream);

/* Per-level sums of the kinetic-energy diagnostic (src/kinetic_energy.f90
 * get_e_kin, :119-160, the loop over radial levels and lm before the radial
 * integrals): for nlev levels of w, dw, z (complex, level stride ld_lm,
 * host or device) and per level or2, orho1 (host), writes sums[k*nlev + i]
 * (host or device) for k = e_p_r, e_t_r, e_p_as_r, e_t_as_r, e_p_es_r,
 * e_t_es_r, e_p_eas_r, e_t_eas_r.  The caller integrates them over radius
 * (rInt_R, src/integration.f90:56-103) and multiplies by eScale/2.
 * At most 65535 levels per call.  Synchronous; FP64. */
int mhx_get_e_kin(mhx_plan *p, int nlev, const double *w, const double *dw, const double *z,
                  size_t ld_lm, const double *or2, const double *orho1, double *sums);

/* Per-level sums of the magnetic-energy diagnostic of the outer core
 * (src/magnetic_energy.f90 get_e_mag, :264-340): for nlev levels of b, db,
 * aj (complex, level stride ld_lm, host or device) and per level or2
 * (host), writes sums[k*nlev + i] (host or device) for k = e_p_r, e_t_r,
 * e_p_as_r, e_t_as_r, e_p_es_r, e_t_es_r, e_p_eas_r, e_t_eas_r,
 * e_dipole_r, e_dipole_ax_r.  The caller integrates over radius and
 * multiplies by LFfac eScale / 2 (els_r = (e_p_r + e_t_r) orho1 sigma).
 * At most 65535 levels per call.  Synchronous; FP64. */
int mhx_get_e_mag(mhx_plan *p, int nlev, const double *b, const double *db, const double *aj,
                  size_t ld_lm, const double *or2, double *sums);

/* The kinetic-energy sums of mhx_get_e_kin without any allocation.  This
 * entry has no counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  It calls no hipMalloc, hipFree or
 * allocating library routine: the kernels' partial sums live in a fixed
 * device array of the library (2.6 MB per device, held until the process
 * ends).  Every array is device memory: w, dw, z (complex, level stride
 * ld_lm), or2 and orho1 (nlev each) and sums (8 x nlev, the layout of
 * mhx_get_e_kin); host arrays are refused (MHX_E_INVALID), as are more than
 * 2048 levels per call.  The plan's device must be current.  Asynchronous:
 * the kernels are queued on `stream` (null: the null stream) and the call
 * returns; sums are valid once the stream reaches that point.  Calls that
 * overlap on one device share its array and race: order them on one stream
 * or wait between them. */
int mhx_get_e_kin_noalloc(mhx_plan *p, int nlev, const double *w, const double *dw,
                          const double *z, size_t ld_lm, const double *or2, const double *orho1,
                          double *sums, void *stream);

/* ---- uneven batches of independent transform jobs (split form) ----
 * A job is the round trip of nfields fields of one kind at one truncation
 * over nlev radial levels: synthesis to grids and analysis back, i.e. per
 * level and field the reference's scal_to_spat + scal_to_SH (scalar) or
 * torpol_to_spat + spat_to_qst (vector) of src/sht_native.f90:41-123,
 * 358-364, which its callers loop over (src/rIter.f90:483-707).  l_max odd,
 * at most 255; m_max = l_max, minc = 1, n_theta = 3 (l_max + 1) / 2,
 * n_phi = 2 n_theta, lcut = l_max.  Arrays are device memory laid out
 * [comp][field * nlev + lev][lm_max] complex: scalar comp s; vector in =
 * W, dW, Z and out = Q, S, T.  mhx_batch_create builds one plan per
 * distinct l_max and the jobs' workspaces and captures the batch into a
 * graph: by default one persistent launch that runs every job as a work
 * list of synthesis items, row items (inverse and forward row FFTs in
 * LDS, no grid in device memory) and analysis items, with the tile shapes
 * of every l_max chosen by timing the whole launch (autotuner); with the
 * environment variable MHX_BATCH_FORM=graph, or where a plan does not admit
 * that kernel, the jobs' per-job launches on parallel graph branches, their
 * kernel forms chosen per l_max by timing its largest job.
 * The timing runs the jobs: during mhx_batch_create (which returns when they
 * are done) the jobs' input arrays are read and their output arrays written.
 * mhx_batch_execute replays the graph (with a stream: returns at once,
 * ordered on the stream; without: returns when done).  The jobs' arrays
 * must stay allocated.  The reference runs these calls one level and field
 * at a time; the batch entries have no counterpart there. */
#define MHX_JOB_SCALAR 0
#define MHX_JOB_VECTOR 1
typedef struct mhx_job {
    int l_max, nlev, kind, nfields;
    const double *in;
    double *out;
} mhx_job;
typedef struct mhx_batch mhx_batch;
int mhx_batch_create(const mhx_job *jobs, int njobs, int precision, mhx_batch **batch);
int mhx_batch_execute(mhx_batch *batch, void *stream);
/* the tuned forms per l_max group (n entries at most): synthesis and
 * analysis form (1 whole-K / whole-theta, 2 chunked) and row-FFT rows per
 * block (persistent form: theta rows per row item); returns the number of
 * groups */
int mhx_batch_tuning(const mhx_batch *batch, int *l_max, int *syn_form, int *ana_form,
                     int *fft_rows, int n);
/* the launch form (*form: 1 one persistent launch, 0 per-job launches) and,
 * per l_max group (n entries at most), the tile shapes the autotuner chose
 * for the persistent launch: orders per synthesis/analysis item (1: one m;
 * 2: m and m_max - m), level groups of 8 rows (field x level) per item for
 * the orders below m_split (order index) and groups_high from it on, theta
 * rows per row item (graph form: 1, 0, 0, 0, row-FFT rows per block);
 * returns the number of groups */
int mhx_batch_shapes(const mhx_batch *batch, int *form, int *l_max, int *orders, int *groups,
                     int *groups_high, int *m_split, int *theta_rows, int n);
/* waits until the batch's executions issued so far are done; MHX_E_HIP if
 * in one of them a work item of the persistent launch gave up waiting for
 * its inputs (its outputs are then not valid), MHX_OK otherwise */
int mhx_batch_status(mhx_batch *batch);
int mhx_batch_destroy(mhx_batch *batch);

/* Whether the nonlinear step (mhx_nl_step, mhx_nl_step_td,
 * mhx_nl_step_courant) computes the du/dr grids: the synthesis of dw, ddw,
 * dz (torpol_to_spat into dvrdrc, dvtdrc, dvpdrc) and their transforms.  On
 * by default; the products and forward outputs do not use them, and with
 * curl advection the reference computes them only on steps that write
 * outputs (src/rIter.f90:527-538).  Off, the step skips that synthesis and
 * the grids' writes (FP64: identical outputs; split form: the tile grid
 * stage transforms empty halves in their place, which changes FP32
 * rounding only).  mhx_nl_step_grids, and a plan timing its launches, keep
 * them.  No counterpart in the reference, whose loop decides per step. */
int mhx_set_dudr(mhx_plan *p, int on);

/* durations (ms) of the stages of the last mhx_nl_step on this plan,
 * measured with HIP events on the launch stream: [0] synthesis Legendre,
 * [1] fused grid stage (inverse FFTs, grid products, forward FFTs), [2] and
 * [3] zero (stages fused into [1]), [4] analysis Legendre.
 * Requires mhx_set_timing(p, 1) before the step. */
int mhx_set_timing(mhx_plan *p, int on);
int mhx_stage_times(const mhx_plan *p, double *ms5);
/* durations (ms) of the individual launches of the last timed mhx_nl_step,
 * in launch order: synthesis (s + gradient), synthesis QST x5 (u, curl u,
 * du/dr, B, curl B), fused grid stage, analysis QST x3 (Adv, VS, VxB),
 * analysis SH (u.grad s).  Returns the number written. */
int mhx_launch_times(const mhx_plan *p, double *ms, int n);

#ifdef __cplusplus
}
#endif
#endif
