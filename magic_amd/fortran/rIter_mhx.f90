!
! The radial loop of MagIC as an rIter_t extension backed by the mhx HIP
! library: the reference's plugin point for the whole loop
! (src/rIteration.f90:15-79, allocated by module radialLoop,
! src/radialLoop.f90:26).  One call of radialLoop runs the interior levels of
! the stage on the GPU in one batch (mhx_rloop: transforms, grid products,
! Courant steps and the lm-space assembly of every interior level at once,
! include/mhx.h) and the two boundary levels through the reference's own
! per-level machinery (an embedded rIter_single_t, src/rIter.f90:188-444).
!
! Stages that compute diagnostics inside the loop (graph and movie frames,
! TO, RMS, power, helicity, fluxes, probes, dtB ...), and runs whose
! switches the batched call does not cover (l_onset, l_RMS, l_full_sphere,
! finite-difference parallel solves, the single-matrix temperature form),
! run the reference's loop unchanged for all levels.
!
module rIter_mhx_mod

   use iso_c_binding
   use precision_mod, only: cp
   use truncation, only: lm_max, lm_maxMag, l_max, m_max, minc, n_theta_max, n_phi_max
   use radial_data, only: n_r_cmb, n_r_icb, nRstart, nRstop, nRstartMag, nRstopMag
   use radial_functions, only: r, or1, or2, or4, orho1, beta, otemp1, visc, lambda, temp0, &
       &                       epscProf, l_R
   use num_param, only: delxr2, delxh2
   use physical_parameters, only: LFfac, CorFac, epsc, epscXi, ViscHeatFac, OhmLossFac, opm, &
       &                          n_r_LCR, ktopv, kbotv, epsPhase, penaltyFac, phaseDiffFac,  &
       &                          tmelt, po, prec_angle, oek, dilution_fac, ra, opr
   use logic, only: l_conv, l_conv_nl, l_heat, l_heat_nl, l_chemical_conv, l_phase_field,    &
       &            l_mag, l_mag_LF, l_mag_nl, l_mag_kin, l_anel, l_anelastic_liquid,         &
       &            l_adv_curl, l_double_curl, l_corr, l_single_matrix, l_precession,         &
       &            l_centrifuge, l_cour_alf_damp, l_onset, l_RMS, l_full_sphere,             &
       &            l_parallel_solve, l_temperature_diff, l_dtB, l_b_nl_cmb, l_b_nl_icb,      &
       &            l_rot_ic, l_cond_ic, l_rot_ma, l_cond_ma
   use fields, only: w_Rloc, dw_Rloc, ddw_Rloc, z_Rloc, dz_Rloc, s_Rloc, xi_Rloc, phi_Rloc,  &
       &             b_Rloc, db_Rloc, ddb_Rloc, aj_Rloc, dj_Rloc, omega_ic, omega_ma
   use time_schemes, only: type_tscheme
   use constants, only: zero
   use rIteration, only: rIter_t
   use rIter_mod, only: rIter_single_t
   use nonlinear_bcs, only: get_br_v_bcs
   use outRot, only: get_lorentz_torque
   use courant_mod, only: courant
   use useful, only: abortRun

   implicit none

   private

   ! mirrors of the C structs of include/mhx.h (mhx_rloop_*)
   type, bind(C) :: mhx_rloop_phys
      integer(c_int) :: l_conv, l_conv_nl, l_heat, l_heat_nl, l_chemical_conv, l_phase_field
      integer(c_int) :: l_mag, l_mag_LF, l_mag_nl, l_mag_kin, l_anel, l_anelastic_liquid
      integer(c_int) :: l_adv_curl, l_double_curl, l_corr, l_single_matrix, l_precession
      integer(c_int) :: l_centrifuge, l_cour_alf_damp, lPressNext
      integer(c_int) :: n_r_LCR
      real(c_double) :: LFfac, CorFac, epsc, epscXi, ViscHeatFac, OhmLossFac, opm
      real(c_double) :: courfac, alffac
      real(c_double) :: epsPhase, penaltyFac, phaseDiffFac, tmelt
      real(c_double) :: po, prec_angle, oek, timeStage, dilution_fac, ra, opr
   end type mhx_rloop_phys

   type, bind(C) :: mhx_rloop_radial
      type(c_ptr) :: r, or1, or2, or4, orho1, beta, otemp1, visc, lambda
      type(c_ptr) :: temp0, epscProf, delxr2, delxh2
      type(c_ptr) :: nR, l_R
   end type mhx_rloop_radial

   type, bind(C) :: mhx_rloop_fields
      type(c_ptr) :: w, dw, ddw, z, dz, s, xi, phi, b, db, ddb, aj, dj
   end type mhx_rloop_fields

   type, bind(C) :: mhx_rloop_out
      type(c_ptr) :: dwdt, dzdt, dpdt, dsdt, dVSrLM, dxidt, dVXirLM, dphidt
      type(c_ptr) :: dbdt, djdt, dVxBhLM, dVxVhLM
      type(c_ptr) :: dtrkc, dthkc
   end type mhx_rloop_out

   interface
      integer(c_int) function mhx_plan_create(lmax, mmax, mminc, nt, np, prec, nlev, p) &
           &         bind(C, name='mhx_plan_create')
         import :: c_int, c_ptr
         integer(c_int), value :: lmax, mmax, mminc, nt, np, prec, nlev
         type(c_ptr), intent(out) :: p
      end function mhx_plan_create
      integer(c_int) function mhx_plan_destroy(p) bind(C, name='mhx_plan_destroy')
         import :: c_int, c_ptr
         type(c_ptr), value :: p
      end function mhx_plan_destroy
      type(c_ptr) function mhx_last_error() bind(C, name='mhx_last_error')
         import :: c_ptr
      end function mhx_last_error
      integer(c_int) function mhx_rloop(p, phys, nlev, rad, fin, ld_lm, fout, stream) &
           &         bind(C, name='mhx_rloop')
         import :: c_int, c_ptr, c_size_t, mhx_rloop_phys, mhx_rloop_radial, mhx_rloop_fields, &
              &    mhx_rloop_out
         type(c_ptr), value :: p
         type(mhx_rloop_phys), intent(in) :: phys
         integer(c_int), value :: nlev
         type(mhx_rloop_radial), intent(in) :: rad
         type(mhx_rloop_fields), intent(in) :: fin
         integer(c_size_t), value :: ld_lm
         type(mhx_rloop_out), intent(in) :: fout
         type(c_ptr), value :: stream
      end function mhx_rloop
   end interface

   type, public, extends(rIter_t) :: rIter_batched_t
      type(rIter_single_t) :: host
      type(c_ptr) :: plan = c_null_ptr
      integer :: nR0 = 0, nR1 = -1                 ! interior levels of this rank
      integer(c_int), allocatable :: nR_arr(:), lR_arr(:)
      real(c_double), allocatable :: dtr(:), dth(:)
   contains
      procedure :: initialize
      procedure :: finalize
      procedure :: radialLoop
   end type rIter_batched_t

contains

   subroutine check(rc, what)
      integer(c_int), intent(in) :: rc
      character(len=*), intent(in) :: what
      character(kind=c_char), pointer :: msg(:)
      character(len=512) :: text
      integer :: i
      if ( rc == 0 ) return
      text = ''
      call c_f_pointer(mhx_last_error(), msg, [512])
      do i=1,512
         if ( msg(i) == c_null_char ) exit
         text(i:i) = msg(i)
      end do
      call abortRun('! mhx '//what//': '//trim(text))
   end subroutine check

   subroutine initialize(this)
      class(rIter_batched_t) :: this
      integer :: nR, n, plen, pstat
      integer(c_int) :: prec
      character(len=16) :: pname

      call this%host%initialize()
      this%nR0 = max(nRstart, n_r_cmb+1)
      this%nR1 = min(nRstop, n_r_icb-1)
      n = max(0, this%nR1-this%nR0+1)
      allocate( this%nR_arr(max(n,1)), this%lR_arr(max(n,1)), this%dtr(max(n,1)), &
      &         this%dth(max(n,1)) )
      do nR=this%nR0,this%nR1
         this%nR_arr(nR-this%nR0+1) = nR
      end do
      ! the plan's precision: FP64 (MHX_PREC_FP64), or the split form
      ! (MHX_PREC_SPLIT16) with the environment variable MHX_PRECISION=split;
      ! the switch sets the fused step covers then run it in that precision
      prec = 0_c_int
      call get_environment_variable('MHX_PRECISION', pname, plen, pstat)
      if ( pstat == 0 .and. trim(pname) == 'split' ) prec = 1_c_int
      call check(mhx_plan_create(int(l_max,c_int), int(m_max,c_int), int(minc,c_int),   &
           &     int(n_theta_max,c_int), int(n_phi_max,c_int), prec, int(max(n,1),c_int), &
           &     this%plan), 'plan')
   end subroutine initialize

   subroutine finalize(this)
      class(rIter_batched_t) :: this
      integer(c_int) :: rc
      call this%host%finalize()
      if ( c_associated(this%plan) ) rc = mhx_plan_destroy(this%plan)
      this%plan = c_null_ptr
      deallocate( this%nR_arr, this%lR_arr, this%dtr, this%dth )
   end subroutine finalize

   function lo(l) result(i)
      logical, intent(in) :: l
      integer(c_int) :: i
      i = 0
      if ( l ) i = 1
   end function lo

   !-- the batched call over interior levels nR0..nR1 on the arrays whose
   !   addresses (first interior level, level stride lm_max) fin and fout hold
   subroutine gpu_levels(this, timeStage, tscheme, lPressNext, fi, fo)
      class(rIter_batched_t), target :: this
      real(cp), intent(in) :: timeStage
      class(type_tscheme), intent(in) :: tscheme
      logical, intent(in) :: lPressNext
      type(mhx_rloop_fields), intent(in) :: fi
      type(mhx_rloop_out), intent(inout) :: fo
      type(mhx_rloop_phys) :: ph
      type(mhx_rloop_radial) :: rd
      integer :: n0, n1, n
      real(c_double), allocatable, target :: rad_copy(:,:)

      n0 = this%nR0; n1 = this%nR1; n = n1-n0+1
      ph%l_conv = lo(l_conv); ph%l_conv_nl = lo(l_conv_nl); ph%l_heat = lo(l_heat)
      ph%l_heat_nl = lo(l_heat_nl); ph%l_chemical_conv = lo(l_chemical_conv)
      ph%l_phase_field = lo(l_phase_field); ph%l_mag = lo(l_mag); ph%l_mag_LF = lo(l_mag_LF)
      ph%l_mag_nl = lo(l_mag_nl); ph%l_mag_kin = lo(l_mag_kin); ph%l_anel = lo(l_anel)
      ph%l_anelastic_liquid = lo(l_anelastic_liquid); ph%l_adv_curl = lo(l_adv_curl)
      ph%l_double_curl = lo(l_double_curl); ph%l_corr = lo(l_corr)
      ph%l_single_matrix = lo(l_single_matrix); ph%l_precession = lo(l_precession)
      ph%l_centrifuge = lo(l_centrifuge); ph%l_cour_alf_damp = lo(l_cour_alf_damp)
      ph%lPressNext = lo(lPressNext); ph%n_r_LCR = int(n_r_LCR, c_int)
      ph%LFfac = LFfac; ph%CorFac = CorFac; ph%epsc = epsc; ph%epscXi = epscXi
      ph%ViscHeatFac = ViscHeatFac; ph%OhmLossFac = OhmLossFac; ph%opm = opm
      ph%courfac = tscheme%courfac; ph%alffac = tscheme%alffac
      ph%epsPhase = epsPhase; ph%penaltyFac = penaltyFac; ph%phaseDiffFac = phaseDiffFac
      ph%tmelt = tmelt; ph%po = po; ph%prec_angle = prec_angle; ph%oek = oek
      ph%timeStage = timeStage; ph%dilution_fac = dilution_fac; ph%ra = ra; ph%opr = opr

      !-- per-level profiles (contiguous copies of the interior levels)
      allocate( rad_copy(n,13) )
      rad_copy(:,1) = r(n0:n1); rad_copy(:,2) = or1(n0:n1); rad_copy(:,3) = or2(n0:n1)
      rad_copy(:,4) = or4(n0:n1); rad_copy(:,5) = orho1(n0:n1); rad_copy(:,6) = beta(n0:n1)
      rad_copy(:,7) = otemp1(n0:n1); rad_copy(:,8) = visc(n0:n1); rad_copy(:,9) = lambda(n0:n1)
      rad_copy(:,10) = temp0(n0:n1); rad_copy(:,11) = epscProf(n0:n1)
      rad_copy(:,12) = delxr2(n0:n1); rad_copy(:,13) = delxh2(n0:n1)
      this%lR_arr(1:n) = int(l_R(n0:n1), c_int)
      rd%r = c_loc(rad_copy(1,1)); rd%or1 = c_loc(rad_copy(1,2)); rd%or2 = c_loc(rad_copy(1,3))
      rd%or4 = c_loc(rad_copy(1,4)); rd%orho1 = c_loc(rad_copy(1,5))
      rd%beta = c_loc(rad_copy(1,6)); rd%otemp1 = c_loc(rad_copy(1,7))
      rd%visc = c_loc(rad_copy(1,8)); rd%lambda = c_loc(rad_copy(1,9))
      rd%temp0 = c_loc(rad_copy(1,10)); rd%epscProf = c_loc(rad_copy(1,11))
      rd%delxr2 = c_loc(rad_copy(1,12)); rd%delxh2 = c_loc(rad_copy(1,13))
      rd%nR = c_loc(this%nR_arr(1)); rd%l_R = c_loc(this%lR_arr(1))
      fo%dtrkc = c_loc(this%dtr(1)); fo%dthkc = c_loc(this%dth(1))

      call check(mhx_rloop(this%plan, ph, int(n,c_int), rd, fi, int(lm_max,c_size_t), fo, &
           &     c_null_ptr), 'radial loop')
      deallocate( rad_copy )
   end subroutine gpu_levels

   !-- address of column n of an R-distributed array (lm_max x levels)
   function col(a, n) result(p)
      complex(cp), target, intent(in) :: a(lm_max,nRstart:nRstop)
      integer, intent(in) :: n
      type(c_ptr) :: p
      p = c_loc(a(1,n))
   end function col

   !-- the same for the magnetic arrays (lm_maxMag x magnetic levels)
   function colm(a, n) result(p)
      complex(cp), target, intent(in) :: a(lm_maxMag,nRstartMag:nRstopMag)
      integer, intent(in) :: n
      type(c_ptr) :: p
      p = c_loc(a(1,n))
   end function colm

   subroutine radialLoop(this,l_graph,l_frame,time,timeStage,tscheme,dtLast,  &
              &          lTOCalc,lTONext,lTONext2,lHelCalc,lPowerCalc,        &
              &          lRmsCalc,lPressCalc,lPressNext,lViscBcCalc,          &
              &          lFluxProfCalc,lPerpParCalc,lGeosCalc,lHemiCalc,      &
              &          lPhaseCalc,l_probe_out,dsdt,dwdt,dzdt,dpdt,dxidt,    &
              &          dphidt,dbdt,djdt,dVxVhLM,dVxBhLM,dVSrLM,dVXirLM,     &
              &          lorentz_torque_ic,lorentz_torque_ma,br_vt_lm_cmb,    &
              &          br_vp_lm_cmb,br_vt_lm_icb,br_vp_lm_icb,dtrkc,dthkc)

      class(rIter_batched_t) :: this
      logical,             intent(in) :: l_graph,l_frame
      logical,             intent(in) :: lTOcalc,lTONext,lTONext2,lHelCalc
      logical,             intent(in) :: lPowerCalc,lHemiCalc
      logical,             intent(in) :: lViscBcCalc,lFluxProfCalc,lPerpParCalc
      logical,             intent(in) :: lRmsCalc,lGeosCalc,lPhaseCalc
      logical,             intent(in) :: l_probe_out
      logical,             intent(in) :: lPressCalc
      logical,             intent(in) :: lPressNext
      real(cp),            intent(in) :: time,timeStage,dtLast
      class(type_tscheme), intent(in) :: tscheme
      complex(cp), intent(out) :: dwdt(lm_max,nRstart:nRstop)
      complex(cp), intent(out) :: dzdt(lm_max,nRstart:nRstop)
      complex(cp), intent(out) :: dpdt(lm_max,nRstart:nRstop)
      complex(cp), intent(out) :: dsdt(lm_max,nRstart:nRstop)
      complex(cp), intent(out) :: dxidt(lm_max,nRstart:nRstop)
      complex(cp), intent(out) :: dphidt(lm_max,nRstart:nRstop)
      complex(cp), intent(out) :: dVSrLM(lm_max,nRstart:nRstop)
      complex(cp), intent(out) :: dVXirLM(lm_max,nRstart:nRstop)
      complex(cp), intent(out) :: dbdt(lm_maxMag,nRstartMag:nRstopMag)
      complex(cp), intent(out) :: djdt(lm_maxMag,nRstartMag:nRstopMag)
      complex(cp), intent(out) :: dVxVhLM(lm_max,nRstart:nRstop)
      complex(cp), intent(out) :: dVxBhLM(lm_maxMag,nRstartMag:nRstopMag)
      real(cp),    intent(out) :: lorentz_torque_ma,lorentz_torque_ic
      complex(cp), intent(out) :: br_vt_lm_cmb(:)
      complex(cp), intent(out) :: br_vp_lm_cmb(:)
      complex(cp), intent(out) :: br_vt_lm_icb(:)
      complex(cp), intent(out) :: br_vp_lm_icb(:)
      real(cp),    intent(out) :: dtrkc(nRstart:nRstop),dthkc(nRstart:nRstop)

      logical :: lDiag, lMagNlBc
      integer :: nR, nBc, n0, k
      type(mhx_rloop_fields) :: fi
      type(mhx_rloop_out) :: fo

      lDiag = l_graph .or. l_frame .or. lTOCalc .or. lTONext .or. lTONext2 .or. lHelCalc .or. &
      &       lPowerCalc .or. lRmsCalc .or. lPressCalc .or. lViscBcCalc .or. lFluxProfCalc .or. &
      &       lPerpParCalc .or. lGeosCalc .or. lHemiCalc .or. lPhaseCalc .or. l_probe_out .or.   &
      &       l_dtB .or. l_onset .or. l_RMS .or. l_full_sphere .or. l_parallel_solve .or.       &
      &       (l_single_matrix .and. l_temperature_diff) .or. this%nR1 < this%nR0
      if ( lDiag ) then
         !-- the reference's own loop over all levels
         call this%host%radialLoop(l_graph,l_frame,time,timeStage,tscheme,dtLast,  &
              &                    lTOCalc,lTONext,lTONext2,lHelCalc,lPowerCalc,   &
              &                    lRmsCalc,lPressCalc,lPressNext,lViscBcCalc,     &
              &                    lFluxProfCalc,lPerpParCalc,lGeosCalc,lHemiCalc, &
              &                    lPhaseCalc,l_probe_out,dsdt,dwdt,dzdt,dpdt,     &
              &                    dxidt,dphidt,dbdt,djdt,dVxVhLM,dVxBhLM,dVSrLM,  &
              &                    dVXirLM,lorentz_torque_ic,lorentz_torque_ma,    &
              &                    br_vt_lm_cmb,br_vp_lm_cmb,br_vt_lm_icb,         &
              &                    br_vp_lm_icb,dtrkc,dthkc)
         return
      end if

      !-- interior levels: one batched call.  Arrays of switched-off physics
      !   are size-1 placeholders in the reference and are not addressed.
      n0 = this%nR0
      fi%w = c_loc(w_Rloc(1,n0)); fi%dw = c_loc(dw_Rloc(1,n0)); fi%ddw = c_loc(ddw_Rloc(1,n0))
      fi%z = c_loc(z_Rloc(1,n0)); fi%dz = c_loc(dz_Rloc(1,n0))
      fi%s = c_null_ptr; fi%xi = c_null_ptr; fi%phi = c_null_ptr
      fi%b = c_null_ptr; fi%db = c_null_ptr; fi%ddb = c_null_ptr
      fi%aj = c_null_ptr; fi%dj = c_null_ptr
      if ( l_heat ) fi%s = c_loc(s_Rloc(1,n0))
      if ( l_chemical_conv ) fi%xi = c_loc(xi_Rloc(1,n0))
      if ( l_phase_field ) fi%phi = col(phi_Rloc, n0)
      if ( l_mag .or. l_mag_LF ) then
         fi%b = c_loc(b_Rloc(1,n0)); fi%db = c_loc(db_Rloc(1,n0)); fi%ddb = c_loc(ddb_Rloc(1,n0))
         fi%aj = c_loc(aj_Rloc(1,n0)); fi%dj = c_loc(dj_Rloc(1,n0))
      end if
      fo%dwdt = col(dwdt, n0); fo%dzdt = col(dzdt, n0); fo%dpdt = c_null_ptr
      if ( (.not. l_double_curl) .or. lPressNext ) fo%dpdt = col(dpdt, n0)
      fo%dsdt = c_null_ptr; fo%dVSrLM = c_null_ptr; fo%dxidt = c_null_ptr
      fo%dVXirLM = c_null_ptr; fo%dphidt = c_null_ptr; fo%dbdt = c_null_ptr
      fo%djdt = c_null_ptr; fo%dVxBhLM = c_null_ptr; fo%dVxVhLM = c_null_ptr
      if ( l_heat ) then
         fo%dsdt = col(dsdt, n0); fo%dVSrLM = col(dVSrLM, n0)
      end if
      if ( l_chemical_conv ) then
         fo%dxidt = col(dxidt, n0); fo%dVXirLM = col(dVXirLM, n0)
      end if
      if ( l_phase_field ) fo%dphidt = col(dphidt, n0)
      if ( l_mag ) then
         fo%dbdt = colm(dbdt, n0); fo%djdt = colm(djdt, n0); fo%dVxBhLM = colm(dVxBhLM, n0)
      end if
      if ( l_double_curl ) fo%dVxVhLM = col(dVxVhLM, n0)
      call gpu_levels(this, timeStage, tscheme, lPressNext, fi, fo)
      do nR=this%nR0,this%nR1
         k = nR-this%nR0+1
         dtrkc(nR) = this%dtr(k)
         dthkc(nR) = this%dth(k)
      end do

      !-- the boundary levels as the reference's loop treats them
      !   (src/rIter.f90:159-188 before the loop, :190-444 for nBc /= 0,
      !   :446-456 after it): no diagnostics are requested in this stage
      if ( nRstart == n_r_cmb ) then
         dtrkc(n_r_cmb)=1.e10_cp
         dthkc(n_r_cmb)=1.e10_cp
         if ( l_mag ) dVxBhLM(:,n_r_cmb)=zero
         if ( l_double_curl ) dVxVhLM(:,n_r_cmb)=zero
      end if
      if ( nRstop == n_r_icb ) then
         dtrkc(n_r_icb)=1.e10_cp
         dthkc(n_r_icb)=1.e10_cp
         if ( l_mag ) dVxBhLM(:,n_r_icb)=zero
         if ( l_double_curl ) dVxVhLM(:,n_r_icb)=zero
      end if
      lorentz_torque_ma = 0.0_cp
      lorentz_torque_ic = 0.0_cp
      lMagNlBc=.false.
      if ( ( l_mag_nl .or. l_mag_kin ) .and.                          &
           &       ( ktopv == 1 .or. l_cond_ma .or.                   &
           &          ( ktopv == 2 .and. l_rot_ma ) ) .or.            &
           &       ( kbotv == 1 .or. l_cond_ic .or.                   &
           &          ( kbotv == 2 .and. l_rot_ic ) ) )               &
           &     lMagNlBc=.true.

      do nR=nRstart,nRstop
         if ( nR /= n_r_cmb .and. nR /= n_r_icb ) cycle
         if ( nR == n_r_cmb ) then
            nBc = ktopv
         else
            nBc = kbotv
         end if
         dtrkc(nR)=1e10_cp
         dthkc(nR)=1e10_cp
         call this%host%nl_lm%set_zero()
         call this%host%transform_to_grid_space(nR, nBc, .false., .false., .false., .false., &
              &                                 .false., .false., .false., .false., .false.,  &
              &                                 .false., .false., .false.)
         if ( lMagNlBc ) then
            call this%host%gsa%get_nl(timeStage, nR, nBc, .false.)
            call lm_space(this, nR, dVSrLM, dVXirLM, dphidt)
         else if ( l_mag ) then
            this%host%nl_lm%VxBtLM(:)=zero
            this%host%nl_lm%VxBpLM(:)=zero
         end if
         if ( nR == n_r_cmb .and. l_b_nl_cmb ) then
            br_vt_lm_cmb(:)=zero
            br_vp_lm_cmb(:)=zero
            call get_br_v_bcs(this%host%gsa%brc, this%host%gsa%vtc, this%host%gsa%vpc, &
                 &            omega_ma, or2(nR), orho1(nR), br_vt_lm_cmb, br_vp_lm_cmb)
         else if ( nR == n_r_icb .and. l_b_nl_icb ) then
            br_vt_lm_icb(:)=zero
            br_vp_lm_icb(:)=zero
            call get_br_v_bcs(this%host%gsa%brc, this%host%gsa%vtc, this%host%gsa%vpc, &
                 &            omega_ic, or2(nR), orho1(nR), br_vt_lm_icb, br_vp_lm_icb)
         end if
         if ( nR == n_r_icb .and. l_mag_LF .and. l_rot_ic .and. l_cond_ic  ) then
            call get_lorentz_torque(lorentz_torque_ic, this%host%gsa%brc, this%host%gsa%bpc, nR)
         end if
         if ( nR == n_r_cmb .and. l_mag_LF .and. l_rot_ma .and. l_cond_ma ) then
            call get_lorentz_torque(lorentz_torque_ma, this%host%gsa%brc, this%host%gsa%bpc, nR)
         end if
         call courant(nR, dtrkc(nR), dthkc(nR), this%host%gsa%vrc, this%host%gsa%vtc, &
              &       this%host%gsa%vpc, this%host%gsa%brc, this%host%gsa%btc,       &
              &       this%host%gsa%bpc, tscheme%courfac, tscheme%alffac)
         if ( l_conv ) then
            call this%host%nl_lm%get_dzdt(nR, nBc, dzdt(:,nR))
            if ( l_double_curl ) then
               call this%host%nl_lm%get_dwdt_double_curl(nR, nBc, dwdt(:,nR), dVxVhLM(:,nR))
            else
               call this%host%nl_lm%get_dwdt(nR, nBc, dwdt(:,nR))
            end if
         end if
         if ( (.not. l_double_curl) .or. lPressNext ) then
            call this%host%nl_lm%get_dpdt(nR, nBc, dpdt(:,nR))
         end if
         if ( l_heat ) call this%host%nl_lm%get_dsdt(nR, nBc, dsdt(:,nR), dVSrLM(:,nR))
         if ( l_chemical_conv ) then
            call this%host%nl_lm%get_dxidt(nBc, dxidt(:,nR), dVXirLM(:,nR))
         end if
         if ( l_mag ) then
            call this%host%nl_lm%get_dbdt(nR, nBc, dbdt(:,nR), djdt(:,nR), dVxBhLM(:,nR))
         end if
      end do

      if ( nRstart == n_r_cmb ) then
         if ( l_heat ) dVSrLM(:,n_r_cmb) =zero
         if ( l_chemical_conv ) dVXirLM(:,n_r_cmb)=zero
      end if
      if ( nRstop == n_r_icb ) then
         if ( l_heat ) dVSrLM(:,n_r_icb) =zero
         if ( l_chemical_conv ) dVXirLM(:,n_r_icb)=zero
      end if
      lorentz_torque_ma=-lorentz_torque_ma

   end subroutine radialLoop

   !-- transform_to_lm_space of a boundary level; the composition and phase
   !   outputs are size-1 placeholders unless their physics is on
   subroutine lm_space(this, nR, dVSrLM, dVXirLM, dphidt)
      class(rIter_batched_t) :: this
      integer, intent(in) :: nR
      complex(cp), intent(inout) :: dVSrLM(lm_max,nRstart:nRstop)
      complex(cp), intent(inout) :: dVXirLM(:,:), dphidt(:,:)
      complex(cp), allocatable :: dummy(:)
      if ( l_chemical_conv .and. l_phase_field ) then
         call this%host%transform_to_lm_space(nR, .false., dVSrLM(:,nR), dVXirLM(:,nR-nRstart+1), &
              &                               dphidt(:,nR-nRstart+1))
      else
         allocate( dummy(lm_max) )
         if ( l_chemical_conv ) then
            call this%host%transform_to_lm_space(nR, .false., dVSrLM(:,nR), dVXirLM(:,nR-nRstart+1), &
                 &                               dummy)
         else if ( l_phase_field ) then
            call this%host%transform_to_lm_space(nR, .false., dVSrLM(:,nR), dummy, &
                 &                               dphidt(:,nR-nRstart+1))
         else
            call this%host%transform_to_lm_space(nR, .false., dVSrLM(:,nR), dummy, dummy)
         end if
         deallocate( dummy )
      end if
   end subroutine lm_space

end module rIter_mhx_mod
