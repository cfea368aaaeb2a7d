!
! Test infrastructure only: an instrumented module radialLoop with the
! public interface of src/radialLoop.f90, linked into the reference's own
! program (oracle/_ref/magic_dump) in place of that file.  It runs the
! reference's radial loop (rIter_single_t of src/rIter.f90) unchanged and, on the call
! number given by the environment variable MHX_DUMP_CALL, writes that call's
! inputs (the R-distributed fields, the radial profiles, the physics switches
! and parameters the loop reads) and outputs (the explicit terms, the Courant
! steps, the boundary products) to the file MHX_DUMP_FILE.  These are the
! golden vectors of the radial loop under the run's physics switches.
!
! File format (little-endian stream): records of
!   name (character*32), kind (int32: 1 real64, 2 complex128, 3 int32),
!   rank (int32), shape (rank x int32), data (column-major)
! ending with the name 'END'.
!
module radialLoop

   use iso_fortran_env, only: int32, real64
   use precision_mod
   use mem_alloc, only: memWrite, bytes_allocated
   use truncation, only: lm_max, lm_maxMag, l_max, l_maxMag, n_r_max, minc, m_max, &
       &                 n_theta_max, n_phi_max
   use radial_data, only: nRstart, nRstop, nRstartMag, nRstopMag, n_r_cmb, n_r_icb
   use time_schemes, only: type_tscheme
   use rIteration, only: rIter_t
   use rIter_mod, only: rIter_single_t
   use logic
   use fields
   use radial_functions
   use physical_parameters
   use num_param, only: delxr2, delxh2
   use special, only: l_vr_cmb, l_vr_icb
   use constants, only: zero

   implicit none

   private

   public :: initialize_radialLoop, finalize_radialLoop, radialLoopG

   class(rIter_t), pointer :: rIter
   integer :: n_calls = 0

contains

   subroutine initialize_radialLoop
      integer(lip) :: local_bytes_used
      local_bytes_used = bytes_allocated
      allocate( rIter_single_t :: rIter )
      call rIter%initialize()
      local_bytes_used = bytes_allocated-local_bytes_used
      call memWrite('radialLoop.f90', local_bytes_used)
   end subroutine initialize_radialLoop

   subroutine finalize_radialLoop
      call rIter%finalize()
      deallocate(rIter)
   end subroutine finalize_radialLoop

   subroutine rec_head(u, name, kind, shp)
      integer, intent(in) :: u, kind
      character(len=*), intent(in) :: name
      integer, intent(in) :: shp(:)
      character(len=32) :: nm
      nm = name
      write(u) nm, int(kind,int32), int(size(shp),int32), int(shp,int32)
   end subroutine rec_head

   subroutine put_r1(u, name, a)
      integer, intent(in) :: u
      character(len=*), intent(in) :: name
      real(cp), intent(in) :: a(:)
      call rec_head(u, name, 1, [size(a)])
      write(u) a
   end subroutine put_r1

   subroutine put_r0(u, name, a)
      integer, intent(in) :: u
      character(len=*), intent(in) :: name
      real(cp), intent(in) :: a
      call rec_head(u, name, 1, [1])
      write(u) a
   end subroutine put_r0

   subroutine put_i0(u, name, a)
      integer, intent(in) :: u
      character(len=*), intent(in) :: name
      integer, intent(in) :: a
      call rec_head(u, name, 3, [1])
      write(u) int(a,int32)
   end subroutine put_i0

   subroutine put_l0(u, name, a)
      integer, intent(in) :: u
      character(len=*), intent(in) :: name
      logical, intent(in) :: a
      integer :: v
      v = 0
      if ( a ) v = 1
      call put_i0(u, name, v)
   end subroutine put_l0

   subroutine put_i1(u, name, a)
      integer, intent(in) :: u
      character(len=*), intent(in) :: name
      integer, intent(in) :: a(:)
      call rec_head(u, name, 3, [size(a)])
      write(u) int(a,int32)
   end subroutine put_i1

   subroutine put_c2(u, name, a)
      integer, intent(in) :: u
      character(len=*), intent(in) :: name
      complex(cp), intent(in) :: a(:,:)
      call rec_head(u, name, 2, shape(a))
      write(u) a
   end subroutine put_c2

   subroutine put_c1(u, name, a)
      integer, intent(in) :: u
      character(len=*), intent(in) :: name
      complex(cp), intent(in) :: a(:)
      call rec_head(u, name, 2, [size(a)])
      write(u) a
   end subroutine put_c1

   subroutine dump_inputs(u, time, timeStage, tscheme)
      integer, intent(in) :: u
      real(cp), intent(in) :: time, timeStage
      class(type_tscheme), intent(in) :: tscheme
      call put_i0(u, 'l_max', l_max); call put_i0(u, 'm_max', m_max)
      call put_i0(u, 'minc', minc); call put_i0(u, 'lm_max', lm_max)
      call put_i0(u, 'n_theta_max', n_theta_max); call put_i0(u, 'n_phi_max', n_phi_max)
      call put_i0(u, 'n_r_max', n_r_max); call put_i0(u, 'nRstart', nRstart)
      call put_i0(u, 'nRstop', nRstop); call put_i0(u, 'n_r_cmb', n_r_cmb)
      call put_i0(u, 'n_r_icb', n_r_icb); call put_i0(u, 'n_r_LCR', n_r_LCR)
      call put_i0(u, 'ktopv', ktopv); call put_i0(u, 'kbotv', kbotv)
      call put_i0(u, 'ktops', ktops); call put_i0(u, 'kbots', kbots)
      call put_l0(u, 'l_mag', l_mag); call put_l0(u, 'l_mag_LF', l_mag_LF)
      call put_l0(u, 'l_mag_nl', l_mag_nl); call put_l0(u, 'l_mag_kin', l_mag_kin)
      call put_l0(u, 'l_conv', l_conv); call put_l0(u, 'l_conv_nl', l_conv_nl)
      call put_l0(u, 'l_heat', l_heat); call put_l0(u, 'l_heat_nl', l_heat_nl)
      call put_l0(u, 'l_chemical_conv', l_chemical_conv)
      call put_l0(u, 'l_phase_field', l_phase_field)
      call put_l0(u, 'l_precession', l_precession); call put_l0(u, 'l_centrifuge', l_centrifuge)
      call put_l0(u, 'l_anel', l_anel); call put_l0(u, 'l_anelastic_liquid', l_anelastic_liquid)
      call put_l0(u, 'l_adv_curl', l_adv_curl); call put_l0(u, 'l_double_curl', l_double_curl)
      call put_l0(u, 'l_corr', l_corr); call put_l0(u, 'l_single_matrix', l_single_matrix)
      call put_l0(u, 'l_cond_ic', l_cond_ic); call put_l0(u, 'l_cond_ma', l_cond_ma)
      call put_l0(u, 'l_rot_ic', l_rot_ic); call put_l0(u, 'l_rot_ma', l_rot_ma)
      call put_l0(u, 'l_b_nl_cmb', l_b_nl_cmb); call put_l0(u, 'l_b_nl_icb', l_b_nl_icb)
      call put_l0(u, 'l_full_sphere', l_full_sphere); call put_l0(u, 'l_HT', l_HT)
      call put_l0(u, 'l_parallel_solve', l_parallel_solve)
      call put_l0(u, 'l_temperature_diff', l_temperature_diff)
      call put_l0(u, 'l_cour_alf_damp', l_cour_alf_damp)
      call put_l0(u, 'l_vr_cmb', l_vr_cmb); call put_l0(u, 'l_vr_icb', l_vr_icb)
      call put_r0(u, 'time', time); call put_r0(u, 'timeStage', timeStage)
      call put_r0(u, 'courfac', tscheme%courfac); call put_r0(u, 'alffac', tscheme%alffac)
      call put_r0(u, 'LFfac', LFfac); call put_r0(u, 'CorFac', CorFac)
      call put_r0(u, 'epsc', epsc); call put_r0(u, 'epscXi', epscXi)
      call put_r0(u, 'ViscHeatFac', ViscHeatFac); call put_r0(u, 'OhmLossFac', OhmLossFac)
      call put_r0(u, 'opm', opm); call put_r0(u, 'r_cmb', r_cmb); call put_r0(u, 'r_icb', r_icb)
      call put_r0(u, 'omega_ic', omega_ic); call put_r0(u, 'omega_ma', omega_ma)
      call put_r0(u, 'epsPhase', epsPhase); call put_r0(u, 'penaltyFac', penaltyFac)
      call put_r0(u, 'phaseDiffFac', phaseDiffFac); call put_r0(u, 'tmelt', tmelt)
      call put_r0(u, 'po', po); call put_r0(u, 'prec_angle', prec_angle)
      call put_r0(u, 'oek', oek); call put_r0(u, 'dilution_fac', dilution_fac)
      call put_r0(u, 'ra', ra); call put_r0(u, 'opr', opr)
      call put_r1(u, 'r', r); call put_r1(u, 'or1', or1); call put_r1(u, 'or2', or2)
      call put_r1(u, 'or3', or3); call put_r1(u, 'or4', or4)
      call put_r1(u, 'orho1', orho1); call put_r1(u, 'orho2', orho2); call put_r1(u, 'rho0', rho0)
      call put_r1(u, 'beta', beta); call put_r1(u, 'otemp1', otemp1); call put_r1(u, 'temp0', temp0)
      call put_r1(u, 'visc', visc); call put_r1(u, 'lambda', lambda)
      call put_r1(u, 'epscProf', epscProf); call put_i1(u, 'l_R', l_R)
      call put_r1(u, 'delxr2', delxr2); call put_r1(u, 'delxh2', delxh2)
      call put_c2(u, 'w', w_Rloc); call put_c2(u, 'dw', dw_Rloc); call put_c2(u, 'ddw', ddw_Rloc)
      call put_c2(u, 'z', z_Rloc); call put_c2(u, 'dz', dz_Rloc)
      call put_c2(u, 'p', p_Rloc)
      if ( l_heat ) then
         call put_c2(u, 's', s_Rloc); call put_c2(u, 'ds', ds_Rloc)
      end if
      if ( l_chemical_conv ) call put_c2(u, 'xi', xi_Rloc)
      if ( l_phase_field ) call put_c2(u, 'phi', phi_Rloc)
      if ( l_mag ) then
         call put_c2(u, 'b', b_Rloc); call put_c2(u, 'db', db_Rloc); call put_c2(u, 'ddb', ddb_Rloc)
         call put_c2(u, 'aj', aj_Rloc); call put_c2(u, 'dj', dj_Rloc)
      end if
   end subroutine dump_inputs

   subroutine radialLoopG(l_graph,l_frame,time,timeStage,tscheme,dtLast,   &
              &          lTOCalc,lTONext,lTONext2,lHelCalc,lPowerCalc,     &
              &          lRmsCalc,lPressCalc,lPressNext,lViscBcCalc,       &
              &          lFluxProfCalc,lPerpParCalc,lGeosCalc,lHemiCalc,   &
              &          lPhaseCalc,l_probe_out,dsdt,dwdt,dzdt,dpdt,dxidt, &
              &          dphidt,dbdt,djdt,dVxVhLM,dVxBhLM,dVSrLM,dVXirLM,  &
              &          lorentz_torque_ic,lorentz_torque_ma,br_vt_lm_cmb, &
              &          br_vp_lm_cmb,br_vt_lm_icb,br_vp_lm_icb,dtrkc,dthkc)

      logical,             intent(in) :: l_graph,l_frame
      logical,             intent(in) :: lTOcalc,lTONext,lTONext2,lHelCalc
      logical,             intent(in) :: lPowerCalc,lGeosCalc
      logical,             intent(in) :: lViscBcCalc,lFluxProfCalc,lPerpParCalc
      logical,             intent(in) :: lRmsCalc,lHemiCalc,lPhaseCalc
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
      complex(cp), intent(out) :: br_vt_lm_cmb(lm_max)
      complex(cp), intent(out) :: br_vp_lm_cmb(lm_max)
      complex(cp), intent(out) :: br_vt_lm_icb(lm_max)
      complex(cp), intent(out) :: br_vp_lm_icb(lm_max)
      real(cp),    intent(out) :: lorentz_torque_ma,lorentz_torque_ic
      real(cp),    intent(out) :: dtrkc(nRstart:nRstop),dthkc(nRstart:nRstop)

      character(len=512) :: fname, arg
      integer :: want, u, st
      logical :: do_dump

      n_calls = n_calls+1
      want = -1
      call get_environment_variable('MHX_DUMP_CALL', arg, status=st)
      if ( st == 0 ) read(arg,*) want
      do_dump = ( want == n_calls )
      if ( do_dump ) then
         call get_environment_variable('MHX_DUMP_FILE', fname)
         open(newunit=u, file=trim(fname), access='stream', form='unformatted', &
         &    status='replace')
         call dump_inputs(u, time, timeStage, tscheme)
         call put_l0(u, 'lDiag', l_graph .or. l_frame .or. lTOCalc .or. lTONext .or.    &
         &           lTONext2 .or. lHelCalc .or. lPowerCalc .or. lRmsCalc .or.      &
         &           lPressCalc .or. lViscBcCalc .or. lFluxProfCalc .or.            &
         &           lPerpParCalc .or. lGeosCalc .or. lHemiCalc .or. lPhaseCalc .or. &
         &           l_probe_out)
         call put_l0(u, 'lPressNext', lPressNext)
      end if

      !-- outputs the reference leaves unset at some levels start from zero,
      !   so that the dump is defined everywhere; arrays of switched-off
      !   physics are placeholders of size 1 and are not touched
      if ( do_dump ) then
         dwdt(:,:)=zero; dzdt(:,:)=zero; dpdt(:,:)=zero
         if ( l_heat ) then
            dsdt(:,:)=zero; dVSrLM(:,:)=zero
         end if
         if ( l_chemical_conv ) then
            dxidt(:,:)=zero; dVXirLM(:,:)=zero
         end if
         if ( l_phase_field ) dphidt(:,:)=zero
         if ( l_mag ) then
            dbdt(:,:)=zero; djdt(:,:)=zero; dVxBhLM(:,:)=zero
         end if
         if ( l_double_curl ) dVxVhLM(:,:)=zero
         br_vt_lm_cmb(:)=zero; br_vp_lm_cmb(:)=zero; br_vt_lm_icb(:)=zero
         br_vp_lm_icb(:)=zero
      end if

      call rIter%radialLoop(l_graph,l_frame,time,timeStage,tscheme,dtLast,    &
           &             lTOCalc,lTONext,lTONext2,lHelCalc,lPowerCalc,        &
           &             lRmsCalc,lPressCalc,lPressNext,lViscBcCalc,          &
           &             lFluxProfCalc,lPerpParCalc,lGeosCalc,lHemiCalc,      &
           &             lPhaseCalc,l_probe_out,dsdt,dwdt,dzdt,dpdt,dxidt,    &
           &             dphidt,dbdt,djdt,dVxVhLM,dVxBhLM,dVSrLM,dVXirLM,     &
           &             lorentz_torque_ic,lorentz_torque_ma, br_vt_lm_cmb,   &
           &             br_vp_lm_cmb,br_vt_lm_icb,br_vp_lm_icb,dtrkc,dthkc)

      if ( do_dump ) then
         call put_c2(u, 'dwdt', dwdt); call put_c2(u, 'dzdt', dzdt)
         call put_c2(u, 'dpdt', dpdt)
         if ( l_heat ) then
            call put_c2(u, 'dsdt', dsdt); call put_c2(u, 'dVSrLM', dVSrLM)
         end if
         if ( l_chemical_conv ) then
            call put_c2(u, 'dxidt', dxidt); call put_c2(u, 'dVXirLM', dVXirLM)
         end if
         if ( l_phase_field ) call put_c2(u, 'dphidt', dphidt)
         if ( l_mag ) then
            call put_c2(u, 'dbdt', dbdt); call put_c2(u, 'djdt', djdt)
            call put_c2(u, 'dVxBhLM', dVxBhLM)
         end if
         if ( l_double_curl ) call put_c2(u, 'dVxVhLM', dVxVhLM)
         call put_c1(u, 'br_vt_lm_cmb', br_vt_lm_cmb); call put_c1(u, 'br_vp_lm_cmb', br_vp_lm_cmb)
         call put_c1(u, 'br_vt_lm_icb', br_vt_lm_icb); call put_c1(u, 'br_vp_lm_icb', br_vp_lm_icb)
         call put_r0(u, 'lorentz_torque_ic', lorentz_torque_ic)
         call put_r0(u, 'lorentz_torque_ma', lorentz_torque_ma)
         call put_r1(u, 'dtrkc', dtrkc); call put_r1(u, 'dthkc', dthkc)
         call rec_head(u, 'END', 3, [0])
         close(u)
      end if

   end subroutine radialLoopG

end module radialLoop
