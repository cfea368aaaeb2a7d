!
! Driver program around the reference's own native spherical-harmonic
! transform modules (module sht of src/sht_native.f90, src/shtransforms.f90,
! src/plms.f90, src/fft.f90, src/horizontal.f90) and its grid-space products
! (src/get_nl.f90).  It is test infrastructure only: it is linked against
! object files compiled from the reference's sources by oracle/Makefile and
! produces golden vectors and the CPU timing baseline.  None of the product
! code calls it.
!
! It also drives the reference's lm-space assembly of the explicit terms
! (src/get_td.f90, module nonlinear_lm_mod) in mode td and its Courant
! step limits (src/courant.f90) on the step's grids in mode courant, and
! the boundary products of a conducting mantle or inner core
! (src/nonlinear_bcs.f90, get_br_v_bcs and get_b_nl_bcs) in mode brv; mode rloop chains the
! three for the bulk levels of the radial loop (src/rIter.f90:188-470).
!
! Usage:  ref_harness MODE l_max minc nalias infile outfile [nlev nrep]
!   MODE = tables | api | nl | nlstep | bench | td | courant | brv | rloop
!
! File formats are little-endian streams: int32 header words followed by
! float64 arrays in Fortran (column-major) order.
!
program ref_harness

   use iso_fortran_env, only: output_unit, int32, int64, real64
   use omp_lib
   use precision_mod, only: cp
   use parallel_mod, only: parallel
   use logic
   use truncation
   use blocking, only: initialize_blocking, st_map
   use horizontal_data
   use radial_functions, only: or1, or2, or3, or4, orho1, orho2, beta, r, otemp1, &
       &                       visc, lambda, epscProf, l_R
   use physical_parameters, only: LFfac, n_r_LCR, CorFac, epsc, opm, conductance_ma, &
       &                          prmag, sigma_ratio
   use num_param, only: delxr2, delxh2
   use courant_mod, only: courant
   use nonlinear_bcs, only: get_br_v_bcs, get_b_nl_bcs
   use fields, only: w_Rloc, dw_Rloc, ddw_Rloc, z_Rloc, dz_Rloc
   use nonlinear_lm_mod, only: nonlinear_lm_t
   use sht
   use shtransforms, only: Plm, dPlm, wPlm, wdPlm
   use grid_space_arrays_mod, only: grid_space_arrays_t

   implicit none

   character(len=32) :: mode, arg
   character(len=512) :: infile, outfile
   integer :: lmax_in, minc_in, nalias_in, nlev, nrep
   logical :: l_scr

   call get_command_argument(1, mode)
   call get_command_argument(2, arg); read(arg,*) lmax_in
   call get_command_argument(3, arg); read(arg,*) minc_in
   call get_command_argument(4, arg); read(arg,*) nalias_in
   call get_command_argument(5, infile)
   call get_command_argument(6, outfile)
   nlev = 1
   nrep = 1
   if ( command_argument_count() >= 7 ) then
      call get_command_argument(7, arg); read(arg,*) nlev
   end if
   if ( command_argument_count() >= 8 ) then
      call get_command_argument(8, arg); read(arg,*) nrep
   end if

   !-- switches the reference reads from its namelists; set for a
   !   Boussinesq MHD run with curl-form advection
   l_finite_diff=.false.; l_cond_ic=.false.; l_save_out=.false.
   l_anel=.false.; l_anelastic_liquid=.false.; l_chemical_conv=.false.
   l_phase_field=.false.; l_precession=.false.; l_centrifuge=.false.
   l_adv_curl=.true.; l_heat_nl=.true.; l_mag_nl=.true.; l_conv_nl=.true.
   l_mag_LF=.true.; l_scramble_theta=.true.; l_non_rot=.false.
   l_mag=.true.

   call parallel()

   l_axi=.false.
   l_max=lmax_in; m_max=0; m_min=0; minc=minc_in; nalias=nalias_in
   n_r_max=1; n_r_ic_max=1; lMagMem=1
   call initialize_truncation()
   call initialize_blocking()
   call initialize_horizontal_data()
   call initialize_sht(l_scr)
   l_scramble_theta=l_scr
   call horizontal()

   write(output_unit,'(A,5I7)') ' ref_harness: l_max,m_max,n_theta,n_phi,lm_max=', &
   &     l_max, m_max, n_theta_max, n_phi_max, lm_max

   select case (trim(mode))
   case ('tables')
      call do_tables()
   case ('api')
      call do_api()
   case ('nl')
      call do_nl()
   case ('bench')
      call do_bench()
   case ('nlstep')
      call do_nlstep()
   case ('td')
      call do_td()
   case ('courant')
      call do_courant()
   case ('brv')
      call do_brv()
   case ('rloop')
      call do_rloop()
   case default
      write(output_unit,*) 'unknown mode ', trim(mode)
      stop 2
   end select

contains

   subroutine write_header(u)
      integer, intent(in) :: u
      write(u) int(l_max,int32), int(m_max,int32), int(minc,int32), &
      &        int(n_theta_max,int32), int(n_phi_max,int32), int(lm_max,int32)
   end subroutine write_header

   subroutine do_tables()
      integer :: u
      open(newunit=u, file=trim(outfile), access='stream', form='unformatted', &
      &    status='replace')
      call write_header(u)
      write(u) theta_ord(:), gauss(:), O_sin_theta_E2(:), cosTheta(:), sinTheta(:)
      write(u) Plm(:,:), dPlm(:,:), wPlm(:,:), wdPlm(:,:)
      write(u) real(st_map%lm2l(:),real64), real(st_map%lm2m(:),real64)
      close(u)
   end subroutine do_tables

   subroutine do_api()
      !-- reads nlev levels; per level: lcut, r, r_icb, or2, 11 coefficient
      !   arrays (complex, lm_max) and 3 grids (n_theta, n_phi); writes every
      !   public transform of module sht applied to them.
      integer :: u, v, lev, i
      integer(int32) :: lcut32
      integer :: lcut
      real(cp) :: rr, ricb, o2
      complex(cp), allocatable :: A(:,:), qLM(:), sLM(:), tLM(:)
      real(cp), allocatable :: G(:,:,:), Gw(:,:,:), g1(:,:), g2(:,:), g3(:,:)
      real(cp), allocatable :: ax(:), ax2(:)

      allocate( A(lm_max,11), qLM(lm_max), sLM(lm_max), tLM(lm_max) )
      allocate( G(n_theta_max,n_phi_max,3), Gw(n_theta_max,n_phi_max,3) )
      allocate( g1(n_theta_max,n_phi_max), g2(n_theta_max,n_phi_max) )
      allocate( g3(n_theta_max,n_phi_max), ax(n_theta_max), ax2(n_theta_max) )

      open(newunit=u, file=trim(infile), access='stream', form='unformatted', &
      &    status='old')
      open(newunit=v, file=trim(outfile), access='stream', form='unformatted', &
      &    status='replace')
      call write_header(v)
      do lev=1,nlev
         read(u) lcut32, rr, ricb, o2
         lcut=lcut32
         do i=1,11
            read(u) A(:,i)
         end do
         read(u) G(:,:,:)

         call scal_to_spat(A(:,1), g1, lcut);                       write(v) g1
         call scal_to_grad_spat(A(:,1), g1, g2, lcut);              write(v) g1, g2
         call pol_to_grad_spat(A(:,1), g1, g2, lcut);               write(v) g1, g2
         call torpol_to_spat(A(:,1), A(:,2), A(:,3), g1, g2, g3, lcut)
         write(v) g1, g2, g3
         call pol_to_curlr_spat(A(:,1), g1, lcut);                  write(v) g1
         call torpol_to_curl_spat(o2, A(:,1), A(:,2), A(:,3), A(:,4), &
         &                        g1, g2, g3, lcut)
         write(v) g1, g2, g3
         call torpol_to_dphspat(A(:,1), A(:,2), g1, g2, lcut);      write(v) g1, g2
         call sphtor_to_spat(A(:,1), A(:,2), g1, g2, lcut);         write(v) g1, g2
         call torpol_to_spat_IC(rr, ricb, A(:,1), A(:,2), A(:,3), g1, g2, g3)
         write(v) g1, g2, g3
         call torpol_to_curl_spat_IC(rr, ricb, A(:,1), A(:,2), A(:,3), A(:,4), &
         &                           g1, g2, g3)
         write(v) g1, g2, g3
         call axi_to_spat(A(1:l_max+1,1), ax);                      write(v) ax
         call toraxi_to_spat(A(1:l_max+1,1), ax, ax2, lcut);        write(v) ax, ax2

         Gw(:,:,:)=G(:,:,:)
         call scal_to_SH(Gw(:,:,1), qLM, lcut);                     write(v) qLM
         Gw(:,:,:)=G(:,:,:)
         call spat_to_sphertor(Gw(:,:,1), Gw(:,:,2), qLM, sLM, lcut)
         write(v) qLM, sLM
         Gw(:,:,:)=G(:,:,:)
         call spat_to_qst(Gw(:,:,1), Gw(:,:,2), Gw(:,:,3), qLM, sLM, tLM, lcut)
         write(v) qLM, sLM, tLM
      end do
      close(u)
      close(v)
   end subroutine do_api

   subroutine set_radial(o2, o4, orh)
      real(cp), intent(in) :: o2, o4, orh
      if ( .not. allocated(or2) ) then
         allocate( or1(1), or2(1), or3(1), or4(1), orho1(1), beta(1), r(1) )
         allocate( otemp1(1), visc(1), lambda(1) )
      end if
      or2(1)=o2; or4(1)=o4; orho1(1)=orh
      r(1)=one_over_sqrt(o2); or1(1)=sqrt(o2); or3(1)=o2*sqrt(o2)
      beta(1)=0.0_cp; otemp1(1)=1.0_cp; visc(1)=1.0_cp; lambda(1)=1.0_cp
      n_r_LCR=0
   end subroutine set_radial

   real(cp) function one_over_sqrt(x)
      real(cp), intent(in) :: x
      one_over_sqrt=1.0_cp/sqrt(x)
   end function one_over_sqrt

   subroutine do_nl()
      !-- reads: or2, or4, orho1, LFfac, then 13 grids
      !   (vr vt vp cvr cvt cvp br bt bp cbr cbt cbp s); writes
      !   Adv(3) LF(3) VxB(3) VS(3) from the reference's get_nl.
      type(grid_space_arrays_t) :: gsa
      integer :: u, v
      real(cp) :: o2, o4, orh, lff

      open(newunit=u, file=trim(infile), access='stream', form='unformatted', &
      &    status='old')
      read(u) o2, o4, orh, lff
      call set_radial(o2, o4, orh)
      LFfac=lff
      call gsa%initialize()
      read(u) gsa%vrc, gsa%vtc, gsa%vpc, gsa%cvrc, gsa%cvtc, gsa%cvpc
      read(u) gsa%brc, gsa%btc, gsa%bpc, gsa%cbrc, gsa%cbtc, gsa%cbpc, gsa%sc
      close(u)
      call gsa%get_nl(0.0_cp, 1, 0, .false.)
      open(newunit=v, file=trim(outfile), access='stream', form='unformatted', &
      &    status='replace')
      call write_header(v)
      write(v) gsa%Advr, gsa%Advt, gsa%Advp, gsa%LFr, gsa%LFt, gsa%LFp
      write(v) gsa%VxBr, gsa%VxBt, gsa%VxBp, gsa%VSr, gsa%VSt, gsa%VSp
      close(v)
   end subroutine do_nl

   subroutine nl_level(gsa, C, lcut, O, uGs)
      !-- one bulk level of the nonlinear-term path in the reference's call
      !   order (src/rIter.f90:483-707, l_adv_curl, derivative fields on)
      type(grid_space_arrays_t), intent(inout) :: gsa
      complex(cp), intent(in) :: C(:,:)
      integer, intent(in) :: lcut
      complex(cp), intent(out) :: O(:,:)
      real(cp), intent(inout) :: uGs(:,:)
      integer :: i
      call scal_to_spat(C(:,6), gsa%sc, lcut)
      call scal_to_grad_spat(C(:,6), gsa%dsdtc, gsa%dsdpc, lcut)
      call torpol_to_spat(C(:,1), C(:,2), C(:,4), gsa%vrc, gsa%vtc, gsa%vpc, lcut)
      call torpol_to_curl_spat(or2(1), C(:,1), C(:,3), C(:,4), C(:,5), &
      &                        gsa%cvrc, gsa%cvtc, gsa%cvpc, lcut)
      call torpol_to_spat(C(:,2), C(:,3), C(:,5), gsa%dvrdrc, gsa%dvtdrc, &
      &                   gsa%dvpdrc, lcut)
      call torpol_to_spat(C(:,7), C(:,8), C(:,10), gsa%brc, gsa%btc, gsa%bpc, lcut)
      call torpol_to_curl_spat(or2(1), C(:,7), C(:,9), C(:,10), C(:,11), &
      &                        gsa%cbrc, gsa%cbtc, gsa%cbpc, lcut)
      call gsa%get_nl(0.0_cp, 1, 0, .false.)
      !$omp parallel do default(shared)
      do i=1,n_phi_max
         gsa%Advr(:,i)=gsa%Advr(:,i)+gsa%LFr(:,i)
         gsa%Advt(:,i)=gsa%Advt(:,i)+gsa%LFt(:,i)
         gsa%Advp(:,i)=gsa%Advp(:,i)+gsa%LFp(:,i)
         uGs(:,i)=or2(1)*O_sin_theta_E2(:)*(gsa%vtc(:,i)*gsa%dsdtc(:,i)+ &
         &        gsa%vpc(:,i)*gsa%dsdpc(:,i))
      end do
      !$omp end parallel do
      call spat_to_qst(gsa%Advr, gsa%Advt, gsa%Advp, O(:,1), O(:,2), O(:,3), lcut)
      call spat_to_qst(gsa%VSr, gsa%VSt, gsa%VSp, O(:,4), O(:,5), O(:,6), lcut)
      call spat_to_qst(gsa%VxBr, gsa%VxBt, gsa%VxBp, O(:,7), O(:,8), O(:,9), lcut)
      call scal_to_SH(uGs, O(:,10), lcut)
   end subroutine nl_level

   subroutine do_nlstep()
      !-- reads nlev levels; per level: lcut, or2, or4, orho1, LFfac and 11
      !   coefficient arrays (w dw ddw z dz s b db ddb aj dj); writes the 18
      !   inverse grids and the 10 forward outputs of nl_level
      type(grid_space_arrays_t) :: gsa
      complex(cp), allocatable :: C(:,:), O(:,:)
      real(cp), allocatable :: uGs(:,:)
      integer :: u, v, lev, i
      integer(int32) :: lcut32
      real(cp) :: o2, o4, orh, lff

      allocate( C(lm_max,11), O(lm_max,10), uGs(n_theta_max,n_phi_max) )
      open(newunit=u, file=trim(infile), access='stream', form='unformatted', &
      &    status='old')
      open(newunit=v, file=trim(outfile), access='stream', form='unformatted', &
      &    status='replace')
      call write_header(v)
      do lev=1,nlev
         read(u) lcut32, o2, o4, orh, lff
         call set_radial(o2, o4, orh)
         LFfac=lff
         if ( lev == 1 ) call gsa%initialize()
         do i=1,11
            read(u) C(:,i)
         end do
         call nl_level(gsa, C, int(lcut32), O, uGs)
         write(v) gsa%vrc, gsa%vtc, gsa%vpc, gsa%cvrc, gsa%cvtc, gsa%cvpc
         write(v) gsa%dvrdrc, gsa%dvtdrc, gsa%dvpdrc, gsa%brc, gsa%btc, gsa%bpc
         write(v) gsa%cbrc, gsa%cbtc, gsa%cbpc, gsa%sc, gsa%dsdtc, gsa%dsdpc
         do i=1,10
            write(v) O(:,i)
         end do
      end do
      close(u)
      close(v)
   end subroutine do_nlstep

   subroutine do_td()
      !-- reads nlev levels; per level: lcut, or2, or4, orho1, CorFac, epsc
      !   and 10 coefficient arrays (AdvrLM AdvtLM AdvpLM VStLM VxBrLM VxBtLM
      !   VxBpLM w dw z); runs get_dwdt, get_dzdt, get_dpdt, get_dsdt and
      !   get_dbdt of an interior level (nBc = 0) of a rotating Boussinesq
      !   MHD run (no double curl) and writes dwdt dzdt dpdt dsdt dbdt djdt
      !   dVxBhLM
      type(nonlinear_lm_t) :: td
      complex(cp), allocatable :: O(:,:), dVSrLM(:)
      integer :: u, v, lev, i
      integer(int32) :: lcut32
      real(cp) :: o2, o4, orh, cf, ep

      l_conv=.true.; l_corr=.true.; l_heat=.true.; l_mag_kin=.false.
      l_single_matrix=.false.; l_double_curl=.false.
      call td%initialize(lm_max)
      allocate( O(lm_max,7), dVSrLM(lm_max) )
      allocate( w_Rloc(lm_max,1), dw_Rloc(lm_max,1), ddw_Rloc(lm_max,1) )
      allocate( z_Rloc(lm_max,1), dz_Rloc(lm_max,1) )
      if ( .not. allocated(l_R) ) allocate( l_R(1), epscProf(1) )
      ddw_Rloc(:,1)=0.0_cp; dz_Rloc(:,1)=0.0_cp
      open(newunit=u, file=trim(infile), access='stream', form='unformatted', &
      &    status='old')
      open(newunit=v, file=trim(outfile), access='stream', form='unformatted', &
      &    status='replace')
      call write_header(v)
      do lev=1,nlev
         read(u) lcut32, o2, o4, orh, cf, ep
         call set_radial(o2, o4, orh)
         l_R(1)=int(lcut32); CorFac=cf; epsc=ep; epscProf(1)=1.0_cp
         read(u) td%AdvrLM(1:lm_max), td%AdvtLM(1:lm_max), td%AdvpLM(1:lm_max)
         read(u) td%VStLM(1:lm_max), td%VxBrLM(1:lm_max), td%VxBtLM(1:lm_max)
         read(u) td%VxBpLM(1:lm_max)
         read(u) w_Rloc(:,1), dw_Rloc(:,1), z_Rloc(:,1)
         !-- get_dpdt leaves l = 0 unset (its loop starts at lm_min = 2)
         O(:,:)=0.0_cp
         call td%get_dwdt(1, 0, O(:,1))
         call td%get_dzdt(1, 0, O(:,2))
         call td%get_dpdt(1, 0, O(:,3))
         call td%get_dsdt(1, 0, O(:,4), dVSrLM)
         call td%get_dbdt(1, 0, O(:,5), O(:,6), O(:,7))
         do i=1,7
            write(v) O(:,i)
         end do
      end do
      close(u)
      close(v)
   end subroutine do_td

   subroutine do_courant()
      !-- reads nlev levels; per level: lcut, or2, or4, orho1, LFfac,
      !   courfac, alffac, opm, delxr2, delxh2 and the 11 coefficient arrays
      !   of nlstep; builds the level's grids as nl_level does and writes the
      !   radial and horizontal Courant steps of src/courant.f90 (magnetic
      !   branch, l_cour_alf_damp) starting from 1e30
      type(grid_space_arrays_t) :: gsa
      complex(cp), allocatable :: C(:,:), O(:,:)
      real(cp), allocatable :: uGs(:,:)
      integer :: u, v, lev, i
      integer(int32) :: lcut32
      real(cp) :: o2, o4, orh, lff, cf, af, pm, dr2, dh2, dtr, dth

      l_mag_kin=.false.; l_cour_alf_damp=.true.
      allocate( C(lm_max,11), O(lm_max,10), uGs(n_theta_max,n_phi_max) )
      if ( .not. allocated(orho2) ) allocate( orho2(1) )
      if ( .not. allocated(delxr2) ) allocate( delxr2(1), delxh2(1) )
      open(newunit=u, file=trim(infile), access='stream', form='unformatted', &
      &    status='old')
      open(newunit=v, file=trim(outfile), access='stream', form='unformatted', &
      &    status='replace')
      call write_header(v)
      do lev=1,nlev
         read(u) lcut32, o2, o4, orh, lff, cf, af, pm, dr2, dh2
         call set_radial(o2, o4, orh)
         orho2(1)=orh*orh; LFfac=lff; opm=pm; delxr2(1)=dr2; delxh2(1)=dh2
         if ( lev == 1 ) call gsa%initialize()
         do i=1,11
            read(u) C(:,i)
         end do
         call nl_level(gsa, C, int(lcut32), O, uGs)
         dtr=1.0e30_cp; dth=1.0e30_cp
         call courant(1, dtr, dth, gsa%vrc, gsa%vtc, gsa%vpc, gsa%brc, gsa%btc, &
         &            gsa%bpc, cf, af)
         write(v) dtr, dth
      end do
      close(u)
      close(v)
   end subroutine do_courant

   subroutine do_rloop()
      !-- the bulk levels of the radial loop (src/rIter.f90:188-470, nBc = 0)
      !   in its order: nl_level (transforms and get_nl), courant on the
      !   level's grids (dtrkc, dthkc from 1e10), then get_dwdt, get_dzdt,
      !   get_dpdt, get_dsdt, get_dbdt on the forward outputs.  Reads nlev
      !   levels; per level: lcut, or2, or4, orho1, LFfac, CorFac, epsc,
      !   courfac, alffac, opm, delxr2, delxh2 and the 11 coefficient arrays
      !   of nlstep; writes dwdt dzdt dpdt dsdt dbdt djdt dVxBhLM, dtrkc,
      !   dthkc.
      type(grid_space_arrays_t) :: gsa
      type(nonlinear_lm_t) :: td
      complex(cp), allocatable :: C(:,:), O(:,:), T(:,:), dVSrLM(:)
      real(cp), allocatable :: uGs(:,:)
      integer :: u, v, lev, i
      integer(int32) :: lcut32
      real(cp) :: o2, o4, orh, lff, cof, ep, cf, af, pm, dr2, dh2, dtr, dth

      l_mag_kin=.false.; l_cour_alf_damp=.true.
      l_conv=.true.; l_corr=.true.; l_heat=.true.
      l_single_matrix=.false.; l_double_curl=.false.
      call td%initialize(lm_max)
      allocate( C(lm_max,11), O(lm_max,10), T(lm_max,7), dVSrLM(lm_max) )
      allocate( uGs(n_theta_max,n_phi_max) )
      allocate( w_Rloc(lm_max,1), dw_Rloc(lm_max,1), ddw_Rloc(lm_max,1) )
      allocate( z_Rloc(lm_max,1), dz_Rloc(lm_max,1) )
      if ( .not. allocated(l_R) ) allocate( l_R(1), epscProf(1) )
      if ( .not. allocated(orho2) ) allocate( orho2(1) )
      if ( .not. allocated(delxr2) ) allocate( delxr2(1), delxh2(1) )
      open(newunit=u, file=trim(infile), access='stream', form='unformatted', &
      &    status='old')
      open(newunit=v, file=trim(outfile), access='stream', form='unformatted', &
      &    status='replace')
      call write_header(v)
      do lev=1,nlev
         read(u) lcut32, o2, o4, orh, lff, cof, ep, cf, af, pm, dr2, dh2
         call set_radial(o2, o4, orh)
         orho2(1)=orh*orh; LFfac=lff; CorFac=cof; epsc=ep; epscProf(1)=1.0_cp
         opm=pm; delxr2(1)=dr2; delxh2(1)=dh2; l_R(1)=int(lcut32)
         if ( lev == 1 ) call gsa%initialize()
         do i=1,11
            read(u) C(:,i)
         end do
         call nl_level(gsa, C, int(lcut32), O, uGs)
         dtr=1.0e10_cp; dth=1.0e10_cp
         call courant(1, dtr, dth, gsa%vrc, gsa%vtc, gsa%vpc, gsa%brc, gsa%btc, &
         &            gsa%bpc, cf, af)
         td%AdvrLM(1:lm_max)=O(:,1); td%AdvtLM(1:lm_max)=O(:,2)
         td%AdvpLM(1:lm_max)=O(:,3); td%VStLM(1:lm_max)=O(:,5)
         td%VxBrLM(1:lm_max)=O(:,7); td%VxBtLM(1:lm_max)=O(:,8)
         td%VxBpLM(1:lm_max)=O(:,9)
         w_Rloc(:,1)=C(:,1); dw_Rloc(:,1)=C(:,2); ddw_Rloc(:,1)=C(:,3)
         z_Rloc(:,1)=C(:,4); dz_Rloc(:,1)=C(:,5)
         T(:,:)=0.0_cp
         call td%get_dwdt(1, 0, T(:,1))
         call td%get_dzdt(1, 0, T(:,2))
         call td%get_dpdt(1, 0, T(:,3))
         call td%get_dsdt(1, 0, T(:,4), dVSrLM)
         call td%get_dbdt(1, 0, T(:,5), T(:,6), T(:,7))
         do i=1,7
            write(v) T(:,i)
         end do
         write(v) dtr, dth
      end do
      close(u)
      close(v)
   end subroutine do_rloop

   subroutine do_brv()
      !-- reads nlev levels; per level: omega, O_r_E_2, O_rho, conductance_ma,
      !   prmag, sigma_ratio and the grids br, vt, vp (n_theta, n_phi);
      !   writes br_vt_lm, br_vp_lm of src/nonlinear_bcs.f90 get_br_v_bcs,
      !   then b_nl_bc, aj_nl_bc of get_b_nl_bcs('CMB') and aj_nl_bc of
      !   get_b_nl_bcs('ICB') on those products
      real(cp), allocatable :: G(:,:,:)
      complex(cp), allocatable :: a(:), b(:), bc(:), jc(:), bi(:), ji(:)
      real(cp) :: om, o2, orh
      integer :: u, v, lev

      allocate( G(n_theta_max,n_phi_max,3), a(lm_max), b(lm_max) )
      allocate( bc(lm_max), jc(lm_max), bi(lm_max), ji(lm_max) )
      open(newunit=u, file=trim(infile), access='stream', form='unformatted', &
      &    status='old')
      open(newunit=v, file=trim(outfile), access='stream', form='unformatted', &
      &    status='replace')
      call write_header(v)
      do lev=1,nlev
         read(u) om, o2, orh, conductance_ma, prmag, sigma_ratio
         read(u) G(:,:,:)
         a(:)=(0.0_cp,0.0_cp); b(:)=(0.0_cp,0.0_cp)
         call get_br_v_bcs(G(:,:,1), G(:,:,2), G(:,:,3), om, o2, orh, a, b)
         bc(:)=(0.0_cp,0.0_cp); jc(:)=(0.0_cp,0.0_cp); ji(:)=(0.0_cp,0.0_cp)
         call get_b_nl_bcs('CMB', a, b, bc, jc)
         call get_b_nl_bcs('ICB', a, b, bi, ji)
         write(v) a, b, bc, jc, ji
      end do
      close(u)
      close(v)
   end subroutine do_brv

   subroutine do_bench()
      !-- Times the per-level transform sequence of a Boussinesq MHD bulk
      !   level (src/rIter.f90:483-707 with l_adv_curl and the derivative
      !   fields): 5 vector + 1 scalar + 1 gradient inverse transforms,
      !   get_nl, 3 vector + 1 scalar forward transforms.  Serial over
      !   levels, OpenMP inside each call, as the reference runs it.
      type(grid_space_arrays_t) :: gsa
      complex(cp), allocatable :: C(:,:,:), O(:,:)
      real(cp), allocatable :: uGs(:,:)
      integer :: lm, l, m, i, lev, rep, v
      integer(int64) :: t0, t1, rate
      real(cp) :: tsum, tbest, checksum
      real(cp), allocatable :: tim(:)

      call set_radial(0.7_cp, 0.49_cp, 1.0_cp)
      LFfac=1.0_cp
      call gsa%initialize()
      allocate( C(lm_max,11,nlev), O(lm_max,10), uGs(n_theta_max,n_phi_max) )
      allocate( tim(nrep) )
      do lev=1,nlev
         do i=1,11
            do lm=1,lm_max
               l=st_map%lm2l(lm); m=st_map%lm2m(lm)
               C(lm,i,lev)=cmplx(cos(0.37_cp*lm+1.3_cp*i+0.11_cp*lev), &
               &                 sin(0.71_cp*lm+0.7_cp*i+0.05_cp*lev),cp) / &
               &           real(l+1,cp)**2
               if ( m == 0 ) C(lm,i,lev)=cmplx(real(C(lm,i,lev)),0.0_cp,cp)
               if ( l == 0 .and. i /= 6 ) C(lm,i,lev)=0.0_cp
            end do
         end do
      end do

      checksum=0.0_cp
      do rep=0,nrep
         call system_clock(t0, rate)
         do lev=1,nlev
            call nl_level(gsa, C(:,:,lev), l_max, O, uGs)
            checksum=checksum+abs(O(2,1))
         end do
         call system_clock(t1)
         if ( rep > 0 ) tim(rep)=real(t1-t0,cp)/real(rate,cp)
      end do
      tsum=sum(tim); tbest=minval(tim)
      open(newunit=v, file=trim(outfile), status='replace')
      write(v,'(A,I6,A,I4,A,I4,A,ES14.6,A,ES14.6,A,ES14.6)') 'threads=', &
      &     omp_get_max_threads(), ' nlev=', nlev, ' nrep=', nrep,        &
      &     ' mean_s=', tsum/nrep, ' best_s=', tbest, ' checksum=', checksum
      write(v,'(A)') 'times_s='
      write(v,'(ES14.6)') tim
      close(v)
   end subroutine do_bench

end program ref_harness
