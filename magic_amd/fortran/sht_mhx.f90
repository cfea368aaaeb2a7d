!
! module sht for MagIC, backed by the mhx HIP library (libmhx.so).
!
! Drop-in replacement of the reference's transform module (src/sht_native.f90):
! the same public names, dummy arguments and intents, compiled in MagIC's
! build in place of sht_native.f90 (or shtns.f90).  Every routine forwards to
! the C ABI of include/mhx.h through iso_c_binding; a nonzero return code is
! turned into the reference's fatal-error path (abortRun,
! src/useful.f90:271-302).  Arrays are the caller's host arrays; the library
! stages them through device memory.  Requires cp = real64.
!
! The plan is created by initialize_sht from the truncation module
! (l_max, m_max, minc, n_theta_max, n_phi_max) and released by finalize_sht,
! as the reference's initialize_transforms / finalize_transforms.
!
module sht

   use iso_c_binding
   use precision_mod, only: cp
   use truncation, only: l_max, m_max, minc, n_theta_max, n_phi_max
   use useful, only: abortRun

   implicit none

   private

   public :: initialize_sht, scal_to_spat, scal_to_grad_spat, pol_to_grad_spat, &
   &         torpol_to_spat, pol_to_curlr_spat, torpol_to_curl_spat,            &
   &         torpol_to_dphspat, scal_to_SH, spat_to_sphertor,                   &
   &         torpol_to_spat_IC, torpol_to_curl_spat_IC, axi_to_spat,            &
   &         spat_to_qst, sphtor_to_spat, toraxi_to_spat, finalize_sht

   type(c_ptr), save :: plan = c_null_ptr

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
      integer(c_int) function mhx_scal_to_spat(p, s, f, lcut) bind(C, name='mhx_scal_to_spat')
         import :: c_int, c_ptr, c_double, c_double_complex
         type(c_ptr), value :: p
         complex(c_double_complex), intent(in) :: s(*)
         real(c_double), intent(out) :: f(*)
         integer(c_int), value :: lcut
      end function mhx_scal_to_spat
      integer(c_int) function mhx_scal_to_grad_spat(p, s, gt, gp, lcut) &
           &         bind(C, name='mhx_scal_to_grad_spat')
         import :: c_int, c_ptr, c_double, c_double_complex
         type(c_ptr), value :: p
         complex(c_double_complex), intent(in) :: s(*)
         real(c_double), intent(out) :: gt(*), gp(*)
         integer(c_int), value :: lcut
      end function mhx_scal_to_grad_spat
      integer(c_int) function mhx_pol_to_grad_spat(p, s, gt, gp, lcut) &
           &         bind(C, name='mhx_pol_to_grad_spat')
         import :: c_int, c_ptr, c_double, c_double_complex
         type(c_ptr), value :: p
         complex(c_double_complex), intent(in) :: s(*)
         real(c_double), intent(out) :: gt(*), gp(*)
         integer(c_int), value :: lcut
      end function mhx_pol_to_grad_spat
      integer(c_int) function mhx_torpol_to_spat(p, w, dw, z, vr, vt, vp, lcut) &
           &         bind(C, name='mhx_torpol_to_spat')
         import :: c_int, c_ptr, c_double, c_double_complex
         type(c_ptr), value :: p
         complex(c_double_complex), intent(in) :: w(*), dw(*), z(*)
         real(c_double), intent(out) :: vr(*), vt(*), vp(*)
         integer(c_int), value :: lcut
      end function mhx_torpol_to_spat
      integer(c_int) function mhx_sphtor_to_spat(p, dw, z, vt, vp, lcut) &
           &         bind(C, name='mhx_sphtor_to_spat')
         import :: c_int, c_ptr, c_double, c_double_complex
         type(c_ptr), value :: p
         complex(c_double_complex), intent(in) :: dw(*), z(*)
         real(c_double), intent(out) :: vt(*), vp(*)
         integer(c_int), value :: lcut
      end function mhx_sphtor_to_spat
      integer(c_int) function mhx_torpol_to_curl_spat_IC(p, r, ricb, db, ddb, j, dj, cbr, cbt, cbp) &
           &         bind(C, name='mhx_torpol_to_curl_spat_IC')
         import :: c_int, c_ptr, c_double, c_double_complex
         type(c_ptr), value :: p
         real(c_double), value :: r, ricb
         complex(c_double_complex), intent(in) :: db(*), ddb(*), j(*), dj(*)
         real(c_double), intent(out) :: cbr(*), cbt(*), cbp(*)
      end function mhx_torpol_to_curl_spat_IC
      integer(c_int) function mhx_torpol_to_spat_IC(p, r, ricb, w, dw, z, br, bt, bp) &
           &         bind(C, name='mhx_torpol_to_spat_IC')
         import :: c_int, c_ptr, c_double, c_double_complex
         type(c_ptr), value :: p
         real(c_double), value :: r, ricb
         complex(c_double_complex), intent(in) :: w(*), dw(*), z(*)
         real(c_double), intent(out) :: br(*), bt(*), bp(*)
      end function mhx_torpol_to_spat_IC
      integer(c_int) function mhx_torpol_to_dphspat(p, dw, z, dvt, dvp, lcut) &
           &         bind(C, name='mhx_torpol_to_dphspat')
         import :: c_int, c_ptr, c_double, c_double_complex
         type(c_ptr), value :: p
         complex(c_double_complex), intent(in) :: dw(*), z(*)
         real(c_double), intent(out) :: dvt(*), dvp(*)
         integer(c_int), value :: lcut
      end function mhx_torpol_to_dphspat
      integer(c_int) function mhx_pol_to_curlr_spat(p, q, cvr, lcut) &
           &         bind(C, name='mhx_pol_to_curlr_spat')
         import :: c_int, c_ptr, c_double, c_double_complex
         type(c_ptr), value :: p
         complex(c_double_complex), intent(in) :: q(*)
         real(c_double), intent(out) :: cvr(*)
         integer(c_int), value :: lcut
      end function mhx_pol_to_curlr_spat
      integer(c_int) function mhx_torpol_to_curl_spat(p, or2, b, ddb, j, dj, cvr, cvt, cvp, lcut) &
           &         bind(C, name='mhx_torpol_to_curl_spat')
         import :: c_int, c_ptr, c_double, c_double_complex
         type(c_ptr), value :: p
         real(c_double), value :: or2
         complex(c_double_complex), intent(in) :: b(*), ddb(*), j(*), dj(*)
         real(c_double), intent(out) :: cvr(*), cvt(*), cvp(*)
         integer(c_int), value :: lcut
      end function mhx_torpol_to_curl_spat
      integer(c_int) function mhx_scal_to_SH(p, f, flm, lcut) bind(C, name='mhx_scal_to_SH')
         import :: c_int, c_ptr, c_double, c_double_complex
         type(c_ptr), value :: p
         real(c_double), intent(in) :: f(*)
         complex(c_double_complex), intent(out) :: flm(*)
         integer(c_int), value :: lcut
      end function mhx_scal_to_SH
      integer(c_int) function mhx_spat_to_qst(p, f, g, h, q, s, t, lcut) &
           &         bind(C, name='mhx_spat_to_qst')
         import :: c_int, c_ptr, c_double, c_double_complex
         type(c_ptr), value :: p
         real(c_double), intent(in) :: f(*), g(*), h(*)
         complex(c_double_complex), intent(out) :: q(*), s(*), t(*)
         integer(c_int), value :: lcut
      end function mhx_spat_to_qst
      integer(c_int) function mhx_spat_to_sphertor(p, f, g, flm, glm, lcut) &
           &         bind(C, name='mhx_spat_to_sphertor')
         import :: c_int, c_ptr, c_double, c_double_complex
         type(c_ptr), value :: p
         real(c_double), intent(in) :: f(*), g(*)
         complex(c_double_complex), intent(out) :: flm(*), glm(*)
         integer(c_int), value :: lcut
      end function mhx_spat_to_sphertor
      integer(c_int) function mhx_axi_to_spat(p, fl, f) bind(C, name='mhx_axi_to_spat')
         import :: c_int, c_ptr, c_double, c_double_complex
         type(c_ptr), value :: p
         complex(c_double_complex), intent(in) :: fl(*)
         real(c_double), intent(out) :: f(*)
      end function mhx_axi_to_spat
      integer(c_int) function mhx_toraxi_to_spat(p, fl, ft, fp, lcut) &
           &         bind(C, name='mhx_toraxi_to_spat')
         import :: c_int, c_ptr, c_double, c_double_complex
         type(c_ptr), value :: p
         complex(c_double_complex), intent(in) :: fl(*)
         real(c_double), intent(out) :: ft(*), fp(*)
         integer(c_int), value :: lcut
      end function mhx_toraxi_to_spat
   end interface

contains

   subroutine check(rc, where)
      integer(c_int), intent(in) :: rc
      character(len=*), intent(in) :: where
      character(kind=c_char), pointer :: msg(:)
      character(len=256) :: text
      integer :: i
      if ( rc == 0 ) return
      text = ''
      call c_f_pointer(mhx_last_error(), msg, [256])
      do i=1,256
         if ( msg(i) == c_null_char ) exit
         text(i:i) = msg(i)
      end do
      call abortRun('mhx '//where//': '//trim(text))
   end subroutine check

   subroutine initialize_sht(l_scrambled_theta)
      !-- src/sht_native.f90:24-32
      logical, intent(out) :: l_scrambled_theta
      if ( cp /= c_double ) call abortRun('mhx module sht needs cp = real64')
      l_scrambled_theta=.true.
      if ( .not. c_associated(plan) ) then
         call check(mhx_plan_create(int(l_max,c_int), int(m_max,c_int), int(minc,c_int), &
              &     int(n_theta_max,c_int), int(n_phi_max,c_int), 0_c_int, 1_c_int, plan), &
              &     'initialize_sht')
      end if
   end subroutine initialize_sht

   subroutine finalize_sht()
      integer(c_int) :: rc
      if ( c_associated(plan) ) then
         rc = mhx_plan_destroy(plan)
         plan = c_null_ptr
      end if
   end subroutine finalize_sht

   subroutine scal_to_spat(Slm, fieldc, lcut)
      complex(cp), intent(in) :: Slm(:)
      integer,     intent(in) :: lcut
      real(cp), intent(out) :: fieldc(:,:)
      call check(mhx_scal_to_spat(plan, Slm, fieldc, int(lcut,c_int)), 'scal_to_spat')
   end subroutine scal_to_spat

   subroutine scal_to_grad_spat(Slm, gradtc, gradpc, lcut)
      complex(cp), intent(in) :: Slm(:)
      integer,     intent(in) :: lcut
      real(cp), intent(out) :: gradtc(:,:)
      real(cp), intent(out) :: gradpc(:,:)
      call check(mhx_scal_to_grad_spat(plan, Slm, gradtc, gradpc, int(lcut,c_int)), &
           &     'scal_to_grad_spat')
   end subroutine scal_to_grad_spat

   subroutine pol_to_grad_spat(Slm, gradtc, gradpc, lcut)
      complex(cp), intent(in) :: Slm(:)
      integer,     intent(in) :: lcut
      real(cp), intent(out) :: gradtc(:,:)
      real(cp), intent(out) :: gradpc(:,:)
      call check(mhx_pol_to_grad_spat(plan, Slm, gradtc, gradpc, int(lcut,c_int)), &
           &     'pol_to_grad_spat')
   end subroutine pol_to_grad_spat

   subroutine torpol_to_spat(Wlm, dWlm, Zlm, vrc, vtc, vpc, lcut)
      complex(cp), intent(in) :: Wlm(:), dWlm(:), Zlm(:)
      integer,     intent(in) :: lcut
      real(cp), intent(out) :: vrc(:,:)
      real(cp), intent(out) :: vtc(:,:)
      real(cp), intent(out) :: vpc(:,:)
      call check(mhx_torpol_to_spat(plan, Wlm, dWlm, Zlm, vrc, vtc, vpc, int(lcut,c_int)), &
           &     'torpol_to_spat')
   end subroutine torpol_to_spat

   subroutine sphtor_to_spat(dWlm, Zlm, vtc, vpc, lcut)
      complex(cp), intent(in) :: dWlm(:), Zlm(:)
      integer,     intent(in) :: lcut
      real(cp), intent(out) :: vtc(:,:)
      real(cp), intent(out) :: vpc(:,:)
      call check(mhx_sphtor_to_spat(plan, dWlm, Zlm, vtc, vpc, int(lcut,c_int)), &
           &     'sphtor_to_spat')
   end subroutine sphtor_to_spat

   subroutine torpol_to_curl_spat_IC(r, r_ICB, dBlm, ddBlm, Jlm, dJlm, cbr, cbt, cbp)
      real(cp),    intent(in) :: r, r_ICB
      complex(cp), intent(in) :: dBlm(:), ddBlm(:)
      complex(cp), intent(in) :: Jlm(:), dJlm(:)
      real(cp), intent(out) :: cbr(:,:)
      real(cp), intent(out) :: cbt(:,:)
      real(cp), intent(out) :: cbp(:,:)
      call check(mhx_torpol_to_curl_spat_IC(plan, real(r,c_double), real(r_ICB,c_double), &
           &     dBlm, ddBlm, Jlm, dJlm, cbr, cbt, cbp), 'torpol_to_curl_spat_IC')
   end subroutine torpol_to_curl_spat_IC

   subroutine torpol_to_spat_IC(r, r_ICB, Wlm, dWlm, Zlm, Br, Bt, Bp)
      real(cp),    intent(in) :: r, r_ICB
      complex(cp), intent(in) :: Wlm(:), dWlm(:), Zlm(:)
      real(cp), intent(out) :: Br(:,:)
      real(cp), intent(out) :: Bt(:,:)
      real(cp), intent(out) :: Bp(:,:)
      call check(mhx_torpol_to_spat_IC(plan, real(r,c_double), real(r_ICB,c_double), &
           &     Wlm, dWlm, Zlm, Br, Bt, Bp), 'torpol_to_spat_IC')
   end subroutine torpol_to_spat_IC

   subroutine torpol_to_dphspat(dWlm, Zlm, dvtdp, dvpdp, lcut)
      complex(cp), intent(in) :: dWlm(:), Zlm(:)
      integer,     intent(in) :: lcut
      real(cp), intent(out) :: dvtdp(:,:)
      real(cp), intent(out) :: dvpdp(:,:)
      call check(mhx_torpol_to_dphspat(plan, dWlm, Zlm, dvtdp, dvpdp, int(lcut,c_int)), &
           &     'torpol_to_dphspat')
   end subroutine torpol_to_dphspat

   subroutine pol_to_curlr_spat(Qlm, cvrc, lcut)
      complex(cp), intent(in) :: Qlm(:)
      integer,     intent(in) :: lcut
      real(cp), intent(out) :: cvrc(:,:)
      call check(mhx_pol_to_curlr_spat(plan, Qlm, cvrc, int(lcut,c_int)), 'pol_to_curlr_spat')
   end subroutine pol_to_curlr_spat

   subroutine torpol_to_curl_spat(or2, Blm, ddBlm, Jlm, dJlm, cvrc, cvtc, cvpc, lcut)
      complex(cp), intent(in) :: Blm(:), ddBlm(:)
      complex(cp), intent(in) :: Jlm(:), dJlm(:)
      real(cp),    intent(in) :: or2
      integer,     intent(in) :: lcut
      real(cp), intent(out) :: cvrc(:,:)
      real(cp), intent(out) :: cvtc(:,:)
      real(cp), intent(out) :: cvpc(:,:)
      call check(mhx_torpol_to_curl_spat(plan, real(or2,c_double), Blm, ddBlm, Jlm, dJlm, &
           &     cvrc, cvtc, cvpc, int(lcut,c_int)), 'torpol_to_curl_spat')
   end subroutine torpol_to_curl_spat

   subroutine scal_to_SH(f, fLM, lcut)
      real(cp), intent(inout) :: f(:,:)
      integer,  intent(in) :: lcut
      complex(cp), intent(out) :: fLM(:)
      call check(mhx_scal_to_SH(plan, f, fLM, int(lcut,c_int)), 'scal_to_SH')
   end subroutine scal_to_SH

   subroutine spat_to_qst(f, g, h, qLM, sLM, tLM, lcut)
      real(cp), intent(inout) :: f(:,:)
      real(cp), intent(inout) :: g(:,:)
      real(cp), intent(inout) :: h(:,:)
      integer,  intent(in) :: lcut
      complex(cp), intent(out) :: qLM(:)
      complex(cp), intent(out) :: sLM(:)
      complex(cp), intent(out) :: tLM(:)
      call check(mhx_spat_to_qst(plan, f, g, h, qLM, sLM, tLM, int(lcut,c_int)), 'spat_to_qst')
   end subroutine spat_to_qst

   subroutine spat_to_sphertor(f, g, fLM, gLM, lcut)
      real(cp), intent(inout) :: f(:,:)
      real(cp), intent(inout) :: g(:,:)
      integer,  intent(in) :: lcut
      complex(cp), intent(out) :: fLM(:)
      complex(cp), intent(out) :: gLM(:)
      call check(mhx_spat_to_sphertor(plan, f, g, fLM, gLM, int(lcut,c_int)), 'spat_to_sphertor')
   end subroutine spat_to_sphertor

   subroutine axi_to_spat(fl_ax, f)
      complex(cp), intent(in) :: fl_ax(l_max+1)
      real(cp), intent(out) :: f(:)
      call check(mhx_axi_to_spat(plan, fl_ax, f), 'axi_to_spat')
   end subroutine axi_to_spat

   subroutine toraxi_to_spat(fl_ax, ft, fp, lcut)
      integer, intent(in) :: lcut
      complex(cp), intent(in) :: fl_ax(l_max+1)
      real(cp), intent(out) :: ft(:)
      real(cp), intent(out) :: fp(:)
      call check(mhx_toraxi_to_spat(plan, fl_ax, ft, fp, int(lcut,c_int)), 'toraxi_to_spat')
   end subroutine toraxi_to_spat

end module sht
