! ISO_C_BINDING interface of libccgpu (include/ccgpu.h) for Dalton's
! Fortran CC drivers, and thin wrappers named after the routines they
! replace.  Every wrapper maps a nonzero status to QUIT(text), Dalton's only
! error path (gp/gptrygve.F:256); QUIT is the caller's (Dalton's) routine.
!
! Arrays are Dalton's: T1AM/OMEGA1 (a,i) a fastest, doubles packed over
! ai <= bj (INDEX, cc/ccsd_energy.F:1666), left vectors in Dalton's packed
! convention (CC_TRDRV, cc/cc_trdrv.F:484-486), orbital energies occupied
! first.  Build with the library:
!   flang -c ccgpu_interface.f90 && flang ... ccgpu_interface.o -L<dir> -lccgpu
module ccgpu_interface
  use iso_c_binding
  implicit none
  integer(c_int), parameter :: CCGPU_FP64 = 0, CCGPU_MIXED = 1
  type(c_ptr), save :: ccgpu_ctx = c_null_ptr
  integer(c_int), save :: ccgpu_prec = CCGPU_FP64
  interface
    integer(c_int) function ccgpu_version() bind(c, name="ccgpu_version")
      import :: c_int
    end function
    type(c_ptr) function ccgpu_last_error() bind(c, name="ccgpu_last_error")
      import :: c_ptr
    end function
    integer(c_int) function ccgpu_init(nocc, nvir, naux, eps, b, mxdiis, ctx) bind(c, name="ccgpu_init")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: nocc, nvir, naux, mxdiis
      real(c_double), intent(in) :: eps(*), b(*)
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function ccgpu_init_ao(nocc, nvir, nbas, eps, cmo, eri, thr, mxdiis, ctx, naux) &
        bind(c, name="ccgpu_init_ao")
      import :: c_int, c_long, c_double, c_ptr
      integer(c_int), value :: nocc, nvir, nbas, mxdiis
      real(c_double), intent(in) :: eps(*), cmo(*), eri(*)
      real(c_double), value :: thr
      type(c_ptr), intent(out) :: ctx
      integer(c_long), intent(out) :: naux
    end function
    integer(c_int) function ccgpu_df_factor_ao(nbas, nmo, cmo, eri, thr, maxvec, b, naux) &
        bind(c, name="ccgpu_df_factor_ao")
      import :: c_int, c_long, c_double
      integer(c_int), value :: nbas, nmo
      real(c_double), intent(in) :: cmo(*), eri(*)
      real(c_double), value :: thr
      integer(c_long), value :: maxvec
      real(c_double), intent(out) :: b(*)
      integer(c_long), intent(out) :: naux
    end function
    integer(c_int) function ccgpu_finalize(ctx) bind(c, name="ccgpu_finalize")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function ccgpu_rhs(ctx, t1am, t2am, omega1, omega2, prec) bind(c, name="ccgpu_rhs")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: t1am(*), t2am(*)
      real(c_double), intent(out) :: omega1(*), omega2(*)
      integer(c_int), value :: prec
    end function
    integer(c_int) function ccgpu_lhtr(ctx, prec, nsimtr, ctr1, ctr2, rho1, rho2) bind(c, name="ccgpu_lhtr")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: prec, nsimtr
      real(c_double), intent(in) :: ctr1(*), ctr2(*)
      real(c_double), intent(out) :: rho1(*), rho2(*)
    end function
    integer(c_int) function ccgpu_energy(ctx, t1am, t2am, ecorr) bind(c, name="ccgpu_energy")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: t1am(*), t2am(*)
      real(c_double), intent(out) :: ecorr
    end function
    integer(c_int) function ccgpu_diis(ctx, s, t, niter) bind(c, name="ccgpu_diis")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: s(*)
      real(c_double), intent(inout) :: t(*)
      integer(c_int), value :: niter
    end function
    integer(c_int) function ccgpu_ccsd_guess(ctx, emp2) bind(c, name="ccgpu_ccsd_guess")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: emp2
    end function
    integer(c_int) function ccgpu_ccsd_iterate(ctx, prec, ecorr, omnorm) bind(c, name="ccgpu_ccsd_iterate")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: prec
      real(c_double), intent(out) :: ecorr, omnorm
    end function
    integer(c_int) function ccgpu_ccsd_energy(ctx, maxit, threnr, thrvec, prec, emp2, eccsd, niter, hist) &
        bind(c, name="ccgpu_ccsd_energy")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: maxit, prec
      real(c_double), value :: threnr, thrvec
      real(c_double), intent(out) :: emp2, eccsd, hist(*)
      integer(c_int), intent(out) :: niter
    end function
    integer(c_int) function ccgpu_ccsd_resume(ctx, maxit, threnr, thrvec, prec, estart, eccsd, niter, hist) &
        bind(c, name="ccgpu_ccsd_resume")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: maxit, prec
      real(c_double), value :: threnr, thrvec
      real(c_double), intent(out) :: estart, eccsd, hist(*)
      integer(c_int), intent(out) :: niter
    end function
    integer(c_int) function ccgpu_get_amplitudes(ctx, t1am, t2am) bind(c, name="ccgpu_get_amplitudes")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: t1am(*), t2am(*)
    end function
    integer(c_int) function ccgpu_set_amplitudes(ctx, t1am, t2am) bind(c, name="ccgpu_set_amplitudes")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: t1am(*), t2am(*)
    end function
    integer(c_int) function ccgpu_pt(ctx, prec, e4, e5) bind(c, name="ccgpu_pt")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: prec
      real(c_double), intent(out) :: e4, e5
    end function
    integer(c_int) function ccgpu_eta(ctx, eta1, eta2) bind(c, name="ccgpu_eta")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: eta1(*), eta2(*)
    end function
    integer(c_int) function ccgpu_lambda_solve(ctx, prec, thr, maxit, maxred, l1, l2, niter, resnorm) &
        bind(c, name="ccgpu_lambda_solve")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: prec, maxit, maxred
      real(c_double), value :: thr
      real(c_double), intent(out) :: l1(*), l2(*)
      integer(c_int), intent(out) :: niter
      real(c_double), intent(out) :: resnorm
    end function
    integer(c_int) function ccgpu_onedens(ctx, prec, l1, l2, dens) bind(c, name="ccgpu_onedens")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: prec
      real(c_double), intent(in) :: l1(*), l2(*)
      real(c_double), intent(out) :: dens(*)
    end function
    integer(c_int) function ccgpu_rsp_filename(list, isym, idx, name) bind(c, name="ccgpu_rsp_filename")
      import :: c_int, c_char
      character(kind=c_char), intent(in) :: list(*)   ! NUL-terminated
      integer(c_int), value :: isym, idx
      character(kind=c_char), intent(out) :: name(*)  ! at least 11 characters
    end function
    integer(c_int) function ccgpu_write_e_intermediates(ctx, prec, e1path, e2path, e1, e2) &
        bind(c, name="ccgpu_write_e_intermediates")
      import :: c_int, c_char, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: prec
      character(kind=c_char), intent(in) :: e1path(*), e2path(*)   ! NUL-terminated
      real(c_double), intent(out) :: e1(*), e2(*)
    end function
    integer(c_int) function ccgpu_write_gamma_intermediate(ctx, prec, path, gamma) &
        bind(c, name="ccgpu_write_gamma_intermediate")
      import :: c_int, c_char, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: prec
      character(kind=c_char), intent(in) :: path(*)   ! NUL-terminated
      real(c_double), intent(out) :: gamma(*)         ! NGAMMA values
    end function
    integer(c_int) function ccgpu_save_amplitudes(ctx, path, intsize) bind(c, name="ccgpu_save_amplitudes")
      import :: c_int, c_char, c_ptr
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: path(*)   ! NUL-terminated
      integer(c_int), value :: intsize
    end function
    integer(c_int) function ccgpu_restart_amplitudes(ctx, path, ecorr) bind(c, name="ccgpu_restart_amplitudes")
      import :: c_int, c_char, c_double, c_ptr
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: path(*)
      real(c_double), intent(out) :: ecorr
    end function
  end interface
contains
  ! Nonzero status -> QUIT('<where>: <library message>').
  subroutine ccgpu_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    character(len=256) :: text
    integer :: i
    if (rc == 0) return
    call c_f_pointer(ccgpu_last_error(), msg, [256])
    text = ' '
    do i = 1, 256
      if (msg(i) == c_null_char) exit
      text(i:i) = msg(i)
    end do
    call quit(where//': '//trim(text))
  end subroutine

  ! CCRHSN(OMEGA1,OMEGA2,T1AM,T2AM,WORK,LWORK,APROXR12) (cc/ccsd_rhs.F:20) for
  ! plain closed-shell CCSD in C1 symmetry: WORK is not used.
  subroutine ccgpu_ccrhsn(omega1, omega2, t1am, t2am)
    real(c_double), intent(out) :: omega1(*), omega2(*)
    real(c_double), intent(in) :: t1am(*), t2am(*)
    call ccgpu_check(ccgpu_rhs(ccgpu_ctx, t1am, t2am, omega1, omega2, ccgpu_prec), 'CCRHSN')
  end subroutine

  ! CC_LHTR (cc/cc_lhtr.F:20) for MODEL = 'CCSD': NSIMTR trial vectors CTR1,
  ! CTR2 (read from FC1AM/FC2AM by the caller) to RHO1, RHO2, back to back.
  subroutine ccgpu_cc_lhtr(nsimtr, ctr1, ctr2, rho1, rho2)
    integer, intent(in) :: nsimtr
    real(c_double), intent(in) :: ctr1(*), ctr2(*)
    real(c_double), intent(out) :: rho1(*), rho2(*)
    call ccgpu_check(ccgpu_lhtr(ccgpu_ctx, ccgpu_prec, int(nsimtr, c_int), ctr1, ctr2, rho1, rho2), 'CC_LHTR')
  end subroutine

  ! CCSD_DIIS(S,T,NDIM,NITER) (cc/ccsd_diis.F:251) with the history on the device.
  subroutine ccgpu_ccsd_diis(s, t, niter)
    real(c_double), intent(in) :: s(*)
    real(c_double), intent(inout) :: t(*)
    integer, intent(in) :: niter
    call ccgpu_check(ccgpu_diis(ccgpu_ctx, s, t, int(niter, c_int)), 'CCSD_DIIS')
  end subroutine

  ! The CC1A/CC1B CCRHSN pass and the E4/E5 sums of cc/ccsd_energy.F:1231-1243.
  subroutine ccgpu_ccsd_pt(e4, e5)
    real(c_double), intent(out) :: e4, e5
    call ccgpu_check(ccgpu_pt(ccgpu_ctx, ccgpu_prec, e4, e5), 'CCSD(T)')
  end subroutine
end module ccgpu_interface
