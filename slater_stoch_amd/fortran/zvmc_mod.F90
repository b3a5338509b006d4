! iso_c_binding interfaces of the ZVMC C-ABI (include/zvmc.h).
!
! This is the binding a Fortran host (the reference's two_e_slater, or the
! driver in two_e_slater_hip.F90) uses to reach the MI355X kernels.  Arrays
! are passed in the reference's own layouts: center(3,n), npower(3,n),
! c_gauss(ldg,n,2), g_gauss(ldg,n), samples r(3,m), quartet arrays (nint)
! in the order of src/build_mapping_ijkl.F90.
module zvmc
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: ZV_OK = 0
  integer(c_int), parameter :: ZV_SAMPLES_FIXED = 0, ZV_SAMPLES_PHILOX = 1
  integer(c_int), parameter :: ZV_PREC_FP64_PARITY = 0, ZV_PREC_FP64 = 1, ZV_PREC_FP32 = 2
  integer(c_int), parameter :: ZV_PREC_BF16 = 3
  integer(c_int), parameter :: ZV_RESULT_ERI = 0, ZV_RESULT_MEAN = 1, ZV_RESULT_STDERR = 2
  integer(c_int), parameter :: ZV_SKETCH_FP64 = 0, ZV_SKETCH_FP32 = 1
  integer(c_int), parameter :: ZV_DAVIDSON_PLAIN = 0, ZV_DAVIDSON_SHIFT_INVERT = 1

  interface
    integer(c_int) function zv_create(device, ctx) bind(C, name="zv_create")
      import :: c_int, c_ptr
      integer(c_int), value :: device
      type(c_ptr), intent(out) :: ctx
    end function

    integer(c_int) function zv_destroy(ctx) bind(C, name="zv_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function zv_last_error(ctx, buf, buflen) bind(C, name="zv_last_error")
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(out) :: buf(*)
      integer(c_int), value :: buflen
    end function

    integer(c_int) function zv_set_basis(ctx, n, center, npower, n_sto, i_type, zeta, n_gauss, &
        ldg, c_gauss, g_gauss, a_zv) bind(C, name="zv_set_basis")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: n, ldg
      real(c_double), intent(in) :: center(*), zeta(*), c_gauss(*), g_gauss(*)
      integer(c_int), intent(in) :: npower(*), n_sto(*), i_type(*), n_gauss(*)
      real(c_double), value :: a_zv
    end function

    integer(c_int64_t) function zv_nint(n) bind(C, name="zv_nint")
      import :: c_int, c_int64_t
      integer(c_int), value :: n
    end function

    integer(c_int) function zv_build_mapping(n, is, js, ks, ls) bind(C, name="zv_build_mapping")
      import :: c_int
      integer(c_int), value :: n
      integer(c_int), intent(out) :: is(*), js(*), ks(*), ls(*)
    end function

    integer(c_int) function zv_gauss_eri(ctx, ijkl_gaus, precond, n_screened) &
        bind(C, name="zv_gauss_eri")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: ijkl_gaus(*), precond(*)
      integer(c_int64_t), intent(out) :: n_screened
    end function

    integer(c_int) function zv_exact_monocenter(ctx, ijkl, err, n_mono) &
        bind(C, name="zv_exact_monocenter")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      real(c_double), intent(inout) :: ijkl(*), err(*)
      integer(c_int64_t), intent(out) :: n_mono
    end function

    integer(c_int) function zv_one_e(ctx, nnuc, charge, nuc, s, v, k) bind(C, name="zv_one_e")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nnuc
      real(c_double), intent(in) :: charge(*), nuc(*)
      real(c_double), intent(out) :: s(*), v(*), k(*)
    end function

    integer(c_int) function zv_gauss_ijkl_list(ctx, nq, qi, qk, qj, ql, out) &
        bind(C, name="zv_gauss_ijkl_list")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: nq
      integer(c_int), intent(in) :: qi(*), qk(*), qj(*), ql(*)
      real(c_double), intent(out) :: out(*)
    end function

    integer(c_int) function zv_mc_blocks(ctx, nbl, npts, mode, seed, stream, precision, r1, r2, &
        pi0, mean, err, block_means) bind(C, name="zv_mc_blocks")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: nbl, mode, precision
      integer(c_int64_t), value :: npts, seed, stream
      type(c_ptr), value :: r1, r2, pi0, block_means
      real(c_double), intent(out) :: mean(*), err(*)
    end function

    integer(c_int) function zv_svd_clean(ctx, rank, q, seed, eri, n_removed) &
        bind(C, name="zv_svd_clean")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: rank, q
      integer(c_int64_t), value :: seed
      real(c_double), intent(inout) :: eri(*)
      integer(c_int), intent(out) :: n_removed
    end function

    integer(c_int) function zv_davidson_clean(ctx, eri, n_removed) &
        bind(C, name="zv_davidson_clean")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(inout) :: eri(*)
      integer(c_int), intent(out) :: n_removed
    end function
    integer(c_int) function zv_get_result(ctx, which, host) bind(C, name="zv_get_result")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: which
      real(c_double), intent(out) :: host(*)
    end function
    integer(c_int) function zv_result_max(ctx, which, val) bind(C, name="zv_result_max")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: which
      real(c_double), intent(out) :: val
    end function
    integer(c_int) function zv_set_verbose(ctx, level) bind(C, name="zv_set_verbose")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: level
    end function
    integer(c_int) function zv_set_tuning(ctx, key, val) bind(C, name="zv_set_tuning")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: key
      integer(c_int64_t), value :: val
    end function
    integer(c_int) function zv_mc_refine(ctx, nbl, npts, seed, stream, precision, target, &
        n_refined, max_err) bind(C, name="zv_mc_refine")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: nbl, precision
      integer(c_int64_t), value :: npts, seed, stream
      real(c_double), value :: target
      integer(c_int64_t), intent(out) :: n_refined
      real(c_double), intent(out) :: max_err
    end function
    integer(c_int) function zv_rsvd_clean(ctx, rank, q, seed, sketch, eri, ritz, n_removed) &
        bind(C, name="zv_rsvd_clean")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: rank, q, sketch
      integer(c_int64_t), value :: seed
      real(c_double), intent(inout) :: eri(*)
      type(c_ptr), value :: ritz
      integer(c_int), intent(out) :: n_removed
    end function
    integer(c_int) function zv_davidson_lowest(ctx, nev, tol, max_iter, precond, eri, evals, &
        resid, n_removed, iters) bind(C, name="zv_davidson_lowest")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nev, max_iter, precond
      real(c_double), value :: tol
      real(c_double), intent(inout) :: eri(*)
      real(c_double), intent(out) :: evals(*), resid(*)
      integer(c_int), intent(out) :: n_removed, iters
    end function
  end interface

contains

  ! The reference's error convention: print the message and stop.
  subroutine zv_check(ctx, rc, what)
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: what
    character(kind=c_char) :: buf(512)
    character(512) :: msg
    integer :: t, ierr
    if (rc == ZV_OK) return
    ierr = zv_last_error(ctx, buf, 512_c_int)
    msg = ' '
    do t = 1, 512
      if (buf(t) == c_null_char) exit
      msg(t:t) = buf(t)
    end do
    print *, trim(what), ' failed (', rc, '): ', trim(msg)
    stop 1
  end subroutine

end module zvmc
