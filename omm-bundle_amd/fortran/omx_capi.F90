!==================================================================================================!
! ISO_C_BINDING interfaces to libommhip.so (C-ABI declared in include/omm_hip.h).                  !
! Matrix arguments are device addresses (type(c_ptr)); scalars are passed by value.                !
!==================================================================================================!
module omx_capi
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: OMX_F64=0, OMX_F32=1, OMX_BF16=2, OMX_C128=3, OMX_C64=4
  integer(c_int), parameter :: OMX_PREC_F64=0, OMX_PREC_F32=1, OMX_PREC_BF16=2, OMX_PREC_BF16X3=3, &
                                OMX_PREC_BF16X6=4
  integer(c_int), parameter :: OMX_K_SPMM=0, OMX_K_GEMM=1, OMX_K_COEFF=2, OMX_K_ELEM=3, OMX_K_DOT=4

  interface
    integer(c_int) function omx_device_count(count) bind(c, name='omx_device_count')
      import :: c_int
      integer(c_int), intent(out) :: count
    end function
    integer(c_int) function omx_init(device, ctx) bind(c, name='omx_init')
      import :: c_int, c_ptr
      integer(c_int), value :: device
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function omx_finalize(ctx) bind(c, name='omx_finalize')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function omx_sync(ctx) bind(c, name='omx_sync')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    type(c_ptr) function omx_last_error() bind(c, name='omx_last_error')
      import :: c_ptr
    end function
    integer(c_int) function omx_malloc(ctx, bytes, dptr) bind(c, name='omx_malloc')
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: ctx
      integer(c_size_t), value :: bytes
      type(c_ptr), intent(out) :: dptr
    end function
    integer(c_int) function omx_free(ctx, dptr) bind(c, name='omx_free')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, dptr
    end function
    integer(c_int) function omx_h2d_2d(ctx, dst, ldd, src, lds, m, n, esize) bind(c, name='omx_h2d_2d')
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx, dst, src
      integer(c_int64_t), value :: ldd, lds
      integer(c_int), value :: m, n, esize
    end function
    integer(c_int) function omx_d2h_2d(ctx, dst, ldd, src, lds, m, n, esize) bind(c, name='omx_d2h_2d')
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx, dst, src
      integer(c_int64_t), value :: ldd, lds
      integer(c_int), value :: m, n, esize
    end function
    integer(c_int) function omx_d2d(ctx, dst, src, bytes) bind(c, name='omx_d2d')
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: ctx, dst, src
      integer(c_size_t), value :: bytes
    end function
    integer(c_int) function omx_convert(ctx, m, n, src, dt_src, lds, dst, dt_dst, ldd) bind(c, name='omx_convert')
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx, src, dst
      integer(c_int), value :: m, n, dt_src, dt_dst
      integer(c_int64_t), value :: lds, ldd
    end function
    integer(c_int) function omx_bsr_from_csc(ctx, nb, colptr, rowidx, val, ptrbase, rowbase, out) &
        bind(c, name='omx_bsr_from_csc')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, colptr, rowidx, val
      integer(c_int), value :: nb, ptrbase, rowbase
      type(c_ptr), intent(out) :: out
    end function
    integer(c_int) function omx_bsr_create(ctx, nb, bcolptr, browidx, bval, out) bind(c, name='omx_bsr_create')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, bcolptr, browidx, bval
      integer(c_int), value :: nb
      type(c_ptr), intent(out) :: out
    end function
    integer(c_int) function omx_bsr_from_csc_z(ctx, nb, colptr, rowidx, val, ptrbase, rowbase, out) &
        bind(c, name='omx_bsr_from_csc_z')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, colptr, rowidx, val
      integer(c_int), value :: nb, ptrbase, rowbase
      type(c_ptr), intent(out) :: out
    end function
    integer(c_int) function omx_bsr_create_z(ctx, nb, bcolptr, browidx, bval, out) bind(c, name='omx_bsr_create_z')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, bcolptr, browidx, bval
      integer(c_int), value :: nb
      type(c_ptr), intent(out) :: out
    end function
    integer(c_int) function omx_set_element_z(ctx, dt, C, ldc, i, j, alpha_re, alpha_im, beta) &
        bind(c, name='omx_set_element_z')
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: ctx, C
      integer(c_int), value :: dt, i, j
      integer(c_int64_t), value :: ldc
      real(c_double), value :: alpha_re, alpha_im, beta
    end function
    integer(c_int) function omx_potrf(ctx, dt, n, A, lda, info) bind(c, name='omx_potrf')
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx, A
      integer(c_int), value :: dt, n
      integer(c_int64_t), value :: lda
      integer(c_int), intent(out) :: info
    end function
    integer(c_int) function omx_trsm(ctx, dt, side, trans, m, n, alpha, U, ldu, B, ldb) bind(c, name='omx_trsm')
      import :: c_int, c_ptr, c_int64_t, c_double, c_char
      type(c_ptr), value :: ctx, U, B
      integer(c_int), value :: dt, m, n
      character(kind=c_char), value :: side, trans
      real(c_double), value :: alpha
      integer(c_int64_t), value :: ldu, ldb
    end function
    integer(c_int) function omx_trmm(ctx, dt, side, trans, m, n, alpha, U, ldu, B, ldb) bind(c, name='omx_trmm')
      import :: c_int, c_ptr, c_int64_t, c_double, c_char
      type(c_ptr), value :: ctx, U, B
      integer(c_int), value :: dt, m, n
      character(kind=c_char), value :: side, trans
      real(c_double), value :: alpha
      integer(c_int64_t), value :: ldu, ldb
    end function
    integer(c_int) function omx_sygst(ctx, dt, n, H, ldh, U, ldu) bind(c, name='omx_sygst')
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx, H, U
      integer(c_int), value :: dt, n
      integer(c_int64_t), value :: ldh, ldu
    end function
    integer(c_int) function omx_inverse_spd(ctx, dt, n, A, lda, info) bind(c, name='omx_inverse_spd')
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx, A
      integer(c_int), value :: dt, n
      integer(c_int64_t), value :: lda
      integer(c_int), intent(out) :: info
    end function
    integer(c_int) function omx_bsr_destroy(a) bind(c, name='omx_bsr_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: a
    end function
    integer(c_int) function omx_spmm(ctx, dt, no, nb, alpha, X, ldx, A, beta, Y, ldy) bind(c, name='omx_spmm')
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: ctx, X, A, Y
      integer(c_int), value :: dt, no, nb
      real(c_double), value :: alpha, beta
      integer(c_int64_t), value :: ldx, ldy
    end function
    integer(c_int) function omx_sddmm_bsr(ctx, dt, no, nb, alpha, A, lda, B, ldb, beta, Cp) &
        bind(c, name='omx_sddmm_bsr')
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: ctx, A, B, Cp
      integer(c_int), value :: dt, no, nb
      real(c_double), value :: alpha, beta
      integer(c_int64_t), value :: lda, ldb
    end function
    integer(c_int) function omx_bsr_get_element(a, i, j, out) bind(c, name='omx_bsr_get_element')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: a
      integer(c_int), value :: i, j
      real(c_double), intent(out) :: out
    end function
    integer(c_int) function omx_bsr_set_walk_groups(a, order, group) bind(c, name='omx_bsr_set_walk_groups')
      import :: c_int, c_ptr
      type(c_ptr), value :: a, order
      integer(c_int), value :: group
    end function
    integer(c_int) function omx_bsr_param_grad(ctx, Dp, nparam, seg, ent, coef, grad) &
        bind(c, name='omx_bsr_param_grad')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, Dp, seg, ent, coef
      integer(c_int), value :: nparam
      real(c_double), intent(out) :: grad(*)
    end function
    integer(c_int) function omx_gemm(ctx, prec, dt, opA, opB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc) &
        bind(c, name='omx_gemm')
      import :: c_int, c_ptr, c_int64_t, c_double, c_char
      type(c_ptr), value :: ctx, A, B, C
      integer(c_int), value :: prec, dt, m, n, k
      character(kind=c_char), value :: opA, opB
      real(c_double), value :: alpha, beta
      integer(c_int64_t), value :: lda, ldb, ldc
    end function
    integer(c_int) function omx_gemm_sym(ctx, prec, dt, n, k, alpha, A, lda, B, ldb, C, ldc) &
        bind(c, name='omx_gemm_sym')
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: ctx, A, B, C
      integer(c_int), value :: prec, dt, n, k
      real(c_double), value :: alpha
      integer(c_int64_t), value :: lda, ldb, ldc
    end function
    integer(c_int) function omx_add(ctx, dt, opA, m, n, alpha, A, lda, beta, C, ldc) bind(c, name='omx_add')
      import :: c_int, c_ptr, c_int64_t, c_double, c_char
      type(c_ptr), value :: ctx, A, C
      integer(c_int), value :: dt, m, n
      character(kind=c_char), value :: opA
      real(c_double), value :: alpha, beta
      integer(c_int64_t), value :: lda, ldc
    end function
    integer(c_int) function omx_scale(ctx, dt, m, n, beta, C, ldc) bind(c, name='omx_scale')
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: ctx, C
      integer(c_int), value :: dt, m, n
      real(c_double), value :: beta
      integer(c_int64_t), value :: ldc
    end function
    integer(c_int) function omx_set(ctx, dt, seC, m, n, alpha, beta, C, ldc) bind(c, name='omx_set')
      import :: c_int, c_ptr, c_int64_t, c_double, c_char
      type(c_ptr), value :: ctx, C
      integer(c_int), value :: dt, m, n
      character(kind=c_char), value :: seC
      real(c_double), value :: alpha, beta
      integer(c_int64_t), value :: ldc
    end function
    integer(c_int) function omx_trace(ctx, dt, n, A, lda, out) bind(c, name='omx_trace')
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: ctx, A
      integer(c_int), value :: dt, n
      integer(c_int64_t), value :: lda
      real(c_double), intent(out) :: out(*)
    end function
    integer(c_int) function omx_mm_trace(ctx, dt, m, n, A, lda, B, ldb, out) bind(c, name='omx_mm_trace')
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: ctx, A, B
      integer(c_int), value :: dt, m, n
      integer(c_int64_t), value :: lda, ldb
      real(c_double), intent(out) :: out(*)
    end function
    integer(c_int) function omx_set_element(ctx, dt, C, ldc, i, j, alpha, beta) bind(c, name='omx_set_element')
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: ctx, C
      integer(c_int), value :: dt, i, j
      integer(c_int64_t), value :: ldc
      real(c_double), value :: alpha, beta
    end function
    integer(c_int) function omx_get_element(ctx, dt, C, ldc, i, j, out) bind(c, name='omx_get_element')
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: ctx, C
      integer(c_int), value :: dt, i, j
      integer(c_int64_t), value :: ldc
      real(c_double), intent(out) :: out(*)
    end function
    integer(c_int) function omx_omm_coeff(ctx, dt, n, ld, HW, SW, HWd, SWd, HWdd, SWdd, coeff) &
        bind(c, name='omx_omm_coeff')
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: ctx, HW, SW, HWd, SWd, HWdd, SWdd
      integer(c_int), value :: dt, n
      integer(c_int64_t), value :: ld
      real(c_double), intent(out) :: coeff(5)
    end function
    integer(c_int) function omx_omm_calc_G(ctx, prec, dt, n, m, HW, SW, ldw, HC, SC, G, ldc) &
        bind(c, name='omx_omm_calc_G')
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx, HW, SW, HC, SC, G
      integer(c_int), value :: prec, dt, n, m
      integer(c_int64_t), value :: ldw, ldc
    end function
    integer(c_int) function omx_omm_pr_dots(ctx, dt, n, m, G, G_p, ld, out) bind(c, name='omx_omm_pr_dots')
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: ctx, G, G_p
      integer(c_int), value :: dt, n, m
      integer(c_int64_t), value :: ld
      real(c_double), intent(out) :: out(2)
    end function
    integer(c_int) function omx_omm_cg_direction(ctx, dt, n, m, lambda, G, D, G_p, ld) &
        bind(c, name='omx_omm_cg_direction')
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: ctx, G, D, G_p
      integer(c_int), value :: dt, n, m
      real(c_double), value :: lambda
      integer(c_int64_t), value :: ld
    end function
    integer(c_int) function omx_omm_update_w(ctx, dt, n, ld, x, Wd, Wdd, W) bind(c, name='omx_omm_update_w')
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: ctx, Wd, Wdd, W
      integer(c_int), value :: dt, n
      integer(c_int64_t), value :: ld
      real(c_double), value :: x
    end function
  end interface

contains

  !> Message of the last C-ABI error as a Fortran string.
  function omx_error_string() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: s(:)
    integer :: n
    p = omx_last_error()
    msg = ''
    if (.not. c_associated(p)) return
    call c_f_pointer(p, s, [1024])
    n = 0
    do while (n < 1024)
      if (s(n+1) == c_null_char) exit
      n = n + 1
    end do
    allocate(character(len=n) :: msg)
    msg = transfer(s(1:n), msg)
  end function omx_error_string

end module omx_capi
