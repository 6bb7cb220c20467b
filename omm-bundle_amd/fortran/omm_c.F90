!==================================================================================================!
! C-callable facade over the Fortran MatrixSwitch / omm() drop-in (the CBIND idea of the          !
! reference's omm_wrapper, libOMM/src/omm_wrapper.F90:5-96): matrices live in a handle table and   !
! omm() is called with the reference argument list.  Used by the Python tests and the benchmark;  !
! Fortran callers use the MatrixSwitch module and omm() directly.                                  !
!==================================================================================================!
module omm_c
  use iso_c_binding
  use omx_capi
  use MatrixSwitch
  use omm_hip_ops
  implicit none
  private

  integer, parameter :: max_handles = 4096
  type(matrix), target, save :: table(max_handles)
  logical, save :: used(max_handles) = .false.

  interface
    subroutine omm(m, n, H, S, new_S, e_min, D_min, calc_ED, eta, C_min, init_C, T, scale_T, flavour, np, ip, &
                   cg_tol, long_out, dealloc, m_storage, m_operation)
      import :: matrix, dp
      character(5), intent(in) :: m_storage
      character(3), intent(in) :: m_operation
      logical, intent(in) :: new_S, calc_ED, init_C, long_out, dealloc
      integer, intent(in) :: m, n, flavour, np, ip
      real(dp), intent(in) :: eta, cg_tol, scale_T
      real(dp), intent(out) :: e_min
      type(matrix), intent(inout) :: H, S, D_min, C_min, T
    end subroutine omm
  end interface

contains

  function cstr(s, n) result(f)
    character(kind=c_char), intent(in) :: s(*)
    integer, intent(in) :: n
    character(len=n) :: f
    integer :: i
    f = ' '
    do i = 1, n
      if (s(i) == c_null_char) exit
      f(i:i) = s(i)
    end do
  end function cstr

  integer function new_handle()
    integer :: h
    do h = 1, max_handles
      if (.not. used(h)) then
        used(h) = .true.
        new_handle = h
        return
      end if
    end do
    new_handle = -1
  end function new_handle

  logical function valid(h)
    integer(c_int), intent(in) :: h
    valid = (h >= 1 .and. h <= max_handles)
    if (valid) valid = used(h)
  end function valid

  integer(c_int) function omm_c_setup(device) bind(c, name='omm_c_setup')
    integer(c_int), value :: device
    call ms_hip_setup(int(device))
    omm_c_setup = 0
  end function omm_c_setup

  integer(c_int) function omm_c_finalize() bind(c, name='omm_c_finalize')
    integer :: h
    do h = 1, max_handles
      if (used(h)) then
        if (table(h)%is_initialized) call m_deallocate(table(h))
        used(h) = .false.
      end if
    end do
    call ms_hip_finalize()
    omm_c_finalize = 0
  end function omm_c_finalize

  !> Allocate a zero-filled matrix with a 5-character MatrixSwitch label; returns a handle (> 0).
  integer(c_int) function omm_c_mat_alloc(rows, cols, label) bind(c, name='omm_c_mat_alloc')
    integer(c_int), value :: rows, cols
    character(kind=c_char), intent(in) :: label(*)
    integer :: h
    h = new_handle()
    if (h < 0) then
      omm_c_mat_alloc = -1
      return
    end if
    call m_allocate(table(h), int(rows), int(cols), cstr(label, 5))
    omm_c_mat_alloc = h
  end function omm_c_mat_alloc

  !> Sparse device operator registered in pspBLAS's convention through m_register_pcsc: 1-based
  !> row indices idx1(nnz) and column pointers idx2(nb+1) (idx2(1) = 1), values val(nnz).
  integer(c_int) function omm_c_mat_register_pcsc(nb, nnz, idx1, idx2, val) bind(c, name='omm_c_mat_register_pcsc')
    integer(c_int), value :: nb, nnz
    integer(c_int), intent(in) :: idx1(*), idx2(*)
    real(c_double), intent(in) :: val(*)
    integer :: h, desc(9)
    h = new_handle()
    if (h < 0) then
      omm_c_mat_register_pcsc = -1
      return
    end if
    ! BLACS descriptor of one process holding the whole nb x nb matrix
    desc = (/1, 0, int(nb), int(nb), int(nb), int(nb), 0, 0, int(nb)/)
    call m_register_pcsc(table(h), idx1(1:nnz), idx2(1:nb+1), val(1:nnz), desc)
    omm_c_mat_register_pcsc = h
  end function omm_c_mat_register_pcsc

  !> m_convert of a handle's matrix to a new 5-character label (threshold < 0: none).
  integer(c_int) function omm_c_mat_convert(h, label, threshold) bind(c, name='omm_c_mat_convert')
    integer(c_int), value :: h
    character(kind=c_char), intent(in) :: label(*)
    real(c_double), value :: threshold
    if (.not. valid(h)) then
      omm_c_mat_convert = -1
      return
    end if
    if (threshold >= 0.0_dp) then
      call m_convert(table(h), cstr(label, 5), threshold)
    else
      call m_convert(table(h), cstr(label, 5))
    end if
    omm_c_mat_convert = 0
  end function omm_c_mat_convert

  !> Sparse operator from a 0-based CSC (colptr nb+1, rowidx, val); label 'gdcsc' or 'sdcsc'.
  integer(c_int) function omm_c_mat_from_csc(nb, colptr, rowidx, val, label) bind(c, name='omm_c_mat_from_csc')
    integer(c_int), value :: nb
    integer(c_int), intent(in) :: colptr(*), rowidx(*)
    real(c_double), intent(in) :: val(*)
    character(kind=c_char), intent(in) :: label(*)
    integer :: h, nnz
    character(5) :: lab
    type(matrix) :: tmp
    lab = cstr(label, 5)
    h = new_handle()
    if (h < 0) then
      omm_c_mat_from_csc = -1
      return
    end if
    nnz = colptr(nb+1)
    ! host CSC in MatrixSwitch's 'sdcsc' convention, then m_copy to the requested label
    tmp%str_type = 'csc'
    tmp%is_initialized = .true.
    tmp%is_serial = .true.
    tmp%is_real = .true.
    tmp%is_sparse = .true.
    tmp%is_square = .true.
    tmp%dim1 = nb
    tmp%dim2 = nb
    allocate(tmp%iaux2(1), tmp%iaux3(max(nnz,1)), tmp%iaux4(nb+1), tmp%dval(max(nnz,1),1))
    tmp%iaux2_is_allocated = .true.
    tmp%iaux3_is_allocated = .true.
    tmp%iaux4_is_allocated = .true.
    tmp%dval_is_allocated = .true.
    tmp%iaux2(1) = nnz
    tmp%iaux3(1:nnz) = rowidx(1:nnz) + 1
    tmp%iaux4(1:nb+1) = colptr(1:nb+1)
    tmp%dval(1:nnz,1) = val(1:nnz)
    call m_copy(table(h), tmp, lab)
    call m_deallocate(tmp)
    omm_c_mat_from_csc = h
  end function omm_c_mat_from_csc

  !> Device sparse operator from 0-based BSR-4 block columns (m_register_gbsr); returns a handle.
  integer(c_int) function omm_c_mat_from_bsr(nb, bcolptr, browidx, bval) bind(c, name='omm_c_mat_from_bsr')
    integer(c_int), value :: nb
    integer(c_int), intent(in) :: bcolptr(*), browidx(*)
    real(c_double), intent(in) :: bval(*)
    integer :: h, nbc, nnzb
    h = new_handle()
    if (h < 0) then
      omm_c_mat_from_bsr = -1
      return
    end if
    nbc = (nb + 3)/4
    nnzb = bcolptr(nbc+1)
    call m_register_gbsr(table(h), int(nb), bcolptr(1:nbc+1), browidx(1:max(nnzb,1)), bval(1:max(16*nnzb,1)))
    omm_c_mat_from_bsr = h
  end function omm_c_mat_from_bsr

  !> Block-column walk (0-based order of the nb/4 block columns) with LDS-staged groups of `group`
  !> block columns for a real device operator (m_set_walk).
  integer(c_int) function omm_c_mat_set_walk(h, nbc, order, group) bind(c, name='omm_c_mat_set_walk')
    integer(c_int), value :: h, nbc, group
    integer(c_int), intent(in) :: order(*)
    omm_c_mat_set_walk = -1
    if (.not. valid(h)) return
    call m_set_walk(table(h), order(1:nbc), int(group))
    omm_c_mat_set_walk = 0
  end function omm_c_mat_set_walk

  !> Complex device operator ('gzcsc') from 0-based BSR-4 block columns, 32 doubles per block.
  integer(c_int) function omm_c_mat_from_bsr_z(nb, bcolptr, browidx, bval) bind(c, name='omm_c_mat_from_bsr_z')
    integer(c_int), value :: nb
    integer(c_int), intent(in) :: bcolptr(*), browidx(*)
    real(c_double), intent(in) :: bval(*)
    integer :: h, nbc, nnzb
    h = new_handle()
    if (h < 0) then
      omm_c_mat_from_bsr_z = -1
      return
    end if
    nbc = (nb + 3)/4
    nnzb = bcolptr(nbc+1)
    call m_register_gbsr_z(table(h), int(nb), bcolptr(1:nbc+1), browidx(1:max(nnzb,1)), bval(1:max(32*nnzb,1)))
    omm_c_mat_from_bsr_z = h
  end function omm_c_mat_from_bsr_z

  !> Complex operator from a dense host complex matrix through m_copy (entries |a| > thr).
  integer(c_int) function omm_c_mat_from_dense_z(nb, a, thr, label) bind(c, name='omm_c_mat_from_dense_z')
    integer(c_int), value :: nb
    type(c_ptr), value :: a
    real(c_double), value :: thr
    character(kind=c_char), intent(in) :: label(*)
    complex(dp), pointer :: ap(:,:)
    type(matrix) :: tmp
    integer :: h
    h = new_handle()
    if (h < 0) then
      omm_c_mat_from_dense_z = -1
      return
    end if
    call c_f_pointer(a, ap, [nb, nb])
    call m_allocate(tmp, int(nb), int(nb), 'szden')
    tmp%zval = ap
    call m_copy(table(h), tmp, cstr(label, 5), thr)
    call m_deallocate(tmp)
    omm_c_mat_from_dense_z = h
  end function omm_c_mat_from_dense_z

  integer(c_int) function omm_c_mat_upload_z(h, a) bind(c, name='omm_c_mat_upload_z')
    integer(c_int), value :: h
    type(c_ptr), value :: a
    complex(dp), pointer :: ap(:,:)
    omm_c_mat_upload_z = -1
    if (.not. valid(h)) return
    call c_f_pointer(a, ap, [table(h)%dim1, table(h)%dim2])
    if (table(h)%is_device) then
      call ms_upload(table(h), ap)
    else
      table(h)%zval = ap
    end if
    omm_c_mat_upload_z = 0
  end function omm_c_mat_upload_z

  integer(c_int) function omm_c_mat_download_z(h, a) bind(c, name='omm_c_mat_download_z')
    integer(c_int), value :: h
    type(c_ptr), value :: a
    complex(dp), pointer :: ap(:,:)
    omm_c_mat_download_z = -1
    if (.not. valid(h)) return
    call c_f_pointer(a, ap, [table(h)%dim1, table(h)%dim2])
    if (table(h)%is_device) then
      call ms_download(table(h), ap)
    else
      ap = table(h)%zval
    end if
    omm_c_mat_download_z = 0
  end function omm_c_mat_download_z

  integer(c_int) function omm_c_get_element_z(h, i, j, re, im) bind(c, name='omm_c_get_element_z')
    integer(c_int), value :: h, i, j
    real(c_double), intent(out) :: re, im
    complex(dp) :: z
    omm_c_get_element_z = -1
    if (.not. valid(h)) return
    call m_get_element(table(h), int(i), int(j), z)
    re = real(z, dp)
    im = aimag(z)
    omm_c_get_element_z = 0
  end function omm_c_get_element_z

  integer(c_int) function omm_c_mat_upload(h, a) bind(c, name='omm_c_mat_upload')
    integer(c_int), value :: h
    type(c_ptr), value :: a
    real(dp), pointer :: ap(:,:)
    omm_c_mat_upload = -1
    if (.not. valid(h)) return
    call c_f_pointer(a, ap, [table(h)%dim1, table(h)%dim2])
    if (table(h)%is_device) then
      call ms_upload(table(h), ap)
    else
      table(h)%dval = ap
    end if
    omm_c_mat_upload = 0
  end function omm_c_mat_upload

  integer(c_int) function omm_c_mat_download(h, a) bind(c, name='omm_c_mat_download')
    integer(c_int), value :: h
    type(c_ptr), value :: a
    real(dp), pointer :: ap(:,:)
    omm_c_mat_download = -1
    if (.not. valid(h)) return
    call c_f_pointer(a, ap, [table(h)%dim1, table(h)%dim2])
    if (table(h)%is_device) then
      call ms_download(table(h), ap)
    else
      ap = table(h)%dval
    end if
    omm_c_mat_download = 0
  end function omm_c_mat_download

  integer(c_int) function omm_c_mat_free(h) bind(c, name='omm_c_mat_free')
    integer(c_int), value :: h
    omm_c_mat_free = -1
    if (.not. valid(h)) return
    if (table(h)%is_initialized) call m_deallocate(table(h))
    used(h) = .false.
    omm_c_mat_free = 0
  end function omm_c_mat_free

  !> Device address and leading dimension of a dense device matrix (null for host matrices).
  type(c_ptr) function omm_c_mat_dptr(h, ld) bind(c, name='omm_c_mat_dptr')
    integer(c_int), value :: h
    integer(c_int64_t), intent(out) :: ld
    omm_c_mat_dptr = c_null_ptr
    ld = 0
    if (.not. valid(h)) return
    omm_c_mat_dptr = table(h)%dptr
    ld = table(h)%ld
  end function omm_c_mat_dptr

  !> omm() with handles; hS <= 0: no S (identity overlap); hD <= 0: skip the density matrix.
  integer(c_int) function omm_c_solve(m, n, hH, hS, new_S, e_min, hD, calc_ED, eta, hC, init_C, flavour, np, ip, &
                                      cg_tol, long_out, dealloc, storage, operation) bind(c, name='omm_c_solve')
    integer(c_int), value :: m, n, hH, hS, hD, hC, new_S, calc_ED, init_C, flavour, np, ip, long_out, dealloc
    real(c_double), value :: eta, cg_tol
    real(c_double), intent(out) :: e_min
    character(kind=c_char), intent(in) :: storage(*), operation(*)
    type(matrix), save, target :: S_absent, D_absent, T_absent
    type(matrix), pointer :: Sm, Dm
    omm_c_solve = -1
    if (.not. (valid(hH) .and. valid(hC))) return
    if (hS > 0) then
      if (.not. valid(hS)) return
      Sm => table(hS)
    else
      Sm => S_absent
    end if
    if (hD > 0) then
      if (.not. valid(hD)) return
      Dm => table(hD)
    else
      Dm => D_absent
    end if
    call omm(int(m), int(n), table(hH), Sm, new_S /= 0, e_min, Dm, calc_ED /= 0, eta, table(hC), init_C /= 0, &
             T_absent, 0.0_dp, int(flavour), int(np), int(ip), cg_tol, long_out /= 0, dealloc /= 0, &
             cstr(storage, 5), cstr(operation, 3))
    omm_c_solve = 0
  end function omm_c_solve

  !> Statistics of the last omm() call: icg, converged, line-search failures, low-precision steps.
  integer(c_int) function omm_c_stats(icg, conv, nfail, icg_lowp, TrQS) bind(c, name='omm_c_stats')
    integer(c_int), intent(out) :: icg, conv, nfail, icg_lowp
    real(c_double), intent(out) :: TrQS
    icg = stat_icg
    conv = stat_conv
    nfail = stat_nfail
    icg_lowp = stat_icg_lowp
    TrQS = stat_TrQS
    omm_c_stats = 0
  end function omm_c_stats

  !> Gradient norms of the last call's FP64 refinement ('hmx' with refine_grad > 0): at its start
  !> and at its last stop test (both 0 otherwise).
  integer(c_int) function omm_c_refine_stats(g0, gend) bind(c, name='omm_c_refine_stats')
    real(c_double), intent(out) :: g0, gend
    g0 = stat_g0
    gend = stat_gend
    omm_c_refine_stats = 0
  end function omm_c_refine_stats

  !> Per-step energies of the last call (up to cap); returns the number written.
  integer(c_int) function omm_c_elog(buf, cap) bind(c, name='omm_c_elog')
    real(c_double), intent(out) :: buf(*)
    integer(c_int), value :: cap
    integer :: k
    k = min(int(cap), stat_nlog)
    buf(1:k) = stat_elog(1:k)
    omm_c_elog = k
  end function omm_c_elog

  !> Wall-clock seconds at the start of the CG loop (t0, returned first) and at the end of each CG
  !> step of the last call (up to cap-1); returns the number written.
  integer(c_int) function omm_c_tlog(buf, cap) bind(c, name='omm_c_tlog')
    real(c_double), intent(out) :: buf(*)
    integer(c_int), value :: cap
    integer :: k
    k = min(int(cap) - 1, stat_nlog)
    buf(1) = stat_t0
    buf(2:k+1) = stat_tlog(1:k)
    omm_c_tlog = k + 1
  end function omm_c_tlog

  !> The C-ABI operator (omx_bsr) behind a device sparse matrix ('gdcsc', 'gzcsc'), null otherwise:
  !> for omx_bsr_get_values / omx_bsr_get_pattern from the caller.
  type(c_ptr) function omm_c_mat_bsr(h) bind(c, name='omm_c_mat_bsr')
    integer(c_int), value :: h
    omm_c_mat_bsr = c_null_ptr
    if (.not. valid(h)) return
    if (table(h)%is_device .and. table(h)%is_sparse) omm_c_mat_bsr = table(h)%bsr
  end function omm_c_mat_bsr

  !> The libommhip context used by the 'g' storage (for event timing from the caller).
  type(c_ptr) function omm_c_ctx() bind(c, name='omm_c_ctx')
    omm_c_ctx = ms_hip_ctx()
  end function omm_c_ctx

  integer(c_int) function omm_c_set_option(name, value) bind(c, name='omm_c_set_option')
    character(kind=c_char), intent(in) :: name(*)
    real(c_double), value :: value
    call omm_set_option(trim(cstr(name, 32)), value)
    omm_c_set_option = 0
  end function omm_c_set_option

  !> dE/dtheta from a sparse device D (dE/dH on H's pattern) and a device Jacobian
  !> (omm_param_grad); element (i, j) of any matrix through m_get_element.
  integer(c_int) function omm_c_param_grad(hD, nparam, seg, ent, coef, grad) bind(c, name='omm_c_param_grad')
    integer(c_int), value :: hD, nparam
    type(c_ptr), value :: seg, ent, coef
    real(c_double), intent(out) :: grad(*)
    omm_c_param_grad = -1
    if (.not. valid(hD)) return
    call omm_param_grad(table(hD), int(nparam), seg, ent, coef, grad(1:nparam))
    omm_c_param_grad = 0
  end function omm_c_param_grad

  integer(c_int) function omm_c_get_element(h, i, j, val) bind(c, name='omm_c_get_element')
    integer(c_int), value :: h, i, j
    real(c_double), intent(out) :: val
    omm_c_get_element = -1
    if (.not. valid(h)) return
    call m_get_element(table(h), int(i), int(j), val)
    omm_c_get_element = 0
  end function omm_c_get_element

  integer(c_int) function omm_c_sync() bind(c, name='omm_c_sync')
    call ms_hip_sync()
    omm_c_sync = 0
  end function omm_c_sync

end module omm_c
