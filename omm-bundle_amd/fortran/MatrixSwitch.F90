!==================================================================================================!
! MatrixSwitch drop-in with a device-resident storage format.                                      !
!                                                                                                  !
! Keeps the public API of the reference module (MatrixSwitch/src/MatrixSwitch.F90:265-295):       !
! type(matrix), m_allocate, m_deallocate, m_copy, mm_multiply, m_add, m_trace, mm_trace, m_scale,  !
! m_set, m_set_element, m_get_element, m_register_sden; and adds storage labels whose first       !
! character is 'g' (GPU-resident, HBM of an MI355X):                                               !
!   'gdden'  dense, FP64          'gsden'  dense, FP32                                             !
!   'gdcsc'  sparse operator (built from CSC, stored as BSR-4 blocks), FP64 values                 !
! and the operation labels (third argument of every op, as 'ref'/'lap' on the host):              !
!   'hip'  FP64 products on the matrix cores          'h32'  FP32 products                         !
!   'hbf'  BF16 products with FP32 accumulation       (traces, adds, sets: always FP64 accumulate) !
!   'hb3'  BF16x3 products (hi*hi + hi*lo + lo*hi of FP32 operands split in two BF16 terms)        !
!   'hb6'  BF16x6 products (six term products of a three-term BF16 split; FP32-grade)              !
! Host labels 'sdden' / 'sdcsc' keep the reference semantics ('ref' loops); 'lap' is served by the !
! same loops (no BLAS is linked).  Device matrices cannot alias host memory: m_register_sden stays !
! host-only, and m_copy moves data between host and device labels (upload / download).            !
! Errors call die() with the reference behaviour: message to MatrixSwitch.err, stop 1             !
! (MatrixSwitch/src/MatrixSwitch_ops.F90:232-254).                                                 !
!==================================================================================================!
module MatrixSwitch
  use iso_c_binding
  use omx_capi
  implicit none
  private

  integer, parameter, public :: dp=selected_real_kind(15,300)

  type, public :: matrix
    character(3) :: str_type                      ! 'den' / 'csc'
    logical :: is_initialized=.false.
    logical :: is_serial=.true.
    logical :: is_real=.true.
    logical :: is_square=.false.
    logical :: is_sparse=.false.
    logical :: iaux1_is_allocated=.false.
    logical :: iaux2_is_allocated=.false.
    logical :: iaux3_is_allocated=.false.
    logical :: iaux4_is_allocated=.false.
    logical :: dval_is_allocated=.false.
    logical :: zval_is_allocated=.false.
    integer :: dim1=0
    integer :: dim2=0
    integer, pointer :: iaux1(:) => null()
    integer, pointer :: iaux2(:) => null()
    integer, pointer :: iaux3(:) => null()        ! CSC row indices (1-based)
    integer, pointer :: iaux4(:) => null()        ! CSC column offsets (0-based, dim2+1)
    real(dp), pointer :: dval(:,:) => null()
    complex(dp), pointer :: zval(:,:) => null()
    ! ---- device storage ----
    logical :: is_device=.false.
    integer(c_int) :: dtype=OMX_F64               ! element type of the device buffer
    type(c_ptr) :: dptr=c_null_ptr                ! dense device buffer (column-major)
    integer(c_int64_t) :: ld=0                    ! its leading dimension
    type(c_ptr) :: bsr=c_null_ptr                 ! sparse device operator (BSR-4)
  end type matrix

  interface m_allocate
    module procedure m_allocate_elements
  end interface
  interface m_set_element
    module procedure m_dset_element, m_zset_element
  end interface
  interface m_get_element
    module procedure m_dget_element, m_zget_element
  end interface
  interface ms_upload
    module procedure ms_upload_d, ms_upload_z
  end interface
  interface m_register_pcsc
    module procedure m_register_pdcsc
  end interface
  interface ms_download
    module procedure ms_download_d, ms_download_z
  end interface

  public :: m_allocate, m_deallocate, m_copy, mm_multiply, m_add, m_trace, mm_trace, m_scale, m_set
  public :: m_set_element, m_get_element, m_register_sden, m_register_gbsr, m_register_gbsr_z, m_set_walk
  public :: m_register_pcsc, m_convert
  public :: ms_hip_setup, ms_hip_finalize, ms_hip_sync, ms_hip_ctx, ms_upload, ms_download, ms_convert
  public :: ms_label_prec, die

  type(c_ptr), save :: g_ctx = c_null_ptr
  integer, save :: g_device = 0

contains

  !================================================================================================!
  ! device context                                                                                 !
  !================================================================================================!

  !> Select the HIP device used by 'g' storage (default 0) and create its context.
  subroutine ms_hip_setup(device)
    integer, intent(in), optional :: device
    integer(c_int) :: rc
    if (c_associated(g_ctx)) return
    if (present(device)) g_device = device
    rc = omx_init(int(g_device, c_int), g_ctx)
    if (rc /= 0) call die('ms_hip_setup: '//omx_error_string())
  end subroutine ms_hip_setup

  subroutine ms_hip_finalize()
    integer(c_int) :: rc
    if (.not. c_associated(g_ctx)) return
    rc = omx_finalize(g_ctx)
    g_ctx = c_null_ptr
  end subroutine ms_hip_finalize

  subroutine ms_hip_sync()
    call chk(omx_sync(ctx()), 'ms_hip_sync')
  end subroutine ms_hip_sync

  !> The C-ABI context (created on first use).
  function ms_hip_ctx() result(c)
    type(c_ptr) :: c
    c = ctx()
  end function ms_hip_ctx

  function ctx() result(c)
    type(c_ptr) :: c
    if (.not. c_associated(g_ctx)) call ms_hip_setup()
    c = g_ctx
  end function ctx

  subroutine chk(rc, where)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: where
    if (rc /= 0) call die(where//': '//omx_error_string())
  end subroutine chk

  !> Product precision selected by an operation label.
  function ms_label_prec(label) result(prec)
    character(3), intent(in), optional :: label
    integer(c_int) :: prec
    prec = OMX_PREC_F64
    if (.not. present(label)) return
    select case (label)
    case ('hip', 'ref', 'lap')
      prec = OMX_PREC_F64
    case ('h32')
      prec = OMX_PREC_F32
    case ('hbf')
      prec = OMX_PREC_BF16
    case ('hb3')
      prec = OMX_PREC_BF16X3
    case ('hb6')
      prec = OMX_PREC_BF16X6
    case default
      call die('invalid implementation label '//label)
    end select
  end function ms_label_prec

  subroutine check_device_label(label, where)
    character(3), intent(in), optional :: label
    character(*), intent(in) :: where
    if (present(label)) then
      if (label /= 'hip' .and. label /= 'h32' .and. label /= 'hbf' .and. label /= 'hb3' .and. label /= 'hb6') &
        call die(where//': invalid implementation')
    end if
  end subroutine check_device_label

  subroutine check_host_label(label, where)
    character(3), intent(in), optional :: label
    character(*), intent(in) :: where
    if (present(label)) then
      if (label /= 'ref' .and. label /= 'lap') call die(where//': invalid implementation')
    end if
  end subroutine check_host_label

  integer(c_int) function esize(m)
    type(matrix), intent(in) :: m
    select case (m%dtype)
    case (OMX_F64)
      esize = 8
    case (OMX_F32)
      esize = 4
    case (OMX_C128)
      esize = 16
    case (OMX_C64)
      esize = 8
    case default
      esize = 8
    end select
  end function esize

  logical function is_cplx_dt(dt)
    integer(c_int), intent(in) :: dt
    is_cplx_dt = (dt == OMX_C128 .or. dt == OMX_C64)
  end function is_cplx_dt

  !================================================================================================!
  ! allocation                                                                                     !
  !================================================================================================!

  !> Allocate a zero-filled matrix (MatrixSwitch.F90:312-419); 'g' labels allocate HBM.
  subroutine m_allocate_elements(m_name, i, j, label)
    type(matrix), intent(inout) :: m_name
    integer, intent(in) :: i, j
    character(5), intent(in), optional :: label
    character(1) :: c1, c2
    integer(c_int) :: rc

    m_name%dim1 = i
    m_name%dim2 = j
    m_name%is_square = (i == j)
    if (present(label)) then
      c1 = label(1:1)
      c2 = label(2:2)
      m_name%str_type = label(3:5)
    else
      c1 = 's'
      c2 = 'd'
      m_name%str_type = 'den'
    end if
    select case (c1)
    case ('s')
      m_name%is_serial = .true.
      m_name%is_device = .false.
    case ('g')
      m_name%is_serial = .true.
      m_name%is_device = .true.
    case ('p')
      call die('m_allocate_elements: compile with MPI')
    case default
      call die('m_allocate_elements: invalid label')
    end select
    select case (c2)
    case ('d')
      m_name%is_real = .true.
      m_name%dtype = OMX_F64
    case ('s')
      if (.not. m_name%is_device) call die('m_allocate_elements: invalid label')
      m_name%is_real = .true.
      m_name%dtype = OMX_F32
    case ('z')
      m_name%is_real = .false.
      m_name%dtype = OMX_C128
    case ('c')
      if (.not. m_name%is_device) call die('m_allocate_elements: invalid label')
      m_name%is_real = .false.
      m_name%dtype = OMX_C64
    case default
      call die('m_allocate_elements: invalid label')
    end select

    if (m_name%str_type == 'den') then
      m_name%is_sparse = .false.
      if (m_name%is_device) then
        m_name%ld = max(i, 1)
        rc = omx_malloc(ctx(), int(m_name%ld, c_size_t) * int(max(j, 1), c_size_t) * int(esize(m_name), c_size_t), &
                        m_name%dptr)
        call chk(rc, 'm_allocate_elements')
      else if (m_name%is_real) then
        allocate(m_name%dval(i, j))
        m_name%dval_is_allocated = .true.
        m_name%dval = 0.0_dp
      else
        allocate(m_name%zval(i, j))
        m_name%zval_is_allocated = .true.
        m_name%zval = (0.0_dp, 0.0_dp)
      end if
    else if (m_name%str_type == 'csc') then
      m_name%is_sparse = .true.
      if (.not. m_name%is_device) then
        allocate(m_name%iaux2(1))
        m_name%iaux2_is_allocated = .true.
        m_name%iaux2(1) = 0
      end if
    else
      call die('m_allocate_elements: invalid label')
    end if
    m_name%is_initialized = .true.
  end subroutine m_allocate_elements

  !> Free a matrix (MatrixSwitch.F90:527-601): registered host arrays are only nullified.
  subroutine m_deallocate(m_name)
    type(matrix), intent(inout) :: m_name
    integer(c_int) :: rc
    if (m_name%is_device) then
      if (c_associated(m_name%dptr)) rc = omx_free(ctx(), m_name%dptr)
      if (c_associated(m_name%bsr)) rc = omx_bsr_destroy(m_name%bsr)
      m_name%dptr = c_null_ptr
      m_name%bsr = c_null_ptr
    end if
    if (associated(m_name%iaux1)) then
      if (m_name%iaux1_is_allocated) then
        deallocate(m_name%iaux1)
      else
        nullify(m_name%iaux1)
      end if
    end if
    if (associated(m_name%iaux2)) then
      if (m_name%iaux2_is_allocated) then
        deallocate(m_name%iaux2)
      else
        nullify(m_name%iaux2)
      end if
    end if
    if (associated(m_name%iaux3)) then
      if (m_name%iaux3_is_allocated) then
        deallocate(m_name%iaux3)
      else
        nullify(m_name%iaux3)
      end if
    end if
    if (associated(m_name%iaux4)) then
      if (m_name%iaux4_is_allocated) then
        deallocate(m_name%iaux4)
      else
        nullify(m_name%iaux4)
      end if
    end if
    if (associated(m_name%dval)) then
      if (m_name%dval_is_allocated) then
        deallocate(m_name%dval)
      else
        nullify(m_name%dval)
      end if
    end if
    if (associated(m_name%zval)) then
      if (m_name%zval_is_allocated) then
        deallocate(m_name%zval)
      else
        nullify(m_name%zval)
      end if
    end if
    m_name%iaux1_is_allocated = .false.
    m_name%iaux2_is_allocated = .false.
    m_name%iaux3_is_allocated = .false.
    m_name%iaux4_is_allocated = .false.
    m_name%dval_is_allocated = .false.
    m_name%zval_is_allocated = .false.
    m_name%is_device = .false.
    m_name%is_initialized = .false.
  end subroutine m_deallocate

  !> Wrap a caller's host array without copying (MatrixSwitch_m_register.F90:99-133).
  subroutine m_register_sden(m_name, A)
    type(matrix), intent(inout) :: m_name
    real(dp), intent(in), target :: A(:,:)
    m_name%str_type = 'den'
    m_name%is_serial = .true.
    m_name%is_real = .true.
    m_name%is_sparse = .false.
    m_name%is_device = .false.
    m_name%dim1 = size(A, 1)
    m_name%dim2 = size(A, 2)
    m_name%is_square = (m_name%dim1 == m_name%dim2)
    m_name%dval => A
    m_name%dval_is_allocated = .false.
    m_name%is_initialized = .true.
  end subroutine m_register_sden

  !> Copy a block-sparse operator given as 0-based BSR-4 block columns (block-CSC: bcolptr has
  !> ceil(nb/4)+1 entries, block rows ascending, 16 column-major values per block) into device
  !> storage ('gdcsc').  The natural input format for tight-binding Hamiltonians with 4-orbital
  !> sites; like m_register_pcsc (MatrixSwitch_m_register.F90:394-443) the data are copied.
  subroutine m_register_gbsr(m_name, nb, bcolptr, browidx, bval)
    type(matrix), intent(inout) :: m_name
    integer, intent(in) :: nb
    integer(c_int), intent(in), target :: bcolptr(:), browidx(:)
    real(dp), intent(in), target :: bval(:)
    integer(c_int) :: rc
    m_name%str_type = 'csc'
    m_name%is_serial = .true.
    m_name%is_real = .true.
    m_name%is_sparse = .true.
    m_name%is_device = .true.
    m_name%dtype = OMX_F64
    m_name%dim1 = nb
    m_name%dim2 = nb
    m_name%is_square = .true.
    rc = omx_bsr_create(ctx(), int(nb, c_int), c_loc(bcolptr), c_loc(browidx), c_loc(bval), m_name%bsr)
    call chk(rc, 'm_register_gbsr')
    m_name%is_initialized = .true.
  end subroutine m_register_gbsr

  !> Register a compressed-sparse-column matrix in pspBLAS's convention (MatrixSwitch_m_register.F90:
  !> 394-443 -> psp_register_dspm, pspMat.F90:87-140): idx1 = row indices (1-based), idx2 = column
  !> pointers (idx2(1) = 1, idx2(j+1)-idx2(j) entries in column j), val = the values, desc = the BLACS
  !> descriptor (desc(3), desc(4) = global rows and columns).  As in pspBLAS the data are copied, here
  !> into HBM as the device operator of 'gdcsc' storage (BSR-4; the pattern's blocks, duplicates
  !> added).  One process holds the whole matrix, so the local dimensions are the global ones.
  subroutine m_register_pdcsc(m_name, idx1, idx2, val, desc)
    type(matrix), intent(inout) :: m_name
    integer, intent(in), target :: desc(9)
    integer, intent(in), target :: idx1(:)
    integer, intent(in), target :: idx2(:)
    real(dp), intent(in), target :: val(:)
    integer(c_int), allocatable, target :: cp(:), ri(:)
    real(dp), allocatable, target :: cv(:)
    integer :: nb, nnz
    integer(c_int) :: rc
    if (desc(3) /= desc(4)) call die('m_register_pcsc: device sparse storage needs a square operator')
    nb = desc(4)
    if (size(idx2) < nb + 1) call die('m_register_pcsc: idx2 holds fewer than desc(4)+1 column pointers')
    nnz = idx2(nb+1) - 1
    if (nnz < 0 .or. nnz > size(idx1) .or. nnz > size(val)) call die('m_register_pcsc: idx2 does not match idx1/val')
    allocate(cp(nb+1), ri(max(nnz,1)), cv(max(nnz,1)))
    cp = int(idx2(1:nb+1), c_int)
    if (nnz > 0) then
      ri(1:nnz) = int(idx1(1:nnz), c_int)
      cv(1:nnz) = val(1:nnz)
    end if
    m_name%str_type = 'csc'
    m_name%is_serial = .true.
    m_name%is_real = .true.
    m_name%is_sparse = .true.
    m_name%is_device = .true.
    m_name%dtype = OMX_F64
    m_name%dim1 = desc(3)
    m_name%dim2 = desc(4)
    m_name%is_square = .true.
    rc = omx_bsr_from_csc(ctx(), int(nb, c_int), c_loc(cp), c_loc(ri), c_loc(cv), 1_c_int, 1_c_int, m_name%bsr)
    call chk(rc, 'm_register_pcsc')
    m_name%is_initialized = .true.
  end subroutine m_register_pdcsc

  !> Change a matrix's storage format in place (MatrixSwitch.F90:987-1041): m_copy to a temporary
  !> with the new label (and threshold), then m_deallocate.  The conversions are those of m_copy:
  !> dense host <-> device, FP64 <-> FP32 on the device, dense -> sparse ('sdcsc', 'gdcsc') with
  !> |a| > threshold.  The reference copies the temporary back and frees it; here m_name takes over
  !> the temporary's storage (its components are pointers and handles), which gives the same result
  !> without a second copy, also for a device operator, which m_copy cannot take as a source.
  subroutine m_convert(m_name, label, threshold, threshold_is_soft)
    type(matrix), intent(inout) :: m_name
    character(5), intent(in), optional :: label
    real(dp), intent(in), optional :: threshold
    logical, intent(in), optional :: threshold_is_soft
    type(matrix) :: temp_matrix
    call m_copy(temp_matrix, m_name, label, threshold, threshold_is_soft)
    call m_deallocate(m_name)
    m_name = temp_matrix
  end subroutine m_convert

  !> Schedule hint for a real device operator ('gdcsc'): the order in which products visit its
  !> block columns (0-based permutation of the nb/4 block columns), taken `group` at a time with
  !> LDS-staged gathers when group > 0.  Results are bitwise unchanged.  No reference counterpart.
  subroutine m_set_walk(m_name, order, group)
    type(matrix), intent(inout) :: m_name
    integer(c_int), intent(in), target :: order(:)
    integer, intent(in) :: group
    if (.not. (m_name%is_device .and. m_name%is_sparse)) call die('m_set_walk: not a device sparse operator')
    call chk(omx_bsr_set_walk_groups(m_name%bsr, c_loc(order), int(group, c_int)), 'm_set_walk')
  end subroutine m_set_walk

  !> m_register_gbsr for a complex operator: 32 doubles per block, (re, im) interleaved ('gzcsc').
  subroutine m_register_gbsr_z(m_name, nb, bcolptr, browidx, bval)
    type(matrix), intent(inout) :: m_name
    integer, intent(in) :: nb
    integer(c_int), intent(in), target :: bcolptr(:), browidx(:)
    real(dp), intent(in), target :: bval(:)
    integer(c_int) :: rc
    m_name%str_type = 'csc'
    m_name%is_serial = .true.
    m_name%is_real = .false.
    m_name%is_sparse = .true.
    m_name%is_device = .true.
    m_name%dtype = OMX_C128
    m_name%dim1 = nb
    m_name%dim2 = nb
    m_name%is_square = .true.
    rc = omx_bsr_create_z(ctx(), int(nb, c_int), c_loc(bcolptr), c_loc(browidx), c_loc(bval), m_name%bsr)
    call chk(rc, 'm_register_gbsr_z')
    m_name%is_initialized = .true.
  end subroutine m_register_gbsr_z

  !> Copy a dense device matrix into another allocated one of the same shape, converting the
  !> element type if they differ (e.g. FP64 <-> FP32 working sets).
  subroutine ms_convert(src, dst)
    type(matrix), intent(in) :: src
    type(matrix), intent(inout) :: dst
    integer(c_int) :: rc
    if (.not. (src%is_device .and. dst%is_device) .or. src%is_sparse .or. dst%is_sparse) &
      call die('ms_convert: dense device matrices expected')
    if (src%dim1 /= dst%dim1 .or. src%dim2 /= dst%dim2) call die('ms_convert: shape mismatch')
    if (is_cplx_dt(src%dtype) .neqv. is_cplx_dt(dst%dtype)) call die('ms_convert: real and complex storage')
    if (is_cplx_dt(src%dtype)) then
      ! complex m x n with leading dimension ld = real 2m x n with leading dimension 2*ld
      rc = omx_convert(ctx(), int(2*src%dim1, c_int), int(src%dim2, c_int), src%dptr, real_dt(src%dtype), &
                       2*src%ld, dst%dptr, real_dt(dst%dtype), 2*dst%ld)
    else
      rc = omx_convert(ctx(), int(src%dim1, c_int), int(src%dim2, c_int), src%dptr, src%dtype, src%ld, dst%dptr, &
                       dst%dtype, dst%ld)
    end if
    call chk(rc, 'ms_convert')
  end subroutine ms_convert

  integer(c_int) function real_dt(dt)
    integer(c_int), intent(in) :: dt
    real_dt = dt
    if (dt == OMX_C128) real_dt = OMX_F64
    if (dt == OMX_C64) real_dt = OMX_F32
  end function real_dt

  !> Host <-> device transfers of a dense complex matrix (C128 directly; C64 staged in C128 on the
  !> device and converted there).
  subroutine ms_upload_z(m_name, A)
    type(matrix), intent(inout) :: m_name
    complex(dp), intent(in), target, contiguous :: A(:,:)
    type(c_ptr) :: tmp
    integer(c_int) :: rc
    if (.not. (m_name%is_device .and. m_name%str_type == 'den')) call die('ms_upload: not a dense device matrix')
    if (.not. is_cplx_dt(m_name%dtype)) call die('ms_upload: complex data into real storage')
    if (size(A,1) /= m_name%dim1 .or. size(A,2) /= m_name%dim2) call die('ms_upload: shape mismatch')
    if (m_name%dtype == OMX_C128) then
      rc = omx_h2d_2d(ctx(), m_name%dptr, m_name%ld, c_loc(A), int(max(m_name%dim1,1), c_int64_t), &
                      int(m_name%dim1, c_int), int(m_name%dim2, c_int), 16_c_int)
      call chk(rc, 'ms_upload')
    else
      rc = omx_malloc(ctx(), int(max(m_name%dim1,1), c_size_t) * int(max(m_name%dim2,1), c_size_t) * 16_c_size_t, tmp)
      call chk(rc, 'ms_upload')
      rc = omx_h2d_2d(ctx(), tmp, int(max(m_name%dim1,1), c_int64_t), c_loc(A), int(max(m_name%dim1,1), c_int64_t), &
                      int(m_name%dim1, c_int), int(m_name%dim2, c_int), 16_c_int)
      call chk(rc, 'ms_upload')
      rc = omx_convert(ctx(), int(2*m_name%dim1, c_int), int(m_name%dim2, c_int), tmp, OMX_F64, &
                       int(2*max(m_name%dim1,1), c_int64_t), m_name%dptr, OMX_F32, 2*m_name%ld)
      call chk(rc, 'ms_upload')
      rc = omx_free(ctx(), tmp)
    end if
  end subroutine ms_upload_z

  subroutine ms_download_z(m_name, A)
    type(matrix), intent(in) :: m_name
    complex(dp), intent(inout), target, contiguous :: A(:,:)
    type(c_ptr) :: tmp
    integer(c_int) :: rc
    if (.not. (m_name%is_device .and. m_name%str_type == 'den')) call die('ms_download: not a dense device matrix')
    if (.not. is_cplx_dt(m_name%dtype)) call die('ms_download: real storage into complex data')
    if (size(A,1) /= m_name%dim1 .or. size(A,2) /= m_name%dim2) call die('ms_download: shape mismatch')
    if (m_name%dtype == OMX_C128) then
      rc = omx_d2h_2d(ctx(), c_loc(A), int(max(m_name%dim1,1), c_int64_t), m_name%dptr, m_name%ld, &
                      int(m_name%dim1, c_int), int(m_name%dim2, c_int), 16_c_int)
      call chk(rc, 'ms_download')
    else
      rc = omx_malloc(ctx(), int(max(m_name%dim1,1), c_size_t) * int(max(m_name%dim2,1), c_size_t) * 16_c_size_t, tmp)
      call chk(rc, 'ms_download')
      rc = omx_convert(ctx(), int(2*m_name%dim1, c_int), int(m_name%dim2, c_int), m_name%dptr, OMX_F32, &
                       2*m_name%ld, tmp, OMX_F64, int(2*max(m_name%dim1,1), c_int64_t))
      call chk(rc, 'ms_download')
      rc = omx_d2h_2d(ctx(), c_loc(A), int(max(m_name%dim1,1), c_int64_t), tmp, int(max(m_name%dim1,1), c_int64_t), &
                      int(m_name%dim1, c_int), int(m_name%dim2, c_int), 16_c_int)
      call chk(rc, 'ms_download')
      rc = omx_free(ctx(), tmp)
    end if
  end subroutine ms_download_z

  !> Host <-> device transfers of a dense matrix (real data into complex storage get a zero
  !> imaginary part).
  subroutine ms_upload_d(m_name, A)
    type(matrix), intent(inout) :: m_name
    real(dp), intent(in), target, contiguous :: A(:,:)
    type(c_ptr) :: tmp
    integer(c_int) :: rc
    if (.not. (m_name%is_device .and. m_name%str_type == 'den')) call die('ms_upload: not a dense device matrix')
    if (size(A,1) /= m_name%dim1 .or. size(A,2) /= m_name%dim2) call die('ms_upload: shape mismatch')
    if (is_cplx_dt(m_name%dtype)) then
      call ms_upload_z(m_name, cmplx(A, 0.0_dp, dp))
      return
    end if
    if (m_name%dtype == OMX_F64) then
      rc = omx_h2d_2d(ctx(), m_name%dptr, m_name%ld, c_loc(A), int(max(m_name%dim1,1), c_int64_t), &
                      int(m_name%dim1, c_int), int(m_name%dim2, c_int), 8_c_int)
      call chk(rc, 'ms_upload')
    else
      ! FP64 host data into an FP32 buffer: stage in FP64 on the device and convert there.
      rc = omx_malloc(ctx(), int(max(m_name%dim1,1), c_size_t) * int(max(m_name%dim2,1), c_size_t) * 8_c_size_t, tmp)
      call chk(rc, 'ms_upload')
      rc = omx_h2d_2d(ctx(), tmp, int(max(m_name%dim1,1), c_int64_t), c_loc(A), int(max(m_name%dim1,1), c_int64_t), &
                      int(m_name%dim1, c_int), int(m_name%dim2, c_int), 8_c_int)
      call chk(rc, 'ms_upload')
      rc = omx_convert(ctx(), int(m_name%dim1, c_int), int(m_name%dim2, c_int), tmp, OMX_F64, &
                       int(max(m_name%dim1,1), c_int64_t), m_name%dptr, m_name%dtype, m_name%ld)
      call chk(rc, 'ms_upload')
      rc = omx_free(ctx(), tmp)
    end if
  end subroutine ms_upload_d

  subroutine ms_download_d(m_name, A)
    type(matrix), intent(in) :: m_name
    real(dp), intent(inout), target, contiguous :: A(:,:)
    type(c_ptr) :: tmp
    integer(c_int) :: rc
    if (.not. (m_name%is_device .and. m_name%str_type == 'den')) call die('ms_download: not a dense device matrix')
    if (size(A,1) /= m_name%dim1 .or. size(A,2) /= m_name%dim2) call die('ms_download: shape mismatch')
    if (is_cplx_dt(m_name%dtype)) call die('ms_download: complex storage into real data')
    if (m_name%dtype == OMX_F64) then
      rc = omx_d2h_2d(ctx(), c_loc(A), int(max(m_name%dim1,1), c_int64_t), m_name%dptr, m_name%ld, &
                      int(m_name%dim1, c_int), int(m_name%dim2, c_int), 8_c_int)
      call chk(rc, 'ms_download')
    else
      rc = omx_malloc(ctx(), int(max(m_name%dim1,1), c_size_t) * int(max(m_name%dim2,1), c_size_t) * 8_c_size_t, tmp)
      call chk(rc, 'ms_download')
      rc = omx_convert(ctx(), int(m_name%dim1, c_int), int(m_name%dim2, c_int), m_name%dptr, m_name%dtype, &
                       m_name%ld, tmp, OMX_F64, int(max(m_name%dim1,1), c_int64_t))
      call chk(rc, 'ms_download')
      rc = omx_d2h_2d(ctx(), c_loc(A), int(max(m_name%dim1,1), c_int64_t), tmp, int(max(m_name%dim1,1), c_int64_t), &
                      int(m_name%dim1, c_int), int(m_name%dim2, c_int), 8_c_int)
      call chk(rc, 'ms_download')
      rc = omx_free(ctx(), tmp)
    end if
  end subroutine ms_download_d

  !================================================================================================!
  ! m_copy (MatrixSwitch.F90:630-963): host copies, CSC construction, host <-> device moves       !
  !================================================================================================!
  subroutine m_copy(m_name, A, label, threshold, threshold_is_soft)
    type(matrix), intent(inout) :: m_name
    type(matrix), intent(inout) :: A
    character(5), intent(in), optional :: label
    real(dp), intent(in), optional :: threshold
    logical, intent(in), optional :: threshold_is_soft
    character(5) :: lab
    real(dp) :: thr, soft
    real(dp), allocatable, target :: tmp(:,:)
    integer :: i, j, k
    integer(c_int) :: rc
    integer, allocatable, target :: cp(:), ri(:)
    real(dp), allocatable, target :: cv(:)

    if (present(label)) then
      lab = label
    else
      if (A%is_device) then
        lab(1:1) = 'g'
        select case (A%dtype)
        case (OMX_F32)
          lab(2:2) = 's'
        case (OMX_C128)
          lab(2:2) = 'z'
        case (OMX_C64)
          lab(2:2) = 'c'
        case default
          lab(2:2) = 'd'
        end select
      else
        lab(1:1) = 's'
        lab(2:2) = 'd'
        if (.not. A%is_real) lab(2:2) = 'z'
      end if
      lab(3:5) = A%str_type
    end if
    thr = 0.0_dp
    soft = 0.0_dp
    if (present(threshold)) thr = threshold
    if (present(threshold_is_soft)) then
      if (threshold_is_soft) soft = thr
    end if
    if (.not. A%is_real) then
      call m_copy_z(m_name, A, lab, thr, present(threshold))
      return
    end if
    if (A%is_sparse .and. lab(3:5) == 'den') call die('m_copy: sparse to dense copy is not available')

    if (lab(3:5) == 'den') then
      call m_allocate(m_name, A%dim1, A%dim2, lab)
      if (.not. A%is_device .and. .not. m_name%is_device) then
        m_name%dval = A%dval
        if (present(threshold)) where (abs(m_name%dval) <= thr) m_name%dval = 0.0_dp
      else if (.not. A%is_device .and. m_name%is_device) then
        allocate(tmp(A%dim1, A%dim2))
        tmp = A%dval
        if (present(threshold)) where (abs(tmp) <= thr) tmp = 0.0_dp
        call ms_upload(m_name, tmp)
      else if (A%is_device .and. .not. m_name%is_device) then
        allocate(tmp(A%dim1, A%dim2))
        call ms_download(A, tmp)
        if (present(threshold)) where (abs(tmp) <= thr) tmp = 0.0_dp
        m_name%dval = tmp
      else
        if (A%dtype == m_name%dtype) then
          rc = omx_d2d(ctx(), m_name%dptr, A%dptr, int(A%ld, c_size_t) * int(max(A%dim2,1), c_size_t) &
                       * int(esize(A), c_size_t))
        else
          rc = omx_convert(ctx(), int(A%dim1, c_int), int(A%dim2, c_int), A%dptr, A%dtype, A%ld, m_name%dptr, &
                           m_name%dtype, m_name%ld)
        end if
        call chk(rc, 'm_copy')
      end if
    else if (lab(3:5) == 'csc') then
      ! CSC from a dense (host or device) matrix, keeping |a| > threshold, or from a host CSC
      ! (MatrixSwitch_m_copy.F90:495-629; row indices 1-based, offsets 0-based).
      if (A%is_sparse) then
        if (A%is_device) call die('m_copy: device sparse source is not available')
        allocate(cp(A%dim2+1), ri(A%iaux2(1)), cv(A%iaux2(1)))
        cp = A%iaux4
        ri = A%iaux3
        cv = A%dval(1:A%iaux2(1), 1)
      else
        allocate(tmp(A%dim1, A%dim2))
        if (A%is_device) then
          call ms_download(A, tmp)
        else
          tmp = A%dval
        end if
        k = count(abs(tmp) > thr)
        allocate(cp(A%dim2+1), ri(max(k,1)), cv(max(k,1)))
        k = 0
        cp(1) = 0
        do i = 1, A%dim2
          do j = 1, A%dim1
            if (abs(tmp(j,i)) > thr) then
              k = k + 1
              ri(k) = j
              cv(k) = tmp(j,i) - soft*tmp(j,i)/abs(tmp(j,i))
            end if
          end do
          cp(i+1) = k
        end do
      end if
      m_name%dim1 = A%dim1
      m_name%dim2 = A%dim2
      m_name%is_square = A%is_square
      m_name%is_real = .true.
      m_name%is_serial = .true.
      m_name%is_sparse = .true.
      m_name%str_type = 'csc'
      if (lab(1:1) == 'g') then
        if (A%dim1 /= A%dim2) call die('m_copy: device sparse storage needs a square operator')
        m_name%is_device = .true.
        m_name%dtype = OMX_F64
        rc = omx_bsr_from_csc(ctx(), int(A%dim2, c_int), c_loc(cp), c_loc(ri), c_loc(cv), 0_c_int, 1_c_int, &
                              m_name%bsr)
        call chk(rc, 'm_copy')
      else if (lab(1:1) == 's') then
        m_name%is_device = .false.
        allocate(m_name%iaux2(1), m_name%iaux3(size(ri)), m_name%iaux4(size(cp)), m_name%dval(size(cv),1))
        m_name%iaux2(1) = cp(A%dim2+1)
        m_name%iaux3 = ri
        m_name%iaux4 = cp
        m_name%dval(:,1) = cv
        m_name%iaux2_is_allocated = .true.
        m_name%iaux3_is_allocated = .true.
        m_name%iaux4_is_allocated = .true.
        m_name%dval_is_allocated = .true.
      else
        call die('m_copy: invalid label')
      end if
      m_name%is_initialized = .true.
    else
      call die('m_copy: invalid label')
    end if
  end subroutine m_copy

  !> m_copy of complex matrices: dense host <-> device moves, and dense -> device sparse operator
  !> ('gzcsc', entries with |a| > threshold; MatrixSwitch_m_copy.F90:495-629 for the CSC rules).
  subroutine m_copy_z(m_name, A, lab, thr, use_thr)
    type(matrix), intent(inout) :: m_name
    type(matrix), intent(inout) :: A
    character(5), intent(in) :: lab
    real(dp), intent(in) :: thr
    logical, intent(in) :: use_thr
    complex(dp), allocatable, target :: tmp(:,:)
    integer, allocatable, target :: cp(:), ri(:)
    real(dp), allocatable, target :: cv(:)
    integer :: i, j, k
    integer(c_int) :: rc
    if (lab(2:2) /= 'z' .and. lab(2:2) /= 'c') call die('m_copy: complex source needs a complex label')
    if (A%is_sparse) call die('m_copy: complex sparse source is not available')
    if (lab(3:5) == 'den') then
      call m_allocate(m_name, A%dim1, A%dim2, lab)
      if (.not. A%is_device .and. .not. m_name%is_device) then
        m_name%zval = A%zval
        if (use_thr) where (abs(m_name%zval) <= thr) m_name%zval = (0.0_dp, 0.0_dp)
      else if (.not. A%is_device .and. m_name%is_device) then
        allocate(tmp(A%dim1, A%dim2))
        tmp = A%zval
        if (use_thr) where (abs(tmp) <= thr) tmp = (0.0_dp, 0.0_dp)
        call ms_upload_z(m_name, tmp)
      else if (A%is_device .and. .not. m_name%is_device) then
        allocate(tmp(A%dim1, A%dim2))
        call ms_download_z(A, tmp)
        if (use_thr) where (abs(tmp) <= thr) tmp = (0.0_dp, 0.0_dp)
        m_name%zval = tmp
      else
        call ms_convert(A, m_name)
      end if
    else if (lab(3:5) == 'csc') then
      if (lab(1:1) /= 'g') call die('m_copy: complex host CSC is not available')
      if (A%dim1 /= A%dim2) call die('m_copy: device sparse storage needs a square operator')
      allocate(tmp(A%dim1, A%dim2))
      if (A%is_device) then
        call ms_download_z(A, tmp)
      else
        tmp = A%zval
      end if
      k = count(abs(tmp) > thr)
      allocate(cp(A%dim2+1), ri(max(k,1)), cv(2*max(k,1)))
      k = 0
      cp(1) = 0
      do i = 1, A%dim2
        do j = 1, A%dim1
          if (abs(tmp(j,i)) > thr) then
            k = k + 1
            ri(k) = j
            cv(2*k-1) = real(tmp(j,i), dp)
            cv(2*k) = aimag(tmp(j,i))
          end if
        end do
        cp(i+1) = k
      end do
      m_name%dim1 = A%dim1
      m_name%dim2 = A%dim2
      m_name%is_square = .true.
      m_name%is_real = .false.
      m_name%is_serial = .true.
      m_name%is_sparse = .true.
      m_name%is_device = .true.
      m_name%str_type = 'csc'
      m_name%dtype = OMX_C128
      rc = omx_bsr_from_csc_z(ctx(), int(A%dim2, c_int), c_loc(cp), c_loc(ri), c_loc(cv), 0_c_int, 1_c_int, &
                              m_name%bsr)
      call chk(rc, 'm_copy')
      m_name%is_initialized = .true.
    else
      call die('m_copy: invalid label')
    end if
  end subroutine m_copy_z

  !================================================================================================!
  ! mm_multiply (MatrixSwitch.F90:1046-1390)                                                      !
  !================================================================================================!
  subroutine mm_multiply(A, opA, B, opB, C, alpha, beta, label)
    type(matrix), intent(in) :: A, B
    type(matrix), intent(inout) :: C
    character(1), intent(in) :: opA, opB
    real(dp), intent(in) :: alpha, beta
    character(3), intent(in), optional :: label
    logical :: trA, trB
    integer :: i, j, k, l, kk
    integer(c_int) :: rc

    trA = process_op(opA)
    trB = process_op(opB)
    if ((.not. trA) .and. (.not. trB)) then
      if ((A%dim1/=C%dim1) .or. (B%dim2/=C%dim2) .or. (A%dim2/=B%dim1)) &
        call die('mm_dmultiply: matrices A, B and C are not compatible')
    else if (trA .and. (.not. trB)) then
      if ((A%dim2/=C%dim1) .or. (B%dim2/=C%dim2) .or. (A%dim1/=B%dim1)) &
        call die('mm_dmultiply: matrices A, B and C are not compatible')
    else if ((.not. trA) .and. trB) then
      if ((A%dim1/=C%dim1) .or. (B%dim1/=C%dim2) .or. (A%dim2/=B%dim2)) &
        call die('mm_dmultiply: matrices A, B and C are not compatible')
    else
      if ((A%dim2/=C%dim1) .or. (B%dim1/=C%dim2) .or. (A%dim1/=B%dim2)) &
        call die('mm_dmultiply: matrices A, B and C are not compatible')
    end if
    if (.not. (A%is_real .and. B%is_real .and. C%is_real)) then
      if (.not. (A%is_device .and. B%is_device .and. C%is_device)) &
        call die('mm_zmultiply: complex matrices are available in device storage only')
      if ((A%is_real .neqv. C%is_real) .or. (B%is_real .neqv. C%is_real)) &
        call die('mm_multiply: mixed real and complex operands')
    end if

    if (A%is_device .and. B%is_device .and. C%is_device) then
      call check_device_label(label, 'mm_dmultiply')
      if (C%is_sparse) then
        ! dense x dense -> sparse pattern, op = ('c','n') (MatrixSwitch_mm_multiply.F90:334-398): SDDMM
        if (A%is_sparse .or. B%is_sparse .or. (.not. trA) .or. trB) call die('mm_dmultiply: invalid implementation')
        if (A%dtype /= B%dtype) call die('mm_dmultiply: element types differ')
        rc = omx_sddmm_bsr(ctx(), A%dtype, int(A%dim1, c_int), int(A%dim2, c_int), alpha, A%dptr, A%ld, B%dptr, &
                           B%ld, beta, C%bsr)
        call chk(rc, 'mm_dmultiply')
        return
      end if
      if (A%is_sparse) call die('mm_dmultiply: invalid implementation')
      if (B%is_sparse) then
        ! dense x sparse operator: the SpMM (omm_ops.F90:61)
        if (trA .or. trB) call die('mm_dmultiply: only op = n for the sparse operator product')
        if (A%dtype /= C%dtype) call die('mm_dmultiply: element types differ')
        rc = omx_spmm(ctx(), A%dtype, int(A%dim1, c_int), int(A%dim2, c_int), alpha, A%dptr, A%ld, B%bsr, beta, &
                      C%dptr, C%ld)
        call chk(rc, 'mm_dmultiply')
      else
        if (A%dtype /= B%dtype .or. A%dtype /= C%dtype) call die('mm_dmultiply: element types differ')
        if (trA) then
          k = A%dim1
        else
          k = A%dim2
        end if
        rc = omx_gemm(ctx(), ms_label_prec(label), A%dtype, opA, opB, int(C%dim1, c_int), int(C%dim2, c_int), &
                      int(k, c_int), alpha, A%dptr, A%ld, B%dptr, B%ld, beta, C%dptr, C%ld)
        call chk(rc, 'mm_dmultiply')
      end if
      return
    end if
    if (A%is_device .or. B%is_device .or. C%is_device) call die('mm_dmultiply: invalid implementation')
    call check_host_label(label, 'mm_dmultiply')

    if (.not. A%is_sparse .and. .not. B%is_sparse .and. .not. C%is_sparse) then
      ! MatrixSwitch_mm_multiply.F90:19-79
      C%dval = beta*C%dval
      do j = 1, C%dim2
        do i = 1, C%dim1
          if (trA) then
            kk = A%dim1
          else
            kk = A%dim2
          end if
          do k = 1, kk
            C%dval(i,j) = C%dval(i,j) + alpha*aval(A, trA, i, k)*aval(B, trB, k, j)
          end do
        end do
      end do
    else if (.not. A%is_sparse .and. B%is_sparse .and. .not. C%is_sparse .and. .not. trB) then
      ! dense x CSC (MatrixSwitch_mm_multiply.F90:263-327, op(B) = n)
      C%dval = beta*C%dval
      do j = 1, C%dim2
        do l = B%iaux4(j)+1, B%iaux4(j+1)
          k = B%iaux3(l)
          do i = 1, C%dim1
            C%dval(i,j) = C%dval(i,j) + alpha*aval(A, trA, i, k)*B%dval(l,1)
          end do
        end do
      end do
    else if (.not. A%is_sparse .and. .not. B%is_sparse .and. C%is_sparse) then
      ! dense x dense -> CSC pattern of C (MatrixSwitch_mm_multiply.F90:334-398)
      C%dval = beta*C%dval
      if (trA) then
        kk = A%dim1
      else
        kk = A%dim2
      end if
      do j = 1, C%dim2
        do l = C%iaux4(j)+1, C%iaux4(j+1)
          i = C%iaux3(l)
          do k = 1, kk
            C%dval(l,1) = C%dval(l,1) + alpha*aval(A, trA, i, k)*aval(B, trB, k, j)
          end do
        end do
      end do
    else
      call die('mm_dmultiply: invalid implementation')
    end if
  end subroutine mm_multiply

  pure real(dp) function aval(M, tr, i, j)
    type(matrix), intent(in) :: M
    logical, intent(in) :: tr
    integer, intent(in) :: i, j
    if (tr) then
      aval = M%dval(j,i)
    else
      aval = M%dval(i,j)
    end if
  end function aval

  logical function process_op(op)
    character(1), intent(in) :: op
    select case (op)
    case ('T', 't', 'C', 'c')
      process_op = .true.
    case ('N', 'n')
      process_op = .false.
    case default
      process_op = .false.
      call die('process_lopM: invalid opM')
    end select
  end function process_op

  !================================================================================================!
  ! m_add, m_scale, m_set, traces, elements                                                        !
  !================================================================================================!

  !> C := alpha*op(A) + beta*C (MatrixSwitch.F90:1731-1902; MatrixSwitch_m_add.F90:95-129).
  subroutine m_add(A, opA, C, alpha, beta, label)
    type(matrix), intent(in) :: A
    type(matrix), intent(inout) :: C
    character(1), intent(in) :: opA
    real(dp), intent(in) :: alpha, beta
    character(3), intent(in), optional :: label
    logical :: trA
    integer :: i, j
    integer(c_int) :: rc
    trA = process_op(opA)
    if (.not. trA) then
      if ((A%dim1/=C%dim1) .or. (A%dim2/=C%dim2)) call die('m_dadd: matrices A and C are not compatible')
    else
      if ((A%dim1/=C%dim2) .or. (A%dim2/=C%dim1)) call die('m_dadd: matrices A and C are not compatible')
    end if
    if (A%is_device .and. C%is_device) then
      call check_device_label(label, 'm_dadd')
      if (A%is_sparse .or. C%is_sparse) call die('m_dadd: invalid implementation')
      if (A%dtype /= C%dtype) call die('m_dadd: element types differ')
      rc = omx_add(ctx(), C%dtype, opA, int(C%dim1, c_int), int(C%dim2, c_int), alpha, A%dptr, A%ld, beta, &
                   C%dptr, C%ld)
      call chk(rc, 'm_dadd')
      return
    end if
    if (A%is_device .or. C%is_device) call die('m_dadd: invalid implementation')
    call check_host_label(label, 'm_dadd')
    if (A%is_sparse .or. C%is_sparse) call die('m_dadd: invalid implementation')
    C%dval = beta*C%dval
    if (.not. trA) then
      C%dval = C%dval + alpha*A%dval
    else
      do i = 1, C%dim1
        do j = 1, C%dim2
          C%dval(i,j) = C%dval(i,j) + alpha*A%dval(j,i)
        end do
      end do
    end if
  end subroutine m_add

  !> C := beta*C (MatrixSwitch.F90:2496-2621).
  subroutine m_scale(C, beta, label)
    type(matrix), intent(inout) :: C
    real(dp), intent(in) :: beta
    character(3), intent(in), optional :: label
    integer(c_int) :: rc
    if (C%is_device) then
      call check_device_label(label, 'm_dscale')
      if (C%is_sparse) call die('m_dscale: invalid implementation')
      rc = omx_scale(ctx(), C%dtype, int(C%dim1, c_int), int(C%dim2, c_int), beta, C%dptr, C%ld)
      call chk(rc, 'm_dscale')
      return
    end if
    call check_host_label(label, 'm_dscale')
    if (C%is_sparse) then
      C%dval(1:C%iaux2(1),1) = beta*C%dval(1:C%iaux2(1),1)
    else
      C%dval = beta*C%dval
    end if
  end subroutine m_scale

  !> C_ij := alpha (i /= j), beta (i == j) (MatrixSwitch.F90:2747-2827; MatrixSwitch_m_set.F90:19-75).
  subroutine m_set(C, seC, alpha, beta, label)
    type(matrix), intent(inout) :: C
    character(1), intent(in) :: seC
    real(dp), intent(in) :: alpha, beta
    character(3), intent(in), optional :: label
    integer :: i, j, lu
    integer(c_int) :: rc
    if (C%is_device) then
      call check_device_label(label, 'm_dset')
      if (C%is_sparse) call die('m_dset: invalid implementation')
      rc = omx_set(ctx(), C%dtype, seC, int(C%dim1, c_int), int(C%dim2, c_int), alpha, beta, C%dptr, C%ld)
      call chk(rc, 'm_dset')
      return
    end if
    call check_host_label(label, 'm_dset')
    if (C%is_sparse) call die('m_dset: invalid implementation')
    if (.not. C%is_real) then
      do i = 1, C%dim1
        do j = 1, C%dim2
          if ((seC == 'u' .or. seC == 'U') .and. j < i) cycle
          if ((seC == 'l' .or. seC == 'L') .and. j > i) cycle
          if (i /= j) then
            C%zval(i,j) = cmplx(alpha, 0.0_dp, dp)
          else
            C%zval(i,j) = cmplx(beta, 0.0_dp, dp)
          end if
        end do
      end do
      return
    end if
    lu = 0
    if (seC == 'u' .or. seC == 'U') lu = 1
    if (seC == 'l' .or. seC == 'L') lu = 2
    do i = 1, C%dim1
      do j = 1, C%dim2
        if (lu == 1 .and. j < i) cycle
        if (lu == 2 .and. j > i) cycle
        if (i /= j) then
          C%dval(i,j) = alpha
        else
          C%dval(i,j) = beta
        end if
      end do
    end do
  end subroutine m_set

  !> alpha := tr(A) (MatrixSwitch.F90:2060-2156).
  subroutine m_trace(A, alpha, label)
    type(matrix), intent(in) :: A
    real(dp), intent(out) :: alpha
    character(3), intent(in), optional :: label
    integer :: i
    integer(c_int) :: rc
    real(c_double) :: tr2(2)
    if (.not. A%is_square) call die('m_dtrace: matrix A is not square')
    if (A%is_device) then
      call check_device_label(label, 'm_dtrace')
      rc = omx_trace(ctx(), A%dtype, int(A%dim1, c_int), A%dptr, A%ld, tr2)
      call chk(rc, 'm_dtrace')
      alpha = tr2(1)  ! complex: the real part, as the reference's real-valued interface
      return
    end if
    call check_host_label(label, 'm_dtrace')
    alpha = 0.0_dp
    if (.not. A%is_real) then
      do i = 1, A%dim1
        alpha = alpha + real(A%zval(i,i), dp)
      end do
      return
    end if
    do i = 1, A%dim1
      alpha = alpha + A%dval(i,i)
    end do
  end subroutine m_trace

  !> alpha := sum_ij A_ij*B_ij (MatrixSwitch.F90:2253-2377).
  subroutine mm_trace(A, B, alpha, label)
    type(matrix), intent(in) :: A, B
    real(dp), intent(out) :: alpha
    character(3), intent(in), optional :: label
    integer :: i, j
    integer(c_int) :: rc
    real(c_double) :: tr2(2)
    if ((A%dim1/=B%dim1) .or. (A%dim2/=B%dim2)) call die('mm_dtrace: matrices A and B are not compatible')
    if (A%is_device .and. B%is_device) then
      call check_device_label(label, 'mm_dtrace')
      if (A%dtype /= B%dtype) call die('mm_dtrace: element types differ')
      rc = omx_mm_trace(ctx(), A%dtype, int(A%dim1, c_int), int(A%dim2, c_int), A%dptr, A%ld, B%dptr, B%ld, tr2)
      call chk(rc, 'mm_dtrace')
      alpha = tr2(1)  ! complex: Re sum conj(A_ij) B_ij
      return
    end if
    if (A%is_device .or. B%is_device) call die('mm_dtrace: invalid implementation')
    call check_host_label(label, 'mm_dtrace')
    alpha = 0.0_dp
    if (.not. A%is_real) then
      do i = 1, A%dim1
        do j = 1, A%dim2
          alpha = alpha + real(conjg(A%zval(i,j))*B%zval(i,j), dp)
        end do
      end do
      return
    end if
    do i = 1, A%dim1
      do j = 1, A%dim2
        alpha = alpha + A%dval(i,j)*B%dval(i,j)
      end do
    end do
  end subroutine mm_trace

  !> C_ij := alpha + beta*C_ij (MatrixSwitch.F90:2917-3086).
  subroutine m_dset_element(C, i, j, alpha, beta, label)
    type(matrix), intent(inout) :: C
    integer, intent(in) :: i, j
    real(dp), intent(in) :: alpha, beta
    character(3), intent(in), optional :: label
    integer(c_int) :: rc
    if (i < 1 .or. i > C%dim1 .or. j < 1 .or. j > C%dim2) call die('m_dset_element: element out of range')
    if (C%is_device) then
      if (C%is_sparse) call die('m_dset_element: invalid implementation')
      rc = omx_set_element(ctx(), C%dtype, C%dptr, C%ld, int(i, c_int), int(j, c_int), alpha, beta)
      call chk(rc, 'm_dset_element')
      return
    end if
    if (C%is_sparse) call die('m_dset_element: invalid implementation')
    if (.not. C%is_real) then
      C%zval(i,j) = cmplx(alpha, 0.0_dp, dp) + beta*C%zval(i,j)
      return
    end if
    C%dval(i,j) = alpha + beta*C%dval(i,j)
  end subroutine m_dset_element

  !> C_ij := alpha + beta*C_ij with complex alpha, beta (m_zset_element; real beta on device).
  subroutine m_zset_element(C, i, j, alpha, beta, label)
    type(matrix), intent(inout) :: C
    integer, intent(in) :: i, j
    complex(dp), intent(in) :: alpha, beta
    character(3), intent(in), optional :: label
    integer(c_int) :: rc
    if (i < 1 .or. i > C%dim1 .or. j < 1 .or. j > C%dim2) call die('m_zset_element: element out of range')
    if (C%is_real) call die('m_zset_element: real matrix')
    if (C%is_device) then
      if (C%is_sparse) call die('m_zset_element: invalid implementation')
      if (aimag(beta) /= 0.0_dp) call die('m_zset_element: complex beta on device storage')
      rc = omx_set_element_z(ctx(), C%dtype, C%dptr, C%ld, int(i, c_int), int(j, c_int), real(alpha, dp), &
                             aimag(alpha), real(beta, dp))
      call chk(rc, 'm_zset_element')
      return
    end if
    if (C%is_sparse) call die('m_zset_element: invalid implementation')
    C%zval(i,j) = alpha + beta*C%zval(i,j)
  end subroutine m_zset_element

  !> alpha := C_ij (MatrixSwitch.F90:3309-3417); the real part for complex matrices.
  subroutine m_dget_element(C, i, j, alpha, label)
    type(matrix), intent(in) :: C
    integer, intent(in) :: i, j
    real(dp), intent(out) :: alpha
    character(3), intent(in), optional :: label
    integer(c_int) :: rc
    integer :: l
    complex(dp) :: z
    real(c_double) :: v2(2)
    if (i < 1 .or. i > C%dim1 .or. j < 1 .or. j > C%dim2) call die('m_dget_element: element out of range')
    if (.not. C%is_real) then
      call m_zget_element(C, i, j, z, label)
      alpha = real(z, dp)
      return
    end if
    if (C%is_device) then
      if (C%is_sparse) then
        rc = omx_bsr_get_element(C%bsr, int(i, c_int), int(j, c_int), alpha)
        call chk(rc, 'm_dget_element')
        return
      end if
      rc = omx_get_element(ctx(), C%dtype, C%dptr, C%ld, int(i, c_int), int(j, c_int), v2)
      call chk(rc, 'm_dget_element')
      alpha = v2(1)
      return
    end if
    if (C%is_sparse) then
      alpha = 0.0_dp
      do l = C%iaux4(j)+1, C%iaux4(j+1)
        if (C%iaux3(l) == i) alpha = C%dval(l,1)
      end do
      return
    end if
    alpha = C%dval(i,j)
  end subroutine m_dget_element

  !> alpha := C_ij as a complex number (m_zget_element).
  subroutine m_zget_element(C, i, j, alpha, label)
    type(matrix), intent(in) :: C
    integer, intent(in) :: i, j
    complex(dp), intent(out) :: alpha
    character(3), intent(in), optional :: label
    integer(c_int) :: rc
    real(c_double) :: v2(2)
    real(dp) :: re
    if (i < 1 .or. i > C%dim1 .or. j < 1 .or. j > C%dim2) call die('m_zget_element: element out of range')
    if (C%is_real) then
      call m_dget_element(C, i, j, re, label)
      alpha = cmplx(re, 0.0_dp, dp)
      return
    end if
    if (C%is_device) then
      if (C%is_sparse) call die('m_zget_element: invalid implementation')
      rc = omx_get_element(ctx(), C%dtype, C%dptr, C%ld, int(i, c_int), int(j, c_int), v2)
      call chk(rc, 'm_zget_element')
      alpha = cmplx(v2(1), v2(2), dp)
      return
    end if
    if (C%is_sparse) call die('m_zget_element: invalid implementation')
    alpha = C%zval(i,j)
  end subroutine m_zget_element

  !> Fatal error with the reference behaviour (MatrixSwitch/src/MatrixSwitch_ops.F90:232-254).
  subroutine die(message)
    character(*), intent(in), optional :: message
    integer :: err_unit
    open(newunit=err_unit, file='MatrixSwitch.err', status='replace')
    write(err_unit, '(a)') 'FATAL ERROR in matrix_switch!'
    if (present(message)) write(err_unit, '(a)') message
    close(err_unit)
    stop 1
  end subroutine die

end module MatrixSwitch
