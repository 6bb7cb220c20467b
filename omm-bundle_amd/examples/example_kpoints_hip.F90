!==================================================================================================!
! k-point example on device storage: the call sequence of the reference's example                 !
! (libOMM/examples/example_kpoints.F90:163-302, flavour 0) with complex H(k) and S(k) built on     !
! the host ('szden'), moved to HBM as 'gzcsc' operators, and C, D, ED allocated as 'gzden'.        !
! Two k-points share the solver through np = 2, ip = k (per-k SAVE'd state); two SCF steps each   !
! (the second with H_ii -= 0.001 and init_C), then the energy-weighted density matrices.           !
! Checks e_min, D_11 and ED_11 against the reference's constants (tolerance 1e-4; the imaginary    !
! parts of D_11 and ED_11 against 0).                                                              !
! Usage: example_kpoints_hip <example_gamma.dat> <example_kpoints.dat> [storage, default 'gzden']  !
!==================================================================================================!
program example_kpoints_hip
  use MatrixSwitch
  implicit none
  real(dp), parameter :: e_min_check(4) = (/-24.43825_dp, -24.45849_dp, -24.43941_dp, -24.45965_dp/)
  real(dp), parameter :: D_el_check(4) = (/0.37534_dp, 0.37608_dp, 0.37573_dp, 0.37648_dp/)
  real(dp), parameter :: ED_el_check(2) = (/-0.43215_dp, -0.43221_dp/)
  complex(dp), parameter :: cmplx_0 = (0.0_dp, 0.0_dp)
  integer, parameter :: nk = 2
  character(5) :: m_storage
  character(3) :: m_operation
  character(len=512) :: path_g, path_k, arg
  logical :: new_S, init_C, dealloc
  integer :: m, n, num_elements, i, j, k
  integer, allocatable :: idx(:,:)
  real(dp) :: e_min, tol
  real(dp), allocatable :: vals(:,:)
  complex(dp) :: el
  type(matrix) :: Hh(nk), Sh(nk), H(nk), S(nk), D_min(nk), ED_min(nk), C_min(nk), T(nk)

  call get_command_argument(1, path_g)
  call get_command_argument(2, path_k)
  m_storage = 'gzden'
  if (command_argument_count() >= 3) then
    call get_command_argument(3, arg)
    m_storage = arg(1:5)
  end if
  m_operation = 'hip'
  tol = 1.0e-4_dp

  ! k = 1: the gamma-point data (real); k = 2: complex H and S
  open(10, file=trim(path_g))
  read(10, '(i2,1x,i2,1x,i4)') m, n, num_elements
  allocate(idx(4, num_elements), vals(8, num_elements))
  vals = 0.0_dp
  do i = 1, num_elements
    read(10, '(i2,1x,i2,4(1x,es10.3e2))') idx(1:2, i), vals(1, i), vals(5, i)
  end do
  close(10)
  open(10, file=trim(path_k))
  read(10, '()')
  do i = 1, num_elements
    read(10, '(i2,1x,i2,4(1x,es10.3e2))') idx(3:4, i), vals(3:4, i), vals(7:8, i)
  end do
  close(10)

  do k = 1, nk
    call m_allocate(Hh(k), m, m, 'szden')
    call m_allocate(Sh(k), m, m, 'szden')
    call m_allocate(D_min(k), m, m, m_storage)
    call m_allocate(ED_min(k), m, m, m_storage)
    call m_allocate(C_min(k), n, m, m_storage)
  end do

  do i = 1, 2
    do k = 1, nk
      if (i == 1) then
        do j = 1, num_elements
          call m_set_element(Hh(k), idx((k-1)*2+1, j), idx((k-1)*2+2, j), &
                             cmplx(vals((k-1)*2+1, j), vals((k-1)*2+2, j), dp), cmplx_0)
          call m_set_element(Sh(k), idx((k-1)*2+1, j), idx((k-1)*2+2, j), &
                             cmplx(vals((k-1)*2+5, j), vals((k-1)*2+6, j), dp), cmplx_0)
        end do
        new_S = .true.
        init_C = .false.
        call m_copy(S(k), Sh(k), 'gzcsc', 0.0_dp)
      else
        call m_set(Hh(k), 'a', 0.0_dp, 0.0_dp)
        do j = 1, num_elements
          call m_set_element(Hh(k), idx((k-1)*2+1, j), idx((k-1)*2+2, j), &
                             cmplx(vals((k-1)*2+1, j), vals((k-1)*2+2, j), dp), cmplx_0)
        end do
        do j = 1, m
          call m_set_element(Hh(k), j, j, -0.001_dp, 1.0_dp)
        end do
        new_S = .false.
        init_C = .true.
        call m_deallocate(H(k))
      end if
      call m_copy(H(k), Hh(k), 'gzcsc', 0.0_dp)
      call omm(m, n, H(k), S(k), new_S, e_min, D_min(k), .false., 0.0_dp, C_min(k), init_C, T(k), 0.0_dp, 0, &
               nk, k, -1.0_dp, .true., .false., m_storage, m_operation)
      print '(2(a,i1),a,f21.15)', 'e_min [', i, ',', k, '] : ', e_min
      call check(e_min, e_min_check((k-1)*2+i))
      call m_get_element(D_min(k), 1, 1, el)
      print '(2(a,i1),2(a,f21.15))', 'D_11 [', i, ',', k, ']  : ', real(el, dp), ' , ', aimag(el)
      call check(real(el, dp), D_el_check((k-1)*2+i))
      call check(aimag(el), 0.0_dp)
    end do
  end do

  do k = 1, nk
    dealloc = (k == nk)
    call omm(m, n, H(k), S(k), .false., e_min, ED_min(k), .true., 0.0_dp, C_min(k), .true., T(k), 0.0_dp, 0, &
             nk, k, -1.0_dp, .true., dealloc, m_storage, m_operation)
    call m_get_element(ED_min(k), 1, 1, el)
    print '(a,i1,2(a,f21.15))', 'ED_11 [', k, ']   : ', real(el, dp), ' , ', aimag(el)
    call check(real(el, dp), ED_el_check(k))
    call check(aimag(el), 0.0_dp)
  end do

  do k = 1, nk
    call m_deallocate(C_min(k))
    call m_deallocate(ED_min(k))
    call m_deallocate(D_min(k))
    call m_deallocate(S(k))
    call m_deallocate(H(k))
    call m_deallocate(Sh(k))
    call m_deallocate(Hh(k))
  end do
  call ms_hip_finalize()

contains

  subroutine check(a, b)
    real(dp), intent(in) :: a, b
    if (abs(a - b) > tol) then
      print '(a,2f21.15)', 'CHECK FAILED: ', a, b
      stop 1
    end if
  end subroutine check

end program example_kpoints_hip
