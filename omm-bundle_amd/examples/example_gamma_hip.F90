!==================================================================================================!
! Gamma-point example on device storage: the call sequence of the reference's example             !
! (libOMM/examples/example_gamma.F90:134-245, flavour 0) with H and S moved to HBM as 'gdcsc'       !
! operators and all dense matrices allocated with the 'gdden' label, operations 'hip'.            !
! Two SCF steps (the second with H_ii -= 0.001 and init_C), then the energy-weighted density      !
! matrix.  Checks e_min, D_11 and ED_11 against the reference's constants (tolerance 1e-4, as     !
! libOMM/examples/example_gamma.F90:269, for every operation label).                               !
! Flavours 1 (Cholesky) and 3 (preconditioning) run on dense device H and S ('gdden'), as the     !
! reference's example runs them on 'sdden' (example_gamma.F90:119-132).                            !
! Usage: example_gamma_hip <example_gamma.dat> [operation label, default 'hip'] [flavour, default 0]!
!==================================================================================================!
program example_gamma_hip
  use MatrixSwitch
  implicit none
  real(dp), parameter :: e_min_check(2) = (/-24.43825_dp, -24.45849_dp/)
  real(dp), parameter :: D_el_check(2) = (/0.37534_dp, 0.37608_dp/)
  real(dp), parameter :: ED_el_check = -0.43215_dp
  character(5) :: m_storage
  character(3) :: m_operation
  character(len=512) :: path, arg
  logical :: new_S, init_C
  integer :: m, n, num_elements, i, j, flavour
  integer, allocatable :: idx(:,:)
  real(dp) :: e_min, el, d_tol
  real(dp), allocatable :: vals(:,:)
  type(matrix) :: Hh, Sh, H, S, D_min, ED_min, C_min, T

  call get_command_argument(1, path)
  m_operation = 'hip'
  if (command_argument_count() >= 2) then
    call get_command_argument(2, arg)
    m_operation = arg(1:3)
  end if
  m_storage = 'gdden'
  flavour = 0
  if (command_argument_count() >= 3) then
    call get_command_argument(3, arg)
    read(arg, *) flavour
  end if
  d_tol = 1.0e-4_dp

  open(10, file=trim(path))
  read(10, '(i2,1x,i2,1x,i4)') m, n, num_elements
  allocate(idx(2, num_elements), vals(2, num_elements))
  do i = 1, num_elements
    read(10, '(i2,1x,i2,2(1x,es10.3e2))') idx(1:2, i), vals(1:2, i)
  end do
  close(10)

  call m_allocate(Hh, m, m, 'sdden')
  call m_allocate(Sh, m, m, 'sdden')
  call m_allocate(D_min, m, m, m_storage)
  call m_allocate(ED_min, m, m, m_storage)
  call m_allocate(C_min, n, m, m_storage)

  do i = 1, 2
    if (i == 1) then
      do j = 1, num_elements
        call m_set_element(Hh, idx(1,j), idx(2,j), vals(1,j), 0.0_dp)
        call m_set_element(Sh, idx(1,j), idx(2,j), vals(2,j), 0.0_dp)
      end do
      new_S = .true.
      init_C = .false.
      if (flavour == 0) then
        call m_copy(S, Sh, 'gdcsc', 0.0_dp)
      else
        call m_copy(S, Sh, 'gdden')
      end if
    else
      do j = 1, m
        call m_set_element(Hh, j, j, -0.001_dp, 1.0_dp)
      end do
      new_S = .false.
      init_C = .true.
      call m_deallocate(H)
    end if
    if (flavour == 0) then
      call m_copy(H, Hh, 'gdcsc', 0.0_dp)
    else
      call m_copy(H, Hh, 'gdden')
    end if
    call omm(m, n, H, S, new_S, e_min, D_min, .false., 0.0_dp, C_min, init_C, T, 0.0_dp, flavour, 1, 1, -1.0_dp, &
             .true., .false., m_storage, m_operation)
    print '(a,i1,a,f21.15)', 'e_min [', i, '] : ', e_min
    call check(e_min, e_min_check(i), 1.0e-4_dp)
    call m_get_element(D_min, 1, 1, el)
    print '(a,i1,a,f21.15)', 'D_11 [', i, ']  : ', el
    call check(el, D_el_check(i), d_tol)
  end do

  call omm(m, n, H, S, .false., e_min, ED_min, .true., 0.0_dp, C_min, .true., T, 0.0_dp, 0, 1, 1, -1.0_dp, &
           .true., .true., m_storage, m_operation)
  call m_get_element(ED_min, 1, 1, el)
  print '(a,f21.15)', 'ED_11     : ', el
  call check(el, ED_el_check, d_tol)

  call m_deallocate(C_min)
  call m_deallocate(ED_min)
  call m_deallocate(D_min)
  call m_deallocate(S)
  call m_deallocate(H)
  call m_deallocate(Sh)
  call m_deallocate(Hh)
  call ms_hip_finalize()

contains

  subroutine check(a, b, tol)
    real(dp), intent(in) :: a, b, tol
    if (abs(a - b) > tol) then
      print '(a,2f21.15)', 'CHECK FAILED: ', a, b
      stop 1
    end if
  end subroutine check

end program example_gamma_hip
