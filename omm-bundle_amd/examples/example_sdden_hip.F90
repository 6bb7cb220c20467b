!==================================================================================================!
! MatrixSwitch API example on device storage: the operations of the reference's example_sdden    !
! (MatrixSwitch/examples/example_sdden.F90:61-140) with 'gdden' matrices and the 'hip' label:     !
! tr(A^T B) computed five ways and E_11, checked against the reference constants at 1e-10.        !
!==================================================================================================!
program example_sdden_hip
  use MatrixSwitch
  implicit none
  real(dp), parameter :: res_check = 31.194861937321846_dp
  real(dp), parameter :: el_check = 2.779421970931733_dp
  character(5), parameter :: m_storage = 'gdden'
  character(3), parameter :: m_operation = 'hip'
  integer :: N, M, i, j
  real(dp) :: rn, res1, res2, res3, res4, res5, el
  real(dp), allocatable :: Ah(:,:), Bh(:,:)
  type(matrix) :: A, B, C, D, E

  N = 15
  M = 8
  allocate(Ah(N,M), Bh(N,M))
  rn = 0.1_dp
  do i = 1, N
    do j = 1, M
      rn = mod(4.2_dp*rn, 1.0_dp)
      Ah(i,j) = rn
      rn = mod(4.2_dp*rn, 1.0_dp)
      Bh(i,j) = rn
    end do
  end do
  call m_allocate(A, N, M, m_storage)
  call m_allocate(B, N, M, m_storage)
  call ms_upload(A, Ah)
  call ms_upload(B, Bh)
  call m_allocate(C, M, N, m_storage)
  call m_allocate(D, M, M, m_storage)
  call m_allocate(E, N, N, m_storage)

  call mm_trace(A, B, res1, m_operation)
  print '(a,f21.15)', 'res1 : ', res1
  call check(res1, res_check)
  call mm_multiply(A, 't', B, 'n', D, 1.0_dp, 0.0_dp, m_operation)
  call m_trace(D, res2, m_operation)
  print '(a,f21.15)', 'res2 : ', res2
  call check(res2, res_check)
  call mm_multiply(B, 'n', A, 't', E, 1.0_dp, 0.0_dp, m_operation)
  call m_trace(E, res3, m_operation)
  print '(a,f21.15)', 'res3 : ', res3
  call check(res3, res_check)
  call m_add(A, 't', C, 1.0_dp, 0.0_dp, m_operation)
  call mm_multiply(C, 'n', B, 'n', D, 1.0_dp, 0.0_dp, m_operation)
  call m_trace(D, res4, m_operation)
  print '(a,f21.15)', 'res4 : ', res4
  call check(res4, res_check)
  call m_add(A, 't', C, 1.0_dp, 0.0_dp, m_operation)
  call mm_multiply(B, 'n', C, 'n', E, 1.0_dp, 0.0_dp, m_operation)
  call m_trace(E, res5, m_operation)
  print '(a,f21.15)', 'res5 : ', res5
  call check(res5, res_check)
  call m_get_element(E, 1, 1, el)
  print '(a,f21.15)', 'E_11 : ', el
  call check(el, el_check)
  call m_deallocate(E)
  call m_deallocate(D)
  call m_deallocate(C)
  call m_deallocate(B)
  call m_deallocate(A)
  call ms_hip_finalize()

contains

  subroutine check(a, b)
    real(dp), intent(in) :: a, b
    if (abs(a - b) > 1.0e-10_dp) then
      print '(a,2f21.15)', 'CHECK FAILED: ', a, b
      stop 1
    end if
  end subroutine check

end program example_sdden_hip
