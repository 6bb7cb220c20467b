!==================================================================================================!
! Second driver for the reference serial build in oracle/_ref -- TEST INFRASTRUCTURE ONLY.        !
!                                                                                                  !
! Runs one MatrixSwitch product with a sparse ('sdcsc') result,                                    !
!   C := alpha*op(A)*op(B) + beta*C  on C's stored pattern,                                        !
! through mm_multiply(..., 'ref') -> mm_multiply_sddensddensdcscref                                !
! (MatrixSwitch_mm_multiply.F90:334-398), so the oracle's or_sddmm_csc can be checked against it.  !
!                                                                                                  !
! Input  (stream, native endian):                                                                  !
!   int32 m, n, k, trA, trB                                                                        !
!   real64 alpha, beta                                                                             !
!   real64 A(m,k) (or A(k,m) if trA), B(k,n) (or B(n,k) if trB), C0(m,n)                           !
! C's pattern is the nonzeros of C0 (m_copy to 'sdcsc' with threshold 0,                           !
! MatrixSwitch_m_copy.F90:553-629).                                                                !
! Output (stream): int32 nnz, int32 colptr(n+1) (0-based), int32 rowidx(nnz) (1-based as stored),  !
!                  real64 val(nnz)                                                                 !
!==================================================================================================!
program ref_mm
  use MatrixSwitch
  implicit none
  integer, parameter :: dp=selected_real_kind(15,300)
  character(len=512) :: fin, fout
  integer :: m, n, k, trA, trB, i, j, nnz
  real(dp) :: alpha, beta
  real(dp), allocatable :: abuf(:,:), bbuf(:,:), cbuf(:,:)
  type(matrix) :: A, B, Cd, C
  character(1) :: opA, opB

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
  read(10) m, n, k, trA, trB
  read(10) alpha, beta
  if (trA==1) then
    allocate(abuf(k,m))
    opA='t'
  else
    allocate(abuf(m,k))
    opA='n'
  end if
  if (trB==1) then
    allocate(bbuf(n,k))
    opB='t'
  else
    allocate(bbuf(k,n))
    opB='n'
  end if
  allocate(cbuf(m,n))
  read(10) abuf
  read(10) bbuf
  read(10) cbuf
  call m_allocate(A, size(abuf,1), size(abuf,2), 'sdden')
  call m_allocate(B, size(bbuf,1), size(bbuf,2), 'sdden')
  call m_allocate(Cd, m, n, 'sdden')
  do j=1,size(abuf,2)
    do i=1,size(abuf,1)
      call m_set_element(A, i, j, abuf(i,j), 0.0_dp, 'ref')
    end do
  end do
  do j=1,size(bbuf,2)
    do i=1,size(bbuf,1)
      call m_set_element(B, i, j, bbuf(i,j), 0.0_dp, 'ref')
    end do
  end do
  do j=1,n
    do i=1,m
      call m_set_element(Cd, i, j, cbuf(i,j), 0.0_dp, 'ref')
    end do
  end do
  call m_copy(C, Cd, 'sdcsc', 0.0_dp)
  call mm_multiply(A, opA, B, opB, C, alpha, beta, 'ref')
  nnz=C%iaux4(n+1)
  write(11) nnz
  write(11) C%iaux4(1:n+1)
  write(11) C%iaux3(1:nnz)
  write(11) C%dval(1:nnz,1)
  close(10)
  close(11)
end program ref_mm
