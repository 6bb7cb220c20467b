!==================================================================================================!
! Driver for the reference serial build in oracle/_ref -- TEST INFRASTRUCTURE ONLY.                !
!                                                                                                  !
! Links against libOMM + MatrixSwitch compiled from the reference sources (serial, -DCONV,         !
! -DNORAND; operations label 'ref') and replays a scripted sequence of omm() calls read from a    !
! binary stream file, writing e_min, C_min and D_min after each call.  It follows the call         !
! pattern of libOMM/examples/example_gamma.F90:134-245 (H/S set element-wise, then omm()).          !
!                                                                                                  !
! Input  (stream, native endian):                                                                  !
!   int32 m, n, ncalls, s_present, sparse, cplx                                                    !
!   per call: int32 new_S, init_C, calc_ED, dealloc, give_C, flavour                               !
!             real64 eta, cg_tol                                                                   !
!             H(m,m); if s_present: S(m,m); if give_C: C(n,m)  (real64, or complex128 if cplx=1)    !
! Output (stream): per call real64 e_min, C(n,m), D(m,m)  (C, D complex128 if cplx=1)              !
! sparse=1 stores H and S as 'sdcsc' (m_copy with threshold 0, MatrixSwitch_m_copy.F90:553-629), !
! so the products run through the dense x CSC kernel (MatrixSwitch_mm_multiply.F90:263-327).      !
! cplx=1 runs on 'szden' storage (mm_multiply_szdenref, MatrixSwitch_mm_multiply.F90:85-185); it   !
! is dense only: the complex branch of m_copy_sdenscscref_thre (MatrixSwitch_m_copy.F90:620)       !
! stores the column pointers as iaux4(i)=k+1 instead of iaux4(i+1)=k, so no valid 'szcsc' operator  !
! comes out of it.                                                                                 !
!==================================================================================================!
program ref_driver
  use MatrixSwitch
  implicit none
  integer, parameter :: dp=selected_real_kind(15,300)
  character(len=512) :: fin, fout
  integer :: m, n, ncalls, s_present, sparse, cplx, icall, i, j
  integer :: new_S, init_C, calc_ED, dealloc, give_C, flavour
  real(dp) :: eta, cg_tol, e_min
  real(dp), allocatable :: buf(:,:), cbuf(:,:)
  complex(dp), allocatable :: zbuf(:,:), zcbuf(:,:)
  character(5) :: st
  type(matrix) :: Hd, Sd, H, S, D_min, C_min, T
  logical :: have_sparse

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
  read(10) m, n, ncalls, s_present, sparse, cplx
  if (cplx==1 .and. sparse==1) stop 'complex operators are dense only'
  allocate(buf(m,m), cbuf(n,m), zbuf(m,m), zcbuf(n,m))
  st='sdden'
  if (cplx==1) st='szden'
  call m_allocate(Hd, m, m, st)
  call m_allocate(Sd, m, m, st)
  call m_allocate(D_min, m, m, st)
  call m_allocate(C_min, n, m, st)
  have_sparse=.false.

  do icall=1,ncalls
    read(10) new_S, init_C, calc_ED, dealloc, give_C, flavour
    read(10) eta, cg_tol
    call read_square(Hd)
    if (s_present==1) call read_square(Sd)
    if (give_C==1) then
      if (cplx==1) then
        read(10) zcbuf
        do j=1,m
          do i=1,n
            call m_set_element(C_min, i, j, zcbuf(i,j), (0.0_dp,0.0_dp), 'ref')
          end do
        end do
      else
        read(10) cbuf
        do j=1,m
          do i=1,n
            call m_set_element(C_min, i, j, cbuf(i,j), 0.0_dp, 'ref')
          end do
        end do
      end if
    end if
    if (sparse==1) then
      if (have_sparse) then
        call m_deallocate(H)
        if (s_present==1) call m_deallocate(S)
      end if
      call m_copy(H, Hd, 'sdcsc', 0.0_dp)
      if (s_present==1) call m_copy(S, Sd, 'sdcsc', 0.0_dp)
      have_sparse=.true.
      call run_omm(H, S)
    else
      call run_omm(Hd, Sd)
    end if
    write(11) e_min
    if (cplx==1) then
      do j=1,m
        do i=1,n
          call m_get_element(C_min, i, j, zcbuf(i,j), 'ref')
        end do
      end do
      write(11) zcbuf
      do j=1,m
        do i=1,m
          call m_get_element(D_min, i, j, zbuf(i,j), 'ref')
        end do
      end do
      write(11) zbuf
    else
      do j=1,m
        do i=1,n
          call m_get_element(C_min, i, j, cbuf(i,j), 'ref')
        end do
      end do
      write(11) cbuf
      do j=1,m
        do i=1,m
          call m_get_element(D_min, i, j, buf(i,j), 'ref')
        end do
      end do
      write(11) buf
    end if
  end do
  close(10)
  close(11)

contains

  subroutine read_square(A)
    type(matrix), intent(inout) :: A
    if (cplx==1) then
      read(10) zbuf
      do j=1,m
        do i=1,m
          call m_set_element(A, i, j, zbuf(i,j), (0.0_dp,0.0_dp), 'ref')
        end do
      end do
    else
      read(10) buf
      do j=1,m
        do i=1,m
          call m_set_element(A, i, j, buf(i,j), 0.0_dp, 'ref')
        end do
      end do
    end if
  end subroutine read_square

  subroutine run_omm(HH, SS)
    type(matrix), intent(inout) :: HH, SS
    type(matrix) :: S_absent
    if (s_present==1) then
      call omm(m, n, HH, SS, new_S==1, e_min, D_min, calc_ED==1, eta, C_min, init_C==1, T, 0.0_dp, &
               flavour, 1, 1, cg_tol, .true., dealloc==1, st, 'ref')
    else
      call omm(m, n, HH, S_absent, new_S==1, e_min, D_min, calc_ED==1, eta, C_min, init_C==1, T, 0.0_dp, &
               flavour, 1, 1, cg_tol, .true., dealloc==1, st, 'ref')
    end if
  end subroutine run_omm

end program ref_driver
