!==================================================================================================!
! omm(): orbital-minimization solver, drop-in for libOMM's external subroutine                    !
! (libOMM/src/omm.F90:5-672, identical argument list and meaning).                                 !
!                                                                                                  !
! Minimises the Kim functional E(C) = tr[(2I - SW) HW], HW = C H C^T, SW = C S C^T, by             !
! Polak-Ribiere conjugate gradients with the exact quartic line search, keeping the reference's    !
! per-(spin,k)-point state between calls (HW, HWd, HWdd, SW, SC, SG, work2, QW, CD, x_min) and     !
! its restart / rescaling rules.  Storage is chosen by m_storage: with a 'g' label ('gdden',      !
! 'gsden') every work matrix lives in HBM and each step runs as MFMA products, the BSR-4 SpMM and  !
! fused FP64 reductions; 'sdden' keeps the reference host semantics.                              !
!                                                                                                  !
! m_operation: 'hip' FP64, 'h32' FP32 products, 'hbf' BF16 products (FP32 accumulate), and 'hmx': !
! mixed precision -- a low-precision CG phase on an FP32 working set, then FP64 CG from the        !
! resulting coefficients (all quantities recomputed exactly) to cg_tol.  In the low-precision      !
! phase the line-search Grams (HWd, SWd, HWdd, SWdd) are BF16 products and the gradient a BF16x3  !
! product (hi*hi + hi*lo + lo*hi of the FP32 operands; FP32 with lowp_grad_bf16x3 = 0); HW and SW  !
! are updated incrementally from the BF16 Grams and recomputed from C in FP32 every 'lowp_rewin'   !
! steps, which bounds their drift by a multiple of the step length.  The phase ends when the       !
! energy decrease predicted by a line search (its quartic, independent of the drift of HW/SW)      !
! stays below 'lowp_tol' relative for 'lowp_consec' consecutive steps, or when 'lowp_stall' steps  !
! bring no new lowest energy (the phase's precision floor).  Traces, the line search and the      !
! energy are FP64 in every mode.                                                                   !
!                                                                                                  !
! Flavours 0 (basic), 1 and 2 (Cholesky reduction to standard form, S given or already factored)  !
! and 3 (preconditioner (S + T/scale_T)^-1) follow the reference's branches; their factorisations !
! run on the device ('gdden' H and S).  Flavour 4 needs pspBLAS's kinetic-energy preconditioner   !
! and stops with die(), as the reference does without pspBLAS.                                     !
!==================================================================================================!
subroutine omm(m, n, H, S, new_S, e_min, D_min, calc_ED, eta, C_min, init_C, T, scale_T, flavour, np, ip, cg_tol, &
               long_out, dealloc, m_storage, m_operation)
  use iso_c_binding
  use omx_capi
  use MatrixSwitch
  use omm_hip_ops
  implicit none

  character(5), intent(in) :: m_storage
  character(3), intent(in) :: m_operation
  logical, intent(in) :: new_S, calc_ED, init_C, long_out, dealloc
  integer, intent(in) :: m, n, flavour, np, ip
  real(dp), intent(in) :: eta, cg_tol, scale_T
  real(dp), intent(out) :: e_min
  type(matrix), intent(inout) :: H, S, D_min, T
  type(matrix), intent(inout), target :: C_min

  logical :: no_S, conv, ls_fail, dev, mixed, use_Cholesky, S_is_U, use_precon, std_form
  logical, save :: log_start = .false.
  logical, allocatable, save :: first_call(:)
  integer :: i, j, k, seed, icg, nfail, icg_lowp, log_unit
  integer :: n_step_max_inner
  integer, parameter :: n_step_max_outer = 100
  real(dp) :: rn(2), el, e_diff, TrQS, cg_tol_internal, lambda, lambda_d, lambda_n, e_min_old, coeff(0:4), dots(2)
  real(dp), allocatable, save :: x_min(:)
  real(dp), allocatable :: cinit(:,:)
  character(3) :: prod
  type(matrix), target :: SWd, SWdd, G, G_p, D, HC, HG, work1
  type(matrix), allocatable, save, target :: HW(:), HWd(:), HWdd(:), SW(:), SC(:), SG(:), work2(:), QW(:), CD(:)
  ! Cholesky-transformed coefficients and the preconditioner, per ip (omm.F90:102-106)
  type(matrix), allocatable, save, target :: C_Chl(:), P(:)
  type(matrix), target :: PG, PG_p
  ! working set of the low-precision phase of 'hmx' (FP32 storage)
  type(matrix), target :: lC, lHW, lHWd, lHWdd, lSW, lSC, lSG, lwork2
  ! the matrices the CG phase works on: the FP64 state of this ip, or the low-precision set
  type(matrix), pointer :: pC, pHW, pHWd, pHWdd, pSW, pSC, pSG, pwork2
  integer(c_int) :: rc

  if (log_start) then
    open(newunit=log_unit, file='libOMM.log', position='append')
  else
    open(newunit=log_unit, file='libOMM.log', status='replace')
    log_start = .true.
  end if

  no_S = .not. S%is_initialized
  use_Cholesky = .false.
  S_is_U = .false.
  use_precon = .false.
  if (.not. no_S) then
    ! omm.F90:125-155
    select case (flavour)
    case (0)
    case (1)
      use_Cholesky = .true.
    case (2)
      use_Cholesky = .true.
      S_is_U = .true.
    case (3)
      use_precon = .true.
    case (4)
      call omm_die('omm: flavour 4 needs the kinetic-energy preconditioner of pspBLAS (calc_PW_precon)')
    case default
      call omm_die('omm: illegal value of flavour')
    end select
  end if
  ! standard form: the metric of the CG state is the identity (no S, or S reduced by Cholesky)
  std_form = no_S .or. use_Cholesky
  mixed = (m_operation == 'hmx')
  if (mixed .and. (use_Cholesky .or. use_precon)) call omm_die('omm: mixed precision is available for flavour 0')
  if (mixed) then
    prod = 'hip'
  else
    prod = m_operation
  end if
  if (cg_tol < 0.0_dp) then
    cg_tol_internal = 1.0e-9_dp
  else
    cg_tol_internal = cg_tol
  end if

  if (calc_ED) then
    if ((.not. allocated(first_call)) .or. first_call(ip)) call omm_die('omm: cannot calculate ED matrix before D matrix')
    ! ED = C^T (HW + eta SW) (2I - SW) C, with the HW update the convergence exit skipped (omm.F90:163-251)
    call update_W(HW(ip), HWd(ip), HWdd(ip), x_min(ip))
    if (eta /= 0.0_dp) call m_add(SW(ip), 'n', HW(ip), eta, 1.0_dp, prod)
    call calc_A2(HW(ip), C_min, CD(ip), D_min, work2(ip), prod)
    if ((.not. dealloc) .and. (eta /= 0.0_dp)) call m_add(SW(ip), 'n', HW(ip), -eta, 1.0_dp, prod)
    if (dealloc) call free_all()
    close(log_unit)
    return
  end if

  if (.not. allocated(first_call)) then
    allocate(first_call(np))
    first_call = .true.
  end if
  if (.not. allocated(x_min)) allocate(x_min(np))
  if (.not. allocated(HW)) allocate(HW(np), HWd(np), HWdd(np), SW(np), SC(np), SG(np), work2(np), QW(np), CD(np))
  if (.not. HW(ip)%is_initialized) call m_allocate(HW(ip), n, n, m_storage)
  if (.not. HWd(ip)%is_initialized) call m_allocate(HWd(ip), n, n, m_storage)
  if (.not. HWdd(ip)%is_initialized) call m_allocate(HWdd(ip), n, n, m_storage)
  if (.not. SW(ip)%is_initialized) call m_allocate(SW(ip), n, n, m_storage)
  if (.not. SC(ip)%is_initialized) call m_allocate(SC(ip), n, m, m_storage)
  if (.not. SG(ip)%is_initialized) call m_allocate(SG(ip), n, m, m_storage)
  if (.not. work2(ip)%is_initialized) call m_allocate(work2(ip), n, m, m_storage)
  dev = opt_fused .and. C_min%is_device .and. (m_storage(1:1) == 'g')
  if (use_Cholesky) then
    if (.not. allocated(C_Chl)) allocate(C_Chl(np))
    if (.not. C_Chl(ip)%is_initialized) call m_allocate(C_Chl(ip), n, m, m_storage)
  end if
  if (use_precon) then
    if (.not. allocated(P)) allocate(P(np))
    if (.not. P(ip)%is_initialized) call m_allocate(P(ip), m, m, m_storage)
  end if

  ! reduce the generalized problem to standard form (omm.F90:305-309)
  if (use_Cholesky) then
    if (new_S .and. (.not. S_is_U)) call m_factorize(S, prod)
    call m_reduce(S, H, prod)
  end if
  ! preconditioner P = (S + T/scale_T)^-1 (omm.F90:312-324)
  if (use_precon .and. new_S) then
    if (T%is_initialized) then
      call m_add(T, 'n', P(ip), 1.0_dp, 0.0_dp, prod)
      call m_add(S, 'n', P(ip), 1.0_dp, 1.0_dp/scale_T, prod)
    else
      call m_add(S, 'n', P(ip), 1.0_dp, 0.0_dp, prod)
    end if
    call m_inverse(P(ip), prod)
  end if

  ! random initial guess, normalised (omm.F90:329-343)
  if (first_call(ip) .and. (.not. init_C)) then
    seed = 123456
    allocate(cinit(n, m))
    do i = 1, m
      do j = 1, n
        do k = 1, 2
          call omm_bsd_lcg(seed, rn(k))
        end do
        cinit(j, i) = sign(0.5_dp*rn(1), rn(2) - 0.5_dp)
      end do
    end do
    if (C_min%is_device) then
      call ms_upload(C_min, cinit)
    else
      C_min%dval = cinit
    end if
    deallocate(cinit)
    call m_scale(C_min, 1.0e-2_dp/sqrt(real(m, dp)), scale_label())
  end if
  if (use_Cholesky .and. (init_C .or. new_S)) then
    call m_add(C_min, 'n', C_Chl(ip), 1.0_dp, 0.0_dp, scale_label())
    call m_transform(S, C_Chl(ip), prod)
  end if

  if (mpi_rank0()) then
    write(log_unit, '(a)') '+---------------------------------------------+'
    write(log_unit, '(a)') '| libOMM                                      |'
    write(log_unit, '(a)') '+---------------------------------------------+'
    if (long_out) write(log_unit, '(a)') '|             e_min            e_diff         |'
  end if

  icg = 0
  nfail = 0
  icg_lowp = 0
  stat_nlog = 0
  stat_g0 = 0.0_dp
  stat_gend = 0.0_dp
  if (m < 10000) then
    n_step_max_inner = m*n - 1
  else
    n_step_max_inner = 10000000
  end if

  if (mixed) then
    if (.not. C_min%is_device) call omm_die('omm: mixed precision needs device storage')
    if (.not. C_min%is_real) call omm_die('omm: mixed precision is available for real matrices')
    ! Low-precision phase on an FP32 copy of everything (products in BF16 or FP32), started from
    ! C_min and recomputed from scratch; then FP64 from the resulting coefficients, again with
    ! every quantity recomputed exactly (the incremental updates of the low-precision phase carry
    ! its rounding).
    call m_allocate(lC, n, m, 'gsden')
    call ms_convert(C_min, lC)
    call point_lowp()
    call alloc_locals('gsden')
    call initial_energy(opt_lowp_label, .true.)
    if (opt_lowp_tol < 0.0_dp) then
      call cg(opt_lowp_label, 0.3_dp*cg_tol_internal, opt_lowp_max_iter, .true.)
    else
      call cg(opt_lowp_label, opt_lowp_tol, opt_lowp_max_iter, .true.)
    end if
    icg_lowp = icg
    call ms_convert(lC, C_min)
    call free_locals()
    call free_lowp()
    call point_state()
    call alloc_locals(m_storage)
    call initial_energy('hip', .true.)
    call cg('hip', cg_tol_internal, opt_max_iter, .false., opt_refine_min, opt_refine_grad)
  else
    call point_state()
    call alloc_locals(m_storage)
    call initial_energy(prod, first_call(ip) .or. init_C .or. new_S)
    call cg(prod, cg_tol_internal, opt_max_iter, .false.)
  end if
  if ((mpi_rank0()) .and. long_out) write(log_unit, '(a)') '+---------------------------------------------+'
  call free_locals()

  if (.not. QW(ip)%is_initialized) call m_allocate(QW(ip), n, n, m_storage)
  if (.not. CD(ip)%is_initialized) call m_allocate(CD(ip), n, m, m_storage)

  ! density matrix D = C^T (2I - SW) C (omm.F90:575-593); skipped when D_min is not allocated
  call m_set(QW(ip), 'a', 0.0_dp, 2.0_dp, scale_label())
  call m_add(SW(ip), 'n', QW(ip), -1.0_dp, 1.0_dp, scale_label())
  if (use_Cholesky) then
    ! back to the original basis: C = C_Chl U^-T (omm.F90:578-581)
    call m_add(C_Chl(ip), 'n', C_min, 1.0_dp, 0.0_dp, scale_label())
    call m_back_transform(S, C_min, prod)
  end if
  if (D_min%is_initialized) then
    call calc_A(QW(ip), C_min, D_min, CD(ip), prod)
  else
    call mm_multiply(QW(ip), 'n', C_min, 'n', CD(ip), 1.0_dp, 0.0_dp, prod)
  end if
  call mm_trace(QW(ip), SW(ip), TrQS, scale_label())
  if (mpi_rank0()) then
    write(log_unit, '(a,i5,a)') '| minim: icg           = ', icg, '                |'
    write(log_unit, '(a,f13.7,a)') '| minim: Tr[(2*I-S)*S] = ', TrQS, '        |'
    write(log_unit, '(a)') '+---------------------------------------------+'
  end if
  e_min = e_min + TrQS*eta

  stat_icg = icg
  stat_icg_lowp = icg_lowp
  stat_nfail = nfail
  stat_TrQS = TrQS
  if (conv) then
    stat_conv = 1
  else
    stat_conv = 0
  end if

  if (dealloc) then
    call free_all()
  else
    first_call(ip) = .false.
  end if
  close(log_unit)

contains

  logical function mpi_rank0()
    mpi_rank0 = .true.
  end function mpi_rank0

  !> Label for adds, scales and reductions (never a low-precision product label).
  function scale_label() result(l)
    character(3) :: l
    if (C_min%is_device) then
      l = 'hip'
    else
      l = 'ref'
    end if
  end function scale_label

  !> HW/SW += x (Wd^T + Wd) + x^2 Wdd (omm.F90:496-502, 511-517)
  subroutine update_W(W, Wd, Wdd, x)
    type(matrix), intent(inout) :: W
    type(matrix), intent(in) :: Wd, Wdd
    real(dp), intent(in) :: x
    if (opt_fused .and. W%is_device .and. Wd%is_device .and. Wdd%is_device) then
      rc = omx_omm_update_w(ms_hip_ctx(), W%dtype, int(W%dim1, c_int), W%ld, x, Wd%dptr, Wdd%dptr, W%dptr)
      if (rc /= 0) call omm_die('omm: '//omx_error_string())
    else
      call m_add(Wd, 'c', W, x, 1.0_dp, scale_label_of(W))
      call m_add(Wd, 'n', W, x, 1.0_dp, scale_label_of(W))
      call m_add(Wdd, 'n', W, x**2, 1.0_dp, scale_label_of(W))
    end if
  end subroutine update_W

  function scale_label_of(A) result(l)
    type(matrix), intent(in) :: A
    character(3) :: l
    if (A%is_device) then
      l = 'hip'
    else
      l = 'ref'
    end if
  end function scale_label_of

  !> Energy, gradient and line-search data at the current C (omm.F90:350-413).
  subroutine initial_energy(lab, recompute_S)
    character(3), intent(in) :: lab
    logical, intent(in) :: recompute_S
    call calc_AW(H, pC, pHW, HC, lab)
    if (std_form) then
      if (recompute_S) call mm_multiply(pC, 'n', pC, 'c', pSW, 1.0_dp, 0.0_dp, lab)
    else
      if (recompute_S) then
        call calc_AW(S, pC, pSW, pSC, lab)
      else
        call m_add(pSG, 'n', pSC, x_min(ip), 1.0_dp, scale_label())
      end if
    end if
    if (std_form) then
      call calc_G(pHW, pSW, G, HC, pC, grad_label(lab))
    else
      call calc_G(pHW, pSW, G, HC, pSC, grad_label(lab))
    end if
    if (use_precon) then
      ! preconditioned search direction PG = G P (omm.F90:379-403)
      call mm_multiply(G, 'n', P(ip), 'n', PG, 1.0_dp, 0.0_dp, lab)
      call mm_multiply(HC, 'n', PG, 'c', pHWd, 1.0_dp, 0.0_dp, lab)
      call mm_multiply(pSC, 'n', PG, 'c', SWd, 1.0_dp, 0.0_dp, lab)
      call calc_AW(H, PG, pHWdd, HG, lab)
      call calc_AW(S, PG, SWdd, pSG, lab)
    else
      if (std_form) then
        call mm_multiply(pC, 'n', G, 'c', SWd, 1.0_dp, 0.0_dp, lab)
      else
        call mm_multiply(pSC, 'n', G, 'c', SWd, 1.0_dp, 0.0_dp, lab)
      end if
      call mm_multiply(HC, 'n', G, 'c', pHWd, 1.0_dp, 0.0_dp, lab)
      call calc_AW(H, G, pHWdd, HG, lab)
      if (std_form) then
        call mm_multiply(G, 'n', G, 'c', SWdd, 1.0_dp, 0.0_dp, lab)
      else
        call calc_AW(S, G, SWdd, pSG, lab)
      end if
    end if
    call calc_coeff(pHW, pSW, pHWd, SWd, pHWdd, SWdd, work1, coeff, lab)
    e_min = coeff(0)
  end subroutine initial_energy

  !> Polak-Ribiere CG with exact quartic line search (omm.F90:429-555).
  !> Stops when e_diff <= tol, or after max_steps steps of this phase (0 = reference caps only).
  !> lowp: the low-precision phase of 'hmx' (periodic recomputation of the state from C, and the
  !> stop test on the decrease predicted by the line search).
  !> min_steps, gtol (the FP64 refinement of 'hmx'): e_diff <= tol ends the loop only from the
  !> min_steps-th step on, and, with gtol > 0, only when the step's gradient ||G_p|| is at most gtol
  !> times the gradient at the start of the loop.  The refinement restarts CG (a steepest-descent
  !> step first) from the low-precision phase's coefficients; the first steps of a restarted CG
  !> can lower the energy by less than cg_tol while C, and so the density matrix, still carries that
  !> phase's error (the energy is quadratic in it, the density matrix linear).  After a few steps
  !> the per-step decrease again measures the remaining error, as in the FP64 solve.
  subroutine cg(lab, tol, max_steps, lowp, min_steps, gtol)
    character(3), intent(in) :: lab
    real(dp), intent(in) :: tol
    integer, intent(in) :: max_steps
    logical, intent(in) :: lowp
    integer, intent(in), optional :: min_steps
    real(dp), intent(in), optional :: gtol
    integer :: outer, inner, steps, n_below, n_stall, nmin
    logical :: ls_conv, capped, rewin_now
    real(dp) :: e_pred, e_best, g2_start, g2
    conv = .false.
    capped = .false.
    steps = 0
    n_below = 0
    n_stall = 0
    e_best = huge(1.0_dp)
    nmin = 0
    if (present(min_steps)) nmin = min_steps
    g2_start = -1.0_dp
    if (present(gtol)) then
      if (gtol > 0.0_dp) then
        call mm_trace(G, G, g2_start, scale_label())
        stat_g0 = sqrt(g2_start)
        stat_gend = stat_g0
      end if
    end if
    stat_t0 = wall_seconds()  ! the coefficient readback that precedes this point left the stream idle
    do outer = 1, n_step_max_outer
      lambda = 0.0_dp
      do inner = 1, n_step_max_inner
        if (use_precon) then
          call m_add(PG, 'n', D, 1.0_dp, lambda, scale_label())
          call m_add(G, 'n', G_p, 1.0_dp, 0.0_dp, scale_label())
          call m_add(PG, 'n', PG_p, 1.0_dp, 0.0_dp, scale_label())
        else if (dev) then
          rc = omx_omm_cg_direction(ms_hip_ctx(), G%dtype, int(n, c_int), int(m, c_int), lambda, G%dptr, D%dptr, &
                                    G_p%dptr, G%ld)
          if (rc /= 0) call omm_die('omm: '//omx_error_string())
        else
          call m_add(G, 'n', D, 1.0_dp, lambda, scale_label())
          call m_add(G, 'n', G_p, 1.0_dp, 0.0_dp, scale_label())
        end if
        e_min_old = e_min
        ! after the first step of a phase, line-search data along the new direction (omm.F90:448-464)
        if (steps > 0) then
          call mm_multiply(HC, 'n', D, 'c', pHWd, 1.0_dp, 0.0_dp, lab)
          if (std_form) then
            call mm_multiply(pC, 'n', D, 'c', SWd, 1.0_dp, 0.0_dp, lab)
          else
            call mm_multiply(pSC, 'n', D, 'c', SWd, 1.0_dp, 0.0_dp, lab)
          end if
          call calc_AW(H, D, pHWdd, HG, lab)
          if (std_form) then
            call mm_multiply(D, 'n', D, 'c', SWdd, 1.0_dp, 0.0_dp, lab)
          else
            call calc_AW(S, D, SWdd, pSG, lab)
          end if
          call calc_coeff(pHW, pSW, pHWd, SWd, pHWdd, SWdd, work1, coeff, lab)
        end if
        call omm_solve_quartic(coeff(0:4), x_min(ip), ls_fail)
        if (ls_fail) then
          if (mpi_rank0()) write(log_unit, '(a)') '| WARNING: Rescaling coefficients!            |'
          e_min = 3.0_dp*e_min
          call m_scale(pC, 0.5_dp, scale_label())
          ls_conv = .false.
          nfail = nfail + 1
        else
          e_min = coeff(4)*x_min(ip)**4 + coeff(3)*x_min(ip)**3 + coeff(2)*x_min(ip)**2 + coeff(1)*x_min(ip) + coeff(0)
          call m_add(D, 'n', pC, x_min(ip), 1.0_dp, scale_label())
          ls_conv = .true.
        end if
        if (ls_fail) then
          call m_scale(pSW, 0.25_dp, scale_label())
        else
          call update_W(pSW, SWd, SWdd, x_min(ip))
        end if
        e_diff = 2.0_dp*abs((e_min - e_min_old)/(e_min + e_min_old))
        if (mpi_rank0() .and. long_out) write(log_unit, '(a,2(1x,i5),2(1x,es15.7e3),1x,a)') '|', outer, inner, &
          e_min, e_diff, '|'
        if (opt_progress > 0) then
          if (mod(icg + 1, opt_progress) == 0) then
            write(0, '(a,a3,a,i6,a,es22.14,a,es10.3)') 'omm: ', lab, ' icg=', icg + 1, ' e_min=', e_min, &
              ' e_diff=', e_diff
            flush(0)
          end if
        end if
        if (stat_nlog < stat_log_cap) then
          stat_nlog = stat_nlog + 1
          stat_elog(stat_nlog) = e_min
          stat_tlog(stat_nlog) = wall_seconds()  ! after the coefficient readback: the stream is idle
        end if
        icg = icg + 1
        steps = steps + 1
        if (lowp) then
          ! relative decrease predicted by this step's quartic from the energy of the current state,
          ! below tol on opt_lowp_consec consecutive steps
          if (.not. ls_fail) then
            e_pred = 2.0_dp*abs((e_min - coeff(0))/(e_min + coeff(0)))
            if (e_pred <= tol) then
              n_below = n_below + 1
            else
              n_below = 0
            end if
            if (n_below >= opt_lowp_consec) conv = .true.
            ! precision floor of the phase: no new lowest energy for opt_lowp_stall steps (while
            ! CG progresses nearly every step lowers it); the FP64 refinement takes over there
            if (e_min < e_best) then
              e_best = e_min
              n_stall = 0
            else
              n_stall = n_stall + 1
            end if
            if (opt_lowp_stall > 0 .and. n_stall >= opt_lowp_stall) conv = .true.
          end if
        else if (e_diff <= tol .and. steps >= nmin) then
          conv = .true.
          if (g2_start > 0.0_dp) then
            call mm_trace(G_p, G_p, g2, scale_label())
            stat_gend = sqrt(g2)
            conv = (g2 <= gtol*gtol*g2_start)
          end if
        end if
        if (conv) exit
        if (max_steps > 0 .and. steps >= max_steps) then
          capped = .true.
          exit
        end if
        rewin_now = lowp .and. (.not. ls_fail) .and. opt_lowp_rewin > 0
        if (rewin_now) rewin_now = (mod(steps, opt_lowp_rewin) == 0)
        if (ls_fail) then
          call m_scale(pHW, 0.25_dp, scale_label())
        else if (.not. rewin_now) then
          call update_W(pHW, pHWd, pHWdd, x_min(ip))
        end if
        if (ls_fail) then
          call m_scale(HC, 0.5_dp, scale_label())
          if (.not. no_S) call m_scale(pSC, 0.5_dp, scale_label())
          call m_add(G_p, 'n', G, 1.0_dp, 0.0_dp, scale_label())
          call m_add(HC, 'n', G, 1.5_dp, 1.0_dp, scale_label())
        else
          if (rewin_now) then
            ! HC = C H, HW = HC C^T, SC = C S, SW = SC C^T at the state's precision (FP32)
            call calc_AW(H, pC, pHW, HC, state_label(lab))
            if (std_form) then
              call mm_multiply(pC, 'n', pC, 'c', pSW, 1.0_dp, 0.0_dp, state_label(lab))
            else
              call calc_AW(S, pC, pSW, pSC, state_label(lab))
            end if
          else
            call m_add(HG, 'n', HC, x_min(ip), 1.0_dp, scale_label())
            if (.not. std_form) call m_add(pSG, 'n', pSC, x_min(ip), 1.0_dp, scale_label())
          end if
          if (std_form) then
            call calc_G(pHW, pSW, G, HC, pC, grad_label(lab))
          else
            call calc_G(pHW, pSW, G, HC, pSC, grad_label(lab))
          end if
        end if
        if (use_precon) call mm_multiply(G, 'n', P(ip), 'n', PG, 1.0_dp, 0.0_dp, lab)
        if (ls_conv) then
          if (use_precon) then
            call m_add(G, 'n', pwork2, 1.0_dp, 0.0_dp, scale_label())
            call m_add(G_p, 'n', pwork2, -1.0_dp, 1.0_dp, scale_label())
            call mm_trace(PG, pwork2, lambda_n, scale_label())
            call mm_trace(PG_p, G_p, lambda_d, scale_label())
          else if (dev) then
            rc = omx_omm_pr_dots(ms_hip_ctx(), G%dtype, int(n, c_int), int(m, c_int), G%dptr, G_p%dptr, G%ld, dots)
            if (rc /= 0) call omm_die('omm: '//omx_error_string())
            lambda_n = dots(1)
            lambda_d = dots(2)
          else
            call m_add(G, 'n', pwork2, 1.0_dp, 0.0_dp, scale_label())
            call m_add(G_p, 'n', pwork2, -1.0_dp, 1.0_dp, scale_label())
            call mm_trace(G, pwork2, lambda_n, scale_label())
            call mm_trace(G_p, G_p, lambda_d, scale_label())
          end if
          lambda = lambda_n/lambda_d
        else
          exit
        end if
      end do
      if (conv .or. capped) exit
    end do
    if (outer > n_step_max_outer) then
      if (mpi_rank0()) write(log_unit, '(a)') '| WARNING: OMM failed to converge!            |'
    end if
  end subroutine cg

  !> Point the CG working set at the FP64 state of this ip and at C_min.
  subroutine point_state()
    if (use_Cholesky) then
      pC => C_Chl(ip)
    else
      pC => C_min
    end if
    pHW => HW(ip)
    pHWd => HWd(ip)
    pHWdd => HWdd(ip)
    pSW => SW(ip)
    pSC => SC(ip)
    pSG => SG(ip)
    pwork2 => work2(ip)
  end subroutine point_state

  !> Point the CG working set at a fresh FP32 set (the low-precision phase of 'hmx').
  subroutine point_lowp()
    call m_allocate(lHW, n, n, 'gsden')
    call m_allocate(lHWd, n, n, 'gsden')
    call m_allocate(lHWdd, n, n, 'gsden')
    call m_allocate(lSW, n, n, 'gsden')
    call m_allocate(lSC, n, m, 'gsden')
    call m_allocate(lSG, n, m, 'gsden')
    call m_allocate(lwork2, n, m, 'gsden')
    pC => lC
    pHW => lHW
    pHWd => lHWd
    pHWdd => lHWdd
    pSW => lSW
    pSC => lSC
    pSG => lSG
    pwork2 => lwork2
  end subroutine point_lowp

  subroutine free_lowp()
    call m_deallocate(lHW)
    call m_deallocate(lHWd)
    call m_deallocate(lHWdd)
    call m_deallocate(lSW)
    call m_deallocate(lSC)
    call m_deallocate(lSG)
    call m_deallocate(lwork2)
    call m_deallocate(lC)
  end subroutine free_lowp

  !> Per-call work matrices (omm.F90:292-303), zero-filled as m_allocate does.
  subroutine alloc_locals(label)
    character(5), intent(in) :: label
    call m_allocate(SWd, n, n, label)
    call m_allocate(SWdd, n, n, label)
    call m_allocate(G, n, m, label)
    call m_allocate(G_p, n, m, label)
    call m_allocate(D, n, m, label)
    call m_allocate(HC, n, m, label)
    call m_allocate(HG, n, m, label)
    call m_allocate(work1, n, n, label)
    if (use_precon) then
      call m_allocate(PG, n, m, label)
      call m_allocate(PG_p, n, m, label)
    end if
  end subroutine alloc_locals

  subroutine free_locals()
    if (PG_p%is_initialized) call m_deallocate(PG_p)
    if (PG%is_initialized) call m_deallocate(PG)
    if (work1%is_initialized) call m_deallocate(work1)
    if (HG%is_initialized) call m_deallocate(HG)
    if (HC%is_initialized) call m_deallocate(HC)
    if (D%is_initialized) call m_deallocate(D)
    if (G_p%is_initialized) call m_deallocate(G_p)
    if (G%is_initialized) call m_deallocate(G)
    if (SWdd%is_initialized) call m_deallocate(SWdd)
    if (SWd%is_initialized) call m_deallocate(SWd)
  end subroutine free_locals

  subroutine free_all()
    integer :: q
    if (allocated(HW)) then
      do q = 1, size(HW)
        if (CD(q)%is_initialized) call m_deallocate(CD(q))
        if (QW(q)%is_initialized) call m_deallocate(QW(q))
        if (work2(q)%is_initialized) call m_deallocate(work2(q))
        if (SG(q)%is_initialized) call m_deallocate(SG(q))
        if (SC(q)%is_initialized) call m_deallocate(SC(q))
        if (SW(q)%is_initialized) call m_deallocate(SW(q))
        if (HWdd(q)%is_initialized) call m_deallocate(HWdd(q))
        if (HWd(q)%is_initialized) call m_deallocate(HWd(q))
        if (HW(q)%is_initialized) call m_deallocate(HW(q))
      end do
      deallocate(HW, HWd, HWdd, SW, SC, SG, work2, QW, CD)
    end if
    if (allocated(C_Chl)) then
      do q = 1, size(C_Chl)
        if (C_Chl(q)%is_initialized) call m_deallocate(C_Chl(q))
      end do
      deallocate(C_Chl)
    end if
    if (allocated(P)) then
      do q = 1, size(P)
        if (P(q)%is_initialized) call m_deallocate(P(q))
      end do
      deallocate(P)
    end if
    if (allocated(x_min)) deallocate(x_min)
    if (allocated(first_call)) deallocate(first_call)
  end subroutine free_all

end subroutine omm
