!==================================================================================================!
! libOMM operator kernels over MatrixSwitch matrices, with fused device paths.                    !
!                                                                                                  !
! calc_G, calc_AW, calc_A, calc_A2 and calc_coeff keep the meaning of libOMM/src/omm_ops.F90:17-219 !
! (same arguments).  When every operand is device-resident ('g' storage) the line-search          !
! coefficients come from one fused FP64 reduction and the gradient from one MFMA product with a   !
! 2n-long contraction; otherwise the reference sequence of MatrixSwitch calls runs.               !
! Also: the line-search quartic solver (omm_quartic.F90:60-153), the BSD LCG of the random       !
! initial guess (omm_rand.F90:19-55), solver options beyond the omm() signature and the last-call !
! statistics.                                                                                      !
!==================================================================================================!
module omm_hip_ops
  use iso_c_binding
  use omx_capi
  use MatrixSwitch
  implicit none

  real(dp), parameter :: Pi = 3.141592653589793238462643383279502884197_dp

  ! ---- options (set with omm_set_option; defaults reproduce the reference) ----
  integer, save :: opt_max_iter = 0          ! > 0: cap on CG steps per call (icg)
  ! mixed precision: predicted relative energy decrease of one line-search step that ends the
  ! low-precision phase (< 0: 0.3*cg_tol)
  real(dp), save :: opt_lowp_tol = -1.0_dp
  integer, save :: opt_lowp_max_iter = 0     ! mixed precision: cap on low-precision steps (0 = none)
  character(3), save :: opt_lowp_label = 'hbf'
  logical, save :: opt_fused = .true.        ! use fused device kernels when all operands are on device
  integer, save :: opt_progress = 0          ! > 0: one line on stderr every opt_progress CG steps
  logical, save :: opt_grad_fp32 = .true.    ! BF16 phase: gradient and state products in FP32
  logical, save :: opt_grad_b3 = .true.      ! BF16 phase: gradient in BF16x3 (state products in FP32)
  logical, save :: opt_state_b3 = .false.    ! BF16 phase: periodic state recomputation in BF16x3
  logical, save :: opt_state_b6 = .false.    ! BF16 phase: periodic state recomputation in BF16x6
  integer, save :: opt_lowp_stall = 60       ! low-precision phase ends after this many steps without a
                                             ! new lowest energy (0 = never)
  integer, save :: opt_lowp_consec = 3       ! low-precision phase: consecutive steps below lowp_tol that end it
  integer, save :: opt_sym_gram = 1          ! symmetric calc_AW Grams (omx_gemm_sym): 0 never, 1 FP64
                                             ! and complex products ('hip'), 2 every device product,
                                             ! 3 those of 1 and FP32 ('h32') products
  integer, save :: opt_lowp_rewin = 5        ! low-precision phase: recompute HC, SC, HW, SW from C every
                                             ! this many steps (0 = incremental updates only)
  ! mixed precision: the FP64 refinement applies the stop test e_diff <= cg_tol from this step on
  integer, save :: opt_refine_min = 8
  ! mixed precision: with > 0, the FP64 refinement also needs the gradient norm at this fraction of
  ! its value at the start of the refinement (0: off)
  real(dp), save :: opt_refine_grad = 0.0_dp

  ! ---- statistics of the last omm() call ----
  integer, save :: stat_icg = 0, stat_conv = 0, stat_nfail = 0, stat_icg_lowp = 0
  real(dp), save :: stat_TrQS = 0.0_dp
  ! gradient norms of the last FP64 refinement ('hmx', with refine_grad > 0): at its start and at
  ! its last stop test
  real(dp), save :: stat_g0 = 0.0_dp, stat_gend = 0.0_dp
  integer, parameter :: stat_log_cap = 100000
  real(dp), save :: stat_elog(stat_log_cap)
  real(dp), save :: stat_tlog(stat_log_cap)   ! wall-clock seconds at the end of each CG step
  real(dp), save :: stat_t0 = 0.0_dp           ! wall-clock seconds when the (last) CG phase started
  integer, save :: stat_nlog = 0

contains

  !> Monotonic wall clock in seconds.
  real(dp) function wall_seconds()
    integer(c_int64_t) :: cnt, rate
    call system_clock(cnt, rate)
    wall_seconds = real(cnt, dp)/real(rate, dp)
  end function wall_seconds

  logical function on_device(A)
    type(matrix), intent(in) :: A
    on_device = A%is_device
  end function on_device

  !> G := 4HC - 2 SW HC - 2 HW SC  (omm_ops.F90:17-39)
  subroutine calc_G(HW, SW, G, HC, SC, m_operation)
    character(3), intent(in) :: m_operation
    type(matrix), intent(in) :: HW, SW
    type(matrix), intent(inout) :: G, HC, SC
    integer(c_int) :: rc
    if (opt_fused .and. on_device(HW) .and. on_device(SW) .and. on_device(G) .and. on_device(HC) &
        .and. on_device(SC)) then
      rc = omx_omm_calc_G(ms_hip_ctx(), ms_label_prec(m_operation), G%dtype, int(G%dim1, c_int), &
                          int(G%dim2, c_int), HW%dptr, SW%dptr, HW%ld, HC%dptr, SC%dptr, G%dptr, G%ld)
      if (rc /= 0) call die('calc_G: '//omx_error_string())
      return
    end if
    call m_add(HC, 'n', G, 4.0_dp, 0.0_dp, m_operation)
    call mm_multiply(SW, 'n', HC, 'n', G, -2.0_dp, 1.0_dp, m_operation)
    call mm_multiply(HW, 'n', SC, 'n', G, -2.0_dp, 1.0_dp, m_operation)
  end subroutine calc_G

  !> AC := C*A; AW := AC*C^T  (omm_ops.F90:44-64)
  !> AW = C*A*C^H is symmetric or Hermitian (A = H or S is): on device storage only its
  !> lower-triangle tiles are computed and mirrored (omx_gemm_sym).
  subroutine calc_AW(A, C, AW, AC, m_operation)
    character(3), intent(in) :: m_operation
    type(matrix), intent(in) :: A, C
    type(matrix), intent(inout) :: AC, AW
    integer(c_int) :: rc
    call mm_multiply(C, 'n', A, 'n', AC, 1.0_dp, 0.0_dp, m_operation)
    if (opt_fused .and. on_device(AC) .and. on_device(C) .and. on_device(AW) .and. AC%dtype == AW%dtype &
        .and. C%dtype == AW%dtype .and. (AW%dtype == OMX_F64 .or. AW%dtype == OMX_F32 .or. &
        AW%dtype == OMX_C128 .or. AW%dtype == OMX_C64) .and. &
        (opt_sym_gram == 2 .or. ((opt_sym_gram == 1 .or. opt_sym_gram == 3) .and. m_operation == 'hip') .or. &
         (opt_sym_gram == 3 .and. m_operation == 'h32'))) then
      rc = omx_gemm_sym(ms_hip_ctx(), ms_label_prec(m_operation), AW%dtype, int(AW%dim1, c_int), &
                        int(C%dim2, c_int), 1.0_dp, AC%dptr, AC%ld, C%dptr, C%ld, AW%dptr, AW%ld)
      if (rc /= 0) call die('calc_AW: '//omx_error_string())
      return
    end if
    call mm_multiply(AC, 'n', C, 'c', AW, 1.0_dp, 0.0_dp, m_operation)
  end subroutine calc_AW

  !> CAW := AW*C1; A := C1^T*CAW  (omm_ops.F90:104-124)
  subroutine calc_A(AW, C1, A, CAW, m_operation)
    character(3), intent(in) :: m_operation
    type(matrix), intent(in) :: AW, C1
    type(matrix), intent(inout) :: A, CAW
    call mm_multiply(AW, 'n', C1, 'n', CAW, 1.0_dp, 0.0_dp, m_operation)
    call mm_multiply(C1, 'c', CAW, 'n', A, 1.0_dp, 0.0_dp, m_operation)
  end subroutine calc_A

  !> CAW := AW*C2; A := C1^T*CAW  (omm_ops.F90:129-150)
  subroutine calc_A2(AW, C1, C2, A, CAW, m_operation)
    character(3), intent(in) :: m_operation
    type(matrix), intent(in) :: AW, C1, C2
    type(matrix), intent(inout) :: A, CAW
    call mm_multiply(AW, 'n', C2, 'n', CAW, 1.0_dp, 0.0_dp, m_operation)
    call mm_multiply(C1, 'c', CAW, 'n', A, 1.0_dp, 0.0_dp, m_operation)
  end subroutine calc_A2

  !> Quartic line-search coefficients (omm_ops.F90:156-219).
  subroutine calc_coeff(HW, SW, HWd, SWd, HWdd, SWdd, SWdH, coeff, m_operation)
    character(3), intent(in) :: m_operation
    type(matrix), intent(in) :: HW, SW, HWd, SWd, HWdd, SWdd
    type(matrix), intent(inout) :: SWdH
    real(dp), intent(inout) :: coeff(0:4)
    real(dp) :: TrH, TrHS, TrHd, TrHdS, TrHSd, TrHdd, TrHddS, TrHSdd, TrHdSd, TrHdSdH, TrHddSd, TrHdSdd, TrHddSdd
    real(dp) :: c5(5)
    integer(c_int) :: rc
    if (opt_fused .and. on_device(HW) .and. on_device(SW) .and. on_device(HWd) .and. on_device(SWd) &
        .and. on_device(HWdd) .and. on_device(SWdd)) then
      rc = omx_omm_coeff(ms_hip_ctx(), HW%dtype, int(HW%dim1, c_int), HW%ld, HW%dptr, SW%dptr, HWd%dptr, &
                         SWd%dptr, HWdd%dptr, SWdd%dptr, c5)
      if (rc /= 0) call die('calc_coeff: '//omx_error_string())
      coeff(0:4) = c5
      return
    end if
    call m_trace(HW, TrH, trace_label(m_operation))
    call mm_trace(HW, SW, TrHS, trace_label(m_operation))
    call m_trace(HWd, TrHd, trace_label(m_operation))
    call mm_trace(SW, HWd, TrHdS, trace_label(m_operation))
    call mm_trace(HW, SWd, TrHSd, trace_label(m_operation))
    call m_trace(HWdd, TrHdd, trace_label(m_operation))
    call mm_trace(HWdd, SW, TrHddS, trace_label(m_operation))
    call mm_trace(HW, SWdd, TrHSdd, trace_label(m_operation))
    call m_add(SWd, 'c', SWdH, 1.0_dp, 0.0_dp, trace_label(m_operation))
    call mm_trace(SWdH, HWd, TrHdSd, trace_label(m_operation))
    call mm_trace(SWd, HWd, TrHdSdH, trace_label(m_operation))
    call mm_trace(HWdd, SWd, TrHddSd, trace_label(m_operation))
    call mm_trace(SWdd, HWd, TrHdSdd, trace_label(m_operation))
    call mm_trace(HWdd, SWdd, TrHddSdd, trace_label(m_operation))
    coeff(0) = 2.0_dp*TrH - TrHS
    coeff(1) = 2.0_dp*(2.0_dp*TrHd - TrHdS - TrHSd)
    coeff(2) = 2.0_dp*(TrHdd - TrHdSd - TrHdSdH) - TrHddS - TrHSdd
    coeff(3) = -2.0_dp*(TrHddSd + TrHdSdd)
    coeff(4) = -TrHddSdd
  end subroutine calc_coeff

  !> Product label of the gradient (calc_G) for a phase whose other products use `lab`.
  function grad_label(lab) result(g)
    character(3), intent(in) :: lab
    character(3) :: g
    g = lab
    if (lab == 'hbf' .and. opt_grad_fp32) g = 'h32'
    if (lab == 'hbf' .and. opt_grad_b3) g = 'hb3'
  end function grad_label

  !> Label of the periodic recomputation of HC, SC, HW, SW in the low-precision phase: FP32 unless
  !> the gradient itself runs in plain BF16; BF16x3 with lowp_state_bf16x3, BF16x6 (FP32-grade
  !> products on BF16 MFMA) with lowp_state_bf16x6.
  function state_label(lab) result(g)
    character(3), intent(in) :: lab
    character(3) :: g
    g = lab
    if (lab == 'hbf' .and. (opt_grad_fp32 .or. opt_grad_b3)) g = 'h32'
    if (lab == 'hbf' .and. opt_state_b3) g = 'hb3'
    if (lab == 'hbf' .and. opt_state_b6) g = 'hb6'
  end function state_label

  !> Reductions always accumulate in FP64; low-precision product labels map to 'hip' for them.
  function trace_label(lab) result(t)
    character(3), intent(in) :: lab
    character(3) :: t
    t = lab
    if (lab == 'h32' .or. lab == 'hbf' .or. lab == 'hb3' .or. lab == 'hb6') t = 'hip'
  end function trace_label

  !> Minimum of the quartic line-search polynomial (omm_quartic.F90:60-153, Numerical Recipes cubic).
  subroutine omm_solve_quartic(c, x_min, fail)
    real(dp), intent(in) :: c(0:4)
    real(dp), intent(out) :: x_min
    logical, intent(out) :: fail
    integer :: x_order(3)
    real(dp) :: t(3), z(3), a, b, d, Q, R, theta, S, U
    fail = .false.
    a = 3.0_dp*c(3)/(4.0_dp*c(4))
    b = 2.0_dp*c(2)/(4.0_dp*c(4))
    if ((abs(b) >= 1.0e11_dp) .or. (abs(c(4)) <= 1.0e-11_dp)) then
      x_min = -0.5_dp*c(1)/c(2)
      return
    end if
    d = c(1)/(4.0_dp*c(4))
    Q = (a**2 - 3.0_dp*b)/9.0_dp
    R = (2.0_dp*a**3 - 9.0_dp*a*b + 27.0_dp*d)/54.0_dp
    if (R**2 < Q**3) then
      theta = acos(R/sqrt(Q**3))
      t(1) = -2.0_dp*sqrt(Q)*cos(theta/3.0_dp) - a/3.0_dp
      t(2) = -2.0_dp*sqrt(Q)*cos((theta + 2.0_dp*Pi)/3.0_dp) - a/3.0_dp
      t(3) = -2.0_dp*sqrt(Q)*cos((theta - 2.0_dp*Pi)/3.0_dp) - a/3.0_dp
      z(1:3) = c(4)*t(1:3)**4 + c(3)*t(1:3)**3 + c(2)*t(1:3)**2 + c(1)*t(1:3) + c(0)
      if (c(4) > 0.0_dp) then
        if (all(z(1) >= z(2:3))) then
          x_order(1:3) = (/1, 2, 3/)
        else if (z(2) > z(3)) then
          x_order(1:3) = (/2, 3, 1/)
        else
          x_order(1:3) = (/3, 1, 2/)
        end if
        if ((0.0_dp <= t(x_order(1))) .and. (t(x_order(2)) <= t(x_order(1)))) then
          x_min = t(x_order(2))
        else
          x_min = t(x_order(3))
        end if
      else
        if (all(z(1) <= z(2:3))) then
          x_min = t(1)
        else if (z(2) < z(3)) then
          x_min = t(2)
        else
          x_min = t(3)
        end if
      end if
    else
      S = -sign(1.0_dp, R)*(abs(R) + sqrt(R**2 - Q**3))**(1.0_dp/3.0_dp)
      if (S == 0.0_dp) then
        U = 0.0_dp
      else
        U = Q/S
      end if
      x_min = (S + U) - (a/3.0_dp)
      if (c(4) < 0.0_dp) fail = .true.
    end if
  end subroutine omm_solve_quartic

  !> x <- (1103515245 x + 12345) mod 2^31, r = x / 2^31  (omm_rand.F90:37-55)
  subroutine omm_bsd_lcg(x, r)
    integer, intent(inout) :: x
    real(dp), intent(out) :: r
    integer(c_int64_t), parameter :: a = 1103515245_c_int64_t, m = 2147483648_c_int64_t
    x = int(mod(a*int(x, c_int64_t) + 12345_c_int64_t, m))
    r = real(x, dp)/real(m, dp)
  end subroutine omm_bsd_lcg

  !> Fatal error in libOMM (omm_ops.F90:1336-1358): message to libOMM.err, stop 1.
  subroutine omm_die(message)
    character(*), intent(in), optional :: message
    integer :: err_unit
    open(newunit=err_unit, file='libOMM.err', status='replace')
    write(err_unit, '(a)') 'FATAL ERROR in libOMM!'
    if (present(message)) write(err_unit, '(a)') message
    close(err_unit)
    stop 1
  end subroutine omm_die

  !================================================================================================!
  ! Dense factorisations of the Cholesky and preconditioned flavours (omm_ops.F90:225-1220).  The  !
  ! device versions act on 'gdden' matrices; host storage needs LAPACK, which this build does not   !
  ! link, and stops as the reference does when compiled without it.                                !
  !================================================================================================!

  !> C := U with C = U^T U, upper (m_factorize, omm_ops.F90:246-315, dpotrf 'u').
  subroutine m_factorize(C, label)
    type(matrix), intent(inout) :: C
    character(3), intent(in), optional :: label
    integer(c_int) :: rc, info
    call dense_device(C, 'm_dfactorize')
    rc = omx_potrf(ms_hip_ctx(), C%dtype, int(C%dim1, c_int), C%dptr, C%ld, info)
    if (rc /= 0) call omm_die('m_dfactorize: '//omx_error_string())
    if (info /= 0) call omm_die('m_dfactorize: error in dpotrf')
  end subroutine m_factorize

  !> C := U^-T C U^-1 with A = U from m_factorize, both triangles (m_reduce, omm_ops.F90:417-524).
  subroutine m_reduce(A, C, label)
    type(matrix), intent(in) :: A
    type(matrix), intent(inout) :: C
    character(3), intent(in), optional :: label
    integer(c_int) :: rc
    call dense_device(A, 'm_dreduce')
    call dense_device(C, 'm_dreduce')
    rc = omx_sygst(ms_hip_ctx(), C%dtype, int(C%dim1, c_int), C%dptr, C%ld, A%dptr, A%ld)
    if (rc /= 0) call omm_die('m_dreduce: '//omx_error_string())
  end subroutine m_reduce

  !> C := C U^T (m_transform, omm_ops.F90:662-730, dtrmm('r','u','t','n')).
  subroutine m_transform(A, C, label)
    type(matrix), intent(in) :: A
    type(matrix), intent(inout) :: C
    character(3), intent(in), optional :: label
    integer(c_int) :: rc
    call dense_device(A, 'm_dtransform')
    call dense_device(C, 'm_dtransform')
    rc = omx_trmm(ms_hip_ctx(), C%dtype, 'r', 't', int(C%dim1, c_int), int(C%dim2, c_int), 1.0_dp, A%dptr, A%ld, &
                  C%dptr, C%ld)
    if (rc /= 0) call omm_die('m_dtransform: '//omx_error_string())
  end subroutine m_transform

  !> C := C U^-T (m_back_transform, omm_ops.F90:829-897, dtrsm('r','u','t','n')).
  subroutine m_back_transform(A, C, label)
    type(matrix), intent(in) :: A
    type(matrix), intent(inout) :: C
    character(3), intent(in), optional :: label
    integer(c_int) :: rc
    call dense_device(A, 'm_dback_transform')
    call dense_device(C, 'm_dback_transform')
    rc = omx_trsm(ms_hip_ctx(), C%dtype, 'r', 't', int(C%dim1, c_int), int(C%dim2, c_int), 1.0_dp, A%dptr, A%ld, &
                  C%dptr, C%ld)
    if (rc /= 0) call omm_die('m_dback_transform: '//omx_error_string())
  end subroutine m_back_transform

  !> C := C^-1 for the symmetric positive definite preconditioner (m_inverse, omm_ops.F90:992-1106).
  subroutine m_inverse(C, label)
    type(matrix), intent(inout) :: C
    character(3), intent(in), optional :: label
    integer(c_int) :: rc, info
    call dense_device(C, 'm_dinverse')
    rc = omx_inverse_spd(ms_hip_ctx(), C%dtype, int(C%dim1, c_int), C%dptr, C%ld, info)
    if (rc /= 0) call omm_die('m_dinverse: '//omx_error_string())
    if (info /= 0) call omm_die('m_dinverse: error in the factorisation')
  end subroutine m_inverse

  subroutine dense_device(A, where)
    type(matrix), intent(in) :: A
    character(*), intent(in) :: where
    if (.not. A%is_device) call omm_die(where//': compile with LAPACK')
    if (A%is_sparse .or. A%dtype /= OMX_F64) call omm_die(where//': invalid implementation')
  end subroutine dense_device

  !> Adjoint of the converged energy for a Hamiltonian linear in parameters theta:
  !> grad(p) = sum_e J_p(e) * Dm(e), with Dm = dE/dH on H's pattern (the 'gdcsc' D_min of a
  !> converged omm() call, C^T (2I - SW) C) and the Jacobian J as device arrays seg (nparam+1),
  !> ent, coef in the layout of omx_lattice_theta_jacobian.  libOMM has no derivatives; this is
  !> the backward pass of omm() with respect to the Hamiltonian's parameters.
  subroutine omm_param_grad(Dm, nparam, seg, ent, coef, grad)
    type(matrix), intent(in) :: Dm
    integer, intent(in) :: nparam
    type(c_ptr), intent(in) :: seg, ent, coef
    real(dp), intent(out) :: grad(nparam)
    integer(c_int) :: rc
    if (.not. (Dm%is_device .and. Dm%is_sparse)) call omm_die('omm_param_grad: D must be a device sparse matrix')
    rc = omx_bsr_param_grad(ms_hip_ctx(), Dm%bsr, int(nparam, c_int), seg, ent, coef, grad)
    if (rc /= 0) call omm_die('omm_param_grad: '//omx_error_string())
  end subroutine omm_param_grad

  !> Set a solver option by name (extensions beyond the omm() argument list).
  subroutine omm_set_option(name, value)
    character(*), intent(in) :: name
    real(dp), intent(in) :: value
    select case (name)
    case ('max_iter')
      opt_max_iter = nint(value)
    case ('lowp_tol')
      opt_lowp_tol = value
    case ('lowp_max_iter')
      opt_lowp_max_iter = nint(value)
    case ('lowp_bf16')
      if (value /= 0.0_dp) then
        opt_lowp_label = 'hbf'
      else
        opt_lowp_label = 'h32'
      end if
    case ('fused')
      opt_fused = (value /= 0.0_dp)
    case ('progress')
      opt_progress = nint(value)
    case ('lowp_consec')
      opt_lowp_consec = max(1, nint(value))
    case ('lowp_rewin')
      opt_lowp_rewin = nint(value)
    case ('lowp_grad_fp32')
      opt_grad_fp32 = (value /= 0.0_dp)
    case ('lowp_grad_bf16x3')
      opt_grad_b3 = (value /= 0.0_dp)
    case ('lowp_state_bf16x3')
      opt_state_b3 = (value /= 0.0_dp)
    case ('lowp_state_bf16x6')
      opt_state_b6 = (value /= 0.0_dp)
    case ('sym_gram')
      opt_sym_gram = max(0, min(3, nint(value)))
    case ('lowp_stall')
      opt_lowp_stall = max(0, nint(value))
    case ('refine_grad')
      opt_refine_grad = value
    case ('refine_min')
      opt_refine_min = max(0, nint(value))
    case default
      call omm_die('omm_set_option: unknown option '//name)
    end select
  end subroutine omm_set_option

end module omm_hip_ops
