! ehip_linsolve.F90 -- the whole linear solve on the GPU as Elmer's
! "Before Linsolve" hook: SolveSystem calls it and skips its own
! SolveLinearSystem when it returns nonzero (fem/src/SolverUtils.src:
! 7605-7608, Load.c:538-553).  Build it against ElmerSolver's modules and
! the ISO_C_BINDING module of libehip, e.g.
!   elmerf90 -o ehip_linsolve.so ehip_linsolve.F90 elmerfem_amd/fortran/ehip.F90 \
!            -L/path/to/elmerfem_amd -lehip
! and select it in the .sif with
!   Before Linsolve = "ehip_linsolve" "EhipLinsolve"
! The keywords are decoded the way IterSolver does (IterSolve.src:221-253,
! 371-400): a method or preconditioner this library does not provide
! (GMRES, ILUT, multigrid, ...) returns 0 and Elmer solves as usual.
FUNCTION EhipLinsolve(Model, Solver, A, b, x, n, DOFs, Norm) RESULT(handled)
  USE DefUtils
  USE ehip
  IMPLICIT NONE
  TYPE(Model_t) :: Model
  TYPE(Solver_t) :: Solver
  TYPE(Matrix_t), POINTER :: A
  INTEGER :: n, DOFs, handled
  REAL(KIND=dp) :: x(n), b(n), Norm
  TYPE(ehip_params) :: p
  TYPE(ehip_report) :: r
  CHARACTER(LEN=MAX_NAME_LEN) :: str
  LOGICAL :: Found, Abort
  INTEGER :: rc, k
  REAL(KIND=dp) :: order

  handled = 0
  str = ListGetString(Solver % Values, 'Linear System Iterative Method', Found)
  IF (.NOT. Found) str = 'bicgstab'
  IF (str(1:9) == 'bicgstab2' .OR. str(1:9) == 'bicgstabl') RETURN
  IF (str(1:8) == 'bicgstab') THEN
    p % method = EHIP_BICGSTAB
  ELSE IF (str(1:3) == 'cgs') THEN
    RETURN
  ELSE IF (str(1:2) == 'cg') THEN
    p % method = EHIP_CG
  ELSE IF (str(1:5) == 'tfqmr' .OR. str(1:5) == 'gmres' .OR. str(1:3) == 'sgs' .OR. str(1:6) == 'jacobi' &
           .OR. str(1:10) == 'richardson' .OR. str(1:3) == 'gcr') THEN
    RETURN
  ELSE
    p % method = EHIP_BICGSTAB                          ! IterSolver's fallback
  END IF

  str = ListGetString(Solver % Values, 'Linear System Preconditioning', Found)
  IF (.NOT. Found) str = 'none'
  IF (ListGetLogical(Solver % Values, 'Linear System Symmetric ILU', Found)) RETURN
  IF (str(1:4) == 'none') THEN
    p % precond = EHIP_PRECOND_NONE
  ELSE IF (str(1:8) == 'diagonal') THEN
    p % precond = EHIP_PRECOND_DIAGONAL
  ELSE IF (str(1:4) == 'ilut') THEN
    RETURN
  ELSE IF (str(1:3) == 'ilu') THEN
    order = ListGetCReal(Solver % Values, 'Linear System ILU Order', Found)
    IF (Found) THEN
      k = NINT(order)
    ELSE
      k = ICHAR(str(4:4)) - ICHAR('0')
    END IF
    IF (k < 0 .OR. k > 9) k = 0
    p % precond = EHIP_PRECOND_ILU0 + k
  ELSE IF (str(1:4) == 'bilu' .OR. str(1:9) == 'multigrid' .OR. str(1:5) == 'vanka') THEN
    RETURN
  ELSE
    p % precond = EHIP_PRECOND_NONE                     ! unknown: IterSolver disables it
  END IF

  p % tol = ListGetCReal(Solver % Values, 'Linear System Convergence Tolerance')
  p % maxit = ListGetInteger(Solver % Values, 'Linear System Max Iterations', minv=1)
  p % scaling = 1
  IF (.NOT. ListGetLogical(Solver % Values, 'Linear System Scaling', Found) .AND. Found) p % scaling = 0
  p % divtol = ListGetCReal(Solver % Values, 'Linear System Divergence Tolerance', Found)
  IF (.NOT. Found) p % divtol = 0.0_dp
  p % precision = EHIP_FP64
  p % mode = EHIP_FAITHFUL
  IF (ListGetLogical(Solver % Values, 'Linear System Mixed Precision', Found) .AND. &
      p % precond <= EHIP_PRECOND_DIAGONAL) THEN
    p % precision = EHIP_MIXED_IR                       ! FP32 inner, FP64 refinement
    p % mode = EHIP_FAST
  END IF
  p % inner_tol = 0.0_dp
  p % inner_maxit = 0
  p % check_every = 0
  p % record_history = 0
  p % profile_every = 0
  p % true_residual_stop = 0

  rc = ehip_solve_csr(n, A % Rows, A % Cols, A % Diag, A % Values, b, x, p, r)
  IF (rc /= EHIP_OK) CALL Fatal('EhipLinsolve', ehip_error_message())
  IF (r % info /= EHIP_HUTI_CONVERGENCE) THEN
    Abort = ListGetLogical(Solver % Values, 'Linear System Abort Not Converged', Found)
    IF (Abort .OR. .NOT. Found) CALL Fatal('EhipLinsolve', 'Failed convergence tolerances.')  ! IterSolve.src:784-792
  END IF
  Norm = r % norm                                       ! ComputeNorm, SolverUtils.src:4887-5006
  handled = 1                                           ! skip the built-in SolveLinearSystem
END FUNCTION EhipLinsolve
