! Test driver: Elmer-style call of the Before-Linsolve shim on a Matrix_t
! holding the 5-point Laplacian of an m x m grid (1-based CRS, Diag set),
! b = 1, x = 0, with the method and preconditioner keywords from the
! command line.  Prints whether the shim took the solve, ComputeNorm, the
! relative residual recomputed here in FP64, and the solution.
! Usage: linsolve_driver m method precond tol
program linsolve_driver
  use DefUtils
  implicit none
  interface
    function EhipLinsolve(Model, Solver, A, b, x, n, DOFs, Norm) result(handled)
      import :: Model_t, Solver_t, Matrix_t, dp
      type(Model_t) :: Model
      type(Solver_t) :: Solver
      type(Matrix_t), pointer :: A
      integer :: n, DOFs, handled
      real(dp) :: x(n), b(n), Norm
    end function
  end interface
  type(Model_t) :: model
  type(Solver_t) :: solver
  type(Matrix_t), pointer :: A
  integer :: m, n, nnz, i, j, row, k, h
  real(dp), allocatable :: b(:), x(:)
  real(dp) :: nrm, tol, s, rr, bb
  character(len=64) :: arg, meth, prec

  call get_command_argument(1, arg)
  read(arg, *) m
  call get_command_argument(2, meth)
  call get_command_argument(3, prec)
  call get_command_argument(4, arg)
  read(arg, *) tol
  n = m * m
  allocate(A)
  allocate(A % Rows(n + 1), A % Cols(5 * n), A % Diag(n), A % Values(5 * n))
  nnz = 0
  A % Rows(1) = 1
  do j = 1, m
    do i = 1, m
      row = (j - 1) * m + i
      if (j > 1) call put(row - m, -1.0_dp)
      if (i > 1) call put(row - 1, -1.0_dp)
      call put(row, 4.0_dp)
      A % Diag(row) = nnz
      if (i < m) call put(row + 1, -1.0_dp)
      if (j < m) call put(row + m, -1.0_dp)
      A % Rows(row + 1) = nnz + 1
    end do
  end do
  allocate(b(n), x(n))
  b = 1.0_dp
  x = 0.0_dp
  call ListAddString(solver % Values, 'Linear System Iterative Method', trim(meth))
  call ListAddString(solver % Values, 'Linear System Preconditioning', trim(prec))
  call ListAddConstReal(solver % Values, 'Linear System Convergence Tolerance', tol)
  call ListAddInteger(solver % Values, 'Linear System Max Iterations', 5000)
  h = EhipLinsolve(model, solver, A, b, x, n, 1, nrm)
  rr = 0.0_dp
  bb = 0.0_dp
  do row = 1, n
    s = 0.0_dp
    do k = A % Rows(row), A % Rows(row + 1) - 1
      s = s + A % Values(k) * x(A % Cols(k))
    end do
    rr = rr + (b(row) - s)**2
    bb = bb + b(row)**2
  end do
  print '(a,i0)', 'handled ', h
  print '(a,es24.16)', 'norm ', nrm
  print '(a,es12.4)', 'res ', sqrt(rr / bb)
  print '(a)', 'x'
  do row = 1, n
    print '(es24.16)', x(row)
  end do
contains
  subroutine put(c, v)
    integer :: c
    real(dp) :: v
    nnz = nnz + 1
    A % Cols(nnz) = c
    A % Values(nnz) = v
  end subroutine
end program linsolve_driver
