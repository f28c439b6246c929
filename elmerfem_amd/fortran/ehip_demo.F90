! ehip_demo.F90 -- calls libehip.so from Fortran the way Elmer's host code
! would, with Matrix_t-style 1-based CRS arrays:
!   1. "Matrix Vector Proc" ABI: v = A u through ehip_matvec, compared with
!      the CRS_MatrixVectorProd loop (fem/src/CRSMatrix.src:3565-3573) run
!      here on the host (same order, so the results must be identical);
!   2. the whole-solve entry (ehip_solve_csr) with CG + Jacobi to 1e-10,
!      the residual recomputed here in FP64.
! Usage: ehip_demo [m]  (2-D 5-point Laplacian on an m x m grid, default 64)
! Exit status 0 only if both checks pass.
program ehip_demo
  use ehip
  implicit none
  integer :: m, n, nnz, i, j, k, row, rc, arg_len
  integer(c_int), allocatable :: rows(:), cols(:), diag(:)
  real(c_double), allocatable :: vals(:), u(:), v(:), vref(:), b(:), x(:), r(:)
  type(c_ptr) :: spmv_handle
  integer(c_int) :: nn, reinit
  type(ehip_params) :: p
  type(ehip_report) :: rep
  character(len=32) :: arg
  real(c_double) :: s, rnorm, bnorm
  logical :: ok

  m = 64
  if (command_argument_count() >= 1) then
     call get_command_argument(1, arg, arg_len)
     read(arg, *) m
  end if
  n = m * m
  allocate(rows(n + 1), cols(5 * n), diag(n), vals(5 * n))
  ! 5-point Laplacian, unknowns numbered x fastest; columns ascending per row
  nnz = 0
  rows(1) = 1
  do j = 1, m
     do i = 1, m
        row = (j - 1) * m + i
        if (j > 1) call put(row - m, -1.0d0)
        if (i > 1) call put(row - 1, -1.0d0)
        call put(row, 4.0d0)
        diag(row) = nnz
        if (i < m) call put(row + 1, -1.0d0)
        if (j < m) call put(row + m, -1.0d0)
        rows(row + 1) = nnz + 1
     end do
  end do

  rc = ehip_init(0_c_int)
  if (rc /= EHIP_OK) then
     print *, "ehip_init failed: ", ehip_error_message()
     stop 1
  end if

  ! 1. Matrix Vector Proc
  allocate(u(n), v(n), vref(n))
  do i = 1, n
     u(i) = sin(0.37d0 * i) + 0.5d0
  end do
  spmv_handle = c_null_ptr
  nn = n
  reinit = 0
  call ehip_matvec(spmv_handle, nn, rows, cols, vals, u, v, reinit)
  do i = 1, n
     s = 0.0d0
     do k = rows(i), rows(i + 1) - 1
        s = s + vals(k) * u(cols(k))
     end do
     vref(i) = s
  end do
  ok = all(v == vref)
  print '(a,l1)', "matvec bit-exact: ", ok
  call ehip_matvec_free(spmv_handle)

  ! 2. whole solve
  allocate(b(n), x(n), r(n))
  b = 1.0d0
  x = 0.0d0
  p%method = EHIP_CG
  p%precond = EHIP_PRECOND_DIAGONAL
  p%tol = 1.0d-10
  p%maxit = 10000
  rc = ehip_solve_csr(int(n, c_int), rows, cols, diag, vals, b, x, p, rep)
  if (rc /= EHIP_OK) then
     print *, "ehip_solve_csr failed: ", ehip_error_message()
     stop 1
  end if
  do i = 1, n
     s = 0.0d0
     do k = rows(i), rows(i + 1) - 1
        s = s + vals(k) * x(cols(k))
     end do
     r(i) = b(i) - s
  end do
  rnorm = sqrt(sum(r * r))
  bnorm = sqrt(sum(b * b))
  print '(a,i0,a,i0,a,es10.3,a,es10.3)', "solve info=", rep%info, " iters=", rep%iters, &
       " res_scaled=", rep%res_scaled, " res_true(host)=", rnorm / bnorm
  ok = ok .and. rep%info == EHIP_HUTI_CONVERGENCE .and. rnorm / bnorm <= 1.0d-9
  if (.not. ok) stop 2
  print '(a)', "ehip_demo: ok"

contains

  subroutine put(col, val)
    integer, intent(in) :: col
    real(c_double), intent(in) :: val
    nnz = nnz + 1
    cols(nnz) = col
    vals(nnz) = val
  end subroutine put

end program ehip_demo
