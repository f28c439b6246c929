! ehip.F90 -- ISO_C_BINDING interface to libehip.so (include/ehip.h).
!
! The Fortran side of the drop-in: Elmer's host code is Fortran, and these
! interfaces let it call the MI355X library with its own arrays.  Conventions
! follow Matrix_t (fem/src/Types.src:155-228): 1-based CRS Rows/Cols/Diag as
! default INTEGER, Values/RHS as REAL(dp).  Every function returns EHIP_OK (0)
! or a negative EHIP_E* code; ehip_error_message() gives the text.
module ehip
  use, intrinsic :: iso_c_binding
  implicit none
  private :: c_strlen

  integer(c_int), parameter :: EHIP_OK = 0, EHIP_EARG = -1, EHIP_EHIP = -2, &
       EHIP_ENOMEM = -3, EHIP_EUNSUP = -4, EHIP_ECAP = -5, EHIP_ESTATE = -6
  ! Linear System Iterative Method (IterSolve.src:221-253)
  integer(c_int), parameter :: EHIP_CG = 0, EHIP_BICGSTAB = 1
  ! Linear System Preconditioning (IterSolve.src:372-411)
  integer(c_int), parameter :: EHIP_PRECOND_NONE = 0, EHIP_PRECOND_DIAGONAL = 1
  ! ILU(k): EHIP_PRECOND_ILU0 + k, k = 0..9
  integer(c_int), parameter :: EHIP_PRECOND_ILU0 = 2
  integer(c_int), parameter :: EHIP_FP64 = 0, EHIP_MIXED_IR = 1
  integer(c_int), parameter :: EHIP_FAITHFUL = 0, EHIP_FAST = 1
  ! HUTI_INFO codes (hutiter/src/huti_fdefs.h)
  integer(c_int), parameter :: EHIP_HUTI_CONVERGENCE = 1, EHIP_HUTI_MAXITER = 2, &
       EHIP_HUTI_DIVERGENCE = 3, EHIP_HUTI_CG_RHO = 20, EHIP_HUTI_BICGSTAB_RHO = 35, &
       EHIP_HUTI_BICGSTAB_SNORM = 36, EHIP_HUTI_BICGSTAB_OMEGA = 37
  integer(c_int), parameter :: EHIP_PHYS_MODELPDE = 0, EHIP_PHYS_ELASTIC = 1

  ! ehip_params: the Linear System keywords.  Defaults are IterSolver's:
  ! BiCGStab, no preconditioning, scaling on.
  type, bind(C) :: ehip_params
     integer(c_int) :: method = EHIP_BICGSTAB
     integer(c_int) :: precond = EHIP_PRECOND_NONE
     real(c_double) :: tol = 1.0d-10
     integer(c_int) :: maxit = 10000
     integer(c_int) :: scaling = 1
     integer(c_int) :: precision = EHIP_FP64
     integer(c_int) :: mode = EHIP_FAITHFUL
     real(c_double) :: inner_tol = 0.0d0
     integer(c_int) :: inner_maxit = 0
     integer(c_int) :: check_every = 0
     real(c_double) :: divtol = 0.0d0
     integer(c_int) :: record_history = 0
     integer(c_int) :: profile_every = 0
     integer(c_int) :: true_residual_stop = 0
  end type ehip_params

  type, bind(C) :: ehip_report
     integer(c_int) :: info = 0
     integer(c_int) :: iters = 0
     integer(c_int) :: outer_iters = 0
     integer(c_int) :: spmv_count = 0
     real(c_double) :: res_scaled = 0.0d0
     real(c_double) :: res_true = 0.0d0
     real(c_double) :: norm = 0.0d0
     real(c_double) :: t_total_ms = 0.0d0
     real(c_double) :: t_krylov_ms = 0.0d0
     real(c_double) :: spmv_us = 0.0d0
     integer(c_int) :: spmv_samples = 0
     integer(c_int) :: spmv_precision = 0
  end type ehip_report

  interface
     integer(c_int) function ehip_init(device) bind(C, name="ehip_init")
       import :: c_int
       integer(c_int), value :: device
     end function ehip_init

     type(c_ptr) function ehip_last_error() bind(C, name="ehip_last_error")
       import :: c_ptr
     end function ehip_last_error

     integer(c_int) function ehip_version() bind(C, name="ehip_version")
       import :: c_int
     end function ehip_version

     ! "Matrix Vector Proc" ABI (fem/src/Load.c:734-747): every argument by
     ! reference; spmv is Matrix_t%SpMV.
     subroutine ehip_matvec(spmv, n, rows, cols, vals, u, v, reinit) bind(C, name="ehip_matvec")
       import :: c_ptr, c_int, c_double
       type(c_ptr) :: spmv
       integer(c_int) :: n, reinit
       integer(c_int) :: rows(*), cols(*)
       real(c_double) :: vals(*), u(*), v(*)
     end subroutine ehip_matvec

     subroutine ehip_matvec_free(spmv) bind(C, name="ehip_matvec_free")
       import :: c_ptr
       type(c_ptr) :: spmv
     end subroutine ehip_matvec_free

     integer(c_int) function ehip_system_create(n, nnz, rows, cols, diag, sys) &
          bind(C, name="ehip_system_create")
       import :: c_int, c_int64_t, c_ptr
       integer(c_int), value :: n
       integer(c_int64_t), value :: nnz
       integer(c_int), intent(in) :: rows(*), cols(*), diag(*)
       type(c_ptr), intent(out) :: sys
     end function ehip_system_create

     integer(c_int) function ehip_system_from_mesh(nnodes, nelem, npe, elements, nbnd, npe_bnd, &
          bnd_elements, dofs, optimize, sys) bind(C, name="ehip_system_from_mesh")
       import :: c_int, c_ptr
       integer(c_int), value :: nnodes, nelem, npe, nbnd, npe_bnd, dofs, optimize
       integer(c_int), intent(in) :: elements(*)
       type(c_ptr), value :: bnd_elements
       type(c_ptr), intent(out) :: sys
     end function ehip_system_from_mesh

     integer(c_int) function ehip_system_free(sys) bind(C, name="ehip_system_free")
       import :: c_int, c_ptr
       type(c_ptr), value :: sys
     end function ehip_system_free

     integer(c_int) function ehip_system_set_values(sys, values, rhs) bind(C, name="ehip_system_set_values")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: sys
       real(c_double), intent(in) :: values(*), rhs(*)
     end function ehip_system_set_values

     integer(c_int) function ehip_system_get_values(sys, values, rhs) bind(C, name="ehip_system_get_values")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: sys
       real(c_double), intent(out) :: values(*), rhs(*)
     end function ehip_system_get_values

     integer(c_int) function ehip_assemble(sys, physics, elem_code, cdim, coords, elem_coef, elem_conv, &
          elem_velo, elem_load, poisson, t_ms) bind(C, name="ehip_assemble")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: sys
       integer(c_int), value :: physics, elem_code, cdim
       real(c_double), intent(in) :: coords(*)
       type(c_ptr), value :: elem_coef, elem_conv, elem_velo, elem_load
       real(c_double), value :: poisson
       real(c_double), intent(out) :: t_ms
     end function ehip_assemble

     integer(c_int) function ehip_dirichlet(sys, nbc, bc_rows, bc_vals, symmetric, scaling) &
          bind(C, name="ehip_dirichlet")
       import :: c_int, c_ptr
       type(c_ptr), value :: sys
       integer(c_int), value :: nbc, symmetric, scaling
       integer(c_int), intent(in) :: bc_rows(*)
       type(c_ptr), value :: bc_vals
     end function ehip_dirichlet

     integer(c_int) function ehip_solve(sys, params, x, report, hist) bind(C, name="ehip_solve")
       import :: c_int, c_ptr, c_double, ehip_params, ehip_report
       type(c_ptr), value :: sys
       type(ehip_params), intent(in) :: params
       real(c_double), intent(inout) :: x(*)
       type(ehip_report), intent(out) :: report
       type(c_ptr), value :: hist
     end function ehip_solve

     integer(c_int) function ehip_get_solution(sys, x) bind(C, name="ehip_get_solution")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: sys
       real(c_double), intent(out) :: x(*)
     end function ehip_get_solution

     ! One-shot whole solve: what a Before Linsolve hook returning 1 runs.
     integer(c_int) function ehip_solve_csr(n, rows, cols, diag, values, b, x, params, report) &
          bind(C, name="ehip_solve_csr")
       import :: c_int, c_double, ehip_params, ehip_report
       integer(c_int), value :: n
       integer(c_int), intent(in) :: rows(*), cols(*), diag(*)
       real(c_double), intent(in) :: values(*), b(*)
       real(c_double), intent(inout) :: x(*)
       type(ehip_params), intent(in) :: params
       type(ehip_report), intent(out) :: report
     end function ehip_solve_csr

     ! The same solve with no device allocation (FP64; limits in ehip.h).
     integer(c_int) function ehip_solve_csr_noalloc(n, rows, cols, diag, values, b, x, params, report) &
          bind(C, name="ehip_solve_csr_noalloc")
       import :: c_int, c_double, ehip_params, ehip_report
       integer(c_int), value :: n
       integer(c_int), intent(in) :: rows(*), cols(*), diag(*)
       real(c_double), intent(in) :: values(*), b(*)
       real(c_double), intent(inout) :: x(*)
       type(ehip_params), intent(in) :: params
       type(ehip_report), intent(out) :: report
     end function ehip_solve_csr_noalloc

     integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
       import :: c_size_t, c_ptr
       type(c_ptr), value :: s
     end function c_strlen
  end interface

contains

  ! Text of the last error (ehip_last_error), as a Fortran string.
  function ehip_error_message() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: cs(:)
    integer :: i, n
    p = ehip_last_error()
    if (.not. c_associated(p)) then
       msg = ""
       return
    end if
    n = int(c_strlen(p))
    call c_f_pointer(p, cs, [n])
    allocate(character(len=n) :: msg)
    do i = 1, n
       msg(i:i) = cs(i)
    end do
  end function ehip_error_message

end module ehip
