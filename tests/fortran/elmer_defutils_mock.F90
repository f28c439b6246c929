! Test infrastructure only: a stand-in for the handful of ElmerSolver
! entities the Before-Linsolve shim (elmerfem_amd/fortran/ehip_linsolve.F90)
! uses -- dp, MAX_NAME_LEN, Model_t, Solver_t % Values, Matrix_t's CRS
! arrays, the ListGet* keyword readers with the reference's argument lists
! (fem/src/Lists.src:2006, 2093, 2114, 2217) and Fatal -- so that the shim
! is compiled and run here.  An Elmer build provides the real module.
module DefUtils
  implicit none
  integer, parameter :: dp = kind(1.0d0)
  integer, parameter :: MAX_NAME_LEN = 128
  type ValueList_t
    character(len=MAX_NAME_LEN) :: name = ''
    character(len=MAX_NAME_LEN) :: s = ''
    real(dp) :: r = 0.0_dp
    integer :: i = 0, kind = 0      ! 1 string, 2 real, 3 integer, 4 logical
    logical :: l = .false.
    type(ValueList_t), pointer :: next => null()
  end type
  type Model_t
    integer :: dummy = 0
  end type
  type Solver_t
    type(ValueList_t), pointer :: Values => null()
  end type
  type Matrix_t
    integer, pointer, contiguous :: Rows(:) => null(), Cols(:) => null(), Diag(:) => null()
    real(dp), pointer, contiguous :: Values(:) => null()
  end type
contains
  function lower(s) result(t)
    character(len=*), intent(in) :: s
    character(len=len(s)) :: t
    integer :: k
    t = s
    do k = 1, len(s)
      if (t(k:k) >= 'A' .and. t(k:k) <= 'Z') t(k:k) = achar(iachar(t(k:k)) + 32)
    end do
  end function
  function find(list, name) result(p)
    type(ValueList_t), pointer :: list, p
    character(len=*) :: name
    p => list
    do while (associated(p))
      if (trim(p % name) == trim(lower(name))) return
      p => p % next
    end do
  end function
  subroutine add(list, name, p)
    type(ValueList_t), pointer :: list, p
    character(len=*) :: name
    allocate(p)
    p % name = lower(name)
    p % next => list
    list => p
  end subroutine
  subroutine ListAddString(list, name, v)
    type(ValueList_t), pointer :: list, p
    character(len=*) :: name, v
    call add(list, name, p)
    p % kind = 1
    p % s = lower(v)
  end subroutine
  subroutine ListAddConstReal(list, name, v)
    type(ValueList_t), pointer :: list, p
    character(len=*) :: name
    real(dp) :: v
    call add(list, name, p)
    p % kind = 2
    p % r = v
  end subroutine
  subroutine ListAddInteger(list, name, v)
    type(ValueList_t), pointer :: list, p
    character(len=*) :: name
    integer :: v
    call add(list, name, p)
    p % kind = 3
    p % i = v
  end subroutine
  subroutine ListAddLogical(list, name, v)
    type(ValueList_t), pointer :: list, p
    character(len=*) :: name
    logical :: v
    call add(list, name, p)
    p % kind = 4
    p % l = v
  end subroutine
  function ListGetString(List, Name, Found) result(S)
    type(ValueList_t), pointer :: List, p
    character(len=*) :: Name
    logical, optional :: Found
    character(len=MAX_NAME_LEN) :: S
    p => find(List, Name)
    S = ''
    if (associated(p)) S = p % s
    if (present(Found)) Found = associated(p)
  end function
  function ListGetCReal(List, Name, Found) result(r)
    type(ValueList_t), pointer :: List, p
    character(len=*) :: Name
    logical, optional :: Found
    real(dp) :: r
    p => find(List, Name)
    r = 0.0_dp
    if (associated(p)) r = p % r
    if (present(Found)) Found = associated(p)
  end function
  function ListGetInteger(List, Name, Found, minv, maxv) result(L)
    type(ValueList_t), pointer :: List, p
    character(len=*) :: Name
    logical, optional :: Found
    integer, optional :: minv, maxv
    integer :: L
    p => find(List, Name)
    L = 0
    if (associated(p)) L = p % i
    if (present(Found)) Found = associated(p)
    if (present(minv)) then
      if (associated(p) .and. L < minv) call Fatal('ListGetInteger', 'value below minv')
    end if
  end function
  function ListGetLogical(List, Name, Found) result(L)
    type(ValueList_t), pointer :: List, p
    character(len=*) :: Name
    logical, optional :: Found
    logical :: L
    p => find(List, Name)
    L = .false.
    if (associated(p)) L = p % l
    if (present(Found)) Found = associated(p)
  end function
  subroutine Fatal(Caller, String)
    character(len=*) :: Caller, String
    print '(a)', 'FATAL ' // trim(Caller) // ': ' // trim(String)
    stop 2
  end subroutine
end module DefUtils
