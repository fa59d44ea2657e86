! ref_fd_bind.f90 -- C-callable wrappers around the reference's own stencil
! module, tools/finite_differences.f90 of FlavioRosales/KG, compiled from its
! source under /root/reference by oracle/Makefile into oracle/_ref/.
!
! TEST INFRASTRUCTURE ONLY: the tests call these wrappers to check the C
! restatement in oracle/kg_oracle.c bit for bit.  The live 2nd-order set takes
! complex(8) arrays in the reference's padded layout with bounds (-g:ni-1-g,
! -g:nj-1-g, 0:nk-1), as the reference passes them through index_bounds
! (lines_method_lib.f90:103-108); the 4th-order set takes real(8) arrays with
! lower bound 0, as declared there.
module kgref_fd_bind
  use iso_c_binding
  use finite_differences
  implicit none
contains

  subroutine set_bounds(ni, nj, nk, g, bounds)
    integer(c_int), intent(in) :: ni, nj, nk, g
    integer, intent(out) :: bounds(3,2)
    bounds(1,1) = -g; bounds(1,2) = ni - 1 - g
    bounds(2,1) = -g; bounds(2,2) = nj - 1 - g
    bounds(3,1) = 0;  bounds(3,2) = nk - 1
  end subroutine set_bounds

  subroutine kgref_dx1(ni, nj, nk, g, h, f, df) bind(C, name="kgref_dx1")
    integer(c_int), value :: ni, nj, nk, g
    real(c_double), value :: h
    complex(c_double_complex), intent(in) :: f(-g:ni-1-g, -g:nj-1-g, 0:nk-1)
    complex(c_double_complex), intent(out) :: df(-g:ni-1-g, -g:nj-1-g, 0:nk-1)
    integer :: bounds(3,2)
    real(kind=8) :: hh
    call set_bounds(ni, nj, nk, g, bounds)
    hh = h
    df = first_derivative_x_2_uniform(f, hh, bounds)
  end subroutine kgref_dx1

  subroutine kgref_dx2(ni, nj, nk, g, h, f, df) bind(C, name="kgref_dx2")
    integer(c_int), value :: ni, nj, nk, g
    real(c_double), value :: h
    complex(c_double_complex), intent(in) :: f(-g:ni-1-g, -g:nj-1-g, 0:nk-1)
    complex(c_double_complex), intent(out) :: df(-g:ni-1-g, -g:nj-1-g, 0:nk-1)
    integer :: bounds(3,2)
    real(kind=8) :: hh
    call set_bounds(ni, nj, nk, g, bounds)
    hh = h
    df = first_derivative_y_2(f, hh, bounds, .false.)
  end subroutine kgref_dx2

  subroutine kgref_dx3(ni, nj, nk, g, h, f, df) bind(C, name="kgref_dx3")
    integer(c_int), value :: ni, nj, nk, g
    real(c_double), value :: h
    complex(c_double_complex), intent(in) :: f(-g:ni-1-g, -g:nj-1-g, 0:nk-1)
    complex(c_double_complex), intent(out) :: df(-g:ni-1-g, -g:nj-1-g, 0:nk-1)
    integer :: bounds(3,2)
    real(kind=8) :: hh
    call set_bounds(ni, nj, nk, g, bounds)
    hh = h
    df = first_derivative_z_2(f, hh, bounds)
  end subroutine kgref_dx3

  subroutine kgref_advec_x(ni, nj, nk, g, h, f, beta, df) bind(C, name="kgref_advec_x")
    integer(c_int), value :: ni, nj, nk, g
    real(c_double), value :: h
    complex(c_double_complex), intent(in) :: f(-g:ni-1-g, -g:nj-1-g, 0:nk-1)
    real(c_double), intent(in) :: beta(-g:ni-1-g, -g:nj-1-g, 0:nk-1)
    complex(c_double_complex), intent(out) :: df(-g:ni-1-g, -g:nj-1-g, 0:nk-1)
    integer :: bounds(3,2)
    real(kind=8) :: hh
    call set_bounds(ni, nj, nk, g, bounds)
    hh = h
    df = advec_x_uniform(f, beta, hh, bounds)
  end subroutine kgref_advec_x

  ! The 4th-order set (real arrays, lower bound 0 on every axis).
  subroutine kgref_dx1_4(ni, nj, nk, h, f, df) bind(C, name="kgref_dx1_4")
    integer(c_int), value :: ni, nj, nk
    real(c_double), value :: h
    real(c_double), intent(in) :: f(0:ni-1, 0:nj-1, 0:nk-1)
    real(c_double), intent(out) :: df(0:ni-1, 0:nj-1, 0:nk-1)
    df = first_derivative_x_4(f, h)
  end subroutine kgref_dx1_4

  subroutine kgref_dx2_4(ni, nj, nk, h, f, df) bind(C, name="kgref_dx2_4")
    integer(c_int), value :: ni, nj, nk
    real(c_double), value :: h
    real(c_double), intent(in) :: f(0:ni-1, 0:nj-1, 0:nk-1)
    real(c_double), intent(out) :: df(0:ni-1, 0:nj-1, 0:nk-1)
    df = first_derivative_y_4(f, h)
  end subroutine kgref_dx2_4

  subroutine kgref_dx3_4(ni, nj, nk, h, f, df) bind(C, name="kgref_dx3_4")
    integer(c_int), value :: ni, nj, nk
    real(c_double), value :: h
    real(c_double), intent(in) :: f(0:ni-1, 0:nj-1, 0:nk-1)
    real(c_double), intent(out) :: df(0:ni-1, 0:nj-1, 0:nk-1)
    df = first_derivative_z_4(f, h)
  end subroutine kgref_dx3_4

  subroutine kgref_advec_x_4(ni, nj, nk, h, f, beta, df) bind(C, name="kgref_advec_x_4")
    integer(c_int), value :: ni, nj, nk
    real(c_double), value :: h
    real(c_double), intent(in) :: f(0:ni-1, 0:nj-1, 0:nk-1), beta(0:ni-1, 0:nj-1, 0:nk-1)
    real(c_double), intent(out) :: df(0:ni-1, 0:nj-1, 0:nk-1)
    df = advec_x_4(f, beta, h)
  end subroutine kgref_advec_x_4

end module kgref_fd_bind
