! Drop-in for the reference's module derivs (src/derivs.f90:15-21): the same
! public routines and argument shapes -- f is the halo array
! (0:nx1, js-2:je+2, ks-2:ke+2), the result the interior (0:nx1, js:je, ks:ke).
! The derivatives are computed by the HIP kernels; the halo planes of f are
! not read (boundaries are applied by in-kernel indexing).
module derivs
  use iso_c_binding
  use parameters
  use gpe_hip_c
  use gpe_context
  implicit none

  private
  public :: deriv_x, deriv_y, deriv_z, deriv_xx, deriv_yy, deriv_zz, deriv_lap

contains

  subroutine apply(op, f, fx, what)
    integer(c_int), intent(in) :: op
    complex (pr), dimension(0:nx1,js-2:je+2,ks-2:ke+2), intent(in) :: f
    complex (pr), dimension(0:nx1,js:je,ks:ke), intent(out) :: fx
    character(len=*), intent(in) :: what
    integer(c_int64_t) :: nya, yo, zo
    call halo_args(nya, yo, zo)
    call check(gpe_deriv(the_ctx(), op, f, nya, yo, zo, fx, host_kind()), what)
  end subroutine apply

  subroutine deriv_x(f, fx)
    complex (pr), dimension(0:nx1,js-2:je+2,ks-2:ke+2), intent(in) :: f
    complex (pr), dimension(0:nx1,js:je,ks:ke), intent(out) :: fx
    call apply(GPE_D_X, f, fx, 'deriv_x')
  end subroutine deriv_x

  subroutine deriv_y(f, fy)
    complex (pr), dimension(0:nx1,js-2:je+2,ks-2:ke+2), intent(in) :: f
    complex (pr), dimension(0:nx1,js:je,ks:ke), intent(out) :: fy
    call apply(GPE_D_Y, f, fy, 'deriv_y')
  end subroutine deriv_y

  subroutine deriv_z(f, fz)
    complex (pr), dimension(0:nx1,js-2:je+2,ks-2:ke+2), intent(in) :: f
    complex (pr), dimension(0:nx1,js:je,ks:ke), intent(out) :: fz
    call apply(GPE_D_Z, f, fz, 'deriv_z')
  end subroutine deriv_z

  subroutine deriv_xx(f, fxx)
    complex (pr), dimension(0:nx1,js-2:je+2,ks-2:ke+2), intent(in) :: f
    complex (pr), dimension(0:nx1,js:je,ks:ke), intent(out) :: fxx
    call apply(GPE_D_XX, f, fxx, 'deriv_xx')
  end subroutine deriv_xx

  subroutine deriv_yy(f, fyy)
    complex (pr), dimension(0:nx1,js-2:je+2,ks-2:ke+2), intent(in) :: f
    complex (pr), dimension(0:nx1,js:je,ks:ke), intent(out) :: fyy
    call apply(GPE_D_YY, f, fyy, 'deriv_yy')
  end subroutine deriv_yy

  subroutine deriv_zz(f, fzz)
    complex (pr), dimension(0:nx1,js-2:je+2,ks-2:ke+2), intent(in) :: f
    complex (pr), dimension(0:nx1,js:je,ks:ke), intent(out) :: fzz
    call apply(GPE_D_ZZ, f, fzz, 'deriv_zz')
  end subroutine deriv_zz

  ! dxx + dyy + dzz in one kernel (used by variables::laplacian)
  subroutine deriv_lap(f, lap)
    complex (pr), dimension(0:nx1,js-2:je+2,ks-2:ke+2), intent(in) :: f
    complex (pr), dimension(0:nx1,js:je,ks:ke), intent(out) :: lap
    call apply(GPE_D_LAP, f, lap, 'laplacian')
  end subroutine deriv_lap

end module derivs
