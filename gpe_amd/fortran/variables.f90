! Drop-in for the hot-path routines of the reference's module variables:
! laplacian (src/variables.f90:55-71), get_norm (537-559), energy (563-587),
! mass (591-611), renormalise (698-709), with the reference's interfaces.
! The work runs on the GPU; reductions accumulate in FP64 in a fixed order.
module variables
  use iso_c_binding
  use parameters
  use gpe_hip_c
  use gpe_context
  use derivs, only : deriv_lap
  implicit none

  private
  public :: laplacian, get_norm, energy, mass, renormalise, state_norm, state_renormalise

  type, public :: var
    complex (pr), allocatable, dimension(:,:,:) :: new
    complex (pr), allocatable, dimension(:,:,:) :: old
    complex (pr), allocatable, dimension(:,:,:) :: old2
  end type var

contains

  function laplacian(in_var)
    complex (pr), dimension(0:nx1,js-2:je+2,ks-2:ke+2), intent(in) :: in_var
    complex (pr), dimension(0:nx1,js:je,ks:ke) :: laplacian
    call deriv_lap(in_var, laplacian)
  end function laplacian

  subroutine put_interior(in_var)
    complex (pr), dimension(0:nx1,js:je,ks:ke), intent(in) :: in_var
    call check(gpe_upload(the_ctx(), in_var, host_kind(), int(je-js+1, c_int64_t), 0_c_int64_t, 0_c_int64_t), &
               'upload')
  end subroutine put_interior

  subroutine get_norm(in_var, norm)
    complex (pr), dimension(0:nx1,js:je,ks:ke), intent(in) :: in_var
    real (pr), intent(out) :: norm
    real(c_double) :: v
    call put_interior(in_var)
    call check(gpe_norm(the_ctx(), v), 'get_norm')
    norm = v
  end subroutine get_norm

  subroutine energy(in_var, E)
    complex (pr), dimension(0:nx1,js-2:je+2,ks-2:ke+2), intent(in) :: in_var
    real (pr), intent(inout) :: E
    real(c_double) :: v
    integer(c_int64_t) :: nya, yo, zo
    call halo_args(nya, yo, zo)
    call check(gpe_upload(the_ctx(), in_var, host_kind(), nya, yo, zo), 'upload')
    call check(gpe_energy(the_ctx(), v), 'energy')
    E = v
  end subroutine energy

  subroutine mass(in_var, M)
    complex (pr), dimension(0:nx1,js:je,ks:ke), intent(in) :: in_var
    real (pr), intent(inout) :: M
    real(c_double) :: v
    call put_interior(in_var)
    call check(gpe_mass(the_ctx(), v), 'mass')
    M = v
  end subroutine mass

  subroutine renormalise(var, norm)
    complex (pr), dimension(0:nx1,js:je,ks:ke), intent(inout) :: var
    real (pr), intent(in) :: norm
    call put_interior(var)
    call check(gpe_renormalise(the_ctx(), real(norm, c_double)), 'renormalise')
    call check(gpe_download(the_ctx(), var, host_kind()), 'download')
  end subroutine renormalise

  ! norm and renormalisation of the device-resident state (no host copy)
  subroutine state_norm(norm)
    real (pr), intent(out) :: norm
    real(c_double) :: v
    call check(gpe_norm(the_ctx(), v), 'get_norm')
    norm = v
  end subroutine state_norm

  subroutine state_renormalise(norm)
    real (pr), intent(in) :: norm
    call check(gpe_renormalise(the_ctx(), real(norm, c_double)), 'renormalise')
  end subroutine state_renormalise

end module variables
