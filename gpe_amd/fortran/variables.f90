! Drop-in for the hot-path routines of the reference's module variables:
! laplacian (src/variables.f90:55-71), get_norm (537-559), energy (563-587),
! mass (591-611), renormalise (698-709), with the reference's interfaces,
! for its diagnostics momentum (615-694), get_vcf (467-507) and
! linelength (879-1145), and for imprint_vortex_line (1148-1162).
! velocity_pdf replaces the sequence get_pdf_velocity + get_pdf x 3 of
! save_velocity_pdf (src/io.f90:236-239) with one call that returns the
! same PDFs, maxima, means and deviations.
! The work runs on the GPU; reductions accumulate in FP64 in a fixed order.
module variables
  use iso_c_binding
  use parameters
  use gpe_hip_c
  use gpe_context
  use derivs, only : deriv_lap
  implicit none

  private
  public :: laplacian, get_norm, energy, mass, renormalise, state_norm, state_renormalise, momentum, get_vcf, &
    velocity_pdf, imprint_vortex_line, state_imprint_vortex_line, state_energy, state_peek, linelength

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

  ! The host-array routines work on the context's auxiliary field: the
  ! array handed in is uploaded there and selected (put_interior, put_halo)
  ! until use_state, so a stale or auxiliary array (the reference's
  ! save_energy(t, psi%old)) never replaces the device-resident state.
  subroutine put_interior(in_var)
    complex (pr), dimension(0:nx1,js:je,ks:ke), intent(in) :: in_var
    call check(gpe_field_upload(the_ctx(), GPE_FIELD_AUX, in_var, host_kind(), int(je-js+1, c_int64_t), &
               0_c_int64_t, 0_c_int64_t), 'upload')
    call check(gpe_select_field(the_ctx(), GPE_FIELD_AUX), 'select')
  end subroutine put_interior

  subroutine use_state()
    call check(gpe_select_field(the_ctx(), GPE_FIELD_STATE), 'select')
  end subroutine use_state

  subroutine get_norm(in_var, norm)
    complex (pr), dimension(0:nx1,js:je,ks:ke), intent(in) :: in_var
    real (pr), intent(out) :: norm
    real(c_double) :: v
    call put_interior(in_var)
    call check(gpe_norm(the_ctx(), v), 'get_norm')
    call use_state()
    norm = v
  end subroutine get_norm

  subroutine energy(in_var, E)
    complex (pr), dimension(0:nx1,js-2:je+2,ks-2:ke+2), intent(in) :: in_var
    real (pr), intent(inout) :: E
    real(c_double) :: v
    call put_halo(in_var)
    call check(gpe_energy(the_ctx(), v), 'energy')
    call use_state()
    E = v
  end subroutine energy

  subroutine put_halo(in_var)
    complex (pr), dimension(0:nx1,js-2:je+2,ks-2:ke+2), intent(in) :: in_var
    integer(c_int64_t) :: nya, yo, zo
    call halo_args(nya, yo, zo)
    call check(gpe_field_upload(the_ctx(), GPE_FIELD_AUX, in_var, host_kind(), nya, yo, zo), 'upload')
    call check(gpe_select_field(the_ctx(), GPE_FIELD_AUX), 'select')
  end subroutine put_halo

  subroutine momentum(in_var, mom)
    complex (pr), dimension(0:nx1,js-2:je+2,ks-2:ke+2), intent(in) :: in_var
    real (pr), dimension(3), intent(out) :: mom
    real(c_double) :: m(3)
    call put_halo(in_var)
    call check(gpe_momentum(the_ctx(), m), 'momentum')
    call use_state()
    mom = m
  end subroutine momentum

  subroutine get_vcf(in_var, f)
    complex (pr), dimension(0:nx1,js-2:je+2,ks-2:ke+2), intent(in) :: in_var
    real (pr), dimension(0:nx1), intent(out) :: f
    real(c_double) :: fd(0:nx1)
    call put_halo(in_var)
    call check(gpe_vcf(the_ctx(), fd), 'get_vcf')
    call use_state()
    f = fd
  end subroutine get_vcf

  subroutine velocity_pdf(in_var, nb, pdf_vx, pdf_vy, pdf_vz, vmax, vmean, vstdev)
    complex (pr), dimension(0:nx1,js-2:je+2,ks-2:ke+2), intent(in) :: in_var
    integer, intent(in) :: nb
    real (pr), dimension(-nb/2+1:nb/2), intent(out) :: pdf_vx, pdf_vy, pdf_vz
    real (pr), dimension(3), intent(out) :: vmax, vmean, vstdev
    integer(c_int64_t) :: cnt(3)
    real(c_double) :: mx(3), mn(3), sd(3), pdf(nb, 3)
    call put_halo(in_var)
    call check(gpe_velocity_pdf(the_ctx(), int(nb, c_int), cnt, mn, sd, mx, pdf), 'velocity_pdf')
    call use_state()
    pdf_vx = pdf(:, 1)
    pdf_vy = pdf(:, 2)
    pdf_vz = pdf(:, 3)
    vmax = mx
    vmean = mn
    vstdev = sd
  end subroutine velocity_pdf

  subroutine mass(in_var, M)
    complex (pr), dimension(0:nx1,js:je,ks:ke), intent(in) :: in_var
    real (pr), intent(inout) :: M
    real(c_double) :: v
    call put_interior(in_var)
    call check(gpe_mass(the_ctx(), v), 'mass')
    call use_state()
    M = v
  end subroutine mass

  subroutine renormalise(var, norm)
    complex (pr), dimension(0:nx1,js:je,ks:ke), intent(inout) :: var
    real (pr), intent(in) :: norm
    call put_interior(var)
    call check(gpe_renormalise(the_ctx(), real(norm, c_double)), 'renormalise')
    call check(gpe_download(the_ctx(), var, host_kind()), 'download')
    call use_state()
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

  ! linelength (src/variables.f90:879-1145): total length of the zero lines
  real (pr) function linelength(time, in_var)
    real (pr), intent(in) :: time
    complex (pr), dimension(0:nx1,js-2:je+2,ks-2:ke+2), intent(in) :: in_var
    real(c_double) :: v
    call put_halo(in_var)
    call check(gpe_linelength(the_ctx(), v), 'linelength')
    call use_state()
    linelength = v
  end function linelength

  ! imprint_vortex_line (src/variables.f90:1148-1162): |in_var| vortex_line(vl1)
  function imprint_vortex_line(in_var)
    complex (pr), dimension(0:nx1,js:je,ks:ke), intent(in) :: in_var
    complex (pr), dimension(0:nx1,js:je,ks:ke) :: imprint_vortex_line
    call put_interior(in_var)
    call state_imprint_vortex_line()
    call check(gpe_download(the_ctx(), imprint_vortex_line, host_kind()), 'download')
    call use_state()
  end function imprint_vortex_line

  subroutine state_imprint_vortex_line()
    type(gpe_line) :: ln
    integer(c_int32_t) :: d
    select case (vl1%dir)
      case ('x')
        d = 0
      case ('y')
        d = 1
      case default
        d = 2
    end select
    ln = gpe_line(vl1%x0, vl1%y0, vl1%z0, vl1%amp1, vl1%amp2, vl1%ll, vl1%sgn, d, &
                  merge(1_c_int32_t, 0_c_int32_t, vl1%imprint_phase))
    call check(gpe_imprint_line(the_ctx(), ln, real(scal, c_double)), 'imprint_vortex_line')
  end subroutine state_imprint_vortex_line

  ! energy of the device-resident state
  subroutine state_energy(E)
    real (pr), intent(out) :: E
    real(c_double) :: v
    call check(gpe_energy(the_ctx(), v), 'energy')
    E = v
  end subroutine state_energy

  ! value of the device-resident state at (i, j, k)
  complex (pr) function state_peek(i, j, k)
    integer, intent(in) :: i, j, k
    real(c_double) :: v(2)
    call check(gpe_peek(the_ctx(), int(i, c_int32_t), int(j, c_int32_t), int(k, c_int32_t), v), 'peek')
    state_peek = cmplx(v(1), v(2), pr)
  end function state_peek

end module variables
