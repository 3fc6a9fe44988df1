! ISO_C_BINDING interface to libgpehip.so (include/gpe_hip.h).  This thin
! layer is the only place the Fortran host touches the C ABI; the
! reference-compatible modules (solve, derivs, variables) sit on top of it.
module gpe_hip_c
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: GPE_OK = 0
  integer(c_int), parameter :: GPE_C128 = 0, GPE_C64 = 1, GPE_MIXED = 2
  integer(c_int), parameter :: GPE_HOST_C128 = 0, GPE_HOST_C64 = 1
  integer(c_int), parameter :: GPE_FIELD_STATE = 0, GPE_FIELD_AUX = 1
  integer(c_int), parameter :: GPE_EULER = 0, GPE_RK2 = 1, GPE_RK4 = 2, GPE_RK45 = 3
  integer(c_int), parameter :: GPE_D_X = 0, GPE_D_Y = 1, GPE_D_Z = 2, GPE_D_XX = 3, &
                               GPE_D_YY = 4, GPE_D_ZZ = 5, GPE_D_LAP = 6

  type, bind(C) :: gpe_params
    integer(c_int32_t) :: nx, ny, nz
    integer(c_int32_t) :: order, bcs, eqn_to_solve, real_time, storage
    real(c_double) :: xr, yr, zr
    real(c_double) :: Urhs, diss_amp, g, mu, nn, omx, omy, omz
    real(c_double) :: vscale
    real(c_double) :: eps, safety, dt_decrease, dt_increase
  end type gpe_params

  type, bind(C) :: gpe_step_info
    real(c_double) :: t, im_t, dt_re, dt_im, errmax, psi2max
    integer(c_int64_t) :: n_accept, n_reject, n_rhs
  end type gpe_step_info

  type, bind(C) :: gpe_line
    real(c_double) :: x0, y0, z0, amp1, amp2, ll, sgn
    integer(c_int32_t) :: dir, imprint_phase
  end type gpe_line

  interface
    integer(c_int) function gpe_abi_version() bind(C, name='gpe_abi_version')
      import
    end function gpe_abi_version

    type(c_ptr) function gpe_last_error(ctx) bind(C, name='gpe_last_error')
      import
      type(c_ptr), value :: ctx
    end function gpe_last_error

    integer(c_int) function gpe_ctx_create(p, ctx) bind(C, name='gpe_ctx_create')
      import
      type(gpe_params), intent(in) :: p
      type(c_ptr), intent(out) :: ctx
    end function gpe_ctx_create

    subroutine gpe_ctx_destroy(ctx) bind(C, name='gpe_ctx_destroy')
      import
      type(c_ptr), value :: ctx
    end subroutine gpe_ctx_destroy

    integer(c_int) function gpe_upload(ctx, host, host_dtype, ny_alloc, y_off, z_off) &
        bind(C, name='gpe_upload')
      import
      type(c_ptr), value :: ctx
      type(*), dimension(*), intent(in) :: host
      integer(c_int), value :: host_dtype
      integer(c_int64_t), value :: ny_alloc, y_off, z_off
    end function gpe_upload

    integer(c_int) function gpe_field_upload(ctx, field, host, host_dtype, ny_alloc, y_off, z_off) &
        bind(C, name='gpe_field_upload')
      import
      type(c_ptr), value :: ctx
      integer(c_int), value :: field
      type(*), dimension(*), intent(in) :: host
      integer(c_int), value :: host_dtype
      integer(c_int64_t), value :: ny_alloc, y_off, z_off
    end function gpe_field_upload

    integer(c_int) function gpe_select_field(ctx, field) bind(C, name='gpe_select_field')
      import
      type(c_ptr), value :: ctx
      integer(c_int), value :: field
    end function gpe_select_field

    integer(c_int) function gpe_set_real_time(ctx, real_time) bind(C, name='gpe_set_real_time')
      import
      type(c_ptr), value :: ctx
      integer(c_int), value :: real_time
    end function gpe_set_real_time

    integer(c_int) function gpe_download(ctx, host, host_dtype) bind(C, name='gpe_download')
      import
      type(c_ptr), value :: ctx
      type(*), dimension(*) :: host
      integer(c_int), value :: host_dtype
    end function gpe_download

    integer(c_int) function gpe_step(ctx, scheme, nsteps, io) bind(C, name='gpe_step')
      import
      type(c_ptr), value :: ctx
      integer(c_int), value :: scheme, nsteps
      type(gpe_step_info), intent(inout) :: io
    end function gpe_step

    integer(c_int) function gpe_solver(ctx, scheme, var_in, ny_alloc, y_off, z_off, var_out, host_dtype, io) &
        bind(C, name='gpe_solver')
      import
      type(c_ptr), value :: ctx
      integer(c_int), value :: scheme
      type(*), dimension(*), intent(in) :: var_in
      integer(c_int64_t), value :: ny_alloc, y_off, z_off
      type(*), dimension(*) :: var_out
      integer(c_int), value :: host_dtype
      type(gpe_step_info), intent(inout) :: io
    end function gpe_solver

    integer(c_int) function gpe_deriv(ctx, op, f, ny_alloc, y_off, z_off, out, host_dtype) &
        bind(C, name='gpe_deriv')
      import
      type(c_ptr), value :: ctx
      integer(c_int), value :: op
      type(*), dimension(*), intent(in) :: f
      integer(c_int64_t), value :: ny_alloc, y_off, z_off
      type(*), dimension(*) :: out
      integer(c_int), value :: host_dtype
    end function gpe_deriv

    integer(c_int) function gpe_rhs(ctx, f, ny_alloc, y_off, z_off, out, host_dtype) bind(C, name='gpe_rhs')
      import
      type(c_ptr), value :: ctx
      type(*), dimension(*), intent(in) :: f
      integer(c_int64_t), value :: ny_alloc, y_off, z_off
      type(*), dimension(*) :: out
      integer(c_int), value :: host_dtype
    end function gpe_rhs

    integer(c_int) function gpe_norm(ctx, v) bind(C, name='gpe_norm')
      import
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: v
    end function gpe_norm

    integer(c_int) function gpe_energy(ctx, v) bind(C, name='gpe_energy')
      import
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: v
    end function gpe_energy

    integer(c_int) function gpe_imprint_line(ctx, vl, scal) bind(C, name='gpe_imprint_line')
      import
      type(c_ptr), value :: ctx
      type(gpe_line), intent(in) :: vl
      real(c_double), value :: scal
    end function gpe_imprint_line

    integer(c_int) function gpe_linelength(ctx, length) bind(C, name='gpe_linelength')
      import
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: length
    end function gpe_linelength

    integer(c_int) function gpe_condensed_particles(ctx, row, eta, nharm, f) &
        bind(C, name='gpe_condensed_particles')
      import
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: row(7)
      type(c_ptr), value :: eta, nharm, f
    end function gpe_condensed_particles

    integer(c_int) function gpe_peek(ctx, i, j, k, v) bind(C, name='gpe_peek')
      import
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: i, j, k
      real(c_double), intent(out) :: v(2)
    end function gpe_peek

    integer(c_int) function gpe_momentum(ctx, mom) bind(C, name='gpe_momentum')
      import
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: mom(3)
    end function gpe_momentum

    integer(c_int) function gpe_velocity_pdf(ctx, nbins, cnt, vmean, vstdev, vmax, pdf) &
        bind(C, name='gpe_velocity_pdf')
      import
      type(c_ptr), value :: ctx
      integer(c_int), value :: nbins
      integer(c_int64_t), intent(out) :: cnt(3)
      real(c_double), intent(out) :: vmean(3), vstdev(3), vmax(3)
      real(c_double), intent(out) :: pdf(*)
    end function gpe_velocity_pdf

    integer(c_int) function gpe_vcf(ctx, f) bind(C, name='gpe_vcf')
      import
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: f(*)
    end function gpe_vcf

    integer(c_int) function gpe_mass(ctx, v) bind(C, name='gpe_mass')
      import
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: v
    end function gpe_mass

    integer(c_int) function gpe_renormalise(ctx, norm) bind(C, name='gpe_renormalise')
      import
      type(c_ptr), value :: ctx
      real(c_double), value :: norm
    end function gpe_renormalise
  end interface

contains

  ! Copy a NUL-terminated C string into a Fortran string.
  function c_string(p) result(s)
    type(c_ptr), intent(in) :: p
    character(len=:), allocatable :: s
    character(kind=c_char), dimension(:), pointer :: a
    integer :: n
    if (.not. c_associated(p)) then
      s = ''
      return
    end if
    call c_f_pointer(p, a, [4096])
    n = 0
    do while (n < 4096)
      if (a(n+1) == c_null_char) exit
      n = n + 1
    end do
    allocate(character(len=n) :: s)
    s = transfer(a(1:n), s)
  end function c_string

end module gpe_hip_c
