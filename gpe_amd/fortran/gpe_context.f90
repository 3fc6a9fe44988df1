! The device context of the Fortran host: created from module parameters
! on first use, owned for the rest of the run.  Every failed C-ABI call
! ends the run through emergency_stop with the library's message.
module gpe_context
  use iso_c_binding
  use gpe_hip_c
  use parameters
  use error, only : emergency_stop
  implicit none
  private
  public :: the_ctx, check, host_kind, halo_args, close_ctx, gpe_storage, gpe_vscale

  type(c_ptr), save :: ctx = c_null_ptr
  ! device options with no counterpart in the reference's module parameters;
  ! set them before the first call that creates the context
  character(5), save :: gpe_storage = 'c128'   ! 'c128', 'c64' or 'mixed' device fields
  real(c_double), save :: gpe_vscale = 0.0_c_double ! V term in eqn 1 (0 = reference)

contains

  function the_ctx() result(c)
    type(c_ptr) :: c
    type(gpe_params) :: gp
    if (.not. c_associated(ctx)) then
      gp%nx = nx; gp%ny = ny; gp%nz = nz
      gp%order = order; gp%bcs = bcs; gp%eqn_to_solve = eqn_to_solve
      gp%real_time = merge(1, 0, real_time)
      select case (trim(gpe_storage))
      case ('c64')
        gp%storage = GPE_C64
      case ('mixed')
        gp%storage = GPE_MIXED
      case default
        gp%storage = GPE_C128
      end select
      gp%xr = xr; gp%yr = yr; gp%zr = zr
      gp%Urhs = Urhs; gp%diss_amp = diss_amp; gp%g = g; gp%mu = mu; gp%nn = nn
      gp%omx = omx; gp%omy = omy; gp%omz = omz; gp%vscale = gpe_vscale
      gp%eps = eps; gp%safety = safety; gp%dt_decrease = dt_decrease; gp%dt_increase = dt_increase
      if (gpe_ctx_create(gp, ctx) /= GPE_OK) &
        call emergency_stop('ERROR: '//c_string(gpe_last_error(c_null_ptr)))
    end if
    c = ctx
  end function the_ctx

  subroutine check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc /= GPE_OK) call emergency_stop('ERROR: '//what//': '//c_string(gpe_last_error(ctx)))
  end subroutine check

  integer(c_int) function host_kind()
    host_kind = merge(GPE_HOST_C128, GPE_HOST_C64, pr == c_double)
  end function host_kind

  ! layout of the reference's halo array (0:nx1, js-2:je+2, ks-2:ke+2) with one rank
  subroutine halo_args(nya, yo, zo)
    integer(c_int64_t), intent(out) :: nya, yo, zo
    nya = int(je - js + 5, c_int64_t)
    yo = 2_c_int64_t
    zo = 2_c_int64_t
  end subroutine halo_args

  subroutine close_ctx()
    if (c_associated(ctx)) call gpe_ctx_destroy(ctx)
    ctx = c_null_ptr
  end subroutine close_ctx

end module gpe_context
