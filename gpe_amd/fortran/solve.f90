! Drop-in for the reference's module solve (src/solve.f90:15-21): the one
! public routine solver(var_in, var_out), with the reference's side effects
! on t, im_t and dt and the rk45 timestep.dat record.  The step itself runs in
! libgpehip.so: fused RK stage kernels for euler/rk2/rk4, Cash-Karp attempts
! with the rkqs controller for rk45.
!
! solver_resident(nsteps) is the performance path: it advances the device
! state without host copies (state_put/state_get move it across).
module solve
  use iso_c_binding
  use parameters
  use gpe_hip_c
  use gpe_context
  use error, only : emergency_stop
  implicit none

  private
  public :: solver, solver_resident, state_put, state_get, scheme_code

contains

  integer(c_int) function scheme_code()
    select case (scheme)
      case ('euler')
        scheme_code = GPE_EULER
      case ('rk2')
        scheme_code = GPE_RK2
      case ('rk4')
        scheme_code = GPE_RK4
      case ('rk45')
        scheme_code = GPE_RK45
      case default
        scheme_code = -1
        call emergency_stop('ERROR: Unrecognised time stepping scheme.')
    end select
  end function scheme_code

  ! real_time changes when imaginary-time relaxation converges
  ! (src/gpe.f90:251-271); rkqs reads it on a retry (src/solve.f90:195-201)
  subroutine pass_real_time()
    call check(gpe_set_real_time(the_ctx(), merge(1_c_int, 0_c_int, real_time)), 'solver')
  end subroutine pass_real_time

  subroutine io_from_state(io)
    type(gpe_step_info), intent(out) :: io
    io%t = t
    io%im_t = im_t
    io%dt_re = real(dt, pr)
    io%dt_im = aimag(dt)
    io%errmax = 0
    io%psi2max = 0
    io%n_accept = 0
    io%n_reject = 0
    io%n_rhs = 0
  end subroutine io_from_state

  subroutine state_from_io(io)
    type(gpe_step_info), intent(in) :: io
    integer :: u
    t = io%t
    im_t = io%im_t
    dt = cmplx(io%dt_re, io%dt_im, pr)
    if (scheme == 'rk45' .and. myrank == 0) then
      ! src/solve.f90:234-241
      if (mod(p, save_rate) == 0) then
        open (newunit=u, status='unknown', position='append', file='timestep.dat')
        write (u, '(3e17.9,i10)') t, im_t, abs(dt), p
        close (u)
      end if
    end if
  end subroutine state_from_io

  subroutine solver(var_in, var_out)
    complex (pr), dimension(0:nx1,js-2:je+2,ks-2:ke+2), intent(in) :: var_in
    complex (pr), dimension(0:nx1,js:je,ks:ke), intent(out) :: var_out
    type(gpe_step_info) :: io
    integer(c_int) :: code
    integer(c_int64_t) :: nya, yo, zo
    code = scheme_code()
    call io_from_state(io)
    call pass_real_time()
    call halo_args(nya, yo, zo)
    call check(gpe_solver(the_ctx(), code, var_in, nya, yo, zo, var_out, host_kind(), io), 'solver')
    call state_from_io(io)
  end subroutine solver

  subroutine solver_resident(nsteps)
    integer, intent(in) :: nsteps
    type(gpe_step_info) :: io
    integer(c_int) :: code
    code = scheme_code()
    call io_from_state(io)
    call pass_real_time()
    call check(gpe_step(the_ctx(), code, int(nsteps, c_int), io), 'solver')
    call state_from_io(io)
  end subroutine solver_resident

  subroutine state_put(psi)
    complex (pr), dimension(0:nx1,js:je,ks:ke), intent(in) :: psi
    call check(gpe_upload(the_ctx(), psi, host_kind(), int(je-js+1, c_int64_t), 0_c_int64_t, 0_c_int64_t), &
               'upload')
  end subroutine state_put

  subroutine state_get(psi)
    complex (pr), dimension(0:nx1,js:je,ks:ke), intent(out) :: psi
    call check(gpe_download(the_ctx(), psi, host_kind()), 'download')
  end subroutine state_get

end module solve
