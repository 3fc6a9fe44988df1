! Fortran host program of the GPE time-stepping path on one MI355X.  It
! follows the reference driver (src/gpe.f90:33-312) for the parts that feed
! the hot path: run.in namelists, the dt convention (76-80), the initial
! state read from end_state00.dat in the reference's restart format
! (src/ic.f90:117-186), the main loop `do while (t+im_t <= end_time)` with
! the RUNNING sentinel (134-159), the eqn-4 norm / imaginary-time logic
! (241-271) and the final end_state.dat (src/io.f90:840-891).  With
! save_pdf / save_vcf (io_params) or save_mom (grid_params) it writes the
! final state's velocity PDF, velocity correlation function and momentum in
! the formats of save_velocity_pdf, save_vel_corr and save_momentum
! (src/io.f90:222-311, 426-445) to pdf_vel<p>.dat, vcf<p>.dat (p as seven
! digits, as itos forms it) and momentum.dat, from the device diagnostics.
!
! host_mode = 'parity' copies psi%new -> psi%old and calls
! solver(psi%old, psi%new) every step exactly as the reference does;
! host_mode = 'resident' keeps psi on the GPU for the whole run.
program gpe_hip_main
  use iso_c_binding, only : c_double, c_int, c_int64_t
  use parameters
  use error
  use gpe_context, only : close_ctx, gpe_storage, gpe_vscale, the_ctx, check
  use gpe_hip_c, only : gpe_momentum, gpe_velocity_pdf, gpe_vcf
  use solve
  use variables
  implicit none

  type(var) :: psi
  real (pr) :: norm, prev_norm, relnorm
  logical :: state_exist, running, resident
  integer :: u, ios, nx_prev, ny_prev, nz_prev, nsteps
  integer(8) :: c0, c1, crate
  complex (pr) :: dt_prev
  real (pr) :: secs

  open (newunit=u, file='run.in', status='old', action='read', iostat=ios)
  if (ios /= 0) call emergency_stop('ERROR: cannot open run.in')
  read (u, nml=run_params, iostat=ios)
  if (ios /= 0) call emergency_stop('ERROR: cannot read namelist run_params')
  rewind (u)
  read (u, nml=eqn_params, iostat=ios)
  if (ios /= 0) call emergency_stop('ERROR: cannot read namelist eqn_params')
  rewind (u)
  read (u, nml=io_params, iostat=ios)
  rewind (u)
  read (u, nml=misc_params, iostat=ios)
  rewind (u)
  read (u, nml=grid_params, iostat=ios)
  close (u)
  call set_extents()
  resident = host_mode /= 'parity'
  gpe_storage = storage
  gpe_vscale = vscale

  errcon = (5.0_pr/safety)**(1.0_pr/dt_increase)
  if (real_time) then
    dt = cmplx(tau, 0.0_pr, pr)
  else
    dt = cmplx(0.0_pr, -tau, pr)
  end if

  ! get_grid (src/ic.f90:32-69): only the spacings are needed on the host
  xl = -xr; yl = -yr; zl = -zr
  dx = (xr-xl)/nx1; dy = (yr-yl)/ny1; dz = (zr-zl)/nz1
  dx2 = dx**2; dy2 = dy**2; dz2 = dz**2

  allocate(psi%new(0:nx1,js:je,ks:ke))
  allocate(psi%old(0:nx1,js-2:je+2,ks-2:ke+2))

  open (newunit=u, file='RUNNING')
  close (u)

  ! initial state: the reference's restart route (src/ic.f90:92-101, 117-186)
  if (.not. restart) call emergency_stop('ERROR: this host reads its initial &
    &state from end_state00.dat; set restart=.true.')
  inquire (file='end_state00.dat', exist=state_exist)
  if (.not. state_exist) call emergency_stop('ERROR: restart=.true. &
    &but end_state00.dat does not exist.')
  open (newunit=u, file='end_state00.dat', access='stream', status='old', action='read')
  read (u) nx_prev
  read (u) ny_prev
  read (u) nz_prev
  read (u) p
  read (u) t, im_t
  read (u) dt_prev
  if (nx_prev /= nx .or. ny_prev /= ny .or. nz_prev /= nz) &
    call emergency_stop('ERROR: end_state00.dat grid does not match grid_params')
  read (u) psi%new
  close (u)
  if (real_time) then
    if (real(dt_prev, pr) == 0.0_pr) then
      dt_prev = cmplx(-aimag(dt_prev), real(dt_prev, pr), pr)
      im_t = 0.0_pr
    end if
  else
    if (aimag(dt_prev) == 0.0_pr) then
      dt_prev = cmplx(aimag(dt_prev), -real(dt_prev, pr), pr)
      t = 0.0_pr
    end if
  end if
  select case (scheme)
    case ('rk45')
      dt = dt_prev
    case default
      if (dt_prev /= dt) p = int(p * dt_prev / dt)
  end select

  if (eqn_to_solve == 4 .and. .not. real_time) then
    call get_norm(psi%new, prev_norm)
    if (renorm) call renormalise(psi%new, prev_norm)
  else
    prev_norm = 1.0_pr
  end if
  norm = prev_norm

  psi%old = 0.0_pr
  if (resident) call state_put(psi%new)

  nsteps = 0
  call system_clock(c0, crate)
  do while (t+im_t <= end_time)
    inquire (file='RUNNING', exist=running)
    if (.not. running) exit

    if (resident) then
      call solver_resident(1)
    else
      psi%old(:,js:je,ks:ke) = psi%new
      call solver(psi%old, psi%new)
    end if

    if (eqn_to_solve == 4) then
      if (resident) then
        call state_norm(norm)
      else
        call get_norm(psi%new, norm)
      end if
      relnorm = abs((norm-prev_norm)/prev_norm)
      if (.not. real_time) then
        if (relnorm < 1e-12_pr) then
          if (stop_imag) exit
          real_time = .true.
          im_t = 0.0_pr
          dt = cmplx(abs(aimag(dt)), abs(real(dt, pr)), pr)
          print *, 'Switching to real time'
        end if
        if (renorm) then
          if (resident) then
            call state_renormalise(norm)
          else
            call renormalise(psi%new, norm)
          end if
        end if
      end if
      prev_norm = norm
    end if
    p = p + 1
    nsteps = nsteps + 1
  end do
  if (resident) call state_get(psi%new)
  call system_clock(c1)
  secs = real(c1 - c0, pr) / real(crate, pr)

  ! end_state (src/io.f90:852-862)
  open (newunit=u, file='end_state.dat', access='stream', status='replace')
  write (u) nx
  write (u) ny
  write (u) nz
  write (u) p
  write (u) t, im_t
  write (u) dt
  write (u) psi%new
  close (u)

  if (save_pdf .or. save_vcf .or. save_mom) call final_diagnostics()

  inquire (file='RUNNING', exist=running)
  if (running) then
    open (newunit=u, file='RUNNING')
    close (u, status='delete')
  end if
  write (*, '(a,i0,a,es12.5,a,es12.5,a,f10.4,a)') 'steps=', nsteps, ' t=', t, ' im_t=', im_t, ' wall=', secs, ' s'
  if (secs > 0) write (*, '(a,f12.4)') 'Gcell-steps/s=', real(nx,pr)*ny*nz*nsteps/secs/1e9_pr
  call close_ctx()

contains

  ! The device diagnostics of the final state (state_put makes it the
  ! context's state in either host mode).
  subroutine final_diagnostics()
    real(c_double) :: mom(3), vmean(3), vstdev(3), vmax(3), f(0:nx-1)
    real(c_double), allocatable :: pdf(:,:)
    integer(c_int64_t) :: cnt(3)
    real (pr) :: bins(3), g(3)
    integer :: i, a
    character(7) :: tag
    write (tag, '(i7.7)') p
    call state_put(psi%new)
    if (save_mom) then
      call check(gpe_momentum(the_ctx(), mom), 'momentum')
      open (newunit=u, file='momentum.dat', status='unknown', position='append')
      write (u, '(4e17.9)') t, mom(1), mom(2), mom(3)
      close (u)
    end if
    if (save_pdf) then
      allocate(pdf(nbins, 3))
      call check(gpe_velocity_pdf(the_ctx(), int(nbins, c_int), cnt, vmean, vstdev, vmax, pdf), 'velocity_pdf')
      open (newunit=u, file='pdf_vel'//tag//'.dat', status='unknown')
      write (u, '(a10,e17.9)') '# t:', t
      write (u, '(a10,3e17.9)') '# Mean:', vmean(1), vmean(2), vmean(3)
      write (u, '(a10,3e17.9)') '# StDev:', vstdev(1), vstdev(2), vstdev(3)
      write (u, '(a10,3e17.9)') '# Sum:', sum(pdf(:, 1)), sum(pdf(:, 2)), sum(pdf(:, 3))
      do i = -nbins/2+1, nbins/2
        do a = 1, 3
          bins(a) = 2.0_pr*real(i, pr)*vmax(a)/real(nbins, pr)
          g(a) = 1.0_pr/(vstdev(a)*sqrt(2.0_pr*pi)) * exp((-0.5_pr*(bins(a)-vmean(a))**2) / vstdev(a)**2)
        end do
        write (u, '(9e17.9)') (bins(a), pdf(i+nbins/2, a), g(a), a = 1, 3)
      end do
      close (u)
    end if
    if (save_vcf) then
      call check(gpe_vcf(the_ctx(), f), 'get_vcf')
      open (newunit=u, file='vcf'//tag//'.dat', status='unknown')
      write (u, '(a4,e17.9)') '# t:', t
      do i = 0, nx-1
        write (u, '(i5,e17.9)') i, f(i)
      end do
      close (u)
    end if
  end subroutine final_diagnostics

end program gpe_hip_main
