! Fortran host program of the GPE time-stepping path on one MI355X.  It
! follows the reference driver (src/gpe.f90:33-312): run.in namelists, the
! dt convention (76-80), the initial state (with restart the checkpoint
! end_state00.dat, src/ic.f90:117-186, else ic_state.dat), and the main loop
! `do while (t+im_t <= end_time)` (134-278) in the reference's order:
!   * the RUNNING sentinel (136-159);
!   * every save_rate steps (modulo(t+im_t, |dt| save_rate) < |dt|) the
!     time series energy.dat and psi_time.dat (save_energy, save_time,
!     src/io.f90:147-218), linelength.dat with save_ll (save_linelength,
!     src/io.f90:1406-1438) and momentum.dat with save_mom;
!   * at save_rate3 intervals condensed_particles and the velocity PDF and
!     correlation function (save_pdf, save_vcf: pdf/pdf_vel<p>.dat,
!     vcf/vcf<p>.dat, src/io.f90:222-311);
!   * at save_rate2 intervals the 3D state with save_3d (idl_surface,
!     proc00/dens<p>.dat, src/io.f90:489-529);
!   * at p_save intervals the periodic checkpoint proc00/end_state.dat and
!     save.dat (src/io.f90:840-891);
!   * the solver step, then the SAVE sentinel (save_run, src/io.f90:1566-1595);
!   * for eqn 4 the norm, norm.dat, the imaginary-time switch, renorm and
!     imprint_vl (gpe.f90:241-271).
! condensed_particles (mass.dat; with save_spectrum spectrum/, idf/ and
! eta_time.dat, src/io.f90:315-404, 632-725) runs before the loop and at
! save_rate3 intervals, as in the reference.  After the loop (at end_time,
! or when RUNNING was removed): proc00/end_state.dat (RUNNING removed) and,
! with save_3d, the final dens file; a restart reads that checkpoint back as
! end_state00.dat (scripts/run.sh -r copies it so).  p is written as
! seven digits, as itos forms it.  Not built: zeros, contours, averages and
! filtered surfaces.
!
! host_mode = 'parity' copies psi%new -> psi%old and calls
! solver(psi%old, psi%new) every step exactly as the reference does;
! host_mode = 'resident' keeps psi on the GPU for the whole run.
program gpe_hip_main
  use iso_c_binding, only : c_double, c_int, c_int64_t, c_loc, c_null_ptr
  use parameters
  use error
  use gpe_context, only : close_ctx, gpe_storage, gpe_vscale, the_ctx, check
  use gpe_hip_c, only : gpe_momentum, gpe_velocity_pdf, gpe_vcf, gpe_condensed_particles, gpe_linelength
  use solve
  use variables
  implicit none

  type(var) :: psi
  real (pr) :: norm, prev_norm, relnorm
  logical :: state_exist, running, resident
  integer :: u, ios, nx_prev, ny_prev, nz_prev, nsteps, ps, n, m, stop_at_p
  character(32) :: envv
  logical :: dirs_made(4) = .false.
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

  ! initial state (ics, src/ic.f90:84-113).  restart: the checkpoint
  ! end_state00.dat as state_restart reads it (src/ic.f90:117-186), times
  ! proc00/end_state_filtered.dat with saved_restart (143-159).  Otherwise
  ! the state in ic_state.dat (same layout; its counters are not used): the
  ! reference composes its initial state from the compile-time ic.in, which
  ! this host receives as a file written by the IC generators.
  if (restart) then
    inquire (file='end_state00.dat', exist=state_exist)
    if (.not. state_exist) call emergency_stop('ERROR: restart=.true. &
      &but end_state00.dat does not exist.')
    print *, 'Getting restart conditions'
    call read_state('end_state00.dat', .true.)
    if (saved_restart) call apply_filtered('proc00/end_state_filtered.dat')
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
  else
    inquire (file='ic_state.dat', exist=state_exist)
    if (.not. state_exist) call emergency_stop('ERROR: restart=.false. &
      &but the initial state ic_state.dat does not exist.')
    call read_state('ic_state.dat', .false.)
  end if

  relnorm = 0.0_pr
  psi%old = 0.0_pr
  if (resident) call state_put(psi%new)
  if (save_3d) call idl_surface()
  if (eqn_to_solve == 4 .and. .not. real_time) then
    if (resident) then
      call state_norm(prev_norm)
      if (renorm) call state_renormalise(prev_norm)
    else
      call get_norm(psi%new, prev_norm)
      if (renorm) call renormalise(psi%new, prev_norm)
    end if
    call save_norm(t, prev_norm, relnorm)
  else
    prev_norm = 1.0_pr
  end if
  norm = prev_norm
  call condensed_particles(t)
  ! src/gpe.f90:116-124
  if (.not. restart) then
    inquire (file='end_state00.dat', exist=state_exist)
    if (state_exist) call emergency_stop('ERROR: restart=.false. but end_state00.dat exists.')
  end if

  ps = int(t/p_save)
  n = int(t/save_rate2)
  m = int(t/save_rate3)

  ! GPE_REMOVE_RUNNING_AT_P=k: the host removes RUNNING itself at the top of
  ! iteration p = k, as an operator would, so that the stop and checkpoint
  ! path runs at a chosen step (tests)
  stop_at_p = -1
  call get_environment_variable('GPE_REMOVE_RUNNING_AT_P', envv, status=ios)
  if (ios == 0) read (envv, *, iostat=ios) stop_at_p

  nsteps = 0
  call system_clock(c0, crate)
  do while (t+im_t <= end_time)
    if (p == stop_at_p) then
      open (newunit=u, file='RUNNING')
      close (u, status='delete')
    end if
    inquire (file='RUNNING', exist=running)
    if (.not. running) then
      print *, 'Stop requested.'
      exit
    end if
    if (.not. resident) psi%old(:,js:je,ks:ke) = psi%new

    if (modulo(t+im_t, abs(dt)*save_rate) < abs(dt)) call time_series()
    if (t+im_t >= save_rate3*m) then
      call condensed_particles(t)
      if (save_pdf) call save_velocity_pdf()
      if (save_vcf) call save_vel_corr()
      m = m+1
    end if
    if (t+im_t >= save_rate2*n) then
      if (save_3d) call idl_surface()
      n = n+1
    end if
    if (t+im_t >= p_save*ps) then
      call end_state(0)
      ps = ps+1
    end if

    if (resident) then
      call solver_resident(1)
    else
      call solver(psi%old, psi%new)
    end if

    ! save_run (src/io.f90:1566-1595): a SAVE file asks for a 3D save
    inquire (file='SAVE', exist=state_exist)
    if (state_exist) then
      call idl_surface()
      open (newunit=u, file='SAVE')
      close (u, status='delete')
    end if

    if (eqn_to_solve == 4) then
      if (resident) then
        call state_norm(norm)
      else
        call get_norm(psi%new, norm)
      end if
      relnorm = abs((norm-prev_norm)/prev_norm)
      if (modulo(t+im_t, abs(dt)*save_rate) < abs(dt)) call save_norm(t, norm, relnorm)
    end if
    if (eqn_to_solve == 4 .and. .not. real_time) then
      if (relnorm < 1e-12_pr) then
        if (stop_imag) exit
        real_time = .true.
        im_t = 0.0_pr
        ps = int(t/p_save)
        n = int(t/save_rate2)
        m = int(t/save_rate3)
        if (save_3d) call idl_surface()
        dt = cmplx(abs(aimag(dt)), abs(real(dt, pr)), pr)
        print *, 'Switching to real time'
        print *, 't =', im_t
        print *, 'p =', p
      end if
      if (resident) then
        if (renorm) call state_renormalise(norm)
        if (imprint_vl) call state_imprint_vortex_line()
      else
        if (renorm) call renormalise(psi%new, norm)
        if (imprint_vl) psi%new = imprint_vortex_line(psi%new)
      end if
    end if
    prev_norm = norm
    p = p + 1
    nsteps = nsteps + 1
  end do
  if (resident) call state_get(psi%new)
  call system_clock(c1)
  secs = real(c1 - c0, pr) / real(crate, pr)

  call end_state(1)
  if (save_3d) call idl_surface()

  write (*, '(a,i0,a,es12.5,a,es12.5,a,f10.4,a)') 'steps=', nsteps, ' t=', t, ' im_t=', im_t, ' wall=', secs, ' s'
  if (secs > 0) write (*, '(a,f12.4)') 'Gcell-steps/s=', real(nx,pr)*ny*nz*nsteps/secs/1e9_pr
  call close_ctx()

contains

  ! A state file in the end_state layout (int32 nx, ny, nz, p; t, im_t;
  ! complex dt; psi, x fastest: src/io.f90:852-862).  counters: take p, t,
  ! im_t and dt_prev from it (state_restart, src/ic.f90:130-141).
  subroutine read_state(fname, counters)
    character(*), intent(in) :: fname
    logical, intent(in) :: counters
    integer :: p_file
    real (pr) :: t_file, im_t_file
    open (newunit=u, file=fname, access='stream', status='old', action='read')
    read (u) nx_prev
    read (u) ny_prev
    read (u) nz_prev
    read (u) p_file
    read (u) t_file, im_t_file
    read (u) dt_prev
    if (nx_prev /= nx .or. ny_prev /= ny .or. nz_prev /= nz) &
      call emergency_stop('ERROR: '//fname//' grid does not match grid_params')
    read (u) psi%new
    close (u)
    if (counters) then
      p = p_file
      t = t_file
      im_t = im_t_file
    end if
  end subroutine read_state

  ! saved_restart: psi times the filtered state of the previous run
  ! (src/ic.f90:143-159; the file holds 4 integers and 2 reals before it)
  subroutine apply_filtered(fname)
    character(*), intent(in) :: fname
    integer :: undef_int, i
    real (pr) :: undef_real
    complex (pr), allocatable :: filt(:,:,:)
    inquire (file=fname, exist=state_exist)
    if (.not. state_exist) call emergency_stop('ERROR: saved_restart=.true. but '//fname//' does not exist.')
    allocate(filt(0:nx1,js:je,ks:ke))
    open (newunit=u, file=fname, access='stream', status='old', action='read')
    do i = 1, 4
      read (u) undef_int
    end do
    read (u) undef_real
    read (u) undef_real
    read (u) filt
    close (u)
    psi%new = psi%new*filt
    deallocate(filt)
  end subroutine apply_filtered

  ! p as seven digits (itos, src/io.f90:29-47)
  character(7) function itos(k)
    integer, intent(in) :: k
    write (itos, '(i7.7)') k
  end function itos

  subroutine make_dir(ix, name)
    integer, intent(in) :: ix
    character(*), intent(in) :: name
    if (.not. dirs_made(ix)) then
      call execute_command_line('mkdir -p '//name)
      dirs_made(ix) = .true.
    end if
  end subroutine make_dir

  ! the current state on the device (parity mode keeps it on the host)
  subroutine on_device()
    if (.not. resident) call state_put(psi%new)
  end subroutine on_device

  ! the current state in psi%new (resident mode keeps it on the device)
  subroutine on_host()
    if (resident) call state_get(psi%new)
  end subroutine on_host

  ! save_energy, save_time (src/io.f90:147-218) and momentum.dat
  subroutine time_series()
    real (pr) :: E, mom(3)
    real(c_double) :: md(3), ll
    complex (pr) :: c
    call on_device()
    call state_energy(E)
    open (newunit=u, file='energy.dat', status='unknown', position='append')
    write (u, '(3e17.9)') t, im_t, E/real(nx*ny*nz, pr)
    close (u)
    c = state_peek(nx/2, ny/2, nz/2)
    open (newunit=u, file='psi_time.dat', status='unknown', position='append')
    write (u, '(6e17.9)') t, im_t, real(c, pr), aimag(c), abs(c)**2, atan2(aimag(c)+1.0e-6_pr, real(c, pr))
    close (u)
    if (save_ll) then
      ! save_linelength (src/io.f90:1406-1438)
      call check(gpe_linelength(the_ctx(), ll), 'linelength')
      open (newunit=u, file='linelength.dat', status='unknown', position='append')
      write (u, '(3e17.9)') t, im_t, ll
      close (u)
    end if
    if (save_mom) then
      call check(gpe_momentum(the_ctx(), md), 'momentum')
      mom = md
      open (newunit=u, file='momentum.dat', status='unknown', position='append')
      write (u, '(4e17.9)') t, mom(1), mom(2), mom(3)
      close (u)
    end if
  end subroutine time_series

  ! condensed_particles (src/io.f90:315-404): mass.dat, and with
  ! save_spectrum the spectrum (src/io.f90:632-725): spectrum/spectrum<p>.dat,
  ! idf/idf<p>.dat, eta_time.dat
  subroutine condensed_particles(time)
    real (pr), intent(in) :: time
    real(c_double) :: row(7)
    real(c_double), target :: eta(7), f(nx/2)
    integer(c_int64_t), target :: nharm(7)
    integer :: i
    call on_device()
    if (save_spectrum) then
      call check(gpe_condensed_particles(the_ctx(), row, c_loc(eta), c_loc(nharm), c_loc(f)), 'condensed_particles')
    else
      call check(gpe_condensed_particles(the_ctx(), row, c_null_ptr, c_null_ptr, c_null_ptr), 'condensed_particles')
    end if
    open (newunit=u, file='mass.dat', status='unknown', position='append')
    write (u, '(9e17.9)') time, im_t, (row(i), i = 1, 7)
    close (u)
    if (.not. save_spectrum) return
    call make_dir(4, 'spectrum idf')
    open (newunit=u, file='spectrum/spectrum'//itos(p)//'.dat', status='unknown')
    do i = 1, 7
      write (u, '(2i9,e17.9)') i, int(nharm(i)), eta(i)
    end do
    close (u)
    open (newunit=u, file='idf/idf'//itos(p)//'.dat', status='unknown')
    do i = 1, nx/2
      write (u, '(i9,e17.9)') i, f(i)
    end do
    close (u)
    open (newunit=u, file='eta_time.dat', status='unknown', position='append')
    write (u, '(4e17.9)') t, eta(1), eta(2), eta(3)
    close (u)
  end subroutine condensed_particles

  ! save_norm (src/io.f90:406-422)
  subroutine save_norm(time, nrm, rel)
    real (pr), intent(in) :: time, nrm, rel
    open (newunit=u, file='norm.dat', status='unknown', position='append')
    write (u, '(4e17.9)') time, im_t, nrm, rel
    close (u)
  end subroutine save_norm

  ! save_velocity_pdf (src/io.f90:222-290) from the device statistics
  subroutine save_velocity_pdf()
    real(c_double) :: vmean(3), vstdev(3), vmax(3)
    real(c_double), allocatable :: pdf(:,:)
    integer(c_int64_t) :: cnt(3)
    real (pr) :: bins(3), g(3)
    integer :: i, a
    call on_device()
    allocate(pdf(nbins, 3))
    call check(gpe_velocity_pdf(the_ctx(), int(nbins, c_int), cnt, vmean, vstdev, vmax, pdf), 'velocity_pdf')
    call make_dir(1, 'pdf')
    open (newunit=u, file='pdf/pdf_vel'//itos(p)//'.dat', status='unknown')
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
  end subroutine save_velocity_pdf

  ! save_vel_corr (src/io.f90:294-311)
  subroutine save_vel_corr()
    real(c_double) :: f(0:nx-1)
    integer :: i
    call on_device()
    call check(gpe_vcf(the_ctx(), f), 'get_vcf')
    call make_dir(2, 'vcf')
    open (newunit=u, file='vcf/vcf'//itos(p)//'.dat', status='unknown')
    write (u, '(a4,e17.9)') '# t:', t
    do i = 0, nx-1
      write (u, '(i5,e17.9)') i, f(i)
    end do
    close (u)
  end subroutine save_vel_corr

  ! idl_surface (src/io.f90:489-529): t+im_t, the grid, the (single) rank's
  ! block and the field, then the coordinates, in the stream layout the
  ! reference's readers expect; p appended to p_saved.dat
  subroutine idl_surface()
    integer :: i
    call on_host()
    call make_dir(3, 'proc00')
    open (newunit=u, file='proc00/dens'//itos(p)//'.dat', access='stream', status='replace')
    write (u) t+im_t
    write (u) nx, ny, nz
    write (u) 1, 1
    write (u) js, je, ks, ke
    write (u) psi%new
    write (u) (xl + real(i, pr)*dx, i = 0, nx1)
    write (u) (yl + real(i, pr)*dy, i = 0, ny1)
    write (u) (zl + real(i, pr)*dz, i = 0, nz1)
    close (u)
    open (newunit=u, file='p_saved.dat', status='unknown', position='append')
    write (u, '(i10)') p
    close (u)
  end subroutine idl_surface

  ! end_state (src/io.f90:840-891): the restart file, save.dat, and with
  ! flag = 1 the removal of RUNNING
  subroutine end_state(flag)
    integer, intent(in) :: flag
    call on_host()
    call make_dir(3, 'proc00')
    open (newunit=u, file='proc00/end_state.dat', access='stream', status='replace')
    write (u) nx
    write (u) ny
    write (u) nz
    write (u) p
    write (u) t, im_t
    write (u) dt
    write (u) psi%new
    close (u)
    open (newunit=u, file='save.dat', status='replace')
    write (u, *) 'Periodically saved state'
    write (u, *) 't =', t
    write (u, *) 'im_t =', im_t
    write (u, *) 'dt =', dt
    write (u, *) 'nx =', nx
    write (u, *) 'ny =', ny
    write (u, *) 'nz =', nz
    write (u, *) 'p =', p
    close (u)
    if (flag == 1) then
      inquire (file='RUNNING', exist=running)
      if (running) then
        open (newunit=u, file='RUNNING')
        close (u, status='delete')
      end if
    end if
  end subroutine end_state

end program gpe_hip_main
