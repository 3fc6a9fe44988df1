! Harness around the reference's own modules (compiled unmodified from
! /root/reference by oracle/build_ref.sh).  It performs the start-up and the
! per-step sequence of the reference driver that touch the time-stepping
! path (src/gpe.f90:33-68, 134-174, 228, 277): MPI start-up, the
! decomposition tables, the double-precision MPI datatypes of
! mpi_constants_precision (src/mpi.f90:24-32), get_grid, then
!   psi%old = psi%new ; ghost exchange ; solver(psi%old, psi%new) ; p = p+1
! while t+im_t <= end_time.  The initial state is read in the reference's
! end_state layout (src/io.f90:852-862) from the file named on the command
! line; the final state, its norm and energy and a per-step (t, im_t, dt)
! trace are written for the tests.
program ref_driver
  use parameters
  use variables
  use ic, only : get_grid
  use solve
  implicit none

  type (var) :: psi
  integer :: u, nxp, nyp, nzp, nst
  real (pr) :: norm, E
  complex (pr) :: dt_file
  character(len=512) :: fin, fout, ftrace

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  call get_command_argument(3, ftrace)

  call MPI_INIT(ierr)
  call MPI_COMM_SIZE(MPI_COMM_WORLD, nprocs, ierr)
  call MPI_COMM_RANK(MPI_COMM_WORLD, myrank, ierr)

  open (newunit=u, file='run.in', status='old')
  read (u, nml=run_params)
  rewind (u)
  read (u, nml=eqn_params)
  rewind (u)
  read (u, nml=io_params)
  rewind (u)
  read (u, nml=misc_params)
  close (u)

  gpe_mpi_real     = MPI_DOUBLE_PRECISION
  gpe_mpi_2real    = MPI_2DOUBLE_PRECISION
  gpe_mpi_complex  = MPI_DOUBLE_COMPLEX

  call setup_itable()
  call para_range(0, nz1, nzprocs, myrankz, ks, ke)
  call para_range(0, ny1, nyprocs, myranky, js, je)
  call array_len()
  call neighbours()

  allocate(psi%old(0:nx1,js-2:je+2,ks-2:ke+2))
  allocate(psi%new(0:nx1,js:je,ks:ke))
  allocate(works1(0:nx1,2,kks:kke))
  allocate(works2(0:nx1,2,kks:kke))
  allocate(workr1(0:nx1,2,kks:kke))
  allocate(workr2(0:nx1,2,kks:kke))

  if (real_time) then
    dt = cmplx(tau, 0.0_pr, pr)
  else
    dt = cmplx(0.0_pr, -tau, pr)
  end if
  call get_grid()

  open (newunit=u, file=trim(fin), access='stream', status='old')
  read (u) nxp
  read (u) nyp
  read (u) nzp
  read (u) p
  read (u) t, im_t
  read (u) dt_file
  read (u) psi%new
  close (u)
  if (scheme == 'rk45') dt = dt_file

  psi%old = 0.0_pr
  open (newunit=nst, file=trim(ftrace), access='stream', status='replace')
  do while (t+im_t <= end_time)
    psi%old(:,js:je,ks:ke) = psi%new
    call send_recv_z(psi%old)
    call pack_y(psi%old)
    call send_recv_y()
    call unpack_y(psi%old)
    call solver(psi%old, psi%new)
    p = p+1
    write (nst) t, im_t, dt
  end do
  close (nst)

  call get_norm(psi%new, norm)
  psi%old(:,js:je,ks:ke) = psi%new
  call send_recv_z(psi%old)
  call pack_y(psi%old)
  call send_recv_y()
  call unpack_y(psi%old)
  call energy(psi%old, E)

  open (newunit=u, file=trim(fout), access='stream', status='replace')
  write (u) nx
  write (u) ny
  write (u) nz
  write (u) p
  write (u) t, im_t
  write (u) dt
  write (u) psi%new
  write (u) norm, E
  close (u)

  call MPI_FINALIZE(ierr)
end program ref_driver
