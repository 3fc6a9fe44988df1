! Timing harness around the reference's own modules (compiled unmodified
! from /root/reference by oracle/build_ref.sh) for the multi-rank CPU
! baseline: MPI start-up, the decomposition tables and datatypes as in
! src/gpe.f90:33-68 and src/mpi.f90:24-32, get_grid, then nsteps of the
! reference's per-step sequence (src/gpe.f90:134-174, 228)
!   psi%old = psi%new ; ghost exchange ; solver(psi%old, psi%new)
! timed with MPI_WTIME between barriers.  Every rank fills its own block of
! the initial state from the global indices (a smooth field of unit mean
! modulus), so no file is read.  Rank 0 prints "ranks cells steps seconds
! norm" (the norm of the final state, which does not depend on the ranks).
! Usage: mpiexec -n <nyprocs*nzprocs> ref_bench <nsteps>   (run.in in the cwd)
program ref_bench
  use parameters
  use variables
  use ic, only : get_grid
  use solve
  implicit none

  type (var) :: psi
  integer :: u, nsteps, i, j, k, s
  real (pr) :: t0, t1, ph, norm
  character(len=32) :: arg

  call get_command_argument(1, arg)
  read (arg, *) nsteps

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

  do k = ks, ke
    do j = js, je
      do i = 0, nx1
        ph = 0.3_pr * sin(6.283185307179586_pr * real(i, pr) / real(nx, pr)) + &
             0.2_pr * cos(6.283185307179586_pr * real(2 * j + k, pr) / real(ny, pr))
        psi%new(i,j,k) = (1.0_pr + 0.1_pr * cos(ph)) * cmplx(cos(ph), sin(ph), pr)
      end do
    end do
  end do
  psi%old = 0.0_pr
  t = 0.0_pr
  im_t = 0.0_pr
  p = 0

  call MPI_BARRIER(MPI_COMM_WORLD, ierr)
  t0 = MPI_WTIME()
  do s = 1, nsteps
    psi%old(:,js:je,ks:ke) = psi%new
    call send_recv_z(psi%old)
    call pack_y(psi%old)
    call send_recv_y()
    call unpack_y(psi%old)
    call solver(psi%old, psi%new)
    p = p+1
  end do
  call MPI_BARRIER(MPI_COMM_WORLD, ierr)
  t1 = MPI_WTIME()
  call get_norm(psi%new, norm)
  if (myrank == 0) print '(i6,1x,i14,1x,i6,1x,es16.8,1x,es24.16)', nprocs, int(nx, 8) * ny * nz, nsteps, &
    t1 - t0, norm

  call MPI_FINALIZE(ierr)
end program ref_bench
