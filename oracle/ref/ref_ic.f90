! Harness around the reference's own ic module (compiled unmodified from
! /root/reference by oracle/build_ref.sh, once per initial condition, with
! that condition's expression as the ic.in the reference includes in ics).
! It performs the start-up of the reference driver that the initial
! condition needs (src/gpe.f90:33-68, 89: run.in, the decomposition of one
! rank, get_grid, get_kc_amp) and calls ics with restart = .false.; the
! field is written as raw complex128 (x fastest) to the file named on the
! command line.
program ref_ic
  use parameters
  use variables, only : para_range
  use ic, only : get_grid, get_kc_amp, ics
  implicit none

  complex (pr), allocatable :: psi(:,:,:)
  integer :: u
  character(len=512) :: fout

  call get_command_argument(1, fout)

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

  call para_range(0, nz1, nzprocs, myrankz, ks, ke)
  call para_range(0, ny1, nyprocs, myranky, js, je)
  allocate(psi(0:nx1,js:je,ks:ke))

  call get_grid()
  call get_kc_amp()
  call ics(psi)

  open (newunit=u, file=trim(fout), access='stream', status='replace')
  write (u) psi
  close (u)

  call MPI_FINALIZE(ierr)
end program ref_ic
