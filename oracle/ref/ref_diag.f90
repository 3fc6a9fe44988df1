! Harness around the reference's own diagnostics (compiled unmodified from
! /root/reference by oracle/build_ref.sh): reads run.in and a state in the
! reference's end_state layout (src/io.f90:852-862), fills the ghost planes
! as the main loop does (src/gpe.f90:165-174), then calls
!   momentum (src/variables.f90:615-694),
!   get_pdf_velocity and get_pdf per component (src/variables.f90:377-463,
!     as save_velocity_pdf does, src/io.f90:222-285),
!   get_vcf (src/variables.f90:467-507),
!   linelength (src/variables.f90:879-1145)
! and writes the results (stream): mom(3), vmean(3), vstdev(3), vmax(3),
! the valid counts (3 int32), pdf_vx, pdf_vy, pdf_vz (nbins each), f(0:nx1),
! the line length.
program ref_diag
  use parameters
  use variables
  use ic, only : get_grid
  implicit none

  type (var) :: psi
  integer :: u, nxp, nyp, nzp
  complex (pr) :: dt_file
  real (pr), dimension(3) :: mom, vmean, vstdev, vmax
  real (pr) :: ll
  real (pr), allocatable, dimension(:) :: vx, vy, vz, pdf_vx, pdf_vy, pdf_vz, f
  character(len=512) :: fin, fout

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)

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
  allocate(pdf_vx(-nbins/2+1:nbins/2), pdf_vy(-nbins/2+1:nbins/2), pdf_vz(-nbins/2+1:nbins/2))
  allocate(f(0:nx1))
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

  psi%old = 0.0_pr
  psi%old(:,js:je,ks:ke) = psi%new
  call send_recv_z(psi%old)
  call pack_y(psi%old)
  call send_recv_y()
  call unpack_y(psi%old)

  call momentum(psi%old, mom)
  call get_pdf_velocity(psi%old, vx, vy, vz, vmean, vstdev)
  call get_pdf(vx, pdf_vx, vmax(1))
  call get_pdf(vy, pdf_vy, vmax(2))
  call get_pdf(vz, pdf_vz, vmax(3))
  call get_vcf(psi%old, f)
  ll = linelength(0.0_pr, psi%old)

  open (newunit=u, file=trim(fout), access='stream', status='replace')
  write (u) mom, vmean, vstdev, vmax
  write (u) size(vx), size(vy), size(vz)
  write (u) pdf_vx, pdf_vy, pdf_vz
  write (u) f
  write (u) ll
  close (u)

  call MPI_FINALIZE(ierr)
end program ref_diag
