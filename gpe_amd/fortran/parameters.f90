! Module state of the Fortran host, with the public names of the reference's
! module parameters (src/parameters.f90:15-131) that the time-stepping path
! reads.  The reference fixes the grid size and precision at compile time
! (src/parameters.in:22-29); here the grid is read at run time from the
! extra namelist grid_params, so one binary serves every workload.
module parameters
  use iso_c_binding, only : c_double
  implicit none
  save

  integer, parameter :: pr = c_double

  ! run_params (src/parameters.f90:25-40)
  real (pr)    :: tau = 0.01_pr
  real (pr)    :: end_time = 1.0_pr
  real (pr)    :: xr = 1.0_pr, yr = 1.0_pr, zr = 1.0_pr
  character(5) :: scheme = 'rk4'
  integer      :: eqn_to_solve = 1
  integer      :: bcs = 1
  integer      :: order = 2
  logical      :: restart = .true.
  logical      :: saved_restart = .false.
  logical      :: renorm = .false.
  logical      :: imprint_vl = .false.
  logical      :: stop_imag = .false.
  logical      :: real_time = .true.

  namelist /run_params/ tau, end_time, xr, yr, zr, scheme, eqn_to_solve, &
    bcs, order, restart, saved_restart, renorm, imprint_vl, stop_imag, &
    real_time

  ! eqn_params (src/parameters.f90:43-55)
  real (pr) :: Urhs = 0.0_pr
  real (pr) :: diss_amp = 0.0_pr
  real (pr) :: scal = 1.0_pr
  real (pr) :: nv = 0.0_pr
  real (pr) :: enerv = 0.0_pr
  real (pr) :: g = 1.0_pr
  real (pr) :: mu = 1.0_pr
  real (pr) :: nn = 1.0_pr
  real (pr) :: omx = 0.0_pr, omy = 0.0_pr, omz = 0.0_pr

  namelist /eqn_params/ Urhs, diss_amp, scal, nv, enerv, g, mu, nn, &
    omx, omy, omz

  ! io_params subset read by the path (save_rate, src/solve.f90:234)
  integer   :: save_rate = 10
  real (pr) :: save_rate2 = 1.0e30_pr, save_rate3 = 1.0e30_pr, p_save = 1.0e30_pr
  logical   :: save_contour = .false., save_3d = .false., save_filter = .false.
  real (pr) :: filter_kc = 0.0_pr
  logical   :: save_average = .false., save_spectrum = .false., save_pdf = .false.
  logical   :: save_vcf = .false., save_ll = .false., save_zeros = .false.

  namelist /io_params/ save_rate, save_rate2, save_rate3, p_save, &
    save_contour, save_3d, save_filter, filter_kc, save_average, &
    save_spectrum, save_pdf, save_vcf, save_ll, save_zeros

  ! misc_params subset: the velocity-PDF bin count (src/parameters.f90:77-86)
  integer   :: nbins = 256
  logical   :: diagnostic = .false., pp_filtered_surface = .false.
  integer   :: nlines = 3, nfilter = 1
  real (pr) :: fscale = 1.0_pr

  namelist /misc_params/ nbins, diagnostic, pp_filtered_surface, nlines, &
    nfilter, fscale

  ! grid_params: the compile-time values of the reference as run-time
  ! input, plus the device options of this build
  integer      :: nx = 64, ny = 64, nz = 64
  character(5) :: storage = 'c128'   ! 'c128', 'c64' or 'mixed' device fields
  real (pr)    :: vscale = 0.0_pr     ! V term in eqn 1 (0 = reference)
  character(8) :: host_mode = 'resident' ! 'resident' or 'parity' (copy per step)
  logical      :: save_mom = .false.  ! momentum.dat with the save_rate time series

  ! vortex line of imprint_vl: line_param of the reference's parameters.in
  ! (a compile-time constant there), here read with grid_params
  type :: line_param
    real (pr)    :: x0, y0, z0, amp1, amp2, ll, sgn
    character(1) :: dir
    logical      :: imprint_phase
  end type line_param
  type (line_param) :: vl1 = line_param(-1.5_pr, 0.0_pr, 0.0_pr, 0.0_pr, 0.0_pr, 0.0_pr, 1.0_pr, 'z', .true.)

  namelist /grid_params/ nx, ny, nz, storage, vscale, host_mode, save_mom, vl1

  ! Parameters for adaptive time stepping (src/parameters.f90:89-93)
  real (pr), parameter :: eps         = 1e-6_pr
  real (pr), parameter :: safety      = 0.9_pr
  real (pr), parameter :: dt_decrease = -0.25_pr
  real (pr), parameter :: dt_increase = -0.20_pr
  real (pr)            :: errcon

  ! Grid dimensions and the (single) local block: js:je, ks:ke cover the domain
  integer :: nx1, ny1, nz1
  integer :: js, je, ks, ke
  integer :: myrank = 0

  complex (pr) :: dt
  real (pr) :: t = 0.0_pr
  real (pr) :: im_t = 0.0_pr
  real (pr), parameter :: pi = 3.1415926535897932384626433832795_pr
  real (pr) :: xl, yl, zl
  real (pr) :: dx, dy, dz
  real (pr) :: dx2, dy2, dz2
  complex (pr), parameter :: eye = (0.0_pr,1.0_pr)
  integer :: p = 0

contains

  subroutine set_extents()
    nx1 = nx - 1
    ny1 = ny - 1
    nz1 = nz - 1
    js = 0
    je = ny1
    ks = 0
    ke = nz1
  end subroutine set_extents

end module parameters
