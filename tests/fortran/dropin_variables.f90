! Test program for the variables drop-ins (gpe_amd/fortran/variables.f90) in
! resident mode: the device state is put once; get_norm, energy, mass,
! momentum, renormalise and linelength are then handed a different host
! array (as the reference's save_energy(t, psi%old) hands a stale copy), and
! the state must come back unchanged.  Prints key/value lines that
! tests/test_dropins.py compares with the C ABI.
program dropin_variables
  use iso_c_binding, only : c_double
  use parameters
  use gpe_context, only : close_ctx
  use solve, only : state_put, state_get
  use variables
  implicit none
  complex (pr), allocatable :: psi(:,:,:), other(:,:,:), halo(:,:,:), back(:,:,:), ren(:,:,:)
  real (pr) :: e0, e1, n0, n1, eo, no, mo, ll, mom(3)
  integer :: i, j, k

  nx = 20; ny = 18; nz = 16
  call set_extents()
  xr = 5.0_pr; yr = 4.5_pr; zr = 4.0_pr
  order = 4; bcs = 1; eqn_to_solve = 1
  xl = -xr; yl = -yr; zl = -zr
  dx = (xr-xl)/nx1; dy = (yr-yl)/ny1; dz = (zr-zl)/nz1
  allocate(psi(0:nx1,js:je,ks:ke), other(0:nx1,js:je,ks:ke), back(0:nx1,js:je,ks:ke), ren(0:nx1,js:je,ks:ke))
  allocate(halo(0:nx1,js-2:je+2,ks-2:ke+2))
  do k = ks, ke
    do j = js, je
      do i = 0, nx1
        psi(i,j,k) = cmplx(1.0_pr + 0.3_pr*cos(0.7_pr*i + 0.2_pr*j), 0.2_pr*sin(0.5_pr*j - 0.3_pr*k), pr)
        other(i,j,k) = cmplx(0.5_pr*sin(0.4_pr*i) - 0.2_pr, 0.8_pr + 0.1_pr*cos(0.9_pr*k + 0.1_pr*i), pr)
      end do
    end do
  end do
  halo = 0.0_pr
  halo(:,js:je,ks:ke) = other

  call state_put(psi)
  call state_energy(e0)
  call state_norm(n0)
  call energy(halo, eo)
  call get_norm(other, no)
  call mass(other, mo)
  call momentum(halo, mom)
  ll = linelength(0.0_pr, halo)
  ren = other
  call renormalise(ren, no)
  call state_energy(e1)
  call state_norm(n1)
  call state_get(back)

  write (*, '(a,es25.17)') 'state_energy_before ', e0
  write (*, '(a,es25.17)') 'state_energy_after ', e1
  write (*, '(a,es25.17)') 'state_norm_before ', n0
  write (*, '(a,es25.17)') 'state_norm_after ', n1
  write (*, '(a,es25.17)') 'state_max_change ', maxval(abs(back - psi))
  write (*, '(a,es25.17)') 'other_energy ', eo
  write (*, '(a,es25.17)') 'other_norm ', no
  write (*, '(a,es25.17)') 'other_mass ', mo
  write (*, '(a,3es25.17)') 'other_momentum ', mom
  write (*, '(a,es25.17)') 'other_linelength ', ll
  write (*, '(a,es25.17)') 'renormalised_norm_sq ', sum(abs(ren)**2)
  write (*, '(a,es25.17)') 'other_norm_sq ', sum(abs(other)**2)
  call close_ctx()
end program dropin_variables
