!> Minimal Fortran host: one water molecule in a cc-pVDZ-like single-primitive
!> basis handed to liboqph, then one J/K build through routec_fock_jk.  Prints
!> info and a checksum of f; run on a GPU node by the tests.
program bridge_driver
  use, intrinsic :: iso_c_binding
  use oqph_bridge
  implicit none
  type(c_ptr) :: s
  integer(c_int) :: rc, info, nbf, nfocks
  integer(c_int) :: am(4) = [0, 0, 1, 0], harm(4) = [0, 0, 0, 0], atom(4) = [0, 0, 0, 1], nprim(4) = [1, 1, 1, 1]
  real(c_double) :: ex(4) = [10.0d0, 0.8d0, 0.9d0, 0.5d0], cc(4)
  real(c_double) :: xyz(6) = [0d0, 0d0, 0d0, 0d0, 0d0, 1.4d0], zn(2) = [8d0, 1d0]
  real(c_double) :: d(21), f(21), se, sc
  integer :: i
  ! primitive normalisation of x^l Gaussians
  cc(1) = (2d0*ex(1)/acos(-1d0))**0.75d0
  cc(2) = (2d0*ex(2)/acos(-1d0))**0.75d0
  cc(3) = (2d0*ex(3)/acos(-1d0))**0.75d0 * sqrt(4d0*ex(3))
  cc(4) = (2d0*ex(4)/acos(-1d0))**0.75d0
  rc = oqph_session_create(s, 0_c_int)
  if (rc /= 0) then
    print '(a,i0)', 'session_create rc=', rc
    stop 1
  end if
  rc = oqph_session_set_basis(s, 4_c_int, am, harm, atom, nprim, ex, cc)
  rc = rc + oqph_session_set_geometry(s, 2_c_int, xyz, zn)
  rc = rc + oqph_set_default_session(s)
  nbf = 6
  nfocks = 1
  se = 1d0
  sc = 1d0
  do i = 1, 21
    d(i) = 0.01d0 * i
  end do
  f = 0d0
  call routec_fock_jk(d, f, nbf, nfocks, se, sc, info)
  print '(a,i0,a,es24.16)', 'info=', info, ' sumf=', sum(f)
  rc = oqph_session_destroy(s)
end program bridge_driver
