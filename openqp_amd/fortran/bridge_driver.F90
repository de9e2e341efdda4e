!> Minimal Fortran host: one water molecule in a cc-pVDZ-like single-primitive
!> basis handed to liboqph, then a J/K build through routec_fock_jk, an
!> incremental build through oqph_fock_jk_inc (from a first density to the
!> second) and a mixed-precision build.  Prints info codes and every element
!> of the results; run on a GPU node by the tests.
program bridge_driver
  use, intrinsic :: iso_c_binding
  use oqph_bridge
  implicit none
  type(c_ptr) :: s
  integer(c_int) :: rc, info, nbf, nfocks
  integer(c_int) :: am(4) = [0, 0, 1, 0], harm(4) = [0, 0, 0, 0], atom(4) = [0, 0, 0, 1], nprim(4) = [1, 1, 1, 1]
  real(c_double) :: ex(4) = [10.0d0, 0.8d0, 0.9d0, 0.5d0], cc(4)
  real(c_double) :: xyz(6) = [0d0, 0d0, 0d0, 0d0, 0d0, 1.4d0], zn(2) = [8d0, 1d0]
  real(c_double) :: d(21), f(21), se, sc, d_old(21), f_old(21), d1(21), finc(21), fmx(21)
  type(oqph_stats) :: st
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
  do i = 1, 21
    print '(a,i0,a,es24.16)', 'f(', i, ')=', f(i)
  end do
  ! incremental: first build at d1, then at d, as fock_jk with f_old does
  do i = 1, 21
    d1(i) = 0.02d0 * mod(i, 5)
  end do
  d_old = 0d0
  f_old = 0d0
  rc = oqph_fock_jk_inc(s, d1, d_old, finc, f_old, 1_c_int, 1d0, 1d0, st)
  rc = rc + oqph_fock_jk_inc(s, d, d_old, finc, f_old, 1_c_int, 1d0, 1d0, st)
  print '(a,i0)', 'inc_rc=', rc
  do i = 1, 21
    print '(a,i0,a,es24.16)', 'finc(', i, ')=', finc(i)
  end do
  rc = oqph_session_set_precision(s, 1_c_int, 1d-2)
  rc = rc + oqph_fock_jk(s, d, fmx, 1_c_int, 1d0, 1d0, st)
  print '(a,i0,a,i0)', 'mixed_rc=', rc, ' quartets=', st%quartets_computed
  do i = 1, 21
    print '(a,i0,a,es24.16)', 'fmx(', i, ')=', fmx(i)
  end do
  rc = oqph_session_destroy(s)
end program bridge_driver
