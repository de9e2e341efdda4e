!> iso_c_binding interface to liboqph (include/oqph.h) for Fortran hosts.
!>
!> The Route-C entry points keep the by-reference signatures of OpenQP's seam
!> (source/modules/routec_bridge.F90:33-61); the session calls carry the basis
!> and geometry the seam does not pass.  Arrays are passed by reference, scalar
!> arguments of the session API by value.
module oqph_bridge
  use, intrinsic :: iso_c_binding
  implicit none
  private
  public :: routec_fock_jk, routec_grad2
  public :: oqph_session_create, oqph_session_destroy, oqph_session_set_basis, &
            oqph_session_set_geometry, oqph_session_set_thresholds, oqph_set_default_session

  interface
    subroutine routec_fock_jk(d, f, nbf, nfocks, se, sc, info) bind(C, name="routec_fock_jk")
      import :: c_double, c_int
      real(c_double), intent(in)    :: d(*)
      real(c_double), intent(inout) :: f(*)
      integer(c_int), intent(in)    :: nbf, nfocks
      real(c_double), intent(in)    :: se, sc
      integer(c_int), intent(out)   :: info
    end subroutine
    subroutine routec_grad2(d, xyz, de, nbf, natm, hfscale, coulscale, info) bind(C, name="routec_grad2")
      import :: c_double, c_int
      real(c_double), intent(in)    :: d(*), xyz(*)
      real(c_double), intent(inout) :: de(*)
      integer(c_int), intent(in)    :: nbf, natm
      real(c_double), intent(in)    :: hfscale, coulscale
      integer(c_int), intent(out)   :: info
    end subroutine
    integer(c_int) function oqph_session_create(s, device) bind(C, name="oqph_session_create")
      import :: c_ptr, c_int
      type(c_ptr), intent(out) :: s
      integer(c_int), value :: device
    end function
    integer(c_int) function oqph_session_destroy(s) bind(C, name="oqph_session_destroy")
      import :: c_ptr, c_int
      type(c_ptr), value :: s
    end function
    integer(c_int) function oqph_session_set_basis(s, nsh, am, harmonic, atom, nprim, ex, cc) &
        bind(C, name="oqph_session_set_basis")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: s
      integer(c_int), value :: nsh
      integer(c_int), intent(in) :: am(*), harmonic(*), atom(*), nprim(*)
      real(c_double), intent(in) :: ex(*), cc(*)
    end function
    integer(c_int) function oqph_session_set_geometry(s, natm, xyz, zn) bind(C, name="oqph_session_set_geometry")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: s
      integer(c_int), value :: natm
      real(c_double), intent(in) :: xyz(*), zn(*)
    end function
    integer(c_int) function oqph_session_set_thresholds(s, int2e_cutoff, grad_cutoff) &
        bind(C, name="oqph_session_set_thresholds")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: s
      real(c_double), value :: int2e_cutoff, grad_cutoff
    end function
    integer(c_int) function oqph_set_default_session(s) bind(C, name="oqph_set_default_session")
      import :: c_ptr, c_int
      type(c_ptr), value :: s
    end function
  end interface
end module oqph_bridge
