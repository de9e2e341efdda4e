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
  public :: routec_fock_jk, routec_grad2, routec_vxc
  public :: oqph_session_create, oqph_session_destroy, oqph_session_set_basis, &
            oqph_session_set_geometry, oqph_session_set_thresholds, oqph_set_default_session, &
            oqph_session_set_precision, oqph_session_set_share, oqph_session_prepare, &
            oqph_fock_jk, oqph_fock_jk_inc, oqph_fock_jk_cam, oqph_grad2, &
            oqph_session_set_grad_precision, oqph_mrsf_int2, oqph_grad2_mrsf, &
            oqph_session_set_xc, oqph_vxc
  public :: oqph_stats

  !> Build statistics (oqph_stats of include/oqph.h).
  type, bind(C) :: oqph_stats
    integer(c_long_long) :: quartets_computed, quartets_screened, quartets_fp32
    real(c_double) :: kernel_ms, model_flops
    integer(c_long_long) :: prim_quartets
  end type

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
    !> policy 0 = FP64, 1 = mixed (FP32 below fp32_cutoff)
    integer(c_int) function oqph_session_set_precision(s, policy, fp32_cutoff) &
        bind(C, name="oqph_session_set_precision")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: s
      integer(c_int), value :: policy
      real(c_double), value :: fp32_cutoff
    end function
    !> 2e-gradient precision: 0 = FP64, 1 = quartets with Q_ij Q_kl below
    !> fp32_cutoff in FP32 (no counterpart in OpenQP)
    integer(c_int) function oqph_session_set_grad_precision(s, policy, fp32_cutoff) &
        bind(C, name="oqph_session_set_grad_precision")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: s
      integer(c_int), value :: policy
      real(c_double), value :: fp32_cutoff
    end function
    !> MRSF response integral pass (int2_driver%run on int2_mrsf_data_t,
    !> tdhf_mrsf_energy.F90:742-752): d3, f3 as the reference's (nf, nmat, nbf, nbf)
    integer(c_int) function oqph_mrsf_int2(s, nf, nmat, ncoul, d3, f3, se, sc, stats) &
        bind(C, name="oqph_mrsf_int2")
      import :: c_ptr, c_int, c_double, oqph_stats
      type(c_ptr), value :: s
      integer(c_int), value :: nf, nmat, ncoul
      real(c_double), intent(in) :: d3(*)
      real(c_double), intent(out) :: f3(*)
      real(c_double), value :: se, sc
      type(oqph_stats), intent(out) :: stats
    end function
    !> MRSF 2e gradient (routec_grad2_mrsf on session s): d, p (nbf, nbf, 2),
    !> spc (7, nbf, nbf), spcscale(3)
    integer(c_int) function oqph_grad2_mrsf(s, d, p, spc, de, hfscale, hfscale2, coulscale, spcscale, mrst, stats) &
        bind(C, name="oqph_grad2_mrsf")
      import :: c_ptr, c_int, c_double, oqph_stats
      type(c_ptr), value :: s
      real(c_double), intent(in) :: d(*), p(*), spc(*), spcscale(*)
      real(c_double), intent(inout) :: de(*)
      real(c_double), value :: hfscale, hfscale2, coulscale
      integer(c_int), value :: mrst
      type(oqph_stats), intent(out) :: stats
    end function
    !> OpenQP MPI share: rank of size (int2.F90:759-761)
    integer(c_int) function oqph_session_set_share(s, rank, size) bind(C, name="oqph_session_set_share")
      import :: c_ptr, c_int
      type(c_ptr), value :: s
      integer(c_int), value :: rank, size
    end function
    integer(c_int) function oqph_session_prepare(s) bind(C, name="oqph_session_prepare")
      import :: c_ptr, c_int
      type(c_ptr), value :: s
    end function
    integer(c_int) function oqph_fock_jk(s, d, f, nfocks, se, sc, stats) bind(C, name="oqph_fock_jk")
      import :: c_ptr, c_int, c_double, oqph_stats
      type(c_ptr), value :: s
      real(c_double), intent(in) :: d(*)
      real(c_double), intent(inout) :: f(*)
      integer(c_int), value :: nfocks
      real(c_double), value :: se, sc
      type(oqph_stats), intent(out) :: stats
    end function
    !> incremental build (scf_addons.F90:1207-1210): f = G[d - d_old] + f_old, then
    !> f_old <- f and d_old <- d
    integer(c_int) function oqph_fock_jk_inc(s, d, d_old, f, f_old, nfocks, se, sc, stats) &
        bind(C, name="oqph_fock_jk_inc")
      import :: c_ptr, c_int, c_double, oqph_stats
      type(c_ptr), value :: s
      real(c_double), intent(in) :: d(*)
      real(c_double), intent(inout) :: d_old(*), f(*), f_old(*)
      integer(c_int), value :: nfocks
      real(c_double), value :: se, sc
      type(oqph_stats), intent(out) :: stats
    end function
    !> range-separated build (int2_run_cam, int2.F90:538-584)
    integer(c_int) function oqph_fock_jk_cam(s, d, f, nfocks, alpha, beta, mu, alpha_coulomb, beta_coulomb, stats) &
        bind(C, name="oqph_fock_jk_cam")
      import :: c_ptr, c_int, c_double, oqph_stats
      type(c_ptr), value :: s
      real(c_double), intent(in) :: d(*)
      real(c_double), intent(inout) :: f(*)
      integer(c_int), value :: nfocks
      real(c_double), value :: alpha, beta, mu, alpha_coulomb, beta_coulomb
      type(oqph_stats), intent(out) :: stats
    end function
    integer(c_int) function oqph_grad2(s, d, de, hfscale, coulscale, stats) bind(C, name="oqph_grad2")
      import :: c_ptr, c_int, c_double, oqph_stats
      type(c_ptr), value :: s
      real(c_double), intent(in) :: d(*)
      real(c_double), intent(inout) :: de(*)
      real(c_double), value :: hfscale, coulscale
      type(oqph_stats), intent(out) :: stats
    end function
    !> Semilocal Vxc (routec_bridge.F90:43-50), by reference like the seam.
    subroutine routec_vxc(d, fxc, nbf, nfocks, eexc, totele, info) bind(C, name="routec_vxc")
      import :: c_double, c_int
      real(c_double), intent(in) :: d(*)
      real(c_double), intent(inout) :: fxc(*)
      integer(c_int), intent(in) :: nbf, nfocks
      real(c_double), intent(inout) :: eexc, totele
      integer(c_int), intent(out) :: info
    end subroutine
    integer(c_int) function oqph_session_set_xc(s, functional, nrad, ntheta) bind(C, name="oqph_session_set_xc")
      import :: c_ptr, c_int, c_char
      type(c_ptr), value :: s
      character(kind=c_char), intent(in) :: functional(*)  ! NUL-terminated
      integer(c_int), value :: nrad, ntheta
    end function
    integer(c_int) function oqph_vxc(s, d, nfocks, fxc, exc, nel) bind(C, name="oqph_vxc")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: s
      real(c_double), intent(in) :: d(*)
      integer(c_int), value :: nfocks
      real(c_double), intent(inout) :: fxc(*)
      real(c_double), intent(out) :: exc, nel
    end function
  end interface
end module oqph_bridge
