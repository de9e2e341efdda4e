/*
 * oqph.h -- C ABI of liboqph, the MI355X (gfx950) direct-SCF two-electron
 * Fock builder (J/K) and two-electron nuclear gradient.
 *
 * Drop-in boundary: OpenQP's Route-C seam (source/modules/routec_bridge.F90)
 * dlopens $OQP_ROUTEC_LIB / $OQP_ROUTEC_GRAD_LIB and binds routec_fock_jk and
 * routec_grad2 by name; liboqph.so exports both with the seam's exact
 * by-reference signatures.  The seam carries no basis, geometry or
 * thresholds, so a session (oqph_session_*) supplies them out of band; the
 * routec_* entry points use the process-default session.
 *
 * Conventions: AO order, normalisation and packing are OpenQP's
 * (source/constants.F90:34-60, source/basis_tools.F90:277-303,
 * source/integrals/int2.F90:1436-1447): packed lower triangle, row-packed,
 * ij = i(i-1)/2 + j (1-based).  Densities are total RHF densities.
 * Lengths in bohr, energies in hartree.
 *
 * Errors: every entry point returns (or writes to *info) 0 on success and a
 * positive code when it declines: 1 unsupported request or no session,
 * 2 device error, 3 out of memory, 4 invalid argument.  Nothing aborts the
 * process; oqph_last_error() describes the last failure.
 *
 * Threading: a session is not re-entrant; distinct sessions may be used from
 * distinct threads.  Each session owns one HIP stream on its device.
 */
#ifndef OQPH_H
#define OQPH_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Route-C drop-ins ------------------------------------------------ */

/* J/K Fock build.  Replaces routec_fock_jk (source/modules/routec_bridge.F90:33-40,
 * caller source/scf_addons.F90:1138-1146).  d, f: packed (ntri, nfocks); f is
 * returned ready to use, i.e. f = sc*J[D] - 0.5*se*K[D] exactly as fock_jk's
 * native path leaves it after its scaling (source/scf_addons.F90:1178-1185).
 * nfocks = 1 (RHF, d the total density) or 2 (UHF/ROHF, d and f hold the
 * alpha then the beta matrices, int2_urohf_data_t semantics, int2.F90:1488-1578);
 * other values decline with info = 1. */
void routec_fock_jk(const double *d, double *f, const int *nbf, const int *nfocks, const double *se,
                    const double *sc, int *info);

/* 2e nuclear gradient.  Replaces routec_grad2 (source/modules/routec_bridge.F90:51-61,
 * caller source/modules/hf_gradient.F90:274-290).  d: packed total density;
 * xyz: (3, natm) bohr; de: (3, natm) receives dE2e/dR = Coulomb*coulscale +
 * exchange*hfscale at fixed density. */
void routec_grad2(const double *d, const double *xyz, double *de, const int *nbf, const int *natm,
                  const double *hfscale, const double *coulscale, int *info);
/* MRSF 2e gradient, the reference's routec_grad2_mrsf_i (routec_bridge.F90:61-75,
 * called from mrsf_2e_grad, tdhf_mrsf_gradient.F90:281-296): d and p are the
 * raw alpha/beta SCF and relaxed difference densities (nbf, nbf, 2), spc the
 * seven MRSF response densities (7, nbf, nbf), Fortran order; de (3, natm)
 * receives the gradient of grd2_mrsf_compute_data_t's two-particle density
 * (tdhf_mrsf_gradient.F90:404-601); spcscale = (CO-CO, OV-OV, CO-OV) spin-pair
 * scales; mrst 1 singlet / 3 triplet.  Uses the default session. */
void routec_grad2_mrsf(const double *d, const double *p, const double *spc, const double *xyz, double *de,
                       const int *nbf, const int *natm, const double *hfscale, const double *hfscale2,
                       const double *coulscale, const double *spcscale, const int *mrst, int *info);

/* ---- Session (no counterpart in the seam, which passes no basis) ------- */

typedef struct oqph_session oqph_session;

typedef struct {
  long long quartets_computed;   /* shell quartets evaluated on the device          */
  long long quartets_screened;   /* shell quartets skipped by the density screen    */
  long long quartets_fp32;       /* of the computed ones, evaluated in FP32          */
  double kernel_ms;              /* device time of the build (HIP events)             */
  double model_flops;            /* algorithmic FLOPs of the build (Rys model)        */
  long long prim_quartets;       /* primitive quartets that passed the screens        */
} oqph_stats;

/* Create a session on HIP device `device`.  Fields mirror oqp_get_basis
 * (source/c_interop.F90:237-280). */
int oqph_session_create(oqph_session **s, int device);
int oqph_session_destroy(oqph_session *s);

/* Basis: nsh shells; am (0..3: s, p, d, f; larger declines with 1), harmonic
 * (1 = pure for l >= 2), atom (0-based), nprim per shell; ex and cc
 * concatenated per shell in shell order, cc carrying OpenQP's primitive
 * normalisation (basis_tools.F90:277-303). */
int oqph_session_set_basis(oqph_session *s, int nsh, const int *am, const int *harmonic, const int *atom,
                           const int *nprim, const double *ex, const double *cc);
/* Geometry: natm atoms, xyz (3, natm) bohr, nuclear charges. */
int oqph_session_set_geometry(oqph_session *s, int natm, const double *xyz, const double *zn);
/* int2e_cutoff (default 5e-11, source/types.F90:121) and the gradient cutoff
 * (default 1e-10, pyoqp perf_levels.py:37). */
int oqph_session_set_thresholds(oqph_session *s, double int2e_cutoff, double grad_cutoff);
/* policy 0 = FP64 everywhere; 1 = mixed: quartets whose bound
 * Q_ij Q_kl max|D| is below fp32_cutoff run in FP32 (FP64 accumulation). */
int oqph_session_set_precision(oqph_session *s, int policy, double fp32_cutoff);
/* Precision of the 2e gradient (no counterpart in the reference, which works
 * in FP64; for callers that accept an error of about 1e-7 Eh/bohr for speed):
 * policy 0 = FP64 (default); 1 = mixed: quartet classes on the class templates
 * run the quartets with Q_ij Q_kl below fp32_cutoff in FP32 arithmetic (FP64
 * screens and per-atom sums). */
int oqph_session_set_grad_precision(oqph_session *s, int policy, double fp32_cutoff);
/* MRSF response integral pass: replaces int2_driver%run on int2_mrsf_data_t
 * (tdhf_mrsf_energy.F90:742-752; digestion tdhf_mrsf_lib.F90:278-308, pass 1,
 * FP64, Tamm-Dancoff).  d3 and f3 are the reference's d3(nf, nmat, nbf, nbf)
 * and f3(nf, nmat, nbf, nbf) arrays (Fortran order; MRSF: nmat = 7,
 * ncoul = 4): f3(v,c,:,:) = sc J[M] (c < ncoul) - se K[M] for
 * M = d3(v,c,:,:), J[M]_ij = sum_kl (ij|kl) M_kl, K[M]_ik = sum_jl (ij|kl)
 * M_jl, screened on the last component as shell_den_screen_mrsf
 * (tdhf_mrsf_lib.F90:176-204).  f3 is overwritten. */
/* The MRSF 2e gradient of routec_grad2_mrsf on session s (same arguments). */
int oqph_grad2_mrsf(oqph_session *s, const double *d, const double *p, const double *spc, double *de,
                    double hfscale, double hfscale2, double coulscale, const double *spcscale, int mrst,
                    oqph_stats *stats);
int oqph_mrsf_int2(oqph_session *s, int nf, int nmat, int ncoul, const double *d3, double *f3, double se,
                   double sc, oqph_stats *stats);
/* MRSF sigma session, the reference's routec_sig ABI (routec_sig.F90:27-57,
 * bound by routec_sig_load :128-210): (A - B) X in the MO frame for the
 * singlet (kind 1) / triplet (kind 3) MRSF response of two open shells
 * (nocca = noccb + 2), replacing mrsfcbc -> int2 (int2_mrsf_data_t) ->
 * mrsfmntoia + mrsfesum (tdhf_mrsf_energy.F90:686-790).  init takes the MO
 * coefficients and MO Fock matrices (nbf x nbf, column-major) and uses the
 * default session's basis and geometry; 0 = ready.  iter: bvec_mo and
 * sigma_mo are (nocca (nbf - noccb)) x nv_new column-major; info 0 = done. */
int routec_sig_init(const int *nbf, const double *mo_a, const double *mo_b, const double *fmo_a,
                    const double *fmo_b, const int *nocca, const int *noccb, const int *kind);
void routec_sig_set_scale(const double *s);
void routec_sig_iter(const double *bvec_mo, const int *nv_new, double *sigma_mo, int *info);
void routec_sig_free(void);
/* ---- semilocal exchange-correlation (Route-C Vxc seam) ----------------- */

/* Vxc build.  Replaces routec_vxc (source/modules/routec_bridge.F90:43-50,
 * caller calc_jk_xc, source/scf_addons.F90:2104-2111, in place of
 * calc_dft_xc_density, :1857-1942).  d: packed (ntri, nfocks), the total
 * density (nfocks = 1) or the alpha then the beta density (nfocks = 2);
 * fxc: packed Vxc per spin, overwritten; eexc: E_xc; totele: the grid
 * electron count.  Uses the default session's functional and grid
 * (oqph_session_set_xc); info = 1 (decline) when there is none, when nbf
 * does not match or on a device error. */
void routec_vxc(const double *d, double *fxc, const int *nbf, const int *nfocks, double *eexc, double *totele,
                int *info);
/* Functional and grid of the session's Vxc builds: names as the reference's
 * libxc table (source/dftlib/libxc.F90:300-460): slater, svwn5, blyp,
 * b3lyp5 (= b3lypv5), b3lyp (= b3lypv1r, VWN RPA fit), bhhlyp; grid nrad
 * radial shells x (ntheta x 2 ntheta) angular points per atom.  1: unknown
 * name; 4: grid out of range. */
int oqph_session_set_xc(oqph_session *s, const char *functional, int nrad, int ntheta);
/* Exact-exchange fraction and GGA flag of a functional name (1: unknown). */
int oqph_xc_functional_info(const char *functional, double *hfex, int *gga);
/* Session form of routec_vxc (same arguments, host buffers, return code). */
int oqph_vxc(oqph_session *s, const double *d, int nfocks, double *fxc, double *exc, double *nel);
/* The session's XC grid: returns the point count (negative on error) and
 * copies up to cap points as (x, y, z, w) into pts (may be NULL). */
long long oqph_xc_grid(oqph_session *s, double *pts, long long cap);
/* Work of one Vxc build on the session's grid: out[0] points, [1] batches,
 * [2] sum over batches of points x basis functions, [3] of points x basis
 * functions^2 (the two GEMMs per spin take 4 out[3] flops), [4] the largest
 * batch's basis-function count. */
int oqph_xc_grid_stats(oqph_session *s, long long *out);

/* Make s the session used by routec_fock_jk / routec_grad2. */
int oqph_set_default_session(oqph_session *s);
/* Build the geometry-dependent device state now (pair records, Schwarz
 * bounds, the rank share's ownership flags); otherwise the first build does
 * it.  Counterpart of
 * int2_compute_t%init + set_screening (source/integrals/int2.F90:245-289, 473-477). */
int oqph_session_prepare(oqph_session *s);
/* Restrict every following J/K build to the quartets OpenQP's MPI rank `rank`
 * of `size` computes: those whose first pair in OpenQP's loop order has
 * mod(ij_pair, size) == rank, ij_pair its 1-based position in the cost-sorted
 * shell-pair map (source/integrals/int2.F90:759-761, 864-921).  Summing the
 * results of all ranks gives the full build.  size = 1: everything. */
int oqph_session_set_share(oqph_session *s, int rank, int size);

/* Host-buffer builds (copy in, build, copy out). */
int oqph_fock_jk(oqph_session *s, const double *d, double *f, int nfocks, double se, double sc,
                 oqph_stats *stats);
/* Incremental build (source/scf_addons.F90:1207-1210, 2036-2042):
 * f = G[d - d_old] + f_old; then f_old <- f and d_old <- d. */
int oqph_fock_jk_inc(oqph_session *s, const double *d, double *d_old, double *f, double *f_old, int nfocks,
                     double se, double sc, oqph_stats *stats);
int oqph_grad2(oqph_session *s, const double *d, double *de, double hfscale, double coulscale,
               oqph_stats *stats);

/* Range-separated (CAM) build, the two passes of int2_run_cam
 * (source/integrals/int2.F90:538-584) as fock_jk runs them for a CAM
 * functional (scf_addons.F90:1162-1166, scale_exch ignored there):
 * f = G[1/r; coulomb alpha_coulomb, exchange alpha] +
 *     G[erf(mu r)/r; coulomb beta_coulomb, exchange beta],
 * each G with fock_jk's packed scaling; the attenuated pass screens with the
 * attenuated Schwarz bounds (int2.F90:675-681).  fock_jk passes
 * alpha_coulomb = 1, beta_coulomb = 0.  nfocks 1 or 2 as oqph_fock_jk. */
int oqph_fock_jk_cam(oqph_session *s, const double *d, double *f, int nfocks, double alpha, double beta, double mu,
                     double alpha_coulomb, double beta_coulomb, oqph_stats *stats);

/* Device-resident builds: d_dev, f_dev are device pointers (ntri doubles) on
 * the session's device; stream 0 means the session stream.  f_dev is
 * overwritten.  Inputs already in HBM: no host transfer in the call. */
int oqph_fock_jk_device(oqph_session *s, const double *d_dev, double *f_dev, int nfocks, double se, double sc,
                        oqph_stats *stats, void *stream);

/* No-allocation RHF build.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  Same result as
 * oqph_fock_jk_device with nfocks = 1; d_dev, f_dev are device pointers
 * supplied by the caller.  The build's own device memory -- shell-pair
 * density maxima, statistics counters, work lists -- lives in a static array
 * of OQPH_NOALLOC_SCRATCH_BYTES that each GPU the library runs on holds
 * until the process ends; a basis/density that needs more declines with
 * code 1 and a message giving the size.  The call performs no hipMalloc,
 * hipFree or other device allocation: the session must have been prepared
 * first (oqph_session_prepare, which does allocate; also after a change of
 * basis, geometry, thresholds or share).  All calls on one GPU share its
 * array: calls that overlap on one GPU (other streams, other host threads'
 * kernels still running) race; the library serialises its own host side. */
#define OQPH_NOALLOC_SCRATCH_BYTES (64ull << 20)
int oqph_fock_jk_noalloc(oqph_session *s, const double *d_dev, double *f_dev, double se, double sc,
                         oqph_stats *stats, void *stream);

/* Per-launch profile of the next builds (HIP events around each class
 * launch); oqph_last_profile copies up to maxn entries of the last build:
 * class code la*1000+lb*100+lc*10+ld (plus 10000 when the launch used the
 * bra-row kernel k_jkb rather than k_jk, 20000 for the run-time-class kernel
 * k_jkh of the classes with an f shell, 30000 for the grouped bra-row kernel
 * k_jkg, 60000 for its tiled form; a multi-class launch of k_jkh reports
 * 40000 + its number of classes), device ms, model FLOPs, computed
 * quartets, primitive quartets.  Returns the number of launches.
 * No counterpart in the seam: measurement hooks for benchmarks. */
int oqph_session_set_profiling(oqph_session *s, int on);
int oqph_last_profile(oqph_session *s, int maxn, int *cls, double *ms, double *flops, long long *quartets,
                      long long *prims);
/* The raw statistics counters of the last build, 8 per launch in launch
 * order: computed quartets, quartets skipped by the density screen, FP32
 * quartets, model FLOPs, primitive quartets (segmented count), primitive
 * quartets evaluated (the grouped kernel evaluates a chunk's shared primitive
 * quartet once for all members; 0 for the other kernels), ket lanes visited
 * (64 per ket iteration of a wavefront: computed + screened quartets over it
 * is the lane occupancy of the screening loop), 0.  Returns the number of
 * launches copied. */
int oqph_last_counters(oqph_session *s, int maxn, unsigned long long *ctr);

/* Schwarz table (nsh*nsh, symmetric) of the current geometry. */
int oqph_schwarz(oqph_session *s, double *q);

/* One-electron integrals S, T, V (full nbf x nbf, row-major, AO
 * representation as above) for the host SCF loop and test densities;
 * any of S, T, V may be NULL.  Host code (OpenMP); the counterpart of
 * OpenQP's omp_hst (source/integrals/int1.F90:147).  l <= 3. */
int oqph_int1e(int nsh, const int *am, const int *harmonic, const int *atom, const int *nprim, const double *ex,
               const double *cc, int natm, const double *xyz, const double *zn, double *S, double *T, double *V);

/* One-electron part of the RHF nuclear gradient, added to grad[3*natm]
 * (Eh/bohr): sum D dH/dR - sum W dS/dR with D the total density and W the
 * energy-weighted density 2 C_occ diag(eps_occ) C_occ^T, both full
 * nbf x nbf in the AO order above; H = T + V (kinetic, Hellmann-Feynman and
 * basis-function terms of V).  Replaces grad_ee_overlap, grad_ee_kinetic,
 * grad_en_hellman_feynman and grad_en_pulay called from hf_1e_grad
 * (source/modules/hf_gradient.F90:143-224; source/integrals/grd1.F90).
 * Host code (OpenMP), l <= 3; the nuclear repulsion term is the caller's. */
int oqph_grad1e(int nsh, const int *am, const int *harmonic, const int *atom, const int *nprim, const double *ex,
                const double *cc, int natm, const double *xyz, const double *zn, const double *D, const double *W,
                double *grad);

/* Basis dimensions of the session. */
int oqph_session_nbf(oqph_session *s);

/* Compiled resources of one class-templated kernel, read from the code
 * object without launching it: kind 0 = J/K wavefront-task kernel, 1 = J/K
 * bra-row kernel, 2 = 2e-gradient kernel, 3 = Schwarz kernel (class (la lb|la lb));
 * out[4] = per-lane private segment bytes, static LDS bytes, registers, largest
 * block.  Returns 1 when the class has no such kernel.  Classes whose kernel
 * needs more than 8704 bytes of private segment per lane run on the
 * run-time-class kernels instead.  No counterpart in the reference. */
int oqph_kernel_resources(int kind, int la, int lb, int lc, int ld, int pure, int *out);

const char *oqph_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
