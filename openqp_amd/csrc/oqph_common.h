// oqph_common.h -- data layout shared by the host session and the device kernels.
//
// Device-resident layout (all in HBM, built once per basis+geometry):
//   ShellPair  pairs[npair]      one record per significant shell pair, the
//                                 higher-L shell first (swap recorded);
//   PrimJK     prims[nprimpair]  primitive pairs of every shell pair, 48 B each
//                                 (PrimPair, 64 B, for the gradient), screened
//                                 as OpenQP screens them
//                                 (source/integrals/int2_pairs.F90:179-266);
//   double     rys[...]           Rys node/weight fit table (rys_table.inc).
// Per build:
//   double     D[ntri]            packed lower-triangular density, row-packed;
//   double     F[ntri]            packed result (atomics);
//   double     dsh[nsh*nsh]       shell-pair max |D| (OpenQP shlden).
#pragma once
#include <stdint.h>

#define OQPH_LMAX 3            // largest shell angular momentum handled on the device
#define OQPH_MAXNAO 10         // Cartesian components of an f shell

struct ShellPair {
  double A[3];                 // centre of shell a (the higher-L shell of the pair)
  double B[3];                 // centre of shell b
  double Q;                    // Schwarz bound sqrt(max|(ab|ab)|)
  int sa, sb;                  // shell indices (0-based) of a and b
  int ao_a, ao_b;              // AO offsets of a and b
  int prim0, nprim;            // primitive-pair range in prims[]
  int si, sj;                  // reference order: si >= sj (shell indices)
  int la, lb;                  // angular momenta of a, b (la >= lb)
};

struct PrimPair {              // 64 bytes: one cache line per primitive pair
  double g;                    // zeta = alpha + beta
  double Kg;                   // K / zeta, K = sqrt(2) pi^(5/4) c_a c_b exp(-alpha beta/zeta AB^2)
  double P[3];                 // Gaussian product centre
  double a, b;                 // exponents of shell a and shell b
  double h;                    // 1 / (2 zeta)
};

// The fields of a primitive pair that the J/K quadrature reads (48 B): the
// J/K and Schwarz kernels stream these; the gradient keeps PrimPair, whose
// exponents it needs for the derivatives.  The product centre is stored
// relative to the pair's first centre A (X = P - A), so the quadrature forms
// PA = X, QC = X' and PQ = (A - C) + X - X' from small differences and never
// from absolute coordinates: results do not depend on where the molecule sits.
struct PrimJK {
  double g;                    // zeta
  double Kg;                   // K / zeta
  double X[3];                 // P - A
  double h;                    // 1 / (2 zeta)
};

// The same fields in FP32 (24 B), read by the FP32 quadrature of the mixed
// mode: half the bytes of a PrimJK per primitive pair streamed.  The
// differences X = P - A are formed in FP64 and then rounded.
struct PrimF32 {
  float g, Kg, X[3], h;
};

// Wavefront tasks of one class-pair launch (J/K build or gradient): segment
// s pairs bra pair seg_bra[s] with the kets ket[seg_k0[s] .. seg_k0[s] +
// seg_n[s]); it is split into ceil(seg_n/64) chunks of 64 kets, one per wavefront, counted by
// the prefix choff (nseg + 1).  Empty segments are not stored.
struct WaveWork {
  const int* seg_bra;
  const int* seg_k0;
  const int* seg_n;
  const long long* choff;
  const int* ket;
  int nseg;
  long long nwaves;
  // run-time-class kernels: kets of a 64-ket chunk per block (64, 32, 16 or
  // 8; 0 means 64): a launch takes nwaves * 64 / kpb blocks
  int kpb;
  // J/K class kernels: the ket class's block of the ket-side Coulomb buffer
  // (JKParams::Jk) starts at jbase; ket pair id iq owns the OC*OD entries
  // from jbase + (iq - kofs) * OC*OD
  int jbase, kofs;
};

// One quartet class (la lb|lc ld), la >= lb, lc >= ld, for the kernels that
// take the class at run time (classes without a template): extents of every
// intermediate and the LDS layout (offsets in doubles), filled by hl_dims().
struct HLDims {
  int la, lb, lc, ld, nr, pure;
  int ni, nk;                  // 2D-integral table extents per root and axis
  int ne, nf;                  // contracted (e0|f0) extents
  int na, nb, nc, nd;          // Cartesian components
  int oa, ob, oc, od;          // AO components
  int nt;                      // gradient: transferred 2D entries per axis, (la+2)(lb+2)(lc+2)(ld+1)
  int offG, offA, offB;        // LDS (doubles): 2D integrals, buffer A, buffer B
  int offT, offC, offI;        // LDS (doubles): transferred 2D table, HRR coefficients, int tables
  int offD, dstr;              // LDS (doubles): the six density blocks of one quartet (J/K), block stride
  int offR, n4;                // gradient, LDS (doubles): per axis and component exponent tuple
                               // (a <= la, b <= lb, c <= lc, d <= ld; n4 of them) the value and the
                               // A, B, C derivatives of the transferred 2D integral, [3][n4][4]
  int lds;                     // bytes of dynamic LDS
};

// One class of a multi-class launch of the run-time-class kernels: its
// layout, the offsets of its wavefront tasks in the launch's lists, its first
// block wave0, its 64-ket chunks nwaves (nwaves * 64 / kpb blocks).
struct HLTask {
  HLDims h;
  long long wave0, nwaves;  // first block, 64-ket chunks
  int seg0, ch0, nseg, ket0;
  int kpb;                  // kets per block (WaveWork::kpb)
};
struct HLTaskList {
  const HLTask* task;
  int ntask;
  const int* seg;          // seg_bra | seg_k0 | seg_n of all launches (nseg each)
  int nseg;
  const long long* choff;
  const int* ket;          // class lists
};

// Compiled resources of one kernel (hipFuncGetAttributes).
struct KernelRes {
  int private_bytes;  // per-lane private segment (register spills)
  int lds_bytes;      // static LDS per block
  int regs;           // registers per lane reported by the runtime
  int max_threads;    // largest block the kernel can be launched with
};
// Largest per-lane private segment a class-templated kernel may use before its
// class is routed to the run-time-class kernels instead: large private
// segments can fail to get scratch next to other kernels' reservations (a
// 17 KB one aborted its queue with HSA_STATUS_ERROR_OUT_OF_RESOURCES).
#define OQPH_PRIVATE_BUDGET 8704

// Primitive-pair-count buckets of the class lists.
#define OQPH_NBKT 6

// Bra-row work of one J/K class-pair launch: block i takes bra pair
// item[i*(1+OQPH_NBKT)] and, for each ket bucket b, the first
// item[i*(1+OQPH_NBKT) + 1 + b] kets of that bucket
// (ket[kb[b] ...]); blocks are ordered by decreasing cost.
struct BraWork {
  const int* item;
  const int* ket;
  int kb[OQPH_NBKT + 1];
  int nitem;
  int jbase, kofs;  // as WaveWork
};

// Grouped bra-row work (k_jkg).  Segmented shells contracted from one set of
// primitive exponents on one atom (cc-pVDZ oxygen: 1s and 2s over nine s
// exponents) form a family; the shell pairs of one class whose two shells
// come from the same two families share every primitive pair's exponents and
// centres and differ only in the contraction coefficients.  A chunk is up to
// OQPH_JKG_MAXMEM such pairs: its primitive records hold the common fields
// once and one K/zeta per member, so a block evaluates the quadrature of a
// primitive quartet once for all members.
#define OQPH_JKG_MAXMEM 4
// Chunk c starts at pool[hdr[c]]:
//   [0] nmem  [1] first primitive record (offset in doubles)  [2] nprim
//   [3] nrow (K/D rows: the AOs of the members' distinct shells)  [4] nshl
//   (distinct shells)  [5] stride (doubles per primitive record, 6 + nmem)
//   [6] offset of the row AO list  [7] offset of the shell list (both from
//   the chunk start)  [8 + 5m ...] member m: device pair id, first K row of
//   its shell a, first K row of its shell b, shell slot of si, of sj.
// Primitive record: zeta, X = P - A (3), 1/(2 zeta), max_m |K_m/zeta|,
// K_m/zeta per member (0 where the member's own pair screens dropped it).
struct ChunkView {
  const int* hdr;
  const int* pool;
  const double* cp;
  const float* cp32;  // the same records in FP32 (mixed mode)
};
#define OQPH_JKG_HDR 8
#define OQPH_JKG_MEM 5
// Tiled form of the grouped kernel: bra pairs of one class sharing their
// first shell, each with its own primitive pairs, at most OQPH_JKT_MAXP of
// them per tile (the LDS copy).  Its tiles use the chunk header layout above
// ([1], [5] unused, [2] the members' primitive pairs) and live in the
// launch's work lists.
#define OQPH_JKT_MAXP 192

// LDS column space of the K rows of one launch: the AOs of the shells of
// angular momentum lc (C space) and, when ld != lc, of ld (D space, after C).
struct ColMap {
  const int* colidx;  // per shell: AO offset among the shells of its angular momentum
  const int* aoC;     // global AO index of each C-space column
  const int* aoD;     // global AO index of each D-space column
  int ncolC;          // C-space columns
  int ncol;           // all columns (ncolC, or ncolC + D-space columns)
  int dofs;           // first D-space column (0 when lc == ld)
  int drows;          // 1: the density in the K rows and the bra rows of the
                      // shell-pair maxima are staged in LDS as well
};

// Statistics counters of one launch: OQPH_CTR_SLOTS slots of 8 counters, one
// 64-byte line each.  A wavefront adds its totals to slot (wavefront index mod
// OQPH_CTR_SLOTS), so the atomics of millions of short wavefronts spread over
// many L2 lines instead of serialising on one; the host sums the slots.
#define OQPH_CTR_SLOTS 128

// Arguments of one J/K build launch (shared by host and device code).
struct JKParams {
  const ShellPair* pairs;
  const PrimJK* prims;
  const PrimF32* prims32;  // the same primitive pairs in FP32 (mixed mode)
  const double* rys;
  const float* rys32;   // FP32 root fits (degree 6)
  const double* xform;  // [2][2][10][10] canonical Cartesian -> AO transforms (L = 2, 3)
  const double* D;      // packed density (ntri)
  double* F;            // packed G (ntri), accumulated with atomics
  // Ket-side Coulomb buffer: the class kernels add each quartet's J_cd into
  // the ket pair's own block (pairs in class-list order, so the kets of a
  // wavefront hit consecutive addresses and their atomics coalesce), and one
  // pass after the launches adds the blocks into F (with the diagonal
  // doubling).  Random FP64 atomics into F were the J_cd bottleneck.
  double* Jk;
  const double* dsh;    // nsh*nsh shell-pair max |D|
  unsigned long long* counters;  // [0] computed quartets, [1] skipped by fine screen, [2] fp32 quartets
  int nsh;
  double cutoff;        // int2e_cutoff c
  double maxden;
  double sc, se;
  int mixed;            // 1: quartets with bound below fp32_cut run in FP32
  double fp32_cut;
  double inv_mu2;       // 1/mu^2 of erf(mu r)/r attenuated integrals; 0: plain 1/r
  const unsigned char* own;  // per pair: 1 when this share owns quartets whose
                             // reference-order first pair it is; null: all
  // Second exchange density of the single-pass open-shell build: when F2 is
  // set, every quartet also adds its exchange part (scale se, no Coulomb)
  // for the packed density D2 into F2, from the same integrals.
  const double* D2;
  double* F2;
  // Response pass over non-symmetric densities (run-time-class kernel only):
  // when Dg is set, D, F, D2 and F2 are unused and every quartet adds, for
  // each of the gnd full nbf x nbf densities M_v, sc J[M_v] (if gcoul[v]) -
  // se K[M_v] into the full matrix F_v, J[M]_ij = sum_kl (ij|kl) M_kl and
  // K[M]_ik = sum_jl (ij|kl) M_jl.
  const double* Dg;
  double* Fg;
  const int* gcoul;
  int gnd;
  int nbf;
};
