/*
 * oqo.c -- CPU FP64 oracle for the direct-SCF two-electron Fock build (J/K)
 *          and its adjoint, the two-electron nuclear gradient.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is linked by the
 * tests, by the package's smoke check and by the CPU-baseline leg of the
 * benchmark, and by nothing on the product path.  The product (liboqph)
 * never links or calls it.
 *
 * What it restates (OpenQP, files relative to the source tree):
 *   - primitive-pair data and screening      source/integrals/int2_pairs.F90:74-266
 *   - cutoff derivation from int2e_cutoff    source/integrals/int2.F90:266-272
 *   - Schwarz table (ints_exchange)          source/integrals/int2.F90:1582-1737
 *   - shell-pair density (shlden), screens   source/integrals/int2.F90:963-1047, 1338-1346
 *   - canonical quartet loop (int2_twoei)    source/integrals/int2.F90:755-821
 *   - primitive-quartet screen (rotated axis) source/integrals/int_rotaxis.F90:292-306, 629
 *   - element cutoff + canonicalisation      source/integrals/int2.F90:1741-1865
 *   - RHF 6-scatter digestion                source/integrals/int2.F90:1414-1484
 *   - 0.5 scaling and diagonal doubling      source/scf_addons.F90:1178-1185
 *   - Cartesian order / component norms      source/constants.F90:34-60, 110+
 *   - Cartesian -> real solid harmonics      source/integrals/cart2sph.F90:111-160
 *   - 2e gradient contraction                source/modules/hf_gradient.F90:428-495,
 *                                            source/integrals/grd2.F90:145-391
 *
 * The ERI values themselves are computed with the Obara-Saika vertical
 * recursion and the Head-Gordon-Pople horizontal recursion over a tabulated
 * Boys function -- a formula independent of the Rys quadrature that the
 * device kernels use, so that the two agree only if both are right.  The
 * reference's own engines (rotated axis for L<=2, Rys or libint above) give
 * the same integrals to rounding; only the screening decisions listed above
 * shape which integrals enter J/K, and those are restated exactly.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define LMAXB 4          /* largest shell angular momentum accepted      */
#define LINT  10         /* largest internal Cartesian level             */
#define NCUM(L) (((L) + 1) * ((L) + 2) * ((L) + 3) / 6)
#define NCART(L) (((L) + 1) * ((L) + 2) / 2)
#define MAXCOMP NCUM(LINT)

/* ------------------------------------------------------------------ */
/* Cartesian components                                                */
/* ------------------------------------------------------------------ */
static int comp_l[MAXCOMP][3];          /* global index -> (lx,ly,lz)   */
static int gidx_tab[LINT + 1][LINT + 1][LINT + 1];
/* OpenQP component order within a shell (constants.F90:34-60) */
static const int ORDER_D[6][3] = {{2,0,0},{0,2,0},{0,0,2},{1,1,0},{1,0,1},{0,1,1}};
static const int ORDER_F[10][3] = {{3,0,0},{0,3,0},{0,0,3},{2,1,0},{2,0,1},
                                   {1,2,0},{0,2,1},{1,0,2},{0,1,2},{1,1,1}};
static const int ORDER_G[15][3] = {{4,0,0},{0,4,0},{0,0,4},{3,1,0},{3,0,1},
                                   {1,3,0},{0,3,1},{1,0,3},{0,1,3},{2,2,0},
                                   {2,0,2},{0,2,2},{2,1,1},{1,2,1},{1,1,2}};
static int out_comp[LMAXB + 1][15][3];
static double c2s_mat[LMAXB + 1][15][9];  /* [l][cart][sph], unit-normalised Cartesians */
static int tables_ready = 0;
static void boys_init(void);

static int gidx(int x, int y, int z) { return gidx_tab[x][y][z]; }

static double dfact(int n) { /* (n)!! with (-1)!! = 1 */
  double r = 1.0;
  for (int k = n; k > 1; k -= 2) r *= k;
  return r;
}

static double binom(int n, int k) {
  if (k < 0 || k > n) return 0.0;
  double r = 1.0;
  for (int i = 1; i <= k; ++i) r = r * (n - i + 1) / i;
  return r;
}

/* overlap of two same-exponent monomials, relative units */
static double mono_ovl(const int *a, const int *b) {
  double r = 1.0;
  for (int d = 0; d < 3; ++d) {
    int s = a[d] + b[d];
    if (s & 1) return 0.0;
    r *= dfact(s - 1);
  }
  return r;
}

static void build_tables(void) {
  if (tables_ready) return;
  int g = 0;
  for (int L = 0; L <= LINT; ++L)
    for (int lx = L; lx >= 0; --lx)
      for (int ly = L - lx; ly >= 0; --ly) {
        int lz = L - lx - ly;
        comp_l[g][0] = lx; comp_l[g][1] = ly; comp_l[g][2] = lz;
        gidx_tab[lx][ly][lz] = g;
        ++g;
      }
  memset(out_comp, 0, sizeof(out_comp));
  out_comp[0][0][0] = 0;
  out_comp[1][0][0] = 1; out_comp[1][1][1] = 1; out_comp[1][2][2] = 1;
  memcpy(out_comp[2], ORDER_D, sizeof(ORDER_D));
  memcpy(out_comp[3], ORDER_F, sizeof(ORDER_F));
  memcpy(out_comp[4], ORDER_G, sizeof(ORDER_G));
  /* real solid harmonics, m = -l..l, on unit-normalised Cartesians:
   * Helgaker/Jorgensen/Olsen eq. (6.4.47)-(6.4.50); normalised in the
   * intra-shell Cartesian metric (cart2sph.F90:111-160 states B S B^T = I). */
  memset(c2s_mat, 0, sizeof(c2s_mat));
  for (int l = 2; l <= LMAXB; ++l) {
    int nc = NCART(l);
    for (int m = -l; m <= l; ++m) {
      int am = abs(m);
      double row[15] = {0};
      int k0 = (m >= 0) ? 0 : 1;
      for (int t = 0; t <= (l - am) / 2; ++t)
        for (int u = 0; u <= t; ++u)
          for (int k = k0; k <= am; k += 2) {
            int sp = t + (k - k0) / 2;
            double c = ((sp & 1) ? -1.0 : 1.0) * pow(0.25, t) * binom(l, t) *
                       binom(l - t, am + t) * binom(t, u) * binom(am, k);
            if (c == 0.0) continue;
            int ax = 2 * t + am - 2 * u - k, ay = 2 * u + k, az = l - 2 * t - am;
            if (ax < 0 || ay < 0 || az < 0) continue;
            int ci = -1;
            for (int q = 0; q < nc; ++q)
              if (out_comp[l][q][0] == ax && out_comp[l][q][1] == ay && out_comp[l][q][2] == az) ci = q;
            /* monomial = unit component * sqrt(self-overlap of monomial) */
            row[ci] += c * sqrt(dfact(2 * ax - 1) * dfact(2 * ay - 1) * dfact(2 * az - 1));
          }
      double n2 = 0.0;
      for (int p = 0; p < nc; ++p)
        for (int q = 0; q < nc; ++q) {
          double sp_ = mono_ovl(out_comp[l][p], out_comp[l][q]) /
                       sqrt(mono_ovl(out_comp[l][p], out_comp[l][p]) * mono_ovl(out_comp[l][q], out_comp[l][q]));
          n2 += row[p] * row[q] * sp_;
        }
      for (int p = 0; p < nc; ++p) c2s_mat[l][p][m + l] = row[p] / sqrt(n2);
    }
  }
  boys_init();
  tables_ready = 1;
}

/* component norm of a unit-normalised Cartesian relative to x^l (shells_pnrm2) */
static double pnrm2(int l, int q) {
  const int *c = out_comp[l][q];
  return sqrt(dfact(2 * l - 1) / (dfact(2 * c[0] - 1) * dfact(2 * c[1] - 1) * dfact(2 * c[2] - 1)));
}

void oqo_c2s(int l, double *out /* [ncart][2l+1] */) {
  build_tables();
  int nc = NCART(l), ns = 2 * l + 1;
  for (int p = 0; p < nc; ++p)
    for (int s = 0; s < ns; ++s) out[p * ns + s] = (l >= 2) ? c2s_mat[l][p][s] : (p == s ? 1.0 : 0.0);
}

void oqo_cart_order(int l, int *out /* [ncart][3] */) {
  build_tables();
  for (int q = 0; q < NCART(l); ++q)
    for (int d = 0; d < 3; ++d) out[3 * q + d] = out_comp[l][q][d];
}

/* ------------------------------------------------------------------ */
/* Boys function F_m(T) = int_0^1 t^2m exp(-T t^2) dt                  */
/* ------------------------------------------------------------------ */
#define BOYS_MMAX 32
#define BOYS_NTAY 8
#define BOYS_H 0.05
#define BOYS_TMAX 45.0
#define BOYS_NGRID ((int)(BOYS_TMAX / BOYS_H) + 2)
static double *boys_tab = NULL;   /* [NGRID][MMAX+NTAY+1] */

static void boys_series(int mtop, long double T, long double *F) {
  /* F_m(T) = exp(-T) sum_k (2T)^k / ((2m+1)(2m+3)...(2m+2k+1)); all terms positive */
  for (int m = 0; m <= mtop; ++m) {
    long double term = 1.0L / (2 * m + 1), sum = term;
    for (int k = 1; k < 2000; ++k) {
      term *= 2.0L * T / (2 * m + 2 * k + 1);
      sum += term;
      if (term < sum * 1e-21L) break;
    }
    F[m] = expl(-T) * sum;
  }
}

static void boys_init(void) {
  if (boys_tab) return;
  int ncol = BOYS_MMAX + BOYS_NTAY + 1;
  double *tab = (double *)malloc(sizeof(double) * BOYS_NGRID * ncol);
  long double F[BOYS_MMAX + BOYS_NTAY + 2];
  for (int k = 0; k < BOYS_NGRID; ++k) {
    boys_series(ncol - 1, (long double)(k * BOYS_H), F);
    for (int m = 0; m < ncol; ++m) tab[k * ncol + m] = (double)F[m];
  }
  boys_tab = tab;
}

void oqo_boys(int mmax, double T, double *F) {
  boys_init();
  if (T < BOYS_TMAX - 1.0) {
    int ncol = BOYS_MMAX + BOYS_NTAY + 1;
    int k = (int)(T / BOYS_H + 0.5);
    double dT = T - k * BOYS_H;
    const double *row = boys_tab + k * ncol;
    double s = 0.0, fac = 1.0;
    for (int j = 0; j < BOYS_NTAY; ++j) {
      s += row[mmax + j] * fac;
      fac *= -dT / (j + 1);
    }
    F[mmax] = s;
    double e = exp(-T);
    for (int m = mmax; m > 0; --m) F[m - 1] = (2.0 * T * F[m] + e) / (2 * m - 1);
  } else {
    double e = exp(-T);
    F[0] = 0.5 * sqrt(M_PI / T);
    for (int m = 0; m < mmax; ++m) F[m + 1] = ((2 * m + 1) * F[m] - e) / (2.0 * T);
  }
}

/* ------------------------------------------------------------------ */
/* Basis                                                               */
/* ------------------------------------------------------------------ */
typedef struct {
  int nsh, nbf, natm;
  const int *am, *harm, *atom, *nprim, *poff, *aooff;
  const double *ex, *cc, *xyz;   /* xyz: [natm][3] bohr */
} oqo_basis;

static int nao_of(const oqo_basis *b, int s) {
  int l = b->am[s];
  return (b->harm[s] && l >= 2) ? 2 * l + 1 : NCART(l);
}

typedef struct {
  double integral, pair, quartet, expo;
  double dab2, dabcut;   /* gradient: skip primitive quartets with fqz^2 dabmax^2 < dabcut ze */
  double imu2;           /* 1/mu^2 of the erf(mu r)/r operator; 0: 1/r */
} oqo_cut;

static oqo_cut make_cut(double c) { /* int2.F90:266-272 */
  oqo_cut r = {c, 1.0e-2 * c, 1.0e-4 * c, 25.0 * log(10.0), 0.0, 0.0};
  return r;
}

/* primitive pair (int2_pairs.F90:179-266) */
typedef struct {
  double a, b, g, ginv, K, P[3];
} ppair;

static int make_pairs(const oqo_basis *bs, int sa, int sb, const oqo_cut *cut, ppair *out) {
  const double *A = bs->xyz + 3 * bs->atom[sa], *B = bs->xyz + 3 * bs->atom[sb];
  double ab2 = 0.0;
  for (int d = 0; d < 3; ++d) ab2 += (A[d] - B[d]) * (A[d] - B[d]);
  const double sqrtpi52 = sqrt(2.0) * pow(M_PI, 1.25);
  int n = 0;
  for (int p = 0; p < bs->nprim[sa]; ++p)
    for (int q = 0; q < bs->nprim[sb]; ++q) {
      double a = bs->ex[bs->poff[sa] + p], b = bs->ex[bs->poff[sb] + q];
      double g = a + b, e12 = a * b * ab2;
      if (e12 > g * cut->expo) continue;
      double gi = 1.0 / g;
      e12 *= gi;
      double k1 = bs->cc[bs->poff[sa] + p] * bs->cc[bs->poff[sb] + q] * exp(-e12);
      if (fabs(k1) < cut->quartet) continue;
      ppair *pp = out + n++;
      pp->a = a; pp->b = b; pp->g = g; pp->ginv = gi; pp->K = sqrtpi52 * k1;
      for (int d = 0; d < 3; ++d) pp->P[d] = (a * A[d] + b * B[d]) * gi;
    }
  return n;
}

/* ------------------------------------------------------------------ */
/* Contracted shell quartet via OS-VRR + HGP-HRR                       */
/* ------------------------------------------------------------------ */
typedef struct {
  double *W;       /* VRR work [ncum(Lab)][ncum(Lcd)][Mtot+1] */
  double *C;       /* contracted [e][f] */
  double *H1, *H2; /* HRR buffers */
  size_t nW, nC, nH;
} work_t;

static void work_init(work_t *w) {
  int ne = NCUM(LINT - 1);
  w->nW = (size_t)ne * ne * (2 * LINT + 1);
  w->W = (double *)malloc(sizeof(double) * w->nW);
  w->nC = (size_t)ne * ne;
  w->C = (double *)malloc(sizeof(double) * w->nC);
  w->nH = (size_t)ne * ne * 16;
  w->H1 = (double *)malloc(sizeof(double) * w->nH);
  w->H2 = (double *)malloc(sizeof(double) * w->nH);
}
static void work_free(work_t *w) { free(w->W); free(w->C); free(w->H1); free(w->H2); }

/* weight mode for the primitive contraction: 0 -> 1, 1 -> 2a, 2 -> 2b, 3 -> 2c, 4 -> 2d */
static double wfac(int mode, const ppair *p, const ppair *q) {
  switch (mode) {
    case 1: return 2.0 * p->a;
    case 2: return 2.0 * p->b;
    case 3: return 2.0 * q->a;
    case 4: return 2.0 * q->b;
    default: return 1.0;
  }
}

/* first axis along which a component can be lowered */
static int pick_dir(const int *c) { return c[0] > 0 ? 0 : (c[1] > 0 ? 1 : 2); }

static int lower(int g, int d) {
  int c[3] = {comp_l[g][0], comp_l[g][1], comp_l[g][2]};
  c[d]--;
  return gidx(c[0], c[1], c[2]);
}
static int raise_(int g, int d) {
  int c[3] = {comp_l[g][0], comp_l[g][1], comp_l[g][2]};
  c[d]++;
  return gidx(c[0], c[1], c[2]);
}
static int lev(int g) { return comp_l[g][0] + comp_l[g][1] + comp_l[g][2]; }

/*
 * Raw Cartesian block (ab|cd) for angular momenta La..Ld on centres A..D,
 * canonical internal component order, x^l primitive normalisation
 * (i.e. before the shells_pnrm2 factors).  Output out[a][b][c][d] with
 * local canonical indices.  np/nq primitive pairs.  Returns 0 when every
 * primitive quartet was screened.
 */
static int raw_block(int La, int Lb, int Lc, int Ld, const double *A, const double *B,
                     const double *Cc, const double *D, const ppair *pp, int np, const ppair *qq,
                     int nq, const oqo_cut *cut, int mode, work_t *w, double *out) {
  int Lab = La + Lb, Lcd = Lc + Ld, Mt = Lab + Lcd;
  int ne = NCUM(Lab), nf = NCUM(Lcd), nm = Mt + 1;
  double *W = w->W, *Cn = w->C;
  memset(Cn, 0, sizeof(double) * ne * nf);
  double F[4 * LINT + 2];
  int any = 0;
  double qc2 = cut->quartet * cut->quartet;
#define WI(e, f, m) W[((size_t)(e) * nf + (f)) * nm + (m)]
  for (int iq = 0; iq < nq; ++iq) {
    const ppair *q = qq + iq;
    double sq = q->K * q->ginv;
    if (fabs(sq) < cut->quartet) continue;               /* int_rotaxis.F90:306 */
    for (int ip = 0; ip < np; ++ip) {
      const ppair *p = pp + ip;
      double sp = p->K * p->ginv;
      if (fabs(sp) < cut->quartet) continue;             /* int_rotaxis.F90:295 */
      double fqz = sp * sq, zeta = p->g, eta = q->g, ze = zeta + eta;
      /* int_rotaxis.F90:629; attenuated: zeta + eta + zeta eta / mu^2 (int_rys.F90:227-232) */
      if (fqz * fqz < qc2 * (ze + zeta * eta * cut->imu2)) continue;
      if (cut->dabcut > 0.0 && fqz * fqz * cut->dab2 < cut->dabcut * ze) continue; /* grd2_rys.F90:376 */
      any = 1;
      double rho = zeta * eta / ze;
      double Wc[3], PA[3], WP[3], QC[3], WQ[3], pq2 = 0.0;
      for (int d = 0; d < 3; ++d) {
        Wc[d] = (zeta * p->P[d] + eta * q->P[d]) / ze;
        PA[d] = p->P[d] - A[d];
        WP[d] = Wc[d] - p->P[d];
        QC[d] = q->P[d] - Cc[d];
        WQ[d] = Wc[d] - q->P[d];
        pq2 += (p->P[d] - q->P[d]) * (p->P[d] - q->P[d]);
      }
      /* K includes sqrt(2) pi^(5/4) c c exp(); product/(zeta eta sqrt(ze)) = 2 pi^(5/2)... */
      double pref = fqz / sqrt(ze) * wfac(mode, p, q);
      if (cut->imu2 > 0.0) {
        /* erf(mu r)/r: the auxiliary integrals become w^(m+1/2) F_m(w T) with
         * w = mu^2 / (mu^2 + rho), which obey the same Obara-Saika recursion
         * since w^(m+1/2) F_m(w T) = (-d/dT)^m [w^(1/2) F_0(w T)] (Ahlrichs,
         * Phys. Chem. Chem. Phys. 8, 3072 (2006)); 1/mu^2 -> 0 gives 1/r back. */
        double om = 1.0 / (1.0 + rho * cut->imu2);
        oqo_boys(Mt, om * rho * pq2, F);
        double f = sqrt(om);
        for (int m = 0; m <= Mt; ++m) { WI(0, 0, m) = pref * f * F[m]; f *= om; }
      } else {
        oqo_boys(Mt, rho * pq2, F);
        for (int m = 0; m <= Mt; ++m) WI(0, 0, m) = pref * F[m];
      }
      double i2z = 0.5 / zeta, i2e = 0.5 / eta, i2ze = 0.5 / ze, rz = rho / zeta, re = rho / eta;
      /* bra VRR: [e0|00]^(m) */
      for (int e = 1; e < ne; ++e) {
        int d = pick_dir(comp_l[e]);
        int e1 = lower(e, d);
        int ed = comp_l[e][d] - 1;
        int e2 = ed > 0 ? lower(e1, d) : -1;
        int mtop = Mt - lev(e);
        for (int m = 0; m <= mtop; ++m) {
          double v = PA[d] * WI(e1, 0, m) + WP[d] * WI(e1, 0, m + 1);
          if (e2 >= 0) v += ed * i2z * (WI(e2, 0, m) - rz * WI(e2, 0, m + 1));
          WI(e, 0, m) = v;
        }
      }
      /* ket VRR: [e0|f0]^(m) */
      for (int f = 1; f < nf; ++f) {
        int d = pick_dir(comp_l[f]);
        int f1 = lower(f, d);
        int fd = comp_l[f][d] - 1;
        int f2 = fd > 0 ? lower(f1, d) : -1;
        for (int e = 0; e < ne; ++e) {
          int mtop = Mt - lev(e) - lev(f);
          if (mtop < 0) continue;
          int ei = comp_l[e][d];
          int em = ei > 0 ? lower(e, d) : -1;
          for (int m = 0; m <= mtop; ++m) {
            double v = QC[d] * WI(e, f1, m) + WQ[d] * WI(e, f1, m + 1);
            if (f2 >= 0) v += fd * i2e * (WI(e, f2, m) - re * WI(e, f2, m + 1));
            if (em >= 0) v += ei * i2ze * WI(em, f1, m + 1);
            WI(e, f, m) = v;
          }
        }
      }
      for (int e = NCUM(La - 1 < 0 ? -1 : La - 1); e < ne; ++e)
        for (int f = NCUM(Lc - 1 < 0 ? -1 : Lc - 1); f < nf; ++f) Cn[e * nf + f] += WI(e, f, 0);
    }
  }
#undef WI
  if (!any) return 0;
  /* HRR on the bra: (a b| from (e 0|, e in [La, Lab]) */
  double AB[3], CD[3];
  for (int d = 0; d < 3; ++d) { AB[d] = A[d] - B[d]; CD[d] = Cc[d] - D[d]; }
  int e0 = NCUM(La - 1 < 0 ? -1 : La - 1), f0 = NCUM(Lc - 1 < 0 ? -1 : Lc - 1);
  int nec = ne - e0, nfc = nf - f0;
  /* stage buffers: T[a_g - e0][b_local][f - f0] */
  double *T = w->H1, *U = w->H2;
  for (int e = 0; e < nec; ++e)
    for (int f = 0; f < nfc; ++f) T[(size_t)e * nfc + f] = Cn[(e + e0) * nf + (f + f0)];
  int nbcur = 1; /* level-0 b */
  for (int s = 1; s <= Lb; ++s) {
    int nbs = NCART(s), bo = NCUM(s - 1), bpo = NCUM(s - 2 < 0 ? -1 : s - 2);
    int atop = NCUM(Lab - s) - e0; /* a rows valid at this stage */
    for (int a = 0; a < atop; ++a)
      for (int bl = 0; bl < nbs; ++bl) {
        int bg = bo + bl, d = pick_dir(comp_l[bg]);
        int b1 = lower(bg, d) - bpo;
        int ar = raise_(a + e0, d) - e0;
        for (int f = 0; f < nfc; ++f)
          U[((size_t)a * nbs + bl) * nfc + f] =
              T[((size_t)ar * nbcur + b1) * nfc + f] + AB[d] * T[((size_t)a * nbcur + b1) * nfc + f];
      }
    double *t = T; T = U; U = t;
    nbcur = nbs;
  }
  /* now T[a (rows of level La only: first NCART(La) of the La block)][b][f] */
  int na = NCART(La), nb = NCART(Lb), nc = NCART(Lc), nd = NCART(Ld);
  /* ket HRR on each (a,b) */
  int nab = na * nb;
  /* reorganise: V[ab][f] */
  double *V = U;
  for (int a = 0; a < na; ++a)
    for (int b = 0; b < nb; ++b)
      for (int f = 0; f < nfc; ++f) V[((size_t)(a * nb + b)) * nfc + f] = T[((size_t)a * nbcur + b) * nfc + f];
  /* T as ket stage buffer [ab][c_g - f0][d_local] */
  double *S0 = V, *S1 = T;
  int ndcur = 1;
  for (int s = 1; s <= Ld; ++s) {
    int nds = NCART(s), dof = NCUM(s - 1), dpo = NCUM(s - 2 < 0 ? -1 : s - 2);
    int ctop = NCUM(Lcd - s) - f0;
    for (int ab = 0; ab < nab; ++ab)
      for (int c = 0; c < ctop; ++c)
        for (int dl = 0; dl < nds; ++dl) {
          int dg = dof + dl, dd = pick_dir(comp_l[dg]);
          int d1 = lower(dg, dd) - dpo;
          int cr = raise_(c + f0, dd) - f0;
          S1[((size_t)ab * nfc + c) * nds + dl] =
              S0[((size_t)ab * nfc + cr) * ndcur + d1] + CD[dd] * S0[((size_t)ab * nfc + c) * ndcur + d1];
        }
    double *t = S0; S0 = S1; S1 = t;
    ndcur = nds;
  }
  for (int ab = 0; ab < nab; ++ab)
    for (int c = 0; c < nc; ++c)
      for (int d = 0; d < nd; ++d) out[((size_t)ab * nc + c) * nd + d] = S0[((size_t)ab * nfc + c) * ndcur + d];
  return 1;
}

/* map canonical-local index of level l to OpenQP output position */
static int out_pos(int l, int local) {
  int g = NCUM(l - 1 < 0 ? -1 : l - 1) + local;
  for (int q = 0; q < NCART(l); ++q)
    if (out_comp[l][q][0] == comp_l[g][0] && out_comp[l][q][1] == comp_l[g][1] &&
        out_comp[l][q][2] == comp_l[g][2])
      return q;
  return -1;
}

/*
 * Transform a raw canonical block to the final AO representation:
 * OpenQP component order, shells_pnrm2 factors, pure projection for
 * harmonic shells with l >= 2.  dims in/out given per index.
 */
static void finalize_block(const int *L, const int *pure, const double *raw, double *out, double *tmp) {
  int nc[4], no[4];
  for (int i = 0; i < 4; ++i) {
    nc[i] = NCART(L[i]);
    no[i] = (pure[i] && L[i] >= 2) ? 2 * L[i] + 1 : nc[i];
  }
  /* reorder + norms into tmp with OpenQP order */
  int pos[4][15];
  double nrm[4][15];
  for (int i = 0; i < 4; ++i)
    for (int q = 0; q < nc[i]; ++q) {
      pos[i][q] = out_pos(L[i], q);
      nrm[i][q] = pnrm2(L[i], pos[i][q]);
    }
  for (int a = 0; a < nc[0]; ++a)
    for (int b = 0; b < nc[1]; ++b)
      for (int c = 0; c < nc[2]; ++c)
        for (int d = 0; d < nc[3]; ++d) {
          double v = raw[((a * nc[1] + b) * nc[2] + c) * nc[3] + d] * nrm[0][a] * nrm[1][b] * nrm[2][c] * nrm[3][d];
          tmp[((pos[0][a] * nc[1] + pos[1][b]) * nc[2] + pos[2][c]) * nc[3] + pos[3][d]] = v;
        }
  /* successive index projections */
  int cur[4] = {nc[0], nc[1], nc[2], nc[3]};
  double *src = tmp, *dst = out;
  for (int i = 0; i < 4; ++i) {
    if (no[i] == nc[i]) continue;
    int left = 1, right = 1;
    for (int k = 0; k < i; ++k) left *= cur[k];
    for (int k = i + 1; k < 4; ++k) right *= cur[k];
    for (int lft = 0; lft < left; ++lft)
      for (int s = 0; s < no[i]; ++s)
        for (int r = 0; r < right; ++r) {
          double acc = 0.0;
          for (int q = 0; q < nc[i]; ++q) acc += c2s_mat[L[i]][q][s] * src[((size_t)lft * cur[i] + q) * right + r];
          dst[((size_t)lft * no[i] + s) * right + r] = acc;
        }
    cur[i] = no[i];
    double *t = src; src = dst; dst = t;
  }
  if (src != out) memcpy(out, src, sizeof(double) * cur[0] * cur[1] * cur[2] * cur[3]);
}

/* Final-representation ERI block for shells (i j | k l); returns 0 if all screened */
static int eri_shell(const oqo_basis *bs, int i, int j, int k, int l, const oqo_cut *cut, work_t *w,
                     ppair *bufp, ppair *bufq, double *out, double *raw, double *tmp) {
  int np = make_pairs(bs, i, j, cut, bufp);
  int nq = make_pairs(bs, k, l, cut, bufq);
  if (np == 0 || nq == 0) return 0;
  int L[4] = {bs->am[i], bs->am[j], bs->am[k], bs->am[l]};
  int pure[4] = {bs->harm[i], bs->harm[j], bs->harm[k], bs->harm[l]};
  const double *A = bs->xyz + 3 * bs->atom[i], *B = bs->xyz + 3 * bs->atom[j];
  const double *C = bs->xyz + 3 * bs->atom[k], *D = bs->xyz + 3 * bs->atom[l];
  if (!raw_block(L[0], L[1], L[2], L[3], A, B, C, D, bufp, np, bufq, nq, cut, 0, w, raw)) return 0;
  finalize_block(L, pure, raw, out, tmp);
  return 1;
}

static int max_prim(const oqo_basis *bs) {
  int m = 1;
  for (int s = 0; s < bs->nsh; ++s)
    if (bs->nprim[s] > m) m = bs->nprim[s];
  return m;
}

#define BLK (21 * 21 * 21 * 21)

/* Schwarz table: Q_ab = sqrt(max |(ab|ab)|) with the tight cutoffs of
 * int2.F90:1600-1604 (ic 1e-15, pair/quartet 1e-17, exponent 50). */
void oqo_schwarz_mu(int nsh, int nbf, int natm, const int *am, const int *harm, const int *atom,
                    const int *nprim, const int *poff, const int *aooff, const double *ex,
                    const double *cc, const double *xyz, double imu2, double *Q);
void oqo_schwarz(int nsh, int nbf, int natm, const int *am, const int *harm, const int *atom,
                 const int *nprim, const int *poff, const int *aooff, const double *ex,
                 const double *cc, const double *xyz, double *Q) {
  oqo_schwarz_mu(nsh, nbf, natm, am, harm, atom, nprim, poff, aooff, ex, cc, xyz, 0.0, Q);
}

/* The same for the erf(mu r)/r operator, imu2 = 1/mu^2 (ints_exchange with
 * mu2, int2.F90:1597, 1625). */
void oqo_schwarz_mu(int nsh, int nbf, int natm, const int *am, const int *harm, const int *atom,
                    const int *nprim, const int *poff, const int *aooff, const double *ex,
                    const double *cc, const double *xyz, double imu2, double *Q) {
  build_tables();
  oqo_basis bs = {nsh, nbf, natm, am, harm, atom, nprim, poff, aooff, ex, cc, xyz};
  oqo_cut cut = {1.0e-15, 1.0e-17, 1.0e-17, 50.0, 0.0, 0.0, imu2};
  int mp = max_prim(&bs);
#pragma omp parallel
  {
    work_t w;
    work_init(&w);
    ppair *bp = (ppair *)malloc(sizeof(ppair) * mp * mp), *bq = (ppair *)malloc(sizeof(ppair) * mp * mp);
    double *out = (double *)malloc(sizeof(double) * BLK), *raw = (double *)malloc(sizeof(double) * BLK),
           *tmp = (double *)malloc(sizeof(double) * BLK);
#pragma omp for schedule(dynamic, 1)
    for (int a = 0; a < nsh; ++a)
      for (int b = 0; b <= a; ++b) {
        double vmax = 0.0;
        if (eri_shell(&bs, a, b, a, b, &cut, &w, bp, bq, out, raw, tmp)) {
          int n = nao_of(&bs, a) * nao_of(&bs, b);
          for (int x = 0; x < n * n; ++x)
            if (fabs(out[x]) > vmax) vmax = fabs(out[x]);
        }
        Q[a * nsh + b] = Q[b * nsh + a] = sqrt(vmax);
      }
    free(bp); free(bq); free(out); free(raw); free(tmp);
    work_free(&w);
  }
}

/* Full final-representation ERI block for one shell quartet (tests) */
int oqo_eri_quartet_mu(int nsh, int nbf, int natm, const int *am, const int *harm, const int *atom,
                       const int *nprim, const int *poff, const int *aooff, const double *ex,
                       const double *cc, const double *xyz, int i, int j, int k, int l, double cutoff,
                       double imu2, double *out);
int oqo_eri_quartet(int nsh, int nbf, int natm, const int *am, const int *harm, const int *atom,
                    const int *nprim, const int *poff, const int *aooff, const double *ex,
                    const double *cc, const double *xyz, int i, int j, int k, int l, double cutoff,
                    double *out) {
  return oqo_eri_quartet_mu(nsh, nbf, natm, am, harm, atom, nprim, poff, aooff, ex, cc, xyz, i, j, k, l, cutoff,
                            0.0, out);
}

/* The same block for the erf(mu r)/r operator, imu2 = 1/mu^2. */
int oqo_eri_quartet_mu(int nsh, int nbf, int natm, const int *am, const int *harm, const int *atom,
                       const int *nprim, const int *poff, const int *aooff, const double *ex,
                       const double *cc, const double *xyz, int i, int j, int k, int l, double cutoff,
                       double imu2, double *out) {
  build_tables();
  oqo_basis bs = {nsh, nbf, natm, am, harm, atom, nprim, poff, aooff, ex, cc, xyz};
  oqo_cut cut = make_cut(cutoff);
  cut.imu2 = imu2;
  if (cutoff <= 0.0) { cut.pair = cut.quartet = 0.0; cut.expo = 1e300; }
  int mp = max_prim(&bs);
  work_t w;
  work_init(&w);
  ppair *bp = (ppair *)malloc(sizeof(ppair) * mp * mp), *bq = (ppair *)malloc(sizeof(ppair) * mp * mp);
  double *raw = (double *)malloc(sizeof(double) * BLK), *tmp = (double *)malloc(sizeof(double) * BLK);
  /* the projections ping-pong between two full-size buffers: the caller's
   * out holds only the final block */
  double *blk = (double *)malloc(sizeof(double) * BLK);
  int r = eri_shell(&bs, i, j, k, l, &cut, &w, bp, bq, blk, raw, tmp);
  int n = nao_of(&bs, i) * nao_of(&bs, j) * nao_of(&bs, k) * nao_of(&bs, l);
  if (r) memcpy(out, blk, sizeof(double) * n);
  else memset(out, 0, sizeof(double) * n);
  free(bp); free(bq); free(raw); free(tmp); free(blk);
  work_free(&w);
  return r;
}

/* ------------------------------------------------------------------ */
/* Fock build (int2_twoei + storeints + int2_rhf_data_t_update)        */
/* ------------------------------------------------------------------ */
static inline long tri(long i, long j) { return i * (i - 1) / 2 + j; } /* 1-based, i >= j */

typedef struct {
  int n, cap;
  int *id;     /* 4 per entry, 1-based AO */
  double *v;
} ibuf;

static void digest_rhf(const ibuf *b, const double *d, double *f, double se, double sc) {
  double x1 = se, x4 = 4.0 * sc;
  for (int n = 0; n < b->n; ++n) {
    long ii = b->id[4 * n], jj = b->id[4 * n + 1], kk = b->id[4 * n + 2], ll = b->id[4 * n + 3];
    double val = b->v[n];
    long ii2 = ii * (ii - 1) / 2, jj2 = jj * (jj - 1) / 2, kk2 = kk * (kk - 1) / 2;
    long ij = ii2 + jj, ik = ii2 + kk, il = ii2 + ll, jk = jj2 + kk, jl = jj2 + ll, kl = kk2 + ll;
    if (jj < kk) jk = kk2 + jj;
    if (jj < ll) jl = ll * (ll - 1) / 2 + jj;
    double v1 = val * x1, v4 = val * x4;
    /* packed arrays are 1-based in the formulas above */
    f[ij - 1] += v4 * d[kl - 1];
    f[kl - 1] += v4 * d[ij - 1];
    f[ik - 1] -= v1 * d[jl - 1];
    f[jl - 1] -= v1 * d[ik - 1];
    f[il - 1] -= v1 * d[jk - 1];
    f[jk - 1] -= v1 * d[il - 1];
  }
}

/* UHF/ROHF digestion (int2_urohf_data_t_update, int2.F90:1488-1578): both
 * Fock matrices get the Coulomb term of the total density Da + Db, and each
 * the exchange term of its own spin density with scale 2 se. */
static void digest_urohf(const ibuf *b, const double *d, double *f, long ntri, double se, double sc) {
  double x2 = 2.0 * se, x4 = 4.0 * sc;
  const double *da = d, *db = d + ntri;
  double *fa = f, *fb = f + ntri;
  for (int n = 0; n < b->n; ++n) {
    long ii = b->id[4 * n], jj = b->id[4 * n + 1], kk = b->id[4 * n + 2], ll = b->id[4 * n + 3];
    double val = b->v[n];
    long ii2 = ii * (ii - 1) / 2, jj2 = jj * (jj - 1) / 2, kk2 = kk * (kk - 1) / 2;
    long ij = ii2 + jj, ik = ii2 + kk, il = ii2 + ll, jk = jj2 + kk, jl = jj2 + ll, kl = kk2 + ll;
    if (jj < kk) jk = kk2 + jj;
    if (jj < ll) jl = ll * (ll - 1) / 2 + jj;
    double v1 = val * x2, v4 = val * x4;
    double cij = v4 * (da[ij - 1] + db[ij - 1]), ckl = v4 * (da[kl - 1] + db[kl - 1]);
    fa[ij - 1] += ckl; fa[kl - 1] += cij;
    fa[ik - 1] -= v1 * da[jl - 1]; fa[jl - 1] -= v1 * da[ik - 1];
    fa[il - 1] -= v1 * da[jk - 1]; fa[jk - 1] -= v1 * da[il - 1];
    fb[ij - 1] += ckl; fb[kl - 1] += cij;
    fb[ik - 1] -= v1 * db[jl - 1]; fb[jl - 1] -= v1 * db[ik - 1];
    fb[il - 1] -= v1 * db[jk - 1]; fb[jk - 1] -= v1 * db[il - 1];
  }
}

/* MRSF response digestion (int2_mrsf_data_t_update, pass 1 FP64 with the
 * Tamm-Dancoff approximation, tdhf_mrsf_lib.F90:278-308).  d3 and f3 are the
 * reference's d3(nf, nmat, nbf, nbf) / f3 arrays in Fortran order; components
 * c < ncoul (the reference: 1..4 of 7) take the Coulomb term with the
 * symmetrised density ds = d3 + d3^T (precomputed there,
 * tdhf_mrsf_lib.F90:81-90), every component the eight exchange targets. */
typedef struct {
  int nf, nmat, ncoul;
  const double *d3;
} mrsf_t;

static void digest_mrsf(const ibuf *b, const mrsf_t *m, long nbf, double *f3, double se, double sc) {
  const long nf = m->nf, nm = m->nmat;
  const double *d3 = m->d3;
#define OQO_D3(v, c, i, j) (((((j) - 1) * nbf + ((i) - 1)) * nm + (c)) * nf + (v))
  for (int n = 0; n < b->n; ++n) {
    long i = b->id[4 * n], j = b->id[4 * n + 1], k = b->id[4 * n + 2], l = b->id[4 * n + 3];
    double val = b->v[n];
    double xval = val * se, cval = val * sc;
    for (int c = 0; c < m->ncoul; ++c)
      for (int v = 0; v < nf; ++v) {
        double dskl = d3[OQO_D3(v, c, k, l)] + d3[OQO_D3(v, c, l, k)];
        double dsij = d3[OQO_D3(v, c, i, j)] + d3[OQO_D3(v, c, j, i)];
        f3[OQO_D3(v, c, i, j)] += cval * dskl;
        f3[OQO_D3(v, c, j, i)] += cval * dskl;
        f3[OQO_D3(v, c, k, l)] += cval * dsij;
        f3[OQO_D3(v, c, l, k)] += cval * dsij;
      }
    for (int c = 0; c < nm; ++c)
      for (int v = 0; v < nf; ++v) {
        f3[OQO_D3(v, c, i, k)] -= xval * d3[OQO_D3(v, c, j, l)];
        f3[OQO_D3(v, c, k, i)] -= xval * d3[OQO_D3(v, c, l, j)];
        f3[OQO_D3(v, c, i, l)] -= xval * d3[OQO_D3(v, c, j, k)];
        f3[OQO_D3(v, c, l, i)] -= xval * d3[OQO_D3(v, c, k, j)];
        f3[OQO_D3(v, c, j, k)] -= xval * d3[OQO_D3(v, c, i, l)];
        f3[OQO_D3(v, c, k, j)] -= xval * d3[OQO_D3(v, c, l, i)];
        f3[OQO_D3(v, c, j, l)] -= xval * d3[OQO_D3(v, c, i, k)];
        f3[OQO_D3(v, c, l, j)] -= xval * d3[OQO_D3(v, c, k, i)];
      }
  }
#undef OQO_D3
}

/* shell_den_screen_mrsf (tdhf_mrsf_lib.F90:176-204): dsh(i, j) = dsh(j, i) =
 * max |d3(:, nmat, mu, nu)| over mu in shell j, nu in shell i, i >= j (the
 * last component only, and only that block of each shell pair) */
static void shell_density_mrsf(const oqo_basis *bs, const mrsf_t *m, double *dsh) {
  int nsh = bs->nsh;
  long nbf = bs->nbf, nf = m->nf, nm = m->nmat;
  for (int si = 0; si < nsh; ++si) {
    int mini = bs->aooff[si], ni = nao_of(bs, si);
    for (int sj = 0; sj <= si; ++sj) {
      int minj = bs->aooff[sj], nj = nao_of(bs, sj);
      double dm = 0.0;
      for (int nu = mini; nu < mini + ni; ++nu)
        for (int mu = minj; mu < minj + nj; ++mu)
          for (int v = 0; v < nf; ++v) {
            double x = fabs(m->d3[((nu * nbf + mu) * nm + (nm - 1)) * nf + v]);
            if (x > dm) dm = x;
          }
      dsh[si * nsh + sj] = dsh[sj * nsh + si] = dm;
    }
  }
}

/* storeints (int2.F90:1741-1865): element cutoff, canonical AO order, degeneracy halving */
static void store_block(const oqo_basis *bs, const int *ids, const double *blk, double cutoff, ibuf *b,
                        long *nint) {
  int nb[4], off[4];
  for (int i = 0; i < 4; ++i) {
    nb[i] = nao_of(bs, ids[i]);
    off[i] = bs->aooff[ids[i]];  /* 0-based */
  }
  int same = ids[0] == ids[2] && ids[1] == ids[3];
  int iandj = ids[0] == ids[1], kandl = ids[2] == ids[3];
  int nij = 0;
  for (int i = 0; i < nb[0]; ++i) {
    int maxj = iandj ? i + 1 : nb[1];
    for (int j = 0; j < maxj; ++j) {
      ++nij;
      int nkl = nij, maxl = nb[3];
      int stop = 0;
      for (int k = 0; k < nb[2] && !stop; ++k) {
        if (kandl) maxl = k + 1;
        if (same) {
          int itmp = maxl < nkl ? maxl : nkl;
          if (itmp == 0) { stop = 1; break; }
          maxl = itmp;
          nkl -= itmp;
        }
        for (int l = 0; l < maxl; ++l) {
          double val = blk[((i * nb[1] + j) * nb[2] + k) * nb[3] + l];
          if (fabs(val) < cutoff) continue;
          ++*nint;
          int i1 = i + 1 + off[0], j1 = j + 1 + off[1], k1 = k + 1 + off[2], l1 = l + 1 + off[3];
          if (i1 < j1) { int t = i1; i1 = j1; j1 = t; }
          if (k1 < l1) { int t = k1; k1 = l1; l1 = t; }
          int ii = i1, jj = j1, kk = k1, ll = l1;
          if (ii < kk) { ii = k1; jj = l1; kk = i1; ll = j1; }
          else if (ii == kk && jj < ll) { ii = i1; jj = l1; kk = k1; ll = j1; }
          if (ii == jj) val *= 0.5;
          if (kk == ll) val *= 0.5;
          if (ii == kk && jj == ll) val *= 0.5;
          if (b->n == b->cap) {
            b->cap *= 2;
            b->id = (int *)realloc(b->id, sizeof(int) * 4 * b->cap);
            b->v = (double *)realloc(b->v, sizeof(double) * b->cap);
          }
          b->id[4 * b->n] = ii; b->id[4 * b->n + 1] = jj; b->id[4 * b->n + 2] = kk; b->id[4 * b->n + 3] = ll;
          b->v[b->n++] = val;
        }
      }
    }
  }
}

/* shlden (int2.F90:999-1047) */
static void shell_density(const oqo_basis *bs, const double *d, int nfocks, long ntri, double *dsh) {
  int nsh = bs->nsh;
  memset(dsh, 0, sizeof(double) * nsh * nsh);
  for (int fk = 0; fk < nfocks; ++fk)
    for (int si = 0; si < nsh; ++si) {
      int mini = bs->aooff[si] + 1, maxi = mini + nao_of(bs, si) - 1;
      for (int sj = 0; sj <= si; ++sj) {
        int minj = bs->aooff[sj] + 1, maxj = minj + nao_of(bs, sj) - 1;
        double dm = 0.0;
        for (int i = mini; i <= maxi; ++i) {
          int mj = (si == sj) ? i : maxj;
          for (int j = minj; j <= mj; ++j) {
            double v = fabs(d[fk * ntri + tri(i, j) - 1]);
            if (v > dm) dm = v;
          }
        }
        if (dm > dsh[si * nsh + sj]) dsh[si * nsh + sj] = dm;
      }
    }
  for (int si = 0; si < nsh; ++si)
    for (int sj = si + 1; sj < nsh; ++sj) dsh[si * nsh + sj] = dsh[sj * nsh + si];
}

static double max6(double a, double b, double c, double d, double e, double f) {
  double m = a;
  if (b > m) m = b;
  if (c > m) m = c;
  if (d > m) m = d;
  if (e > m) m = e;
  if (f > m) m = f;
  return m;
}

/*
 * fock_jk restatement for RHF (nfocks = 1): returns the packed 2e Fock
 * contribution f = 0.5 * F_raw with the diagonal doubled, as fock_jk hands
 * back (scf_addons.F90:1178-1185).  Q is the Schwarz table (nsh*nsh).
 * stats[0] = nschwz (skipped quartets), stats[1] = computed quartets,
 * stats[2] = stored integrals.
 */
static int fock_jk_nf(int nfocks, int nsh, int nbf, int natm, const int *am, const int *harm, const int *atom,
                      const int *nprim, const int *poff, const int *aooff, const double *ex,
                      const double *cc, const double *xyz, const double *Q, const double *Qmap, const double *d,
                      double *f, double se, double sc, double cutoff, double imu2, long *stats, int rank, int size,
                      const mrsf_t *mr);

int oqo_fock_jk(int nsh, int nbf, int natm, const int *am, const int *harm, const int *atom,
                const int *nprim, const int *poff, const int *aooff, const double *ex,
                const double *cc, const double *xyz, const double *Q, const double *d, double *f,
                double se, double sc, double cutoff, long *stats) {
  return fock_jk_nf(1, nsh, nbf, natm, am, harm, atom, nprim, poff, aooff, ex, cc, xyz, Q, NULL, d, f, se, sc,
                    cutoff, 0.0, stats, 0, 1, NULL);
}

/* Same build restricted to the quartets rank `rank` of `size` computes under
 * OpenQP's MPI decomposition mod(ij_pair, size) == rank (int2.F90:759-761),
 * ij_pair the 1-based position of the outer pair (i, j) in the cost-sorted
 * shell-pair map (int2_build_shell_pair_map, :864-921). */
int oqo_fock_jk_part(int nsh, int nbf, int natm, const int *am, const int *harm, const int *atom,
                     const int *nprim, const int *poff, const int *aooff, const double *ex,
                     const double *cc, const double *xyz, const double *Q, const double *d, double *f,
                     double se, double sc, double cutoff, long *stats, int rank, int size) {
  return fock_jk_nf(1, nsh, nbf, natm, am, harm, atom, nprim, poff, aooff, ex, cc, xyz, Q, NULL, d, f, se, sc,
                    cutoff, 0.0, stats, rank, size, NULL);
}

/* General form: nfocks 1 (RHF) or 2 (UHF/ROHF); Q screens the quartets, Qmap
 * (NULL: Q) orders the shell-pair map -- OpenQP orders it by the regular
 * bounds also in the attenuated pass (int2.F90:883-892); imu2 = 1/mu^2 of the
 * erf(mu r)/r operator (0: 1/r). */
int oqo_fock_jk_ex(int nsh, int nbf, int natm, const int *am, const int *harm, const int *atom,
                   const int *nprim, const int *poff, const int *aooff, const double *ex,
                   const double *cc, const double *xyz, const double *Q, const double *Qmap, const double *d,
                   double *f, int nfocks, double se, double sc, double cutoff, double imu2, long *stats, int rank,
                   int size) {
  return fock_jk_nf(nfocks, nsh, nbf, natm, am, harm, atom, nprim, poff, aooff, ex, cc, xyz, Q, Qmap, d, f, se, sc,
                    cutoff, imu2, stats, rank, size, NULL);
}

/* MRSF response integral pass (int2_driver%run on int2_mrsf_data_t,
 * tdhf_mrsf_energy.F90:742-752, tdhf_mrsf_lib.F90:58-308): f3 (same shape as
 * d3, Fortran order) = sum over the screened quartets of the digestion above;
 * no final scaling (int2_mrsf_data_t_parallel_stop only reduces). */
int oqo_mrsf_int2(int nsh, int nbf, int natm, const int *am, const int *harm, const int *atom,
                  const int *nprim, const int *poff, const int *aooff, const double *ex,
                  const double *cc, const double *xyz, const double *Q, int nf, int nmat, int ncoul,
                  const double *d3, double *f3, double se, double sc, double cutoff, long *stats) {
  mrsf_t m = {nf, nmat, ncoul, d3};
  return fock_jk_nf(1, nsh, nbf, natm, am, harm, atom, nprim, poff, aooff, ex, cc, xyz, Q, NULL, NULL, f3, se, sc,
                    cutoff, 0.0, stats, 0, 1, &m);
}

/* UHF/ROHF build (nfocks = 2): d and f hold the alpha then the beta packed
 * matrices; screening uses the shell maxima over both densities (shlden with
 * nfocks = 2, int2.F90:999-1047), digestion int2_urohf_data_t_update, and the
 * same final 0.5 scaling and diagonal doubling (scf_addons.F90:1178-1185). */
int oqo_fock_jk_uhf(int nsh, int nbf, int natm, const int *am, const int *harm, const int *atom,
                    const int *nprim, const int *poff, const int *aooff, const double *ex,
                    const double *cc, const double *xyz, const double *Q, const double *d, double *f,
                    double se, double sc, double cutoff, long *stats) {
  return fock_jk_nf(2, nsh, nbf, natm, am, harm, atom, nprim, poff, aooff, ex, cc, xyz, Q, NULL, d, f, se, sc,
                    cutoff, 0.0, stats, 0, 1, NULL);
}

/* int2_build_shell_pair_map (int2.F90:864-921): 1-based position of pair
 * (i, j), i >= j (0-based here), in the work list sorted by 64 classes of
 * log2(Q_ij nbf_i nbf_j i), nbf the Cartesian counts, most expensive class
 * first, i descending then j ascending inside a class. */
static long *pair_map_positions(int nsh, const int *am, const double *Q) {
  const int NCLASS = 64, OFFS = 42;
  long cnt[65] = {0}, off[65] = {0};
  long *pos = (long *)malloc(sizeof(long) * (size_t)nsh * (nsh + 1) / 2);
#define OQO_CLS(i, j)                                                                       \
  ({                                                                                        \
    double sw_ = Q[(i) * nsh + (j)] > 1.0e-30 ? Q[(i) * nsh + (j)] : 1.0e-30;               \
    double cost_ = sw_ * (double)(NCART(am[i]) * NCART(am[j])) * (double)((i) + 1);         \
    int t_ = (int)(log(cost_) / log(2.0)) + OFFS;                                           \
    t_ = t_ < 0 ? 0 : (t_ > NCLASS ? NCLASS : t_);                                          \
    NCLASS - t_;                                                                            \
  })
  for (int i = nsh - 1; i >= 0; --i)
    for (int j = 0; j <= i; ++j) ++cnt[OQO_CLS(i, j)];
  for (int c = 1; c <= NCLASS; ++c) off[c] = off[c - 1] + cnt[c - 1];
  for (int i = nsh - 1; i >= 0; --i)
    for (int j = 0; j <= i; ++j) pos[(long)i * (i + 1) / 2 + j] = ++off[OQO_CLS(i, j)];
#undef OQO_CLS
  return pos;
}

/* the pair-map position of every pair, for tests */
void oqo_pair_map(int nsh, const int *am, const double *Q, long *out) {
  long *pos = pair_map_positions(nsh, am, Q);
  memcpy(out, pos, sizeof(long) * (size_t)nsh * (nsh + 1) / 2);
  free(pos);
}

static int fock_jk_nf(int nfocks, int nsh, int nbf, int natm, const int *am, const int *harm, const int *atom,
                      const int *nprim, const int *poff, const int *aooff, const double *ex,
                      const double *cc, const double *xyz, const double *Q, const double *Qmap, const double *d,
                      double *f, double se, double sc, double cutoff, double imu2, long *stats, int rank, int size,
                      const mrsf_t *mr) {
  build_tables();
  oqo_basis bs = {nsh, nbf, natm, am, harm, atom, nprim, poff, aooff, ex, cc, xyz};
  oqo_cut cut = make_cut(cutoff);
  cut.imu2 = imu2;
  long ntri = (long)nbf * (nbf + 1) / 2;
  double *dsh = (double *)malloc(sizeof(double) * nsh * nsh);
  if (mr) shell_density_mrsf(&bs, mr, dsh);
  else shell_density(&bs, d, nfocks, ntri, dsh);
  double maxden = 0.0;
  for (int x = 0; x < nsh * nsh; ++x)
    if (fabs(dsh[x]) > maxden) maxden = fabs(dsh[x]);
  long *pos = size > 1 ? pair_map_positions(nsh, am, Qmap ? Qmap : Q) : NULL;
  int mp = max_prim(&bs);
  long nschwz = 0, nquart = 0, nints = 0;
  int nthr = 1;
#ifdef _OPENMP
  nthr = omp_get_max_threads();
#endif
  const long nf = mr ? (long)mr->nf * mr->nmat * nbf * nbf : (long)nfocks * ntri;
  double *fthr = (double *)calloc((size_t)nthr * nf, sizeof(double));
#pragma omp parallel reduction(+ : nschwz, nquart, nints)
  {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    double *fl = fthr + (size_t)tid * nf;
    work_t w;
    work_init(&w);
    ppair *bp = (ppair *)malloc(sizeof(ppair) * mp * mp), *bq = (ppair *)malloc(sizeof(ppair) * mp * mp);
    double *out = (double *)malloc(sizeof(double) * BLK), *raw = (double *)malloc(sizeof(double) * BLK),
           *tmp = (double *)malloc(sizeof(double) * BLK);
    ibuf b = {0, 4096, (int *)malloc(sizeof(int) * 4 * 4096), (double *)malloc(sizeof(double) * 4096)};
#pragma omp for schedule(dynamic, 1)
    for (int i = 0; i < nsh; ++i)
      for (int j = 0; j <= i; ++j) {
        if (pos && (pos[(long)i * (i + 1) / 2 + j] % size) != rank) continue;
        /* screen_ij (int2.F90:963-971) */
        if (Q[i * nsh + j] * maxden < cutoff) {
          nschwz += (long)(i + 1) * i / 2 + (j + 1);
          continue;
        }
        for (int k = 0; k <= i; ++k) {
          int jork = (k == i) ? j : k;
          for (int l = 0; l <= jork; ++l) {
            /* screen_ijkl (int2.F90:975-986) */
            double t = Q[i * nsh + j] * Q[k * nsh + l] *
                       max6(4 * dsh[i * nsh + j], 4 * dsh[k * nsh + l], dsh[j * nsh + l], dsh[j * nsh + k],
                            dsh[i * nsh + l], dsh[i * nsh + k]);
            if (t < cutoff) { ++nschwz; continue; }
            ++nquart;
            if (!eri_shell(&bs, i, j, k, l, &cut, &w, bp, bq, out, raw, tmp)) continue;
            int ids[4] = {i, j, k, l};
            b.n = 0;
            store_block(&bs, ids, out, cutoff, &b, &nints);
            if (mr) digest_mrsf(&b, mr, nbf, fl, se, sc);
            else if (nfocks == 1) digest_rhf(&b, d, fl, se, sc);
            else digest_urohf(&b, d, fl, ntri, se, sc);
          }
        }
      }
    free(b.id); free(b.v);
    free(bp); free(bq); free(out); free(raw); free(tmp);
    work_free(&w);
  }
  for (long x = 0; x < nf; ++x) {
    double s = 0.0;
    for (int t = 0; t < nthr; ++t) s += fthr[(size_t)t * nf + x];
    f[x] = mr ? s : 0.5 * s;
  }
  if (!mr)
    for (int fk = 0; fk < nfocks; ++fk)
      for (long i = 1; i <= nbf; ++i) f[fk * ntri + tri(i, i) - 1] *= 2.0;
  free(fthr);
  free(dsh);
  free(pos);
  if (stats) { stats[0] = nschwz; stats[1] = nquart; stats[2] = nints; }
  return 0;
}

/* Full unscreened ERI tensor (nbf^4, final AO representation), tests only */
void oqo_eri_tensor(int nsh, int nbf, int natm, const int *am, const int *harm, const int *atom,
                    const int *nprim, const int *poff, const int *aooff, const double *ex,
                    const double *cc, const double *xyz, double *eri) {
  build_tables();
  oqo_basis bs = {nsh, nbf, natm, am, harm, atom, nprim, poff, aooff, ex, cc, xyz};
  oqo_cut cut = {0.0, 0.0, 0.0, 1e300, 0.0, 0.0};
  int mp = max_prim(&bs);
  work_t w;
  work_init(&w);
  ppair *bp = (ppair *)malloc(sizeof(ppair) * mp * mp), *bq = (ppair *)malloc(sizeof(ppair) * mp * mp);
  double *out = (double *)malloc(sizeof(double) * BLK), *raw = (double *)malloc(sizeof(double) * BLK),
         *tmp = (double *)malloc(sizeof(double) * BLK);
  long n = nbf;
  memset(eri, 0, sizeof(double) * n * n * n * n);
  for (int i = 0; i < nsh; ++i)
    for (int j = 0; j < nsh; ++j)
      for (int k = 0; k < nsh; ++k)
        for (int l = 0; l < nsh; ++l) {
          if (!eri_shell(&bs, i, j, k, l, &cut, &w, bp, bq, out, raw, tmp)) continue;
          int na = nao_of(&bs, i), nb = nao_of(&bs, j), nc = nao_of(&bs, k), nd = nao_of(&bs, l);
          for (int a = 0; a < na; ++a)
            for (int b = 0; b < nb; ++b)
              for (int c = 0; c < nc; ++c)
                for (int d = 0; d < nd; ++d) {
                  long A = aooff[i] + a, B = aooff[j] + b, C = aooff[k] + c, D = aooff[l] + d;
                  eri[((A * n + B) * n + C) * n + D] = out[((a * nb + b) * nc + c) * nd + d];
                }
        }
  free(bp); free(bq); free(out); free(raw); free(tmp);
  work_free(&w);
}

/* ------------------------------------------------------------------ */
/* One-electron integrals S, T, V (Obara-Saika), full nbf x nbf        */
/* ------------------------------------------------------------------ */
/* 1D overlap table s[i][j], i <= imax, j <= jmax (without the Gaussian prefactor) */
static void ovl1d(int imax, int jmax, double PA, double PB, double g, double s[][LINT + 3]) {
  double i2g = 0.5 / g;
  s[0][0] = 1.0;
  for (int i = 0; i <= imax; ++i)
    for (int j = 0; j <= jmax; ++j) {
      if (i == 0 && j == 0) continue;
      double v;
      if (i > 0) {
        v = PA * s[i - 1][j];
        if (i > 1) v += (i - 1) * i2g * s[i - 2][j];
        if (j > 0) v += j * i2g * s[i - 1][j - 1];
      } else {
        v = PB * s[i][j - 1];
        if (j > 1) v += (j - 1) * i2g * s[i][j - 2];
      }
      s[i][j] = v;
    }
}

static void int1e_shell(const oqo_basis *bs, int sa, int sb, const double *zn, double *S, double *T, double *V,
                        double *tmp) {
  int la = bs->am[sa], lb = bs->am[sb];
  int na = NCART(la), nb = NCART(lb);
  const double *A = bs->xyz + 3 * bs->atom[sa], *B = bs->xyz + 3 * bs->atom[sb];
  double AB[3], ab2 = 0.0;
  for (int d = 0; d < 3; ++d) { AB[d] = A[d] - B[d]; ab2 += AB[d] * AB[d]; }
  for (int x = 0; x < na * nb; ++x) S[x] = T[x] = V[x] = 0.0;
  int L = la + lb;
  static __thread double vr[NCUM(LINT - 1)][2 * LINT + 2];
  for (int p = 0; p < bs->nprim[sa]; ++p)
    for (int q = 0; q < bs->nprim[sb]; ++q) {
      double a = bs->ex[bs->poff[sa] + p], b = bs->ex[bs->poff[sb] + q], g = a + b;
      double cab = bs->cc[bs->poff[sa] + p] * bs->cc[bs->poff[sb] + q];
      double E = exp(-a * b / g * ab2);
      double P[3], PA[3], PB[3];
      for (int d = 0; d < 3; ++d) {
        P[d] = (a * A[d] + b * B[d]) / g;
        PA[d] = P[d] - A[d];
        PB[d] = P[d] - B[d];
      }
      double s[3][LINT + 3][LINT + 3];
      for (int d = 0; d < 3; ++d) ovl1d(la + 1, lb + 2, PA[d], PB[d], g, s[d]);
      double s0 = cab * E * pow(M_PI / g, 1.5);
      int a0 = NCUM(la - 1 < 0 ? -1 : la - 1), b0 = NCUM(lb - 1 < 0 ? -1 : lb - 1);
      for (int ia = 0; ia < na; ++ia)
        for (int ib = 0; ib < nb; ++ib) {
          const int *ca = comp_l[a0 + ia], *cb = comp_l[b0 + ib];
          double sx = s[0][ca[0]][cb[0]], sy = s[1][ca[1]][cb[1]], sz = s[2][ca[2]][cb[2]];
          S[ia * nb + ib] += s0 * sx * sy * sz;
          double t1[3];
          for (int d = 0; d < 3; ++d) {
            int i = ca[d], j = cb[d];
            double v = -2.0 * b * b * s[d][i][j + 2] + b * (2 * j + 1) * s[d][i][j];
            if (j >= 2) v -= 0.5 * j * (j - 1) * s[d][i][j - 2];
            t1[d] = v;
          }
          T[ia * nb + ib] += s0 * (t1[0] * sy * sz + sx * t1[1] * sz + sx * sy * t1[2]);
        }
      /* nuclear attraction */
      double F[2 * LINT + 2];
      for (int c = 0; c < bs->natm; ++c) {
        if (zn[c] == 0.0) continue;
        const double *Cn = bs->xyz + 3 * c;
        double PC[3], pc2 = 0.0;
        for (int d = 0; d < 3; ++d) { PC[d] = P[d] - Cn[d]; pc2 += PC[d] * PC[d]; }
        oqo_boys(L, g * pc2, F);
        double pre = -zn[c] * cab * E * 2.0 * M_PI / g;
        int ne = NCUM(L);
        for (int m = 0; m <= L; ++m) vr[0][m] = pre * F[m];
        for (int e = 1; e < ne; ++e) {
          int d = pick_dir(comp_l[e]);
          int e1 = lower(e, d), ed = comp_l[e][d] - 1, e2 = ed > 0 ? lower(e1, d) : -1;
          for (int m = 0; m <= L - lev(e); ++m) {
            double v = PA[d] * vr[e1][m] - PC[d] * vr[e1][m + 1];
            if (e2 >= 0) v += ed * 0.5 / g * (vr[e2][m] - vr[e2][m + 1]);
            vr[e][m] = v;
          }
        }
        /* HRR (a, b+1_i) = (a+1_i, b) + AB_i (a, b) on [e] with e in [la, L] */
        int e0 = a0, nec = ne - e0;
        double *cur = tmp, *nxt = tmp + 4096;
        for (int e = 0; e < nec; ++e) cur[e] = vr[e + e0][0];
        int nbc = 1;
        for (int st = 1; st <= lb; ++st) {
          int nbs = NCART(st), bo = NCUM(st - 1), bpo = NCUM(st - 2 < 0 ? -1 : st - 2);
          int atop = NCUM(L - st) - e0;
          for (int ar = 0; ar < atop; ++ar)
            for (int bl = 0; bl < nbs; ++bl) {
              int bg = bo + bl, d = pick_dir(comp_l[bg]);
              int b1 = lower(bg, d) - bpo, au = raise_(ar + e0, d) - e0;
              nxt[ar * nbs + bl] = cur[au * nbc + b1] + AB[d] * cur[ar * nbc + b1];
            }
          double *t = cur; cur = nxt; nxt = t;
          nbc = nbs;
        }
        for (int ia = 0; ia < na; ++ia)
          for (int ib = 0; ib < nb; ++ib) V[ia * nb + ib] += cur[ia * nbc + ib];
      }
    }
}

static void finalize_2(int la, int pa, int lb, int pb, const double *raw, double *out) {
  int L[4] = {la, lb, 0, 0}, P[4] = {pa, pb, 0, 0};
  double tmp[21 * 21];
  finalize_block(L, P, raw, out, tmp);
}

/* S, T, V as full nbf x nbf matrices; zn nuclear charges per atom */
void oqo_int1e(int nsh, int nbf, int natm, const int *am, const int *harm, const int *atom,
               const int *nprim, const int *poff, const int *aooff, const double *ex,
               const double *cc, const double *xyz, const double *zn, double *S, double *T, double *V) {
  build_tables();
  oqo_basis bs = {nsh, nbf, natm, am, harm, atom, nprim, poff, aooff, ex, cc, xyz};
#pragma omp parallel
  {
    double s[21 * 21], t[21 * 21], v[21 * 21], o[21 * 21], tmp[8192];
#pragma omp for schedule(dynamic, 1)
    for (int a = 0; a < nsh; ++a)
      for (int b = 0; b < nsh; ++b) {
        int1e_shell(&bs, a, b, zn, s, t, v, tmp);
        int na = nao_of(&bs, a), nb = nao_of(&bs, b);
        double *dst[3] = {S, T, V};
        double *srcs[3] = {s, t, v};
        for (int w = 0; w < 3; ++w) {
          finalize_2(am[a], harm[a], am[b], harm[b], srcs[w], o);
          for (int i = 0; i < na; ++i)
            for (int j = 0; j < nb; ++j) dst[w][(long)(aooff[a] + i) * nbf + aooff[b] + j] = o[i * nb + j];
        }
      }
  }
}

/* ------------------------------------------------------------------ */
/* Two-electron nuclear gradient (grd2_driver_gen + hf_2e_grad, RHF)   */
/* ------------------------------------------------------------------ */
/* AO <- raw canonical Cartesian map of one shell: X[i][p] with
 * chi_i = sum_p X[i][p] phi_p, phi_p the x^l-normalised Cartesians. */
static void shell_xmap(int l, int pure, double *X /* [nao][ncart] */) {
  int nc = NCART(l), no = (pure && l >= 2) ? 2 * l + 1 : nc;
  memset(X, 0, sizeof(double) * no * nc);
  for (int p = 0; p < nc; ++p) {      /* p: canonical local index */
    int q = out_pos(l, p);            /* OpenQP position of that component */
    double nrm = pnrm2(l, q);
    if (no == nc) X[q * nc + p] = nrm;
    else
      for (int m = 0; m < no; ++m) X[m * nc + p] = c2s_mat[l][q][m] * nrm;
  }
}

/*
 * dE2/dR at fixed density, E2 = 0.5 sum (ab|cd)[cs D_ab D_cd - hs/2 D_ac D_bd]
 * (Coulomb x coulscale + exchange x hfscale), as hf_2e_grad accumulates it
 * (hf_gradient.F90:428-495 for Gamma, grd2.F90:145-391 for loops/screens):
 *   Gamma_abcd = 4 cs Dc_ab Dc_cd - hs (Dc_ac Dc_bd + Dc_ad Dc_bc)
 * on the Cartesian-effective density Dc = X^T D X; coarse screen
 * Q_ij Q_kl < grad_cutoff; fine screen max|Gamma| Q_ij Q_kl < grad_cutoff/2;
 * pair data with exponent cutoff 20 ln 10 and prefactor cutoff 1e-20,
 * primitive screen fqz^2 max|Gamma|^2 < (1e-11)^2 ze (grd2.F90:219-241,
 * grd2_rys.F90:375-376).  Derivatives: d/dA (a| = 2 alpha (a+1| - a (a-1|
 * per primitive, D by translational invariance.  de: [natm][3].
 */
/* MRSF two-particle density (grd2_mrsf_compute_data_t_get_density,
 * tdhf_mrsf_gradient.F90:404-601): the eleven Cartesian matrices
 * m[0..10] = d2a, d2b, p2a, p2b, ball, bo2v, bo1v, bco1, bco2, o21v, co12
 * (ncf x ncf, M(i,j) = m[i ncf + j]), scales hs, hs2, cs, qf[3], sgnk. */
typedef struct {
  const double *m[11];
  double hs, hs2, cs, qf[3], sgnk;
} mrsf_gam_t;

static double mrsf_gamma(const mrsf_gam_t *g, long ncf, long i, long j, long k, long l) {
#define M_(x, r, c) (g->m[x][(r) * ncf + (c)])
  const int D2A = 0, D2B = 1, P2A = 2, P2B = 3, BALL = 4, BO2V = 5, BO1V = 6, BCO1 = 7, BCO2 = 8, O21V = 9, CO12 = 10;
  double df1 = ((M_(D2A, i, j) + M_(P2A, i, j)) * M_(D2A, k, l) + M_(D2A, i, j) * M_(P2A, k, l)) * 4.0 * g->cs;
  if (g->hs != 0.0 || g->hs2 != 0.0) {
    double dq1 = (M_(D2A, i, k) + M_(P2A, i, k)) * M_(D2A, j, l) + M_(D2A, i, k) * M_(P2A, j, l) +
                 (M_(D2A, i, l) + M_(P2A, i, l)) * M_(D2A, j, k) + M_(D2A, i, l) * M_(P2A, j, k) +
                 (M_(D2B, i, k) + M_(P2B, i, k)) * M_(D2B, j, l) + M_(D2B, i, k) * M_(P2B, j, l) +
                 (M_(D2B, i, l) + M_(P2B, i, l)) * M_(D2B, j, k) + M_(D2B, i, l) * M_(P2B, j, k);
    double dt2 = M_(BALL, i, k) * M_(BALL, j, l) + M_(BALL, k, i) * M_(BALL, l, j) +
                 M_(BALL, i, l) * M_(BALL, j, k) + M_(BALL, l, i) * M_(BALL, k, j);
    df1 = df1 - g->hs * dq1 - g->hs2 * 2.0 * dt2;
  }
  /* the eight-term spin-pair products (:490-586) */
#define PB(x) (M_(x, i, k) * M_(x, l, j) + M_(x, i, l) * M_(x, k, j) + M_(x, j, k) * M_(x, l, i) + \
               M_(x, j, l) * M_(x, k, i) + M_(x, l, j) * M_(x, i, k) + M_(x, k, j) * M_(x, i, l) + \
               M_(x, l, i) * M_(x, j, k) + M_(x, k, i) * M_(x, j, l))
#define PC(x, y) (M_(x, i, k) * M_(y, j, l) + M_(x, i, l) * M_(y, j, k) + M_(x, j, k) * M_(y, i, l) + \
                  M_(x, j, l) * M_(y, i, k) + M_(x, l, j) * M_(y, k, i) + M_(x, k, j) * M_(y, l, i) + \
                  M_(x, l, i) * M_(y, k, j) + M_(x, k, i) * M_(y, l, j))
#define PD(x, y) (M_(x, i, j) * M_(y, l, k) + M_(x, i, j) * M_(y, k, l) + M_(x, j, i) * M_(y, l, k) + \
                  M_(x, j, i) * M_(y, k, l) + M_(x, l, k) * M_(y, i, j) + M_(x, k, l) * M_(y, i, j) + \
                  M_(x, l, k) * M_(y, j, i) + M_(x, k, l) * M_(y, j, i))
  if (g->qf[0] != 0.0) df1 += g->sgnk * g->qf[0] * PB(CO12);
  if (g->qf[1] != 0.0) df1 += g->sgnk * g->qf[1] * PB(O21V);
  if (g->qf[2] != 0.0)
    df1 += g->sgnk * g->qf[2] *
           (-PC(BCO1, BO2V) - PC(BCO2, BO1V) - PC(BO2V, BCO1) - PC(BO1V, BCO2) + PD(BCO1, BO2V) + PD(BCO2, BO1V) +
            PD(BO2V, BCO1) + PD(BO1V, BCO2));
#undef PB
#undef PC
#undef PD
#undef M_
  return df1;
}

static int grad2_core(int nsh, int nbf, int natm, const int *am, const int *harm, const int *atom,
                      const int *nprim, const int *poff, const int *aooff, const double *ex,
                      const double *cc, const double *xyz, const double *Q, const double *d, double *de,
                      double hfscale, double coulscale, double cutoff, long *stats, const double *mrsf_ao,
                      const mrsf_gam_t *mg_in);

int oqo_grad2(int nsh, int nbf, int natm, const int *am, const int *harm, const int *atom,
              const int *nprim, const int *poff, const int *aooff, const double *ex,
              const double *cc, const double *xyz, const double *Q, const double *d, double *de,
              double hfscale, double coulscale, double cutoff, long *stats) {
  return grad2_core(nsh, nbf, natm, am, harm, atom, nprim, poff, aooff, ex, cc, xyz, Q, d, de, hfscale, coulscale,
                    cutoff, stats, NULL, NULL);
}

/* MRSF 2e gradient (mrsf_2e_grad -> grd2_driver with grd2_mrsf_compute_data_t,
 * tdhf_mrsf_gradient.F90:204-337): d, p are the raw alpha/beta SCF and
 * relaxed difference densities (nbf, nbf, 2) and spc the seven response
 * densities (7, nbf, nbf), all in Fortran order; init forms the sums and
 * differences (:341-353), each matrix goes to the Cartesian effective form
 * (:361-390), and Gamma is mrsf_gamma. */
int oqo_grad2_mrsf(int nsh, int nbf, int natm, const int *am, const int *harm, const int *atom,
                   const int *nprim, const int *poff, const int *aooff, const double *ex,
                   const double *cc, const double *xyz, const double *Q, const double *d, const double *p,
                   const double *spc, double *de, double hfscale, double hfscale2, double coulscale,
                   const double *spcscale, int mrst, double cutoff, long *stats) {
  const size_t n2 = (size_t)nbf * nbf;
  double *ao = (double *)malloc(sizeof(double) * 11 * n2);  /* row-major M(i, j) = ao[x n2 + i nbf + j] */
  for (long i = 0; i < nbf; ++i)
    for (long j = 0; j < nbf; ++j) {
      const size_t f = (size_t)i + (size_t)nbf * j, r = (size_t)i * nbf + j;
      const double da = d[f], db = d[f + n2], pa = p[f], pb = p[f + n2];
      ao[0 * n2 + r] = da + db;
      ao[1 * n2 + r] = da - db;
      ao[2 * n2 + r] = pa + pb;
      ao[3 * n2 + r] = pa - pb;
      /* spc(c, i, j): ball = 7, bo2v = 1, bo1v = 2, bco1 = 3, bco2 = 4, o21v = 5, co12 = 6 */
      static const int comp[7] = {6, 0, 1, 2, 3, 4, 5};
      for (int x = 0; x < 7; ++x) ao[(4 + x) * n2 + r] = spc[comp[x] + 7 * f];
    }
  mrsf_gam_t mg = {{0}, hfscale, hfscale2, coulscale, {spcscale[0], spcscale[1], spcscale[2]},
                   mrst == 3 ? -1.0 : 1.0};
  int rc = grad2_core(nsh, nbf, natm, am, harm, atom, nprim, poff, aooff, ex, cc, xyz, Q, NULL, de, hfscale,
                      coulscale, cutoff, stats, ao, &mg);
  free(ao);
  return rc;
}

static int grad2_core(int nsh, int nbf, int natm, const int *am, const int *harm, const int *atom,
                      const int *nprim, const int *poff, const int *aooff, const double *ex,
                      const double *cc, const double *xyz, const double *Q, const double *d, double *de,
                      double hfscale, double coulscale, double cutoff, long *stats, const double *mrsf_ao,
                      const mrsf_gam_t *mg_in) {
  build_tables();
  oqo_basis bs = {nsh, nbf, natm, am, harm, atom, nprim, poff, aooff, ex, cc, xyz};
  /* Cartesian offsets and Dc = X^T D X */
  int *coff = (int *)malloc(sizeof(int) * nsh);
  int ncf = 0;
  for (int s = 0; s < nsh; ++s) { coff[s] = ncf; ncf += NCART(am[s]); }
  double *Dc = (double *)calloc((size_t)ncf * ncf, sizeof(double));
  /* MRSF: the eleven matrices in the same Cartesian form (a general X^T M X) */
  double *Mc = NULL;
  mrsf_gam_t mg;
  if (mrsf_ao) {
    mg = *mg_in;
    Mc = (double *)calloc((size_t)11 * ncf * ncf, sizeof(double));
    double Ya[15 * 15], Yb[15 * 15];
    for (int x = 0; x < 11; ++x) {
      const double *M = mrsf_ao + (size_t)x * nbf * nbf;
      double *C = Mc + (size_t)x * ncf * ncf;
      for (int a = 0; a < nsh; ++a) {
        int na = nao_of(&bs, a), ca = NCART(am[a]);
        shell_xmap(am[a], harm[a], Ya);
        for (int b = 0; b < nsh; ++b) {
          int nb = nao_of(&bs, b), cb = NCART(am[b]);
          shell_xmap(am[b], harm[b], Yb);
          for (int p = 0; p < ca; ++p)
            for (int q = 0; q < cb; ++q) {
              double s = 0.0;
              for (int i = 0; i < na; ++i)
                for (int j = 0; j < nb; ++j)
                  s += Ya[i * ca + p] * M[(size_t)(aooff[a] + i) * nbf + aooff[b] + j] * Yb[j * cb + q];
              C[(size_t)(coff[a] + p) * ncf + coff[b] + q] = s;
            }
        }
      }
      mg.m[x] = C;
    }
  }
  double Xa[15 * 15], Xb[15 * 15];
  for (int a = 0; a < nsh && d; ++a) {
    int na = nao_of(&bs, a), ca = NCART(am[a]);
    shell_xmap(am[a], harm[a], Xa);
    for (int b = 0; b < nsh; ++b) {
      int nb = nao_of(&bs, b), cb = NCART(am[b]);
      shell_xmap(am[b], harm[b], Xb);
      for (int p = 0; p < ca; ++p)
        for (int q = 0; q < cb; ++q) {
          double s = 0.0;
          for (int i = 0; i < na; ++i)
            for (int j = 0; j < nb; ++j) {
              long I = aooff[a] + i, J = aooff[b] + j;
              long ij = I > J ? I * (I + 1) / 2 + J : J * (J + 1) / 2 + I;
              s += Xa[i * ca + p] * d[ij] * Xb[j * cb + q];
            }
          Dc[(size_t)(coff[a] + p) * ncf + coff[b] + q] = s;
        }
    }
  }
  double zbig = 0.0;
  int ntot = 0;
  for (int s = 0; s < nsh; ++s) ntot += nprim[s];
  for (int x = 0; x < ntot; ++x) if (ex[x] > zbig) zbig = ex[x];
  double dabcut = 1.0e-11;
  if (zbig > 1.0e6) dabcut /= 10;
  if (zbig > 1.0e7) dabcut /= 10;
  /* gdat%dabcut holds dabcut squared (grd2_rys.F90:146) */
  oqo_cut base = {dabcut, 1.0e-20, 1.0e-20, 20.0 * log(10.0), 0.0, dabcut * dabcut};
  int mp = max_prim(&bs);
  long skip1 = 0, skip2 = 0, nint = 0;
  int nthr = 1;
#ifdef _OPENMP
  nthr = omp_get_max_threads();
#endif
  double *dethr = (double *)calloc((size_t)nthr * natm * 3, sizeof(double));
#pragma omp parallel reduction(+ : skip1, skip2, nint)
  {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    double *del = dethr + (size_t)tid * natm * 3;
    work_t w;
    work_init(&w);
    ppair *bp = (ppair *)malloc(sizeof(ppair) * mp * mp), *bq = (ppair *)malloc(sizeof(ppair) * mp * mp);
    double *G = (double *)malloc(sizeof(double) * 10000);
    double *Rp = (double *)malloc(sizeof(double) * BLK), *Rm = (double *)malloc(sizeof(double) * BLK);
#pragma omp for schedule(dynamic, 1)
    for (int i = 0; i < nsh; ++i)
      for (int j = 0; j <= i; ++j) {
        int np = make_pairs(&bs, i, j, &base, bp);
        if (np == 0) continue;
        for (int k = 0; k <= i; ++k) {
          int lmx = (k == i) ? j : k;
          for (int l = 0; l <= lmx; ++l) {
            int nq = make_pairs(&bs, k, l, &base, bq);
            if (nq == 0) continue;
            double gmax = Q[i * nsh + j] * Q[k * nsh + l];
            if (gmax < cutoff) { ++skip1; continue; }
            int sh[4] = {i, j, k, l};
            int L[4] = {am[i], am[j], am[k], am[l]};
            int nc[4] = {NCART(L[0]), NCART(L[1]), NCART(L[2]), NCART(L[3])};
            /* Gamma in raw canonical Cartesians */
            double gm = 0.0;
            for (int a = 0; a < nc[0]; ++a)
              for (int b = 0; b < nc[1]; ++b)
                for (int c = 0; c < nc[2]; ++c)
                  for (int e = 0; e < nc[3]; ++e) {
                    size_t A = coff[i] + a, B_ = coff[j] + b, C = coff[k] + c, E = coff[l] + e;
                    double g = Mc ? mrsf_gamma(&mg, ncf, (long)A, (long)B_, (long)C, (long)E)
                                  : 4.0 * coulscale * Dc[A * ncf + B_] * Dc[C * ncf + E] -
                                        hfscale * (Dc[A * ncf + C] * Dc[B_ * ncf + E] + Dc[A * ncf + E] * Dc[B_ * ncf + C]);
                    G[((a * nc[1] + b) * nc[2] + c) * nc[3] + e] = g;
                    if (fabs(g) > gm) gm = fabs(g);
                  }
            if (gm * gmax < 0.5 * cutoff) { ++skip2; continue; }
            ++nint;
            oqo_cut cut = base;
            cut.dab2 = gm * gm;
            double deg = 1.0;
            if (i == j) deg *= 0.5;
            if (k == l) deg *= 0.5;
            if (i == k && j == l) deg *= 0.5;
            const double *cen[4] = {xyz + 3 * atom[i], xyz + 3 * atom[j], xyz + 3 * atom[k], xyz + 3 * atom[l]};
            double fd[4][3] = {{0}};
            for (int ctr = 0; ctr < 3; ++ctr) {   /* derivative centre A, B, C */
              int Lp[4] = {L[0], L[1], L[2], L[3]}, Lm[4] = {L[0], L[1], L[2], L[3]};
              Lp[ctr] += 1;
              Lm[ctr] -= 1;
              int okp = raw_block(Lp[0], Lp[1], Lp[2], Lp[3], cen[0], cen[1], cen[2], cen[3], bp, np, bq, nq,
                                  &cut, ctr + 1, &w, Rp);
              int okm = (Lm[ctr] >= 0) ? raw_block(Lm[0], Lm[1], Lm[2], Lm[3], cen[0], cen[1], cen[2], cen[3], bp,
                                                   np, bq, nq, &cut, 0, &w, Rm) : 0;
              if (!okp && !okm) continue;
              int ncp[4] = {NCART(Lp[0]), NCART(Lp[1]), NCART(Lp[2]), NCART(Lp[3])};
              int ncm[4] = {Lm[0] >= 0 ? NCART(Lm[0]) : 0, Lm[1] >= 0 ? NCART(Lm[1]) : 0,
                            Lm[2] >= 0 ? NCART(Lm[2]) : 0, Lm[3] >= 0 ? NCART(Lm[3]) : 0};
              int lev0 = NCUM(L[ctr] - 1 < 0 ? -1 : L[ctr] - 1);
              int levp = NCUM(L[ctr]);
              int levm = NCUM(L[ctr] - 2 < 0 ? -1 : L[ctr] - 2);
              for (int a = 0; a < nc[0]; ++a)
                for (int b = 0; b < nc[1]; ++b)
                  for (int c = 0; c < nc[2]; ++c)
                    for (int e = 0; e < nc[3]; ++e) {
                      double g = G[((a * nc[1] + b) * nc[2] + c) * nc[3] + e];
                      if (g == 0.0) continue;
                      int idx[4] = {a, b, c, e};
                      int gcomp = lev0 + idx[ctr];
                      for (int dd = 0; dd < 3; ++dd) {
                        int ip[4] = {a, b, c, e};
                        double v = 0.0;
                        if (okp) {
                          ip[ctr] = raise_(gcomp, dd) - levp;
                          v += Rp[((ip[0] * ncp[1] + ip[1]) * ncp[2] + ip[2]) * ncp[3] + ip[3]];
                        }
                        int nlow = comp_l[gcomp][dd];
                        if (okm && nlow > 0) {
                          int im[4] = {a, b, c, e};
                          im[ctr] = lower(gcomp, dd) - levm;
                          v -= nlow * Rm[((im[0] * ncm[1] + im[1]) * ncm[2] + im[2]) * ncm[3] + im[3]];
                        }
                        fd[ctr][dd] += g * v;
                      }
                    }
            }
            for (int dd = 0; dd < 3; ++dd) fd[3][dd] = -(fd[0][dd] + fd[1][dd] + fd[2][dd]);
            for (int ctr = 0; ctr < 4; ++ctr)
              for (int dd = 0; dd < 3; ++dd) del[3 * atom[sh[ctr]] + dd] += deg * fd[ctr][dd];
          }
        }
      }
    free(bp); free(bq); free(G); free(Rp); free(Rm);
    work_free(&w);
  }
  for (int x = 0; x < natm * 3; ++x) {
    double s = 0.0;
    for (int t = 0; t < nthr; ++t) s += dethr[(size_t)t * natm * 3 + x];
    de[x] = s;
  }
  free(dethr); free(Dc); free(coff); free(Mc);
  if (stats) { stats[0] = skip1; stats[1] = skip2; stats[2] = nint; }
  return 0;
}
