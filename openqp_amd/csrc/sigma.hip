// sigma.hip -- the MRSF-TDDFT sigma session of OpenQP's routec_sig seam
// (routec_sig_init / set_scale / iter / free, routec_sig.F90:27-57): for the
// new Davidson trial vectors X (MO occupied-virtual block) it returns
// (A - B) X in the same MO frame, replacing the native triple
// mrsfcbc -> int2_driver%run(int2_mrsf_data_t) -> mrsfmntoia + mrsfesum
// (tdhf_mrsf_energy.F90:686-790) for the singlet (kind 1) and triplet
// (kind 3) MRSF response of a two-open-shell ROHF reference (nocca =
// noccb + 2), Tamm-Dancoff, every spin-pair-coupling scale equal to the HF
// scale (the only case the reference sends to the seam,
// tdhf_mrsf_energy.F90:643-657).  The integral pass runs on the device
// (oqph_mrsf_int2, the default session's basis and geometry); the MO <-> AO
// transformations around it run on the host threads.
//
// Matrices are column-major nbf x nbf (Fortran order), M(i, j) = m[i + j nbf].
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/oqph.h"

namespace oqph {
oqph_session* default_session();
void set_error_text(const char* msg);
}  // namespace oqph

namespace {

struct SigState {
  bool ready = false;
  int nbf = 0, nocca = 0, noccb = 0, kind = 1;
  double scale = 1.0;
  std::vector<double> va, vb, fa, fb;
};
SigState g_sig;
std::mutex g_sig_mtx;
const double ISQ2 = 0.70710678118654752440;

inline size_t ix(int i, int j, int n) { return (size_t)i + (size_t)j * n; }

// C = A B (column-major, m x k times k x n)
void gemm_nn(int m, int n, int k, const double* A, int lda, const double* B, int ldb, double* C, int ldc) {
#pragma omp parallel for schedule(static)
  for (int j = 0; j < n; ++j) {
    double* c = C + (size_t)j * ldc;
    for (int i = 0; i < m; ++i) c[i] = 0.0;
    for (int l = 0; l < k; ++l) {
      const double b = B[l + (size_t)j * ldb];
      if (b == 0.0) continue;
      const double* a = A + (size_t)l * lda;
      for (int i = 0; i < m; ++i) c[i] += a[i] * b;
    }
  }
}
// C = A^T B (A is k x m)
void gemm_tn(int m, int n, int k, const double* A, int lda, const double* B, int ldb, double* C, int ldc) {
#pragma omp parallel for schedule(static)
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < m; ++i) {
      const double* a = A + (size_t)i * lda;
      const double* b = B + (size_t)j * ldb;
      double s = 0.0;
      for (int l = 0; l < k; ++l) s += a[l] * b[l];
      C[i + (size_t)j * ldc] = s;
    }
}
// y = M x (M column-major n x n), y = M^T x
void matvec(int n, const double* M, const double* x, double* y) {
  for (int i = 0; i < n; ++i) y[i] = 0.0;
  for (int j = 0; j < n; ++j) {
    const double xj = x[j];
    if (xj == 0.0) continue;
    for (int i = 0; i < n; ++i) y[i] += M[ix(i, j, n)] * xj;
  }
}
void matvec_t(int n, const double* M, const double* x, double* y) {
  for (int j = 0; j < n; ++j) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += M[ix(i, j, n)] * x[i];
    y[j] = s;
  }
}
// M += alpha u v^T
void outer_add(int n, double alpha, const double* u, const double* v, double* M) {
  for (int j = 0; j < n; ++j) {
    const double s = alpha * v[j];
    if (s == 0.0) continue;
    for (int i = 0; i < n; ++i) M[ix(i, j, n)] += u[i] * s;
  }
}

// iatogen (tdhf_lib.F90:480-498): trial vector -> nbf x nbf matrix, block
// (0..nocca-1, noccb..nbf-1) column-major
void iatogen(const SigState& g, const double* pv, double* X) {
  const int n = g.nbf;
  std::memset(X, 0, sizeof(double) * (size_t)n * n);
  size_t t = 0;
  for (int j = g.noccb; j < n; ++j)
    for (int i = 0; i < g.nocca; ++i) X[ix(i, j, n)] = pv[t++];
}

// mrsfcbc (tdhf_mrsf_lib.F90:940-1273): the seven AO transition densities
// bo2v, bo1v, bco1, bco2, o21v, co12, ball of X into out[c] (each n x n)
void cbc(const SigState& g, const double* X, double* const out[7]) {
  const int n = g.nbf, na = g.nocca, nb = g.noccb, l1 = na - 2, l2 = na - 1;
  const double *va = g.va.data(), *vb = g.vb.data();
  for (int c = 0; c < 7; ++c) std::memset(out[c], 0, sizeof(double) * (size_t)n * n);
  std::vector<double> t(n), t2(n);
  // t = vb(:, na:) X(row, na:)^T
  auto vvir = [&](int row, double* y) {
    for (int i = 0; i < n; ++i) y[i] = 0.0;
    for (int a = na; a < n; ++a) {
      const double x = X[ix(row, a, n)];
      if (x != 0.0)
        for (int i = 0; i < n; ++i) y[i] += vb[ix(i, a, n)] * x;
    }
  };
  // t = va(:, :nb) X(:nb, col)
  auto vocc = [&](int col, double* y) {
    for (int i = 0; i < n; ++i) y[i] = 0.0;
    for (int k = 0; k < nb; ++k) {
      const double x = X[ix(k, col, n)];
      if (x != 0.0)
        for (int i = 0; i < n; ++i) y[i] += va[ix(i, k, n)] * x;
    }
  };
  vvir(l2, t.data());
  outer_add(n, 1.0, va + ix(0, l2, n), t.data(), out[0]);  // bo2v
  outer_add(n, 1.0, t.data(), va + ix(0, l1, n), out[4]);  // o21v, first term
  vvir(l1, t.data());
  outer_add(n, 1.0, va + ix(0, l1, n), t.data(), out[1]);  // bo1v
  outer_add(n, -1.0, t.data(), va + ix(0, l2, n), out[4]);  // o21v, second term
  if (nb > 0) {
    vocc(l1, t.data());
    outer_add(n, 1.0, t.data(), vb + ix(0, l1, n), out[2]);   // bco1
    outer_add(n, 1.0, vb + ix(0, l2, n), t.data(), out[5]);   // co12, first term
    vocc(l2, t2.data());
    outer_add(n, 1.0, t2.data(), vb + ix(0, l2, n), out[3]);  // bco2
    outer_add(n, -1.0, vb + ix(0, l1, n), t2.data(), out[5]);  // co12, second term
  }
  double* ball = out[6];
  for (size_t e = 0; e < (size_t)n * n; ++e) ball[e] = out[0][e] + out[1][e] + out[2][e] + out[3][e];
  if (nb > 0) {
    // ball += va(:, :nb) (vb(:, na:) X(:nb, na:)^T)^T
    std::vector<double> T((size_t)n * nb, 0.0);  // T(mu, i) = sum_a vb(mu, a) X(i, a)
    for (int i = 0; i < nb; ++i)
      for (int a = na; a < n; ++a) {
        const double x = X[ix(i, a, n)];
        if (x != 0.0)
          for (int mu = 0; mu < n; ++mu) T[ix(mu, i, n)] += vb[ix(mu, a, n)] * x;
      }
#pragma omp parallel for schedule(static)
    for (int nu = 0; nu < n; ++nu)
      for (int i = 0; i < nb; ++i) {
        const double s = T[ix(nu, i, n)];
        if (s != 0.0)
          for (int mu = 0; mu < n; ++mu) ball[ix(mu, nu, n)] += va[ix(mu, i, n)] * s;
      }
  }
  const double *a1 = va + ix(0, l1, n), *a2 = va + ix(0, l2, n), *b1 = vb + ix(0, l1, n), *b2 = vb + ix(0, l2, n);
  const double x11 = X[ix(l1, l1, n)] * ISQ2;
  if (g.kind == 1) {
    outer_add(n, X[ix(l2, l1, n)], a2, b1, ball);
    outer_add(n, X[ix(l1, l2, n)], a1, b2, ball);
    outer_add(n, x11, a1, b1, ball);
    outer_add(n, -x11, a2, b2, ball);
  } else {
    outer_add(n, x11, a1, b1, ball);
    outer_add(n, x11, a2, b2, ball);
  }
}

// mrsfmntoia (tdhf_mrsf_lib.F90:1463-1735): the digested matrices
// f[0..6] = ado2v, ado1v, adco1, adco2, ao21v, aco12, agdlr back to the
// trial block, written to out (nocca x (nbf - noccb), column-major)
void mntoia(const SigState& g, double* const f[7], double* out) {
  const int n = g.nbf, na = g.nocca, nb = g.noccb, l1 = na - 2, l2 = na - 1;
  const double *va = g.va.data(), *vb = g.vb.data();
  std::vector<double> T((size_t)n * n), wrk((size_t)n * n), scr;
  gemm_nn(n, n, n, f[6], n, vb, n, T.data(), n);        // agdlr vb
  gemm_tn(n, n, n, va, n, T.data(), n, wrk.data(), n);   // va^T agdlr vb
  scr = wrk;
  std::vector<double> t(n), t2(n), y(n);
  // wrk(:na-2, l2) += va(:, :na-2)^T (ado1v vb(:, l2) + aco12 vb(:, l1))
  matvec(n, f[1], vb + ix(0, l2, n), t.data());
  matvec(n, f[5], vb + ix(0, l1, n), t2.data());
  for (int i = 0; i < n; ++i) t[i] += t2[i];
  for (int i = 0; i < na - 2; ++i) {
    double s = 0.0;
    for (int mu = 0; mu < n; ++mu) s += va[ix(mu, i, n)] * t[mu];
    wrk[ix(i, l2, n)] += s;
  }
  // wrk(:na-2, l1) += va(:, :na-2)^T (ado2v vb(:, l1) - aco12 vb(:, l2))
  matvec(n, f[0], vb + ix(0, l1, n), t.data());
  matvec(n, f[5], vb + ix(0, l2, n), t2.data());
  for (int i = 0; i < n; ++i) t[i] -= t2[i];
  for (int i = 0; i < na - 2; ++i) {
    double s = 0.0;
    for (int mu = 0; mu < n; ++mu) s += va[ix(mu, i, n)] * t[mu];
    wrk[ix(i, l1, n)] += s;
  }
  // wrk(l1, na:) += vb(:, na:)^T (adco2^T va(:, l1) + ao21v^T va(:, l2))
  matvec_t(n, f[3], va + ix(0, l1, n), t.data());
  matvec_t(n, f[4], va + ix(0, l2, n), t2.data());
  for (int i = 0; i < n; ++i) t[i] += t2[i];
  for (int a = na; a < n; ++a) {
    double s = 0.0;
    for (int mu = 0; mu < n; ++mu) s += vb[ix(mu, a, n)] * t[mu];
    wrk[ix(l1, a, n)] += s;
  }
  // wrk(l2, na:) += vb(:, na:)^T (adco1^T va(:, l2) - ao21v^T va(:, l1))
  matvec_t(n, f[2], va + ix(0, l2, n), t.data());
  matvec_t(n, f[4], va + ix(0, l1, n), t2.data());
  for (int i = 0; i < n; ++i) t[i] -= t2[i];
  for (int a = na; a < n; ++a) {
    double s = 0.0;
    for (int mu = 0; mu < n; ++mu) s += vb[ix(mu, a, n)] * t[mu];
    wrk[ix(l2, a, n)] += s;
  }
  if (g.kind == 1) {
    wrk[ix(l1, l1, n)] = (scr[ix(l1, l1, n)] - scr[ix(l2, l2, n)]) * ISQ2;
    wrk[ix(l2, l2, n)] = 0.0;
  } else {
    wrk[ix(l1, l1, n)] = (scr[ix(l1, l1, n)] + scr[ix(l2, l2, n)]) * ISQ2;
    wrk[ix(l2, l1, n)] = wrk[ix(l1, l2, n)] = wrk[ix(l2, l2, n)] = 0.0;
  }
  size_t q = 0;
  for (int j = nb; j < n; ++j)
    for (int i = 0; i < na; ++i) out[q++] = wrk[ix(i, j, n)];
}

// mrsfesum (tdhf_mrsf_lib.F90:1918-2036): the MO Fock part, added to out
void esum(const SigState& g, const double* X, double* out) {
  const int n = g.nbf, na = g.nocca, nb = g.noccb, l1 = na - 2, l2 = na - 1;
  const double *fij = g.fa.data(), *fab = g.fb.data();
  std::vector<double> scr(X, X + (size_t)n * n), w((size_t)n * n, 0.0);
  scr[ix(l1, l1, n)] = 0.0;
  scr[ix(l2, l2, n)] = 0.0;
  // w(:na, nb:) = scr(:na, nb:) fab(nb:, nb:)^T - fij(:na, :na) scr(:na, nb:)
#pragma omp parallel for schedule(static)
  for (int j = nb; j < n; ++j)
    for (int i = 0; i < na; ++i) {
      double s = 0.0;
      for (int b = nb; b < n; ++b) s += scr[ix(i, b, n)] * fab[ix(j, b, n)];
      for (int k = 0; k < na; ++k) s -= fij[ix(i, k, n)] * scr[ix(k, j, n)];
      w[ix(i, j, n)] = s;
    }
  const double xlr = X[ix(l1, l1, n)], p = g.kind == 1 ? -1.0 : 1.0;
  for (int j = nb; j < n; ++j) {
    w[ix(l1, j, n)] += fab[ix(j, l1, n)] * xlr * ISQ2;
    w[ix(l2, j, n)] += p * fab[ix(j, l2, n)] * xlr * ISQ2;
  }
  for (int i = 0; i < na; ++i) {
    w[ix(i, l1, n)] -= fij[ix(i, l1, n)] * xlr * ISQ2;
    w[ix(i, l2, n)] -= p * fij[ix(i, l2, n)] * xlr * ISQ2;
  }
  double d1 = 0.0, d2 = 0.0, d3 = 0.0, d4 = 0.0;
  for (int k = 0; k < na; ++k) {
    d1 += fij[ix(l1, k, n)] * scr[ix(k, l1, n)];
    d2 += fij[ix(l2, k, n)] * scr[ix(k, l2, n)];
  }
  for (int b = nb; b < n; ++b) {
    d3 += fab[ix(l1, b, n)] * scr[ix(l1, b, n)];
    d4 += fab[ix(l2, b, n)] * scr[ix(l2, b, n)];
  }
  const double dumn = -d1 - p * d2 + d3 + p * d4;
  w[ix(l1, l1, n)] = dumn * ISQ2 + xlr * (fab[ix(l1, l1, n)] + fab[ix(l2, l2, n)] - fij[ix(l1, l1, n)] -
                                          fij[ix(l2, l2, n)]) * 0.5;
  if (g.kind == 1) {
    w[ix(l2, l2, n)] = 0.0;
  } else {
    w[ix(l2, l1, n)] = w[ix(l1, l2, n)] = w[ix(l2, l2, n)] = 0.0;
  }
  size_t q = 0;
  for (int j = nb; j < n; ++j)
    for (int i = 0; i < na; ++i) out[q++] += w[ix(i, j, n)];
}

}  // namespace

extern "C" {

int routec_sig_init(const int* nbf, const double* mo_a, const double* mo_b, const double* fmo_a,
                    const double* fmo_b, const int* nocca, const int* noccb, const int* kind) {
  if (!nbf || !mo_a || !mo_b || !fmo_a || !fmo_b || !nocca || !noccb || !kind) return 4;
  oqph_session* s = oqph::default_session();
  if (!s) {
    oqph::set_error_text("routec_sig_init: no default session");
    return 1;
  }
  if (*nbf != oqph_session_nbf(s)) {
    oqph::set_error_text("routec_sig_init: nbf does not match the session basis");
    return 4;
  }
  if ((*kind != 1 && *kind != 3) || *nocca != *noccb + 2 || *noccb < 0 || *nocca > *nbf) {
    oqph::set_error_text("routec_sig_init: only the singlet / triplet MRSF response of two open shells");
    return 1;
  }
  std::lock_guard<std::mutex> lk(g_sig_mtx);
  const size_t n2 = (size_t)*nbf * *nbf;
  g_sig.nbf = *nbf;
  g_sig.nocca = *nocca;
  g_sig.noccb = *noccb;
  g_sig.kind = *kind;
  g_sig.scale = 1.0;
  g_sig.va.assign(mo_a, mo_a + n2);
  g_sig.vb.assign(mo_b, mo_b + n2);
  g_sig.fa.assign(fmo_a, fmo_a + n2);
  g_sig.fb.assign(fmo_b, fmo_b + n2);
  g_sig.ready = true;
  return 0;
}

void routec_sig_set_scale(const double* s) {
  if (!s) return;
  std::lock_guard<std::mutex> lk(g_sig_mtx);
  g_sig.scale = *s;
}

void routec_sig_iter(const double* bvec_mo, const int* nv_new, double* sigma_mo, int* info) {
  if (!info) return;
  *info = 1;
  if (!bvec_mo || !nv_new || !sigma_mo || *nv_new < 1) return;
  std::lock_guard<std::mutex> lk(g_sig_mtx);
  if (!g_sig.ready) {
    oqph::set_error_text("routec_sig_iter: no sigma session (routec_sig_init)");
    return;
  }
  oqph_session* s = oqph::default_session();
  if (!s || oqph_session_nbf(s) != g_sig.nbf) {
    oqph::set_error_text("routec_sig_iter: the default session changed");
    return;
  }
  const int n = g_sig.nbf, nv = *nv_new;
  const size_t n2 = (size_t)n * n, nt = (size_t)g_sig.nocca * (n - g_sig.noccb);
  std::vector<double> X(n2 * nv), comp(7 * n2), d3(7 * n2 * nv), f3(7 * n2 * nv);
  double* cp[7];
  for (int c = 0; c < 7; ++c) cp[c] = comp.data() + c * n2;
  for (int v = 0; v < nv; ++v) {
    iatogen(g_sig, bvec_mo + v * nt, X.data() + v * n2);
    cbc(g_sig, X.data() + v * n2, cp);
    // d3(v, c, mu, nu) in the reference's (nf, 7, nbf, nbf) Fortran order
    for (int c = 0; c < 7; ++c)
      for (size_t e = 0; e < n2; ++e) d3[(size_t)v + nv * (c + 7 * e)] = cp[c][e];
  }
  int rc = oqph_mrsf_int2(s, nv, 7, 4, d3.data(), f3.data(), g_sig.scale, g_sig.scale, nullptr);
  if (rc) {
    *info = rc;
    return;
  }
  for (int v = 0; v < nv; ++v) {
    for (int c = 0; c < 7; ++c) {
      const double sg = (g_sig.kind == 3 && c < 6) ? -1.0 : 1.0;  // tdhf_mrsf_energy.F90:756-761
      for (size_t e = 0; e < n2; ++e) cp[c][e] = sg * f3[(size_t)v + nv * (c + 7 * e)];
    }
    mntoia(g_sig, cp, sigma_mo + v * nt);
    esum(g_sig, X.data() + v * n2, sigma_mo + v * nt);
  }
  *info = 0;
}

void routec_sig_free(void) {
  std::lock_guard<std::mutex> lk(g_sig_mtx);
  g_sig = SigState{};
}

}  // extern "C"
