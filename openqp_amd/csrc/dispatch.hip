// dispatch.hip -- class dispatch of the J/K and Schwarz kernels, the small
// per-build kernels, and the host copies of the Rys tables.
#define OQPH_WANT_RYS_DATA
#include "eri_device.h"
#include "grad_device.h"

namespace oqph {

// Quartet classes with class-templated kernels.  OQPH_JK_CLASSES: J/K
// (k_jk, k_jkb; jk_class.hip) -- the s/p/d classes and the small f classes;
// OQPH_GRAD_CLASSES: 2e gradient (k_grad; grad_class.hip).  The lists match
// CLASSES + F_JK and the gradient list of build.py (a CPU test checks it).
#define OQPH_JK_CLASSES(X)                                                                                      \
  X(0, 0, 0, 0) X(1, 0, 0, 0) X(1, 0, 1, 0) X(1, 1, 0, 0) X(1, 1, 1, 0) X(1, 1, 1, 1) X(2, 0, 0, 0) X(2, 0, 1, 0) \
  X(2, 0, 1, 1) X(2, 0, 2, 0) X(2, 1, 0, 0) X(2, 1, 1, 0) X(2, 1, 1, 1) X(2, 1, 2, 0) X(2, 1, 2, 1) X(2, 2, 0, 0) \
  X(2, 2, 1, 0) X(2, 2, 1, 1) X(2, 2, 2, 0) X(2, 2, 2, 1) X(2, 2, 2, 2) X(3, 0, 0, 0) X(3, 0, 1, 0) X(3, 0, 1, 1) \
  X(3, 0, 2, 0) X(3, 0, 2, 1) X(3, 0, 3, 0) X(3, 1, 0, 0) X(3, 1, 1, 0) X(3, 1, 2, 0) X(3, 2, 0, 0) X(3, 2, 1, 0) \
  X(3, 3, 0, 0)
#define OQPH_GRAD_CLASSES(X)                                                                                    \
  X(0, 0, 0, 0) X(1, 0, 0, 0) X(1, 0, 1, 0) X(1, 1, 0, 0) X(1, 1, 1, 0) X(1, 1, 1, 1) X(2, 0, 0, 0) X(2, 0, 1, 0) \
  X(2, 0, 1, 1) X(2, 0, 2, 0) X(2, 1, 0, 0) X(2, 1, 1, 0) X(2, 1, 1, 1) X(2, 1, 2, 0) X(2, 1, 2, 1) X(2, 2, 0, 0) \
  X(2, 2, 1, 0) X(2, 2, 1, 1) X(2, 2, 2, 0) X(3, 0, 0, 0) X(3, 0, 1, 0) X(3, 0, 1, 1) X(3, 0, 2, 0) X(3, 0, 3, 0) \
  X(3, 1, 0, 0) X(3, 1, 1, 0) X(3, 2, 0, 0) X(3, 3, 0, 0)
#define OQPH_SCHWARZ_CLASSES(X) X(0, 0) X(1, 0) X(1, 1) X(2, 0) X(2, 1) X(2, 2)

#define OQPH_CODE(a, b, c, d) ((a) * 1000 + (b) * 100 + (c) * 10 + (d))

#define OQPH_DECL_JK(a, b, c, d)                                                                            \
  void launch_jk_##a##b##c##d(bool, const JKParams&, const WaveWork&, hipStream_t);                        \
  bool launch_jkb_##a##b##c##d(bool, const JKParams&, const BraWork&, const ColMap&, size_t, hipStream_t); \
  bool launch_jkg_##a##b##c##d(bool, const JKParams&, const BraWork&, const ColMap&, const ChunkView&, size_t,  \
                               hipStream_t);                                                                \
  int jkg_info_##a##b##c##d(bool, size_t*, size_t);                                                         \
  bool launch_jkt_##a##b##c##d(bool, const JKParams&, const BraWork&, const ColMap&, const ChunkView&, size_t,  \
                               hipStream_t);                                                                \
  int jkt_info_##a##b##c##d(bool, size_t*, size_t);                                                         \
  bool jkb_info_##a##b##c##d(bool, int*, size_t*, size_t);                                                 \
  bool jk_res_##a##b##c##d(int, bool, KernelRes*);
OQPH_JK_CLASSES(OQPH_DECL_JK)
#define OQPH_DECL_GR(a, b, c, d)                                              \
  void launch_grad_##a##b##c##d(const GradParams&, const WaveWork&, hipStream_t, bool, bool); \
  bool grad_res_##a##b##c##d(KernelRes*, bool, bool);
OQPH_GRAD_CLASSES(OQPH_DECL_GR)
#define OQPH_DECL_SCH(a, b)                                                                                  \
  void launch_schwarz_##a##b(bool, const ShellPair*, const PrimJK*, const int*, int, const double*, const double*, \
                             double, double, double*, hipStream_t);
OQPH_SCHWARZ_CLASSES(OQPH_DECL_SCH)

__device__ __forceinline__ long long tri_idx_(int i, int j) {
  int a = i > j ? i : j, b = i > j ? j : i;
  return (long long)a * (a + 1) / 2 + b;
}

bool launch_jk(int la, int lb, int lc, int ld, bool pure, const JKParams& p, const WaveWork& cw, hipStream_t s) {
  if (cw.nwaves == 0) return true;
  switch (OQPH_CODE(la, lb, lc, ld)) {
#define OQPH_CASE(a, b, c, d) \
  case OQPH_CODE(a, b, c, d): launch_jk_##a##b##c##d(pure, p, cw, s); return true;
    OQPH_JK_CLASSES(OQPH_CASE)
#undef OQPH_CASE
    default: return false;
  }
}

bool launch_jkb(int la, int lb, int lc, int ld, bool pure, const JKParams& p, const BraWork& bw, const ColMap& cm,
                size_t lds, hipStream_t s) {
  if (bw.nitem == 0) return true;
  switch (OQPH_CODE(la, lb, lc, ld)) {
#define OQPH_CASE(a, b, c, d) \
  case OQPH_CODE(a, b, c, d): return launch_jkb_##a##b##c##d(pure, p, bw, cm, lds, s);
    OQPH_JK_CLASSES(OQPH_CASE)
#undef OQPH_CASE
    default: return false;
  }
}

bool launch_jkg(int la, int lb, int lc, int ld, bool pure, const JKParams& p, const BraWork& bw, const ColMap& cm,
                const ChunkView& cv, size_t lds, hipStream_t s) {
  if (bw.nitem == 0) return true;
  switch (OQPH_CODE(la, lb, lc, ld)) {
#define OQPH_CASE(a, b, c, d) \
  case OQPH_CODE(a, b, c, d): return launch_jkg_##a##b##c##d(pure, p, bw, cm, cv, lds, s);
    OQPH_JK_CLASSES(OQPH_CASE)
#undef OQPH_CASE
    default: return false;
  }
}

bool launch_jkt(int la, int lb, int lc, int ld, bool pure, const JKParams& p, const BraWork& bw, const ColMap& cm,
                const ChunkView& cv, size_t lds, hipStream_t s) {
  if (bw.nitem == 0) return true;
  switch (OQPH_CODE(la, lb, lc, ld)) {
#define OQPH_CASE(a, b, c, d) \
  case OQPH_CODE(a, b, c, d): return launch_jkt_##a##b##c##d(pure, p, bw, cm, cv, lds, s);
    OQPH_JK_CLASSES(OQPH_CASE)
#undef OQPH_CASE
    default: return false;
  }
}

// members per tile of the class's tiled bra-row kernel; 0: none
int jkt_info(int la, int lb, int lc, int ld, bool pure, size_t* static_lds, size_t lds_max) {
  switch (OQPH_CODE(la, lb, lc, ld)) {
#define OQPH_CASE(a, b, c, d) \
  case OQPH_CODE(a, b, c, d): return jkt_info_##a##b##c##d(pure, static_lds, lds_max);
    OQPH_JK_CLASSES(OQPH_CASE)
#undef OQPH_CASE
    default: return 0;
  }
}

// members per chunk of the class's grouped bra-row kernel; 0: none
int jkg_info(int la, int lb, int lc, int ld, bool pure, size_t* static_lds, size_t lds_max) {
  switch (OQPH_CODE(la, lb, lc, ld)) {
#define OQPH_CASE(a, b, c, d) \
  case OQPH_CODE(a, b, c, d): return jkg_info_##a##b##c##d(pure, static_lds, lds_max);
    OQPH_JK_CLASSES(OQPH_CASE)
#undef OQPH_CASE
    default: return 0;
  }
}

// false when the class has no bra-row kernel
bool jkb_info(int la, int lb, int lc, int ld, bool pure, int* nrow, size_t* static_lds, size_t lds_max) {
  switch (OQPH_CODE(la, lb, lc, ld)) {
#define OQPH_CASE(a, b, c, d) \
  case OQPH_CODE(a, b, c, d): return jkb_info_##a##b##c##d(pure, nrow, static_lds, lds_max);
    OQPH_JK_CLASSES(OQPH_CASE)
#undef OQPH_CASE
    default: return false;
  }
}

// Compiled resources of one class-templated kernel: kind 0 = k_jk, 1 = k_jkb,
// 2 = k_grad, 3 = k_schwarz (lc, ld ignored), 4 = k_jkg, 6 = tiled k_jkg; false when there is no such
// kernel.  hipFuncGetAttributes reads the code object, nothing is launched.
// kind | 16 (k_jk, k_jkb, k_jkg): the open-shell variant
bool kernel_resources(int kind, int la, int lb, int lc, int ld, bool pure, KernelRes* r) {
  const int base = kind & 15;
  if (base == 0 || base == 1 || base == 3 || base == 4 || base == 6) {
    if (base == 3) { lc = la; ld = lb; }
    switch (OQPH_CODE(la, lb, lc, ld)) {
#define OQPH_CASE(a, b, c, d) \
  case OQPH_CODE(a, b, c, d): return jk_res_##a##b##c##d(kind, pure, r);
      OQPH_JK_CLASSES(OQPH_CASE)
#undef OQPH_CASE
      default: return false;
    }
  }
  if (base == 2) {  // kind 2 | 16: the FP32 form
    switch (OQPH_CODE(la, lb, lc, ld)) {
#define OQPH_CASE(a, b, c, d) \
  case OQPH_CODE(a, b, c, d): return grad_res_##a##b##c##d(r, (kind & 16) != 0, (kind & 32) != 0);
      OQPH_GRAD_CLASSES(OQPH_CASE)
#undef OQPH_CASE
      default: return false;
    }
  }
  return false;
}

bool launch_schwarz(int la, int lb, bool pure, const ShellPair* pairs, const PrimJK* prims, const int* list, int n,
                    const double* tab, const double* xf, double qcut, double imu2, double* Q, hipStream_t s) {
  if (n == 0) return true;
  switch (la * 10 + lb) {
#define OQPH_CASE(a, b) \
  case (a) * 10 + (b): launch_schwarz_##a##b(pure, pairs, prims, list, n, tab, xf, qcut, imu2, Q, s); return true;
    OQPH_SCHWARZ_CLASSES(OQPH_CASE)
#undef OQPH_CASE
    default: return false;
  }
}

bool launch_grad(int la, int lb, int lc, int ld, const GradParams& p, const WaveWork& ww, hipStream_t s, bool f32,
                 bool mrsf) {
  if (ww.nwaves == 0) return true;
  switch (OQPH_CODE(la, lb, lc, ld)) {
#define OQPH_CASE(a, b, c, d) \
  case OQPH_CODE(a, b, c, d): launch_grad_##a##b##c##d(p, ww, s, f32, mrsf); return true;
    OQPH_GRAD_CLASSES(OQPH_CASE)
#undef OQPH_CASE
    default: return false;
  }
}

// Cartesian-effective density Dc = X^T D X, one thread per shell pair block
// (build_cart_density, source/basis_tools.F90:407-452, with the raw x^l
// normalised Cartesians of the derivative kernels)
// nbf > 0: D is a full row-major nbf x nbf matrix (not necessarily
// symmetric) instead of the packed triangle
__global__ void k_cart_density(const double* D, int nsh, const int* am, const int* pure, const int* ao_off,
                               const int* coff, const double* xf, int ncf, double* Dc, int nbf) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nsh * nsh) return;
  const int a = (int)(t / nsh), b = (int)(t % nsh);
  const int la = am[a], lb = am[b];
  const int ca = (la + 1) * (la + 2) / 2, cb = (lb + 1) * (lb + 2) / 2;
  const int na = (pure[a] && la >= 2) ? 2 * la + 1 : ca, nb = (pure[b] && lb >= 2) ? 2 * lb + 1 : cb;
  for (int p = 0; p < ca; ++p)
    for (int q = 0; q < cb; ++q) {
      double s = 0.0;
      for (int i = 0; i < na; ++i) {
        const double xa = la >= 2 ? xf[(((la - 2) * 2 + (pure[a] ? 1 : 0)) * 10 + i) * 10 + p] : (i == p ? 1.0 : 0.0);
        if (xa == 0.0) continue;
        for (int j = 0; j < nb; ++j) {
          const double xb = lb >= 2 ? xf[(((lb - 2) * 2 + (pure[b] ? 1 : 0)) * 10 + j) * 10 + q] : (j == q ? 1.0 : 0.0);
          if (xb == 0.0) continue;
          const int r = ao_off[a] + i, c = ao_off[b] + j;
          s += xa * (nbf > 0 ? D[(long long)r * nbf + c] : D[tri_idx_(r, c)]) * xb;
        }
      }
      Dc[(long long)(coff[a] + p) * ncf + coff[b] + q] = s;
    }
}

void launch_cart_density(const double* D, int nsh, const int* am, const int* pure, const int* ao_off, const int* coff,
                         const double* xf, int ncf, double* Dc, hipStream_t s) {
  const long long n = (long long)nsh * nsh;
  hipLaunchKernelGGL(k_cart_density, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, D, nsh, am, pure, ao_off,
                     coff, xf, ncf, Dc, 0);
}

// the same transform of a full (general) row-major nbf x nbf matrix
void launch_cart_matrix(const double* M, int nbf, int nsh, const int* am, const int* pure, const int* ao_off,
                        const int* coff, const double* xf, int ncf, double* Mc, hipStream_t s) {
  const long long n = (long long)nsh * nsh;
  hipLaunchKernelGGL(k_cart_density, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, M, nsh, am, pure, ao_off,
                     coff, xf, ncf, Mc, nbf);
}

bool kernels_support(int lmax) { return lmax <= 3; }

// classes with an f shell that have a class-templated kernel (the others run
// on the run-time-class kernels k_jkh / k_gradh); (dd|dd) also runs on
// k_jkh: its streamed template needs 12-14 KB of spill scratch per lane
bool jk_template(int la, int lb, int lc, int ld) {
  switch (la * 1000 + lb * 100 + lc * 10 + ld) {
    case 2222:
      return false;
    case 3000:
    case 3010:
    case 3011:
    case 3020:
    case 3021:
    case 3030:
    case 3100:
    case 3110:
    case 3120:
    case 3200:
    case 3210:
    case 3300:
      return true;
    default: return la < 3 && lc < 3;
  }
}
// (dd|dd) and (dd|dp) are also left to k_gradh: their templates would need
// 9-17 KB of spill scratch per lane, more than the runtime can always
// provide next to the other kernels' scratch
bool grad_template(int la, int lb, int lc, int ld) {
  switch (la * 1000 + lb * 100 + lc * 10 + ld) {
    case 2222:
    case 2221:
      return false;
    case 3000:
    case 3010:
    case 3011:
    case 3020:
    case 3030:
    case 3100:
    case 3110:
    case 3200:
    case 3300:
      return true;
    default: return la < 3 && lc < 3;
  }
}


// shell-pair density maxima (shlden, source/integrals/int2.F90:999-1047)
// shell-pair max |D| (shlden, int2.F90:999-1047); with D2 the max over both
// densities (nfocks = 2)
__global__ void k_shell_density(const double* D, const double* D2, const int* ao_off, const int* nao, int nsh,
                                double* dsh) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const long long np = (long long)nsh * (nsh + 1) / 2;
  if (t >= np) return;
  int i = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while ((long long)i * (i + 1) / 2 > t) --i;
  while ((long long)(i + 1) * (i + 2) / 2 <= t) ++i;
  const int j = t - i * (i + 1) / 2;
  double m = 0.0;
  for (int a = 0; a < nao[i]; ++a) {
    const int x = ao_off[i] + a;
    const int jmax = (i == j) ? a + 1 : nao[j];
    for (int b = 0; b < jmax; ++b) {
      const long long xy = tri_idx_(x, ao_off[j] + b);
      m = fmax(m, fabs(D[xy]));
      if (D2) m = fmax(m, fabs(D2[xy]));
    }
  }
  dsh[i * nsh + j] = m;
  dsh[j * nsh + i] = m;
}

__global__ void k_max_reduce(const double* x, long long n, unsigned long long* out) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double m = 0.0;
  for (; i < n; i += (long long)gridDim.x * blockDim.x) m = fmax(m, fabs(x[i]));
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_down(m, o));
  if (__lane_id() == 0) atomicMax(out, (unsigned long long)__double_as_longlong(m));
}

void launch_shell_density(const double* D, const double* D2, const int* ao_off, const int* nao, int nsh, double* dsh,
                          hipStream_t s) {
  const long long np = (long long)nsh * (nsh + 1) / 2;
  const int bs = 256;
  hipLaunchKernelGGL(k_shell_density, dim3((unsigned)((np + bs - 1) / bs)), dim3(bs), 0, s, D, D2, ao_off, nao, nsh,
                     dsh);
}

// p = c (a + b), m = c (a - b), elementwise over n doubles; p == a and m == b
// (in place) are allowed: each element is read before it is written
__global__ void k_sumdiff(const double* a, const double* b, double* p, double* m, double c, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const double x = a[i], y = b[i];
    p[i] = c * (x + y);
    m[i] = c * (x - y);
  }
}

void launch_sumdiff(const double* a, const double* b, double* p, double* m, double c, long long n, hipStream_t s) {
  hipLaunchKernelGGL(k_sumdiff, dim3(1024), dim3(256), 0, s, a, b, p, m, c, n);
}

// F += the ket-side Coulomb blocks: pair i's block (its AO block, a-major,
// entries [joff[i], joff[i + 1])) into the packed matrix, diagonal entries
// doubled as the J/K digestion doubles them.  The two entries (x, y) and
// (y, x) of a diagonal shell pair land on one packed element, hence atomics.
__global__ void k_jscatter(const ShellPair* pairs, const int* joff, int npair, int pure, const double* Jk, double* F) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npair) return;
  const ShellPair P = pairs[i];
  const int na = (pure && P.la >= 2) ? 2 * P.la + 1 : (P.la + 1) * (P.la + 2) / 2;
  const int nb = (pure && P.lb >= 2) ? 2 * P.lb + 1 : (P.lb + 1) * (P.lb + 2) / 2;
  const double* J = Jk + joff[i];
  for (int a = 0; a < na; ++a)
    for (int b = 0; b < nb; ++b) {
      const double v = J[a * nb + b];
      if (v == 0.0) continue;
      const int x = P.ao_a + a, y = P.ao_b + b;
      unsafeAtomicAdd(F + tri_idx_(x, y), x == y ? 2.0 * v : v);
    }
}

void launch_jscatter(const ShellPair* pairs, const int* joff, int npair, bool pure, const double* Jk, double* F,
                     hipStream_t s) {
  if (npair <= 0) return;
  hipLaunchKernelGGL(k_jscatter, dim3((unsigned)((npair + 255) / 256)), dim3(256), 0, s, pairs, joff, npair,
                     pure ? 1 : 0, Jk, F);
}

void launch_max_reduce(const double* x, long long n, unsigned long long* out, hipStream_t s) {
  hipLaunchKernelGGL(k_max_reduce, dim3(256), dim3(256), 0, s, x, n, out);
}


const double* rys_table_host(size_t* n) {
  *n = RYS_TABLE_SIZE;
  return RYS_TABLE;
}
const float* rys32_table_host(size_t* n) {
  *n = RYS32_TABLE_SIZE;
  return RYS32_TABLE;
}
const double* rys_herm_host(size_t* n) {
  *n = sizeof(RYS_HERM) / sizeof(double);
  return RYS_HERM;
}


}  // namespace oqph
