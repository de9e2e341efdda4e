// dispatch.hip -- class dispatch of the J/K and Schwarz kernels, the small
// per-build kernels, and the host copies of the Rys tables.
#define OQPH_WANT_RYS_DATA
#include "eri_device.h"

namespace oqph {

void launch_jk_0000(bool, const JKParams&, const ClassWork&, hipStream_t);
void launch_jk_1000(bool, const JKParams&, const ClassWork&, hipStream_t);
void launch_jk_1010(bool, const JKParams&, const ClassWork&, hipStream_t);
void launch_jk_1100(bool, const JKParams&, const ClassWork&, hipStream_t);
void launch_jk_1110(bool, const JKParams&, const ClassWork&, hipStream_t);
void launch_jk_1111(bool, const JKParams&, const ClassWork&, hipStream_t);
void launch_jk_2000(bool, const JKParams&, const ClassWork&, hipStream_t);
void launch_jk_2010(bool, const JKParams&, const ClassWork&, hipStream_t);
void launch_jk_2011(bool, const JKParams&, const ClassWork&, hipStream_t);
void launch_jk_2020(bool, const JKParams&, const ClassWork&, hipStream_t);
void launch_jk_2100(bool, const JKParams&, const ClassWork&, hipStream_t);
void launch_jk_2110(bool, const JKParams&, const ClassWork&, hipStream_t);
void launch_jk_2111(bool, const JKParams&, const ClassWork&, hipStream_t);
void launch_jk_2120(bool, const JKParams&, const ClassWork&, hipStream_t);
void launch_jk_2121(bool, const JKParams&, const ClassWork&, hipStream_t);
void launch_jk_2200(bool, const JKParams&, const ClassWork&, hipStream_t);
void launch_jk_2210(bool, const JKParams&, const ClassWork&, hipStream_t);
void launch_jk_2211(bool, const JKParams&, const ClassWork&, hipStream_t);
void launch_jk_2220(bool, const JKParams&, const ClassWork&, hipStream_t);
void launch_jk_2221(bool, const JKParams&, const ClassWork&, hipStream_t);
void launch_jk_2222(bool, const JKParams&, const ClassWork&, hipStream_t);
void launch_schwarz_00(bool, const ShellPair*, const PrimPair*, const int*, int, const double*, const double*, const double*, double, double*, hipStream_t);
void launch_schwarz_10(bool, const ShellPair*, const PrimPair*, const int*, int, const double*, const double*, const double*, double, double*, hipStream_t);
void launch_schwarz_11(bool, const ShellPair*, const PrimPair*, const int*, int, const double*, const double*, const double*, double, double*, hipStream_t);
void launch_schwarz_20(bool, const ShellPair*, const PrimPair*, const int*, int, const double*, const double*, const double*, double, double*, hipStream_t);
void launch_schwarz_21(bool, const ShellPair*, const PrimPair*, const int*, int, const double*, const double*, const double*, double, double*, hipStream_t);
void launch_schwarz_22(bool, const ShellPair*, const PrimPair*, const int*, int, const double*, const double*, const double*, double, double*, hipStream_t);

bool launch_jk(int la, int lb, int lc, int ld, bool pure, const JKParams& p, const ClassWork& cw, hipStream_t s) {
  if (cw.nwork == 0) return true;
  switch (la * 1000 + lb * 100 + lc * 10 + ld) {
    case 0: launch_jk_0000(pure, p, cw, s); return true;
    case 1000: launch_jk_1000(pure, p, cw, s); return true;
    case 1010: launch_jk_1010(pure, p, cw, s); return true;
    case 1100: launch_jk_1100(pure, p, cw, s); return true;
    case 1110: launch_jk_1110(pure, p, cw, s); return true;
    case 1111: launch_jk_1111(pure, p, cw, s); return true;
    case 2000: launch_jk_2000(pure, p, cw, s); return true;
    case 2010: launch_jk_2010(pure, p, cw, s); return true;
    case 2011: launch_jk_2011(pure, p, cw, s); return true;
    case 2020: launch_jk_2020(pure, p, cw, s); return true;
    case 2100: launch_jk_2100(pure, p, cw, s); return true;
    case 2110: launch_jk_2110(pure, p, cw, s); return true;
    case 2111: launch_jk_2111(pure, p, cw, s); return true;
    case 2120: launch_jk_2120(pure, p, cw, s); return true;
    case 2121: launch_jk_2121(pure, p, cw, s); return true;
    case 2200: launch_jk_2200(pure, p, cw, s); return true;
    case 2210: launch_jk_2210(pure, p, cw, s); return true;
    case 2211: launch_jk_2211(pure, p, cw, s); return true;
    case 2220: launch_jk_2220(pure, p, cw, s); return true;
    case 2221: launch_jk_2221(pure, p, cw, s); return true;
    case 2222: launch_jk_2222(pure, p, cw, s); return true;
    default: return false;
  }
}

bool launch_schwarz(int la, int lb, bool pure, const ShellPair* pairs, const PrimPair* prims, const int* list, int n,
                    const double* tab, const double* herm, const double* xf, double qcut, double* Q, hipStream_t s) {
  if (n == 0) return true;
  switch (la * 10 + lb) {
    case 0: launch_schwarz_00(pure, pairs, prims, list, n, tab, herm, xf, qcut, Q, s); return true;
    case 10: launch_schwarz_10(pure, pairs, prims, list, n, tab, herm, xf, qcut, Q, s); return true;
    case 11: launch_schwarz_11(pure, pairs, prims, list, n, tab, herm, xf, qcut, Q, s); return true;
    case 20: launch_schwarz_20(pure, pairs, prims, list, n, tab, herm, xf, qcut, Q, s); return true;
    case 21: launch_schwarz_21(pure, pairs, prims, list, n, tab, herm, xf, qcut, Q, s); return true;
    case 22: launch_schwarz_22(pure, pairs, prims, list, n, tab, herm, xf, qcut, Q, s); return true;
    default: return false;
  }
}

bool kernels_support(int lmax) { return lmax <= 2; }

__device__ __forceinline__ long long tri_idx_(int i, int j) {
  int a = i > j ? i : j, b = i > j ? j : i;
  return (long long)a * (a + 1) / 2 + b;
}

// shell-pair density maxima (shlden, source/integrals/int2.F90:999-1047)
__global__ void k_shell_density(const double* D, const int* ao_off, const int* nao, int nsh, double* dsh) {
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
    for (int b = 0; b < jmax; ++b) m = fmax(m, fabs(D[tri_idx_(x, ao_off[j] + b)]));
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

void launch_shell_density(const double* D, const int* ao_off, const int* nao, int nsh, double* dsh, hipStream_t s) {
  const long long np = (long long)nsh * (nsh + 1) / 2;
  const int bs = 256;
  hipLaunchKernelGGL(k_shell_density, dim3((unsigned)((np + bs - 1) / bs)), dim3(bs), 0, s, D, ao_off, nao, nsh, dsh);
}

void launch_max_reduce(const double* x, long long n, unsigned long long* out, hipStream_t s) {
  hipLaunchKernelGGL(k_max_reduce, dim3(256), dim3(256), 0, s, x, n, out);
}


const double* rys_table_host(size_t* n) {
  *n = RYS_TABLE_SIZE;
  return RYS_TABLE;
}
const double* rys_herm_host(size_t* n) {
  *n = sizeof(RYS_HERM) / sizeof(double);
  return RYS_HERM;
}


}  // namespace oqph
