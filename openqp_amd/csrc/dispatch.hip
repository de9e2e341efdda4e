// dispatch.hip -- class dispatch of the J/K and Schwarz kernels, the small
// per-build kernels, and the host copies of the Rys tables.
#define OQPH_WANT_RYS_DATA
#include "eri_device.h"
#include "grad_device.h"

namespace oqph {

void launch_jk_0000(bool, const JKParams&, const WaveWork&, hipStream_t);
void launch_jk_1000(bool, const JKParams&, const WaveWork&, hipStream_t);
void launch_jk_1010(bool, const JKParams&, const WaveWork&, hipStream_t);
void launch_jk_1100(bool, const JKParams&, const WaveWork&, hipStream_t);
void launch_jk_1110(bool, const JKParams&, const WaveWork&, hipStream_t);
void launch_jk_1111(bool, const JKParams&, const WaveWork&, hipStream_t);
void launch_jk_2000(bool, const JKParams&, const WaveWork&, hipStream_t);
void launch_jk_2010(bool, const JKParams&, const WaveWork&, hipStream_t);
void launch_jk_2011(bool, const JKParams&, const WaveWork&, hipStream_t);
void launch_jk_2020(bool, const JKParams&, const WaveWork&, hipStream_t);
void launch_jk_2100(bool, const JKParams&, const WaveWork&, hipStream_t);
void launch_jk_2110(bool, const JKParams&, const WaveWork&, hipStream_t);
void launch_jk_2111(bool, const JKParams&, const WaveWork&, hipStream_t);
void launch_jk_2120(bool, const JKParams&, const WaveWork&, hipStream_t);
void launch_jk_2121(bool, const JKParams&, const WaveWork&, hipStream_t);
void launch_jk_2200(bool, const JKParams&, const WaveWork&, hipStream_t);
void launch_jk_2210(bool, const JKParams&, const WaveWork&, hipStream_t);
void launch_jk_2211(bool, const JKParams&, const WaveWork&, hipStream_t);
void launch_jk_2220(bool, const JKParams&, const WaveWork&, hipStream_t);
void launch_jk_2221(bool, const JKParams&, const WaveWork&, hipStream_t);
void launch_jk_2222(bool, const JKParams&, const WaveWork&, hipStream_t);
void launch_jk_3000(bool, const JKParams&, const WaveWork&, hipStream_t);
void launch_jk_3010(bool, const JKParams&, const WaveWork&, hipStream_t);
void launch_jk_3011(bool, const JKParams&, const WaveWork&, hipStream_t);
void launch_jk_3020(bool, const JKParams&, const WaveWork&, hipStream_t);
void launch_jk_3021(bool, const JKParams&, const WaveWork&, hipStream_t);
void launch_jk_3030(bool, const JKParams&, const WaveWork&, hipStream_t);
void launch_jk_3100(bool, const JKParams&, const WaveWork&, hipStream_t);
void launch_jk_3110(bool, const JKParams&, const WaveWork&, hipStream_t);
void launch_jk_3120(bool, const JKParams&, const WaveWork&, hipStream_t);
void launch_jk_3200(bool, const JKParams&, const WaveWork&, hipStream_t);
void launch_jk_3210(bool, const JKParams&, const WaveWork&, hipStream_t);
void launch_jk_3300(bool, const JKParams&, const WaveWork&, hipStream_t);
bool launch_jkb_0000(bool, const JKParams&, const BraWork&, const ColMap&, size_t, hipStream_t);
bool jkb_info_0000(bool, int*, size_t*, size_t);
bool launch_jkb_1000(bool, const JKParams&, const BraWork&, const ColMap&, size_t, hipStream_t);
bool jkb_info_1000(bool, int*, size_t*, size_t);
bool launch_jkb_1010(bool, const JKParams&, const BraWork&, const ColMap&, size_t, hipStream_t);
bool jkb_info_1010(bool, int*, size_t*, size_t);
bool launch_jkb_1100(bool, const JKParams&, const BraWork&, const ColMap&, size_t, hipStream_t);
bool jkb_info_1100(bool, int*, size_t*, size_t);
bool launch_jkb_1110(bool, const JKParams&, const BraWork&, const ColMap&, size_t, hipStream_t);
bool jkb_info_1110(bool, int*, size_t*, size_t);
bool launch_jkb_1111(bool, const JKParams&, const BraWork&, const ColMap&, size_t, hipStream_t);
bool jkb_info_1111(bool, int*, size_t*, size_t);
bool launch_jkb_2000(bool, const JKParams&, const BraWork&, const ColMap&, size_t, hipStream_t);
bool jkb_info_2000(bool, int*, size_t*, size_t);
bool launch_jkb_2010(bool, const JKParams&, const BraWork&, const ColMap&, size_t, hipStream_t);
bool jkb_info_2010(bool, int*, size_t*, size_t);
bool launch_jkb_2011(bool, const JKParams&, const BraWork&, const ColMap&, size_t, hipStream_t);
bool jkb_info_2011(bool, int*, size_t*, size_t);
bool launch_jkb_2020(bool, const JKParams&, const BraWork&, const ColMap&, size_t, hipStream_t);
bool jkb_info_2020(bool, int*, size_t*, size_t);
bool launch_jkb_2100(bool, const JKParams&, const BraWork&, const ColMap&, size_t, hipStream_t);
bool jkb_info_2100(bool, int*, size_t*, size_t);
bool launch_jkb_2110(bool, const JKParams&, const BraWork&, const ColMap&, size_t, hipStream_t);
bool jkb_info_2110(bool, int*, size_t*, size_t);
bool launch_jkb_2111(bool, const JKParams&, const BraWork&, const ColMap&, size_t, hipStream_t);
bool jkb_info_2111(bool, int*, size_t*, size_t);
bool launch_jkb_2120(bool, const JKParams&, const BraWork&, const ColMap&, size_t, hipStream_t);
bool jkb_info_2120(bool, int*, size_t*, size_t);
bool launch_jkb_2121(bool, const JKParams&, const BraWork&, const ColMap&, size_t, hipStream_t);
bool jkb_info_2121(bool, int*, size_t*, size_t);
bool launch_jkb_2200(bool, const JKParams&, const BraWork&, const ColMap&, size_t, hipStream_t);
bool jkb_info_2200(bool, int*, size_t*, size_t);
bool launch_jkb_2210(bool, const JKParams&, const BraWork&, const ColMap&, size_t, hipStream_t);
bool jkb_info_2210(bool, int*, size_t*, size_t);
bool launch_jkb_2211(bool, const JKParams&, const BraWork&, const ColMap&, size_t, hipStream_t);
bool jkb_info_2211(bool, int*, size_t*, size_t);
bool launch_jkb_2220(bool, const JKParams&, const BraWork&, const ColMap&, size_t, hipStream_t);
bool jkb_info_2220(bool, int*, size_t*, size_t);
bool launch_jkb_2221(bool, const JKParams&, const BraWork&, const ColMap&, size_t, hipStream_t);
bool jkb_info_2221(bool, int*, size_t*, size_t);
bool launch_jkb_2222(bool, const JKParams&, const BraWork&, const ColMap&, size_t, hipStream_t);
bool launch_jkb_3000(bool, const JKParams&, const BraWork&, const ColMap&, size_t, hipStream_t);
bool launch_jkb_3010(bool, const JKParams&, const BraWork&, const ColMap&, size_t, hipStream_t);
bool launch_jkb_3011(bool, const JKParams&, const BraWork&, const ColMap&, size_t, hipStream_t);
bool launch_jkb_3020(bool, const JKParams&, const BraWork&, const ColMap&, size_t, hipStream_t);
bool launch_jkb_3021(bool, const JKParams&, const BraWork&, const ColMap&, size_t, hipStream_t);
bool launch_jkb_3030(bool, const JKParams&, const BraWork&, const ColMap&, size_t, hipStream_t);
bool launch_jkb_3100(bool, const JKParams&, const BraWork&, const ColMap&, size_t, hipStream_t);
bool launch_jkb_3110(bool, const JKParams&, const BraWork&, const ColMap&, size_t, hipStream_t);
bool launch_jkb_3120(bool, const JKParams&, const BraWork&, const ColMap&, size_t, hipStream_t);
bool launch_jkb_3200(bool, const JKParams&, const BraWork&, const ColMap&, size_t, hipStream_t);
bool launch_jkb_3210(bool, const JKParams&, const BraWork&, const ColMap&, size_t, hipStream_t);
bool launch_jkb_3300(bool, const JKParams&, const BraWork&, const ColMap&, size_t, hipStream_t);
bool jkb_info_2222(bool, int*, size_t*, size_t);
bool jkb_info_3000(bool, int*, size_t*, size_t);
bool jkb_info_3010(bool, int*, size_t*, size_t);
bool jkb_info_3011(bool, int*, size_t*, size_t);
bool jkb_info_3020(bool, int*, size_t*, size_t);
bool jkb_info_3021(bool, int*, size_t*, size_t);
bool jkb_info_3030(bool, int*, size_t*, size_t);
bool jkb_info_3100(bool, int*, size_t*, size_t);
bool jkb_info_3110(bool, int*, size_t*, size_t);
bool jkb_info_3120(bool, int*, size_t*, size_t);
bool jkb_info_3200(bool, int*, size_t*, size_t);
bool jkb_info_3210(bool, int*, size_t*, size_t);
bool jkb_info_3300(bool, int*, size_t*, size_t);
void launch_schwarz_00(bool, const ShellPair*, const PrimJK*, const int*, int, const double*, const double*, double, double*, hipStream_t);
void launch_schwarz_10(bool, const ShellPair*, const PrimJK*, const int*, int, const double*, const double*, double, double*, hipStream_t);
void launch_schwarz_11(bool, const ShellPair*, const PrimJK*, const int*, int, const double*, const double*, double, double*, hipStream_t);
void launch_schwarz_20(bool, const ShellPair*, const PrimJK*, const int*, int, const double*, const double*, double, double*, hipStream_t);
void launch_schwarz_21(bool, const ShellPair*, const PrimJK*, const int*, int, const double*, const double*, double, double*, hipStream_t);
void launch_schwarz_22(bool, const ShellPair*, const PrimJK*, const int*, int, const double*, const double*, double, double*, hipStream_t);

__device__ __forceinline__ long long tri_idx_(int i, int j) {
  int a = i > j ? i : j, b = i > j ? j : i;
  return (long long)a * (a + 1) / 2 + b;
}

bool launch_jk(int la, int lb, int lc, int ld, bool pure, const JKParams& p, const WaveWork& cw, hipStream_t s) {
  if (cw.nwaves == 0) return true;
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
    case 3000: launch_jk_3000(pure, p, cw, s); return true;
    case 3010: launch_jk_3010(pure, p, cw, s); return true;
    case 3011: launch_jk_3011(pure, p, cw, s); return true;
    case 3020: launch_jk_3020(pure, p, cw, s); return true;
    case 3021: launch_jk_3021(pure, p, cw, s); return true;
    case 3030: launch_jk_3030(pure, p, cw, s); return true;
    case 3100: launch_jk_3100(pure, p, cw, s); return true;
    case 3110: launch_jk_3110(pure, p, cw, s); return true;
    case 3120: launch_jk_3120(pure, p, cw, s); return true;
    case 3200: launch_jk_3200(pure, p, cw, s); return true;
    case 3210: launch_jk_3210(pure, p, cw, s); return true;
    case 3300: launch_jk_3300(pure, p, cw, s); return true;
    default: return false;
  }
}

bool launch_jkb(int la, int lb, int lc, int ld, bool pure, const JKParams& p, const BraWork& bw, const ColMap& cm,
                size_t lds, hipStream_t s) {
  if (bw.nitem == 0) return true;
  switch (la * 1000 + lb * 100 + lc * 10 + ld) {
    case 0: return launch_jkb_0000(pure, p, bw, cm, lds, s);
    case 1000: return launch_jkb_1000(pure, p, bw, cm, lds, s);
    case 1010: return launch_jkb_1010(pure, p, bw, cm, lds, s);
    case 1100: return launch_jkb_1100(pure, p, bw, cm, lds, s);
    case 1110: return launch_jkb_1110(pure, p, bw, cm, lds, s);
    case 1111: return launch_jkb_1111(pure, p, bw, cm, lds, s);
    case 2000: return launch_jkb_2000(pure, p, bw, cm, lds, s);
    case 2010: return launch_jkb_2010(pure, p, bw, cm, lds, s);
    case 2011: return launch_jkb_2011(pure, p, bw, cm, lds, s);
    case 2020: return launch_jkb_2020(pure, p, bw, cm, lds, s);
    case 2100: return launch_jkb_2100(pure, p, bw, cm, lds, s);
    case 2110: return launch_jkb_2110(pure, p, bw, cm, lds, s);
    case 2111: return launch_jkb_2111(pure, p, bw, cm, lds, s);
    case 2120: return launch_jkb_2120(pure, p, bw, cm, lds, s);
    case 2121: return launch_jkb_2121(pure, p, bw, cm, lds, s);
    case 2200: return launch_jkb_2200(pure, p, bw, cm, lds, s);
    case 2210: return launch_jkb_2210(pure, p, bw, cm, lds, s);
    case 2211: return launch_jkb_2211(pure, p, bw, cm, lds, s);
    case 2220: return launch_jkb_2220(pure, p, bw, cm, lds, s);
    case 2221: return launch_jkb_2221(pure, p, bw, cm, lds, s);
    case 2222: return launch_jkb_2222(pure, p, bw, cm, lds, s);
    case 3000: return launch_jkb_3000(pure, p, bw, cm, lds, s);
    case 3010: return launch_jkb_3010(pure, p, bw, cm, lds, s);
    case 3011: return launch_jkb_3011(pure, p, bw, cm, lds, s);
    case 3020: return launch_jkb_3020(pure, p, bw, cm, lds, s);
    case 3021: return launch_jkb_3021(pure, p, bw, cm, lds, s);
    case 3030: return launch_jkb_3030(pure, p, bw, cm, lds, s);
    case 3100: return launch_jkb_3100(pure, p, bw, cm, lds, s);
    case 3110: return launch_jkb_3110(pure, p, bw, cm, lds, s);
    case 3120: return launch_jkb_3120(pure, p, bw, cm, lds, s);
    case 3200: return launch_jkb_3200(pure, p, bw, cm, lds, s);
    case 3210: return launch_jkb_3210(pure, p, bw, cm, lds, s);
    case 3300: return launch_jkb_3300(pure, p, bw, cm, lds, s);
    default: return false;
  }
}

// false when the class has no bra-row kernel
bool jkb_info(int la, int lb, int lc, int ld, bool pure, int* nrow, size_t* static_lds, size_t lds_max) {
  switch (la * 1000 + lb * 100 + lc * 10 + ld) {
    case 0: return jkb_info_0000(pure, nrow, static_lds, lds_max);
    case 1000: return jkb_info_1000(pure, nrow, static_lds, lds_max);
    case 1010: return jkb_info_1010(pure, nrow, static_lds, lds_max);
    case 1100: return jkb_info_1100(pure, nrow, static_lds, lds_max);
    case 1110: return jkb_info_1110(pure, nrow, static_lds, lds_max);
    case 1111: return jkb_info_1111(pure, nrow, static_lds, lds_max);
    case 2000: return jkb_info_2000(pure, nrow, static_lds, lds_max);
    case 2010: return jkb_info_2010(pure, nrow, static_lds, lds_max);
    case 2011: return jkb_info_2011(pure, nrow, static_lds, lds_max);
    case 2020: return jkb_info_2020(pure, nrow, static_lds, lds_max);
    case 2100: return jkb_info_2100(pure, nrow, static_lds, lds_max);
    case 2110: return jkb_info_2110(pure, nrow, static_lds, lds_max);
    case 2111: return jkb_info_2111(pure, nrow, static_lds, lds_max);
    case 2120: return jkb_info_2120(pure, nrow, static_lds, lds_max);
    case 2121: return jkb_info_2121(pure, nrow, static_lds, lds_max);
    case 2200: return jkb_info_2200(pure, nrow, static_lds, lds_max);
    case 2210: return jkb_info_2210(pure, nrow, static_lds, lds_max);
    case 2211: return jkb_info_2211(pure, nrow, static_lds, lds_max);
    case 2220: return jkb_info_2220(pure, nrow, static_lds, lds_max);
    case 2221: return jkb_info_2221(pure, nrow, static_lds, lds_max);
    case 2222: return jkb_info_2222(pure, nrow, static_lds, lds_max);
    case 3000: return jkb_info_3000(pure, nrow, static_lds, lds_max);
    case 3010: return jkb_info_3010(pure, nrow, static_lds, lds_max);
    case 3011: return jkb_info_3011(pure, nrow, static_lds, lds_max);
    case 3020: return jkb_info_3020(pure, nrow, static_lds, lds_max);
    case 3021: return jkb_info_3021(pure, nrow, static_lds, lds_max);
    case 3030: return jkb_info_3030(pure, nrow, static_lds, lds_max);
    case 3100: return jkb_info_3100(pure, nrow, static_lds, lds_max);
    case 3110: return jkb_info_3110(pure, nrow, static_lds, lds_max);
    case 3120: return jkb_info_3120(pure, nrow, static_lds, lds_max);
    case 3200: return jkb_info_3200(pure, nrow, static_lds, lds_max);
    case 3210: return jkb_info_3210(pure, nrow, static_lds, lds_max);
    case 3300: return jkb_info_3300(pure, nrow, static_lds, lds_max);
    default: return false;
  }
}

bool launch_schwarz(int la, int lb, bool pure, const ShellPair* pairs, const PrimJK* prims, const int* list, int n,
                    const double* tab, const double* xf, double qcut, double* Q, hipStream_t s) {
  if (n == 0) return true;
  switch (la * 10 + lb) {
    case 0: launch_schwarz_00(pure, pairs, prims, list, n, tab, xf, qcut, Q, s); return true;
    case 10: launch_schwarz_10(pure, pairs, prims, list, n, tab, xf, qcut, Q, s); return true;
    case 11: launch_schwarz_11(pure, pairs, prims, list, n, tab, xf, qcut, Q, s); return true;
    case 20: launch_schwarz_20(pure, pairs, prims, list, n, tab, xf, qcut, Q, s); return true;
    case 21: launch_schwarz_21(pure, pairs, prims, list, n, tab, xf, qcut, Q, s); return true;
    case 22: launch_schwarz_22(pure, pairs, prims, list, n, tab, xf, qcut, Q, s); return true;
    default: return false;
  }
}

void launch_grad_0000(const GradParams&, const WaveWork&, hipStream_t);
void launch_grad_1000(const GradParams&, const WaveWork&, hipStream_t);
void launch_grad_1010(const GradParams&, const WaveWork&, hipStream_t);
void launch_grad_1100(const GradParams&, const WaveWork&, hipStream_t);
void launch_grad_1110(const GradParams&, const WaveWork&, hipStream_t);
void launch_grad_1111(const GradParams&, const WaveWork&, hipStream_t);
void launch_grad_2000(const GradParams&, const WaveWork&, hipStream_t);
void launch_grad_2010(const GradParams&, const WaveWork&, hipStream_t);
void launch_grad_2011(const GradParams&, const WaveWork&, hipStream_t);
void launch_grad_2020(const GradParams&, const WaveWork&, hipStream_t);
void launch_grad_2100(const GradParams&, const WaveWork&, hipStream_t);
void launch_grad_2110(const GradParams&, const WaveWork&, hipStream_t);
void launch_grad_2111(const GradParams&, const WaveWork&, hipStream_t);
void launch_grad_2120(const GradParams&, const WaveWork&, hipStream_t);
void launch_grad_2121(const GradParams&, const WaveWork&, hipStream_t);
void launch_grad_2200(const GradParams&, const WaveWork&, hipStream_t);
void launch_grad_2210(const GradParams&, const WaveWork&, hipStream_t);
void launch_grad_2211(const GradParams&, const WaveWork&, hipStream_t);
void launch_grad_2220(const GradParams&, const WaveWork&, hipStream_t);
void launch_grad_3000(const GradParams&, const WaveWork&, hipStream_t);
void launch_grad_3010(const GradParams&, const WaveWork&, hipStream_t);
void launch_grad_3011(const GradParams&, const WaveWork&, hipStream_t);
void launch_grad_3020(const GradParams&, const WaveWork&, hipStream_t);
void launch_grad_3030(const GradParams&, const WaveWork&, hipStream_t);
void launch_grad_3100(const GradParams&, const WaveWork&, hipStream_t);
void launch_grad_3110(const GradParams&, const WaveWork&, hipStream_t);
void launch_grad_3200(const GradParams&, const WaveWork&, hipStream_t);
void launch_grad_3300(const GradParams&, const WaveWork&, hipStream_t);

bool launch_grad(int la, int lb, int lc, int ld, const GradParams& p, const WaveWork& ww, hipStream_t s) {
  if (ww.nwaves == 0) return true;
  switch (la * 1000 + lb * 100 + lc * 10 + ld) {
    case 0: launch_grad_0000(p, ww, s); return true;
    case 1000: launch_grad_1000(p, ww, s); return true;
    case 1010: launch_grad_1010(p, ww, s); return true;
    case 1100: launch_grad_1100(p, ww, s); return true;
    case 1110: launch_grad_1110(p, ww, s); return true;
    case 1111: launch_grad_1111(p, ww, s); return true;
    case 2000: launch_grad_2000(p, ww, s); return true;
    case 2010: launch_grad_2010(p, ww, s); return true;
    case 2011: launch_grad_2011(p, ww, s); return true;
    case 2020: launch_grad_2020(p, ww, s); return true;
    case 2100: launch_grad_2100(p, ww, s); return true;
    case 2110: launch_grad_2110(p, ww, s); return true;
    case 2111: launch_grad_2111(p, ww, s); return true;
    case 2120: launch_grad_2120(p, ww, s); return true;
    case 2121: launch_grad_2121(p, ww, s); return true;
    case 2200: launch_grad_2200(p, ww, s); return true;
    case 2210: launch_grad_2210(p, ww, s); return true;
    case 2211: launch_grad_2211(p, ww, s); return true;
    case 2220: launch_grad_2220(p, ww, s); return true;
    case 3000: launch_grad_3000(p, ww, s); return true;
    case 3010: launch_grad_3010(p, ww, s); return true;
    case 3011: launch_grad_3011(p, ww, s); return true;
    case 3020: launch_grad_3020(p, ww, s); return true;
    case 3030: launch_grad_3030(p, ww, s); return true;
    case 3100: launch_grad_3100(p, ww, s); return true;
    case 3110: launch_grad_3110(p, ww, s); return true;
    case 3200: launch_grad_3200(p, ww, s); return true;
    case 3300: launch_grad_3300(p, ww, s); return true;
    default: return false;
  }
}

// Cartesian-effective density Dc = X^T D X, one thread per shell pair block
// (build_cart_density, source/basis_tools.F90:407-452, with the raw x^l
// normalised Cartesians of the derivative kernels)
__global__ void k_cart_density(const double* D, int nsh, const int* am, const int* pure, const int* ao_off,
                               const int* coff, const double* xf, int ncf, double* Dc) {
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
          s += xa * D[tri_idx_(ao_off[a] + i, ao_off[b] + j)] * xb;
        }
      }
      Dc[(long long)(coff[a] + p) * ncf + coff[b] + q] = s;
    }
}

void launch_cart_density(const double* D, int nsh, const int* am, const int* pure, const int* ao_off, const int* coff,
                         const double* xf, int ncf, double* Dc, hipStream_t s) {
  const long long n = (long long)nsh * nsh;
  hipLaunchKernelGGL(k_cart_density, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, D, nsh, am, pure, ao_off,
                     coff, xf, ncf, Dc);
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

__global__ void k_add(const double* x, const double* y, double* z, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    z[i] = x[i] + y[i];
}

// z = x + y (n doubles)
void launch_add(const double* x, const double* y, double* z, long long n, hipStream_t s) {
  hipLaunchKernelGGL(k_add, dim3(1024), dim3(256), 0, s, x, y, z, n);
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
