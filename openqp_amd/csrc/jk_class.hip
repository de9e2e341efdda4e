// jk_class.hip -- instantiates the J/K kernels of one quartet class
// (XA XB | XC XD), both Cartesian and pure AO variants, and for the diagonal
// classes (XA XB | XA XB) the Schwarz kernel.  Built once per class with
// -DXA=.. -DXB=.. -DXC=.. -DXD=.. so the classes compile in parallel.
#include "jk_device.h"
#include "jkg_device.h"

#define OQPH_CAT2(a, b, c, d) a##b##c##d
#define OQPH_NAME(p, a, b, c, d) OQPH_CAT2(p, a, b, c) ## d

namespace oqph {

#define OQPH_JK_LAUNCHER_(a, b, c, d) launch_jk_##a##b##c##d
#define OQPH_JK_LAUNCHER(a, b, c, d) OQPH_JK_LAUNCHER_(a, b, c, d)
#define OQPH_SCH_LAUNCHER_(a, b) launch_schwarz_##a##b
#define OQPH_SCH_LAUNCHER(a, b) OQPH_SCH_LAUNCHER_(a, b)

void OQPH_JK_LAUNCHER(XA, XB, XC, XD)(bool pure, const JKParams& p, const WaveWork& ww, hipStream_t s) {
  const int bs = JK_BLOCK;
  const long long nb = (ww.nwaves * 64 + bs - 1) / bs;
  if (pure)
    hipLaunchKernelGGL((k_jk<XA, XB, XC, XD, true>), dim3((unsigned)nb), dim3(bs), 0, s, p, ww);
  else
    hipLaunchKernelGGL((k_jk<XA, XB, XC, XD, false>), dim3((unsigned)nb), dim3(bs), 0, s, p, ww);
}

#define OQPH_JKB_LAUNCHER_(a, b, c, d) launch_jkb_##a##b##c##d
#define OQPH_JKB_LAUNCHER(a, b, c, d) OQPH_JKB_LAUNCHER_(a, b, c, d)
#define OQPH_JKB_INFO_(a, b, c, d) jkb_info_##a##b##c##d
#define OQPH_JKB_INFO(a, b, c, d) OQPH_JKB_INFO_(a, b, c, d)

// Bra-row kernel of this class (not for the streamed large classes):
// K-row count and static LDS bytes, and the launch with `lds` bytes of K rows.
bool OQPH_JKB_INFO(XA, XB, XC, XD)(bool pure, int* nrow, size_t* static_lds, size_t lds_max) {
  if constexpr (Big<XA, XB, XC, XD>::value) {
    return false;
  } else {
    const void* f = pure ? (const void*)k_jkb<XA, XB, XC, XD, true> : (const void*)k_jkb<XA, XB, XC, XD, false>;
    hipFuncAttributes a;
    if (hipFuncGetAttributes(&a, f) != hipSuccess) return false;
    if (lds_max > a.sharedSizeBytes &&
        hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)(lds_max - a.sharedSizeBytes)) !=
            hipSuccess)
      return false;
    *nrow = pure ? AO<XA, true>::N + AO<XB, true>::N : AO<XA, false>::N + AO<XB, false>::N;
    *static_lds = a.sharedSizeBytes;
    return true;
  }
}

bool OQPH_JKB_LAUNCHER(XA, XB, XC, XD)(bool pure, const JKParams& p, const BraWork& bw, const ColMap& cm, size_t lds,
                                       hipStream_t s) {
  if constexpr (Big<XA, XB, XC, XD>::value) {
    return false;
  } else {
    const int bs = JKB<XA, XB, XC, XD>::threads;
    if (pure)
      hipLaunchKernelGGL((k_jkb<XA, XB, XC, XD, true>), dim3((unsigned)bw.nitem), dim3(bs), lds, s, p, bw, cm);
    else
      hipLaunchKernelGGL((k_jkb<XA, XB, XC, XD, false>), dim3((unsigned)bw.nitem), dim3(bs), lds, s, p, bw, cm);
    return true;
  }
}

#define OQPH_JKG_LAUNCHER_(a, b, c, d) launch_jkg_##a##b##c##d
#define OQPH_JKG_LAUNCHER(a, b, c, d) OQPH_JKG_LAUNCHER_(a, b, c, d)
#define OQPH_JKG_INFO_(a, b, c, d) jkg_info_##a##b##c##d
#define OQPH_JKG_INFO(a, b, c, d) OQPH_JKG_INFO_(a, b, c, d)

// Grouped bra-row kernel of this class: members per chunk (0 when the class
// has none, i.e. one member: k_jkb serves it) and static LDS bytes; raises
// the kernel's dynamic-LDS limit to lds_max.
int OQPH_JKG_INFO(XA, XB, XC, XD)(bool pure, size_t* static_lds, size_t lds_max) {
  if constexpr (Big<XA, XB, XC, XD>::value || JKG<XA, XB, XC, XD>::nm < 2) {
    return 0;
  } else {
    const void* f = pure ? (const void*)k_jkg<XA, XB, XC, XD, true> : (const void*)k_jkg<XA, XB, XC, XD, false>;
    hipFuncAttributes a;
    if (hipFuncGetAttributes(&a, f) != hipSuccess) return 0;
    if (a.localSizeBytes > OQPH_PRIVATE_BUDGET) return 0;
    if (lds_max > a.sharedSizeBytes &&
        hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)(lds_max - a.sharedSizeBytes)) !=
            hipSuccess)
      return 0;
    *static_lds = a.sharedSizeBytes;
    return JKG<XA, XB, XC, XD>::nm;
  }
}

bool OQPH_JKG_LAUNCHER(XA, XB, XC, XD)(bool pure, const JKParams& p, const BraWork& bw, const ColMap& cm,
                                       const ChunkView& cv, size_t lds, hipStream_t s) {
  if constexpr (Big<XA, XB, XC, XD>::value || JKG<XA, XB, XC, XD>::nm < 2) {
    return false;
  } else {
    const int bs = JKG<XA, XB, XC, XD>::threads;
    if (pure)
      hipLaunchKernelGGL((k_jkg<XA, XB, XC, XD, true>), dim3((unsigned)bw.nitem), dim3(bs), lds, s, p, bw, cm, cv);
    else
      hipLaunchKernelGGL((k_jkg<XA, XB, XC, XD, false>), dim3((unsigned)bw.nitem), dim3(bs), lds, s, p, bw, cm, cv);
    return true;
  }
}

#define OQPH_JK_RES_(a, b, c, d) jk_res_##a##b##c##d
#define OQPH_JK_RES(a, b, c, d) OQPH_JK_RES_(a, b, c, d)

static bool func_res(const void* f, KernelRes* r) {
  hipFuncAttributes a;
  if (hipFuncGetAttributes(&a, f) != hipSuccess) return false;
  r->private_bytes = (int)a.localSizeBytes;
  r->lds_bytes = (int)a.sharedSizeBytes;
  r->regs = a.numRegs;
  r->max_threads = a.maxThreadsPerBlock;
  return true;
}

// Compiled resources of this class's kernels: kind 0 = k_jk, 1 = k_jkb,
// 3 = k_schwarz (diagonal classes only), 4 = k_jkg.
bool OQPH_JK_RES(XA, XB, XC, XD)(int kind, bool pure, KernelRes* r) {
  if (kind == 0)
    return func_res(pure ? (const void*)k_jk<XA, XB, XC, XD, true> : (const void*)k_jk<XA, XB, XC, XD, false>, r);
  if (kind == 1) {
    if constexpr (Big<XA, XB, XC, XD>::value) {
      return false;
    } else {
      return func_res(pure ? (const void*)k_jkb<XA, XB, XC, XD, true> : (const void*)k_jkb<XA, XB, XC, XD, false>,
                      r);
    }
  }
  if (kind == 4) {
    if constexpr (Big<XA, XB, XC, XD>::value || JKG<XA, XB, XC, XD>::nm < 2) {
      return false;
    } else {
      return func_res(pure ? (const void*)k_jkg<XA, XB, XC, XD, true> : (const void*)k_jkg<XA, XB, XC, XD, false>,
                      r);
    }
  }
#if XA == XC && XB == XD
  if (kind == 3)
    return func_res(pure ? (const void*)k_schwarz<XA, XB, true> : (const void*)k_schwarz<XA, XB, false>, r);
#endif
  return false;
}

#if XA == XC && XB == XD
void OQPH_SCH_LAUNCHER(XA, XB)(bool pure, const ShellPair* pairs, const PrimJK* prims, const int* list, int n,
                               const double* tab, const double* xf, double qcut, double imu2, double* Q,
                               hipStream_t s) {
  const int bs = 64;
  if (pure)
    hipLaunchKernelGGL((k_schwarz<XA, XB, true>), dim3((n + bs - 1) / bs), dim3(bs), 0, s, pairs, prims, list, n, tab,
                       xf, qcut, imu2, Q);
  else
    hipLaunchKernelGGL((k_schwarz<XA, XB, false>), dim3((n + bs - 1) / bs), dim3(bs), 0, s, pairs, prims, list, n,
                       tab, xf, qcut, imu2, Q);
}
#endif

}  // namespace oqph
