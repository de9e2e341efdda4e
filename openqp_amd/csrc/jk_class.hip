// jk_class.hip -- instantiates the J/K kernels of one quartet class
// (XA XB | XC XD), both Cartesian and pure AO variants, and for the diagonal
// classes (XA XB | XA XB) the Schwarz kernel.  Built once per class with
// -DXA=.. -DXB=.. -DXC=.. -DXD=.. so the classes compile in parallel.
#include <algorithm>

#include "jk_device.h"
#include "jkg_device.h"

#define OQPH_CAT2(a, b, c, d) a##b##c##d
#define OQPH_NAME(p, a, b, c, d) OQPH_CAT2(p, a, b, c) ## d

namespace oqph {

#define OQPH_JK_LAUNCHER_(a, b, c, d) launch_jk_##a##b##c##d
#define OQPH_JK_LAUNCHER(a, b, c, d) OQPH_JK_LAUNCHER_(a, b, c, d)
#define OQPH_SCH_LAUNCHER_(a, b) launch_schwarz_##a##b
#define OQPH_SCH_LAUNCHER(a, b) OQPH_SCH_LAUNCHER_(a, b)

// Each kernel comes in four variants: Cartesian or pure AOs, and closed-shell
// or open-shell pass (p.F2 set: the exchange of a second density as well).
template <bool TWO>
static void jk_launch(bool pure, const JKParams& p, const WaveWork& ww, hipStream_t s) {
  const int bs = JK_BLOCK;
  const long long nb = (ww.nwaves * 64 + bs - 1) / bs;
  if (pure)
    hipLaunchKernelGGL((k_jk<XA, XB, XC, XD, true, TWO>), dim3((unsigned)nb), dim3(bs), 0, s, p, ww);
  else
    hipLaunchKernelGGL((k_jk<XA, XB, XC, XD, false, TWO>), dim3((unsigned)nb), dim3(bs), 0, s, p, ww);
}

void OQPH_JK_LAUNCHER(XA, XB, XC, XD)(bool pure, const JKParams& p, const WaveWork& ww, hipStream_t s) {
  if (p.F2) jk_launch<true>(pure, p, ww, s);
  else jk_launch<false>(pure, p, ww, s);
}

static const void* jk_fn(int kind, bool pure, bool two) {
  if (kind == 0) {
    if (two) return pure ? (const void*)k_jk<XA, XB, XC, XD, true, true> : (const void*)k_jk<XA, XB, XC, XD, false, true>;
    return pure ? (const void*)k_jk<XA, XB, XC, XD, true, false> : (const void*)k_jk<XA, XB, XC, XD, false, false>;
  }
  if (kind == 1) {
    if constexpr (Big<XA, XB, XC, XD>::value) {
      return nullptr;
    } else {
      if (two)
        return pure ? (const void*)k_jkb<XA, XB, XC, XD, true, true> : (const void*)k_jkb<XA, XB, XC, XD, false, true>;
      return pure ? (const void*)k_jkb<XA, XB, XC, XD, true, false> : (const void*)k_jkb<XA, XB, XC, XD, false, false>;
    }
  }
  if (kind == 4) {
    if constexpr (Big<XA, XB, XC, XD>::value || JKG<XA, XB, XC, XD>::nm < 2) {
      return nullptr;
    } else {
      if (two)
        return pure ? (const void*)k_jkg<XA, XB, XC, XD, true, true, false>
                    : (const void*)k_jkg<XA, XB, XC, XD, false, true, false>;
      return pure ? (const void*)k_jkg<XA, XB, XC, XD, true, false, false>
                  : (const void*)k_jkg<XA, XB, XC, XD, false, false, false>;
    }
  }
  if (kind == 6) {
    if constexpr (Big<XA, XB, XC, XD>::value || JKT<XA, XB, XC, XD>::nm < 2) {
      return nullptr;
    } else {
      if (two)
        return pure ? (const void*)k_jkg<XA, XB, XC, XD, true, true, true>
                    : (const void*)k_jkg<XA, XB, XC, XD, false, true, true>;
      return pure ? (const void*)k_jkg<XA, XB, XC, XD, true, false, true>
                  : (const void*)k_jkg<XA, XB, XC, XD, false, false, true>;
    }
  }
  return nullptr;
}

// raises the dynamic-LDS limit of both variants of kernel `kind` to what
// lds_max leaves next to its static LDS; returns the static LDS bytes
static bool allow_dyn_lds(int kind, bool pure, size_t lds_max, size_t* static_lds) {
  for (int two = 0; two < 2; ++two) {
    const void* f = jk_fn(kind, pure, two != 0);
    hipFuncAttributes a;
    if (!f || hipFuncGetAttributes(&a, f) != hipSuccess) return false;
    if (lds_max > a.sharedSizeBytes &&
        hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)(lds_max - a.sharedSizeBytes)) !=
            hipSuccess)
      return false;
    *static_lds = std::max(*static_lds, (size_t)a.sharedSizeBytes);
  }
  return true;
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
    size_t st = 0;
    if (!allow_dyn_lds(1, pure, lds_max, &st)) return false;
    *nrow = pure ? AO<XA, true>::N + AO<XB, true>::N : AO<XA, false>::N + AO<XB, false>::N;
    *static_lds = st;
    return true;
  }
}

bool OQPH_JKB_LAUNCHER(XA, XB, XC, XD)(bool pure, const JKParams& p, const BraWork& bw, const ColMap& cm, size_t lds,
                                       hipStream_t s) {
  if constexpr (Big<XA, XB, XC, XD>::value) {
    return false;
  } else {
    const int bs = JKB<XA, XB, XC, XD>::threads;
    if (p.F2) {
      if (pure)
        hipLaunchKernelGGL((k_jkb<XA, XB, XC, XD, true, true>), dim3((unsigned)bw.nitem), dim3(bs), lds, s, p, bw, cm);
      else
        hipLaunchKernelGGL((k_jkb<XA, XB, XC, XD, false, true>), dim3((unsigned)bw.nitem), dim3(bs), lds, s, p, bw, cm);
    } else {
      if (pure)
        hipLaunchKernelGGL((k_jkb<XA, XB, XC, XD, true, false>), dim3((unsigned)bw.nitem), dim3(bs), lds, s, p, bw, cm);
      else
        hipLaunchKernelGGL((k_jkb<XA, XB, XC, XD, false, false>), dim3((unsigned)bw.nitem), dim3(bs), lds, s, p, bw,
                           cm);
    }
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
    for (int two = 0; two < 2; ++two) {
      hipFuncAttributes a;
      if (hipFuncGetAttributes(&a, jk_fn(4, pure, two != 0)) != hipSuccess) return 0;
      if (a.localSizeBytes > OQPH_PRIVATE_BUDGET) return 0;
    }
    size_t st = 0;
    if (!allow_dyn_lds(4, pure, lds_max, &st)) return 0;
    *static_lds = st;
    return JKG<XA, XB, XC, XD>::nm;
  }
}

bool OQPH_JKG_LAUNCHER(XA, XB, XC, XD)(bool pure, const JKParams& p, const BraWork& bw, const ColMap& cm,
                                       const ChunkView& cv, size_t lds, hipStream_t s) {
  if constexpr (Big<XA, XB, XC, XD>::value || JKG<XA, XB, XC, XD>::nm < 2) {
    return false;
  } else {
    const int bs = JKG<XA, XB, XC, XD>::threads;
    const dim3 g((unsigned)bw.nitem), b(bs);
    if (p.F2) {
      if (pure) hipLaunchKernelGGL((k_jkg<XA, XB, XC, XD, true, true, false>), g, b, lds, s, p, bw, cm, cv);
      else hipLaunchKernelGGL((k_jkg<XA, XB, XC, XD, false, true, false>), g, b, lds, s, p, bw, cm, cv);
    } else {
      if (pure) hipLaunchKernelGGL((k_jkg<XA, XB, XC, XD, true, false, false>), g, b, lds, s, p, bw, cm, cv);
      else hipLaunchKernelGGL((k_jkg<XA, XB, XC, XD, false, false, false>), g, b, lds, s, p, bw, cm, cv);
    }
    return true;
  }
}

#define OQPH_JKT_LAUNCHER_(a, b, c, d) launch_jkt_##a##b##c##d
#define OQPH_JKT_LAUNCHER(a, b, c, d) OQPH_JKT_LAUNCHER_(a, b, c, d)
#define OQPH_JKT_INFO_(a, b, c, d) jkt_info_##a##b##c##d
#define OQPH_JKT_INFO(a, b, c, d) OQPH_JKT_INFO_(a, b, c, d)

// Tiled bra-row kernel of this class (bra pairs sharing their first shell):
// members per tile (0: none) and static LDS bytes; raises its dynamic-LDS
// limit to lds_max.
int OQPH_JKT_INFO(XA, XB, XC, XD)(bool pure, size_t* static_lds, size_t lds_max) {
  if constexpr (Big<XA, XB, XC, XD>::value || JKT<XA, XB, XC, XD>::nm < 2) {
    return 0;
  } else {
    for (int two = 0; two < 2; ++two) {
      hipFuncAttributes a;
      if (hipFuncGetAttributes(&a, jk_fn(6, pure, two != 0)) != hipSuccess) return 0;
      if (a.localSizeBytes > OQPH_PRIVATE_BUDGET) return 0;
    }
    size_t st = 0;
    if (!allow_dyn_lds(6, pure, lds_max, &st)) return 0;
    *static_lds = st;
    return JKT<XA, XB, XC, XD>::nm;
  }
}

bool OQPH_JKT_LAUNCHER(XA, XB, XC, XD)(bool pure, const JKParams& p, const BraWork& bw, const ColMap& cm,
                                       const ChunkView& cv, size_t lds, hipStream_t s) {
  if constexpr (Big<XA, XB, XC, XD>::value || JKT<XA, XB, XC, XD>::nm < 2) {
    return false;
  } else {
    const dim3 g((unsigned)bw.nitem), b(JKG<XA, XB, XC, XD>::threads);
    if (p.F2) {
      if (pure) hipLaunchKernelGGL((k_jkg<XA, XB, XC, XD, true, true, true>), g, b, lds, s, p, bw, cm, cv);
      else hipLaunchKernelGGL((k_jkg<XA, XB, XC, XD, false, true, true>), g, b, lds, s, p, bw, cm, cv);
    } else {
      if (pure) hipLaunchKernelGGL((k_jkg<XA, XB, XC, XD, true, false, true>), g, b, lds, s, p, bw, cm, cv);
      else hipLaunchKernelGGL((k_jkg<XA, XB, XC, XD, false, false, true>), g, b, lds, s, p, bw, cm, cv);
    }
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
// 3 = k_schwarz (diagonal classes only), 4 = k_jkg, 6 = tiled k_jkg; kind |
// 16: the open-shell variant.
bool OQPH_JK_RES(XA, XB, XC, XD)(int kind, bool pure, KernelRes* r) {
  const bool two = (kind & 16) != 0;
  kind &= 15;
  if (kind == 0 || kind == 1 || kind == 4 || kind == 6) {
    const void* f = jk_fn(kind, pure, two);
    return f && func_res(f, r);
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
