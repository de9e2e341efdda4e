// hl.hip -- launchers of the run-time-class kernels (the classes without a
// class template) and the host-side sizing of their LDS layout.
#include <algorithm>

#include "hl_device.h"

namespace oqph {

// Extents and LDS layout of class (la lb|lc ld).  grad = 1: the gradient
// layout (2D integrals one level higher, Gamma block, 3 natm per-atom sums);
// else the quartet-block layout (acc / HRR / AO-transform ping-pong buffers).
HLDims hl_dims(int la, int lb, int lc, int ld, int pure, int grad, int natm) {
  HLDims h{};
  h.la = la; h.lb = lb; h.lc = lc; h.ld = ld;
  h.pure = pure ? 1 : 0;
  const int L = la + lb + lc + ld + (grad ? 1 : 0);
  h.nr = L / 2 + 1;
  h.ni = la + lb + 1 + (grad ? 1 : 0);
  h.nk = lc + ld + 1 + (grad ? 1 : 0);
  h.ne = ncum(la + lb) - ncum(la - 1);
  h.nf = ncum(lc + ld) - ncum(lc - 1);
  h.na = ncart(la); h.nb = ncart(lb); h.nc = ncart(lc); h.nd = ncart(ld);
  auto nao = [&](int l) { return (pure && l >= 2) ? 2 * l + 1 : ncart(l); };
  h.oa = nao(la); h.ob = nao(lb); h.oc = nao(lc); h.od = nao(ld);
  const int nG = 3 * h.nr * h.ni * h.nk;
  int sA = 0, sB = 0;
  if (grad) {
    sA = h.na * h.nb * h.nc * h.nd;
    sB = 3 * natm;
  } else {
    // the sequence of hl_block: acc in A, bra HRR into B, then every stage
    // writes the other buffer
    sA = h.ne * h.nf;
    sB = h.na * h.nb * h.nf;
    bool curB = true;
    int oa = h.na, ob = h.nb;
    auto write = [&](int n) {
      if (curB) sA = std::max(sA, n); else sB = std::max(sB, n);
      curB = !curB;
    };
    // (a transform that keeps the component count works in place)
    if (lb >= 2) { ob = h.ob; if (h.ob != h.nb) write(h.na * ob * h.nf); }
    if (la >= 2) { oa = h.oa; if (h.oa != h.na) write(oa * ob * h.nf); }
    write(oa * ob * h.nc * h.nd);
    if (ld >= 2 && h.od != h.nd) write(oa * ob * h.nc * h.od);
    if (lc >= 2 && h.oc != h.nc) write(oa * ob * h.oc * h.od);
  }
  // gradient: 2D integrals transferred to every exponent the derivatives use
  h.nt = grad ? (la + 2) * (lb + 2) * (lc + 2) * (ld + 1) : 0;
  const int nT = 3 * h.nt;
  const int nC = 123;  // HRR coefficient tables
  const int nI = h.ne + h.nf + h.na * h.nb + h.nc * h.nd + h.nt;  // int tables
  h.offG = 0;
  h.offA = nG;
  h.offB = h.offA + sA;
  h.offT = h.offB + sB;
  h.offC = h.offT + nT;
  h.offD = h.offC + nC;
  // J/K: D_ab, D_cd, D_ac, D_ad, D_bc, D_bd, each at most (largest AO count)^2
  const int omax = std::max(std::max(h.oa, h.ob), std::max(h.oc, h.od));
  h.dstr = omax * omax;
  h.n4 = grad ? (la + 1) * (lb + 1) * (lc + 1) * (ld + 1) : 0;
  h.offR = (h.offD + (grad ? 0 : 6 * h.dstr) + 3) & ~3;  // 32-byte aligned records
  h.offI = h.offR + 12 * h.n4;
  h.lds = (int)sizeof(double) * (h.offI + (nI + 1) / 2);
  return h;
}

template <typename K>
static void allow_lds(K kern, int bytes) {
  if (bytes > 64 * 1024) hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

// blocks of a run-time-class launch: 64 / kpb per 64-ket chunk
static unsigned hl_blocks(const WaveWork& ww) {
  const int kpb = ww.kpb > 0 ? ww.kpb : 64;
  return (unsigned)(ww.nwaves * (64 / kpb));
}

// p.Dg set: the response-pass instances
void launch_jkh(const HLDims& h, const JKParams& p, const WaveWork& ww, hipStream_t s) {
  if (ww.nwaves == 0) return;
  if (p.Dg) {
    allow_lds(k_jkh<true>, h.lds);
    hipLaunchKernelGGL(k_jkh<true>, dim3(hl_blocks(ww)), dim3(HL_BLOCK), h.lds, s, p, ww, h);
  } else {
    allow_lds(k_jkh<false>, h.lds);
    hipLaunchKernelGGL(k_jkh<false>, dim3(hl_blocks(ww)), dim3(HL_BLOCK), h.lds, s, p, ww, h);
  }
}

void launch_jkh_multi(const JKParams& p, const HLTaskList& tl, long long nblocks, int lds, hipStream_t s) {
  if (nblocks == 0) return;
  if (p.Dg) {
    allow_lds(k_jkh_multi<true>, lds);
    hipLaunchKernelGGL(k_jkh_multi<true>, dim3((unsigned)nblocks), dim3(HL_BLOCK), lds, s, p, tl);
  } else {
    allow_lds(k_jkh_multi<false>, lds);
    hipLaunchKernelGGL(k_jkh_multi<false>, dim3((unsigned)nblocks), dim3(HL_BLOCK), lds, s, p, tl);
  }
}

void launch_schwarzh(const HLDims& h, const ShellPair* pairs, const PrimJK* prims, const int* list, int n,
                     const double* rys, const double* xf, double qcut, double imu2, double* Q, hipStream_t s) {
  if (n == 0) return;
  allow_lds(k_schwarzh, h.lds);
  hipLaunchKernelGGL(k_schwarzh, dim3((unsigned)n), dim3(HL_BLOCK), h.lds, s, pairs, prims, list, n, rys, xf, qcut,
                     imu2, Q, h);
}

void launch_gradh_multi(const GradParams& p, const HLTaskList& tl, long long nblocks, int lds, hipStream_t s) {
  if (nblocks == 0) return;
  if (p.Mc) {  // the MRSF two-particle density
    allow_lds(k_gradh_multi<true>, lds);
    hipLaunchKernelGGL(k_gradh_multi<true>, dim3((unsigned)nblocks), dim3(HL_BLOCK), lds, s, p, tl);
  } else {
    allow_lds(k_gradh_multi<false>, lds);
    hipLaunchKernelGGL(k_gradh_multi<false>, dim3((unsigned)nblocks), dim3(HL_BLOCK), lds, s, p, tl);
  }
}

void launch_gradh(const HLDims& h, const GradParams& p, const WaveWork& ww, hipStream_t s) {
  if (ww.nwaves == 0) return;
  if (p.Mc) {
    allow_lds(k_gradh<true>, h.lds);
    hipLaunchKernelGGL(k_gradh<true>, dim3(hl_blocks(ww)), dim3(HL_BLOCK), h.lds, s, p, ww, h);
  } else {
    allow_lds(k_gradh<false>, h.lds);
    hipLaunchKernelGGL(k_gradh<false>, dim3(hl_blocks(ww)), dim3(HL_BLOCK), h.lds, s, p, ww, h);
  }
}

}  // namespace oqph
