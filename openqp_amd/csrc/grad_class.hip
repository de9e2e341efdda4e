// grad_class.hip -- instantiates the 2e-gradient kernel of one quartet class
// (XA XB | XC XD); built once per class with -DXA..-DXD.
#include "grad_device.h"

namespace oqph {

#define OQPH_GR_LAUNCHER_(a, b, c, d) launch_grad_##a##b##c##d
#define OQPH_GR_LAUNCHER(a, b, c, d) OQPH_GR_LAUNCHER_(a, b, c, d)

// the MRSF form is instantiated for the classes up to 27 Cartesian
// components (its two-particle density is gathered per element)
constexpr bool MRSF_T = ncart(XA) * ncart(XB) * ncart(XC) * ncart(XD) <= 27;

// f32: the FP32 form (quartets the host assigned to FP32); mrsf: the MRSF
// two-particle density (FP64)
void OQPH_GR_LAUNCHER(XA, XB, XC, XD)(const GradParams& p, const WaveWork& ww, hipStream_t s, bool f32, bool mrsf) {
  const int bs = 256;  // four wavefronts
  const long long nb = (ww.nwaves + 3) / 4;
  if constexpr (MRSF_T)
    if (mrsf) {
      hipLaunchKernelGGL((k_grad<XA, XB, XC, XD, double, true>), dim3((unsigned)nb), dim3(bs),
                         sizeof(double) * 3 * p.natm, s, p, ww);
      return;
    }
  if (f32)
    hipLaunchKernelGGL((k_grad<XA, XB, XC, XD, float>), dim3((unsigned)nb), dim3(bs), sizeof(double) * 3 * p.natm, s,
                       p, ww);
  else
    hipLaunchKernelGGL((k_grad<XA, XB, XC, XD, double>), dim3((unsigned)nb), dim3(bs), sizeof(double) * 3 * p.natm,
                       s, p, ww);
}

#define OQPH_GR_RES_(a, b, c, d) grad_res_##a##b##c##d
#define OQPH_GR_RES(a, b, c, d) OQPH_GR_RES_(a, b, c, d)

// Compiled resources of this class's gradient kernel (f32: its FP32 form;
// mrsf: its MRSF form, false when the class has none).
bool OQPH_GR_RES(XA, XB, XC, XD)(KernelRes* r, bool f32, bool mrsf) {
  hipFuncAttributes a;
  const void* f = f32 ? (const void*)k_grad<XA, XB, XC, XD, float> : (const void*)k_grad<XA, XB, XC, XD, double>;
  if (mrsf) {
    if constexpr (MRSF_T) f = (const void*)k_grad<XA, XB, XC, XD, double, true>;
    else return false;
  }
  if (hipFuncGetAttributes(&a, f) != hipSuccess) return false;
  r->private_bytes = (int)a.localSizeBytes;
  r->lds_bytes = (int)a.sharedSizeBytes;
  r->regs = a.numRegs;
  r->max_threads = a.maxThreadsPerBlock;
  return true;
}

}  // namespace oqph
