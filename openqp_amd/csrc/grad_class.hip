// grad_class.hip -- instantiates the 2e-gradient kernel of one quartet class
// (XA XB | XC XD); built once per class with -DXA..-DXD.
#include "grad_device.h"

namespace oqph {

#define OQPH_GR_LAUNCHER_(a, b, c, d) launch_grad_##a##b##c##d
#define OQPH_GR_LAUNCHER(a, b, c, d) OQPH_GR_LAUNCHER_(a, b, c, d)

// f32: the FP32 form (quartets the host assigned to FP32)
void OQPH_GR_LAUNCHER(XA, XB, XC, XD)(const GradParams& p, const WaveWork& ww, hipStream_t s, bool f32) {
  const int bs = 256;  // four wavefronts
  const long long nb = (ww.nwaves + 3) / 4;
  if (f32)
    hipLaunchKernelGGL((k_grad<XA, XB, XC, XD, float>), dim3((unsigned)nb), dim3(bs), sizeof(double) * 3 * p.natm, s,
                       p, ww);
  else
    hipLaunchKernelGGL((k_grad<XA, XB, XC, XD, double>), dim3((unsigned)nb), dim3(bs), sizeof(double) * 3 * p.natm,
                       s, p, ww);
}

#define OQPH_GR_RES_(a, b, c, d) grad_res_##a##b##c##d
#define OQPH_GR_RES(a, b, c, d) OQPH_GR_RES_(a, b, c, d)

// Compiled resources of this class's gradient kernel (f32: its FP32 form).
bool OQPH_GR_RES(XA, XB, XC, XD)(KernelRes* r, bool f32) {
  hipFuncAttributes a;
  const void* f = f32 ? (const void*)k_grad<XA, XB, XC, XD, float> : (const void*)k_grad<XA, XB, XC, XD, double>;
  if (hipFuncGetAttributes(&a, f) != hipSuccess) return false;
  r->private_bytes = (int)a.localSizeBytes;
  r->lds_bytes = (int)a.sharedSizeBytes;
  r->regs = a.numRegs;
  r->max_threads = a.maxThreadsPerBlock;
  return true;
}

}  // namespace oqph
