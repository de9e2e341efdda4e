// xc.hip -- semilocal exchange-correlation build on the device: Vxc, E_xc and
// the grid electron count for the packed closed-shell or alpha/beta densities
// of OpenQP's Route-C Vxc seam (routec_vxc, source/modules/routec_bridge.F90:43-50;
// native path calc_dft_xc_density, source/scf_addons.F90:1857-1942 ->
// dmatd_density_blk, source/dftlib/dft_gridint_energy.F90:408-516).
//
// Grid: per atom a Treutler-Ahlrichs M4 radial grid (Chebyshev second kind,
// alpha = 0.6, per-element xi) times a Gauss-Legendre (cos theta) x uniform
// (phi) product angular grid, Stratmann-Scuseria-Frisch partition weights
// (a = 0.64) computed on the device, points with weight <= 1e-20 dropped.
// The kept points are sorted into cubic boxes (4 bohr) and cut into batches
// of at most 32768 points; a batch sees only the shells whose most diffuse
// primitive reaches its bounding box (|c| exp(-alpha r^2) >= 1e-12).
// Per-batch matrices are stored point-fastest (column-major npts x nb), so
// the basis-function and per-point kernels (one thread per point) coalesce.
//
// Per batch and spin: phi, grad phi (k_ao) -> X = phi D_b (rocBLAS dgemm on
// the gathered D_b) -> per point rho, grad rho, the functional and its
// derivatives (forward-mode dual numbers), Y = w (vrho/2 phi + v . grad phi)
// (k_xc_point) -> V_b = phi^T Y (dgemm) -> V += V_b + V_b^T (k_scatter),
// on three streams (lanes) with their own buffers.
// Functionals: Slater, Becke 88, VWN (Ceperley-Alder "VWN5" and RPA fits),
// Lee-Yang-Parr, combined as libxc combines them for the names the reference
// accepts (source/dftlib/libxc.F90:300-460).
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/oqph.h"
#include "ao_xform.h"
#include "session_view.h"

namespace oqph {
oqph_session* default_session();
void set_error_text(const char* msg);
}  // namespace oqph

namespace {

constexpr double PI = 3.14159265358979323846;
constexpr double RHO_CUT = 1e-10;    // points below this total density contribute nothing
constexpr double RHO_FLOOR = 1e-30;  // spin densities are floored here
constexpr double SSF_A = 0.64;
constexpr double BOX = 4.0;          // bohr
constexpr int MAX_BATCH = 32768;
constexpr int KCHUNK = 2048;         // points per partial V_b (split of the long GEMM dimension)
constexpr double AO_EPS = 1e-12;

void err(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
void err(const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  oqph::set_error_text(buf);
}

// ------------------------------------------------------------ functionals
enum Comp { C_LDAX = 0, C_B88, C_VWN5, C_VWNRPA, C_LYP };

struct XCFunc {
  int n;
  int comp[4];
  double w[4];
  double hfex;
  int gga;
};

bool lookup(const char* name, XCFunc* f) {
  std::string s(name ? name : "");
  for (auto& c : s) c = (char)tolower((unsigned char)c);
  if (s == "b3lypv5") s = "b3lyp5";
  if (s == "b3lypv1r") s = "b3lyp";
  if (s == "bhandhlyp") s = "bhhlyp";
  auto set = [&](std::initializer_list<std::pair<int, double>> cs, double hf) {
    f->n = 0;
    f->gga = 0;
    for (auto& c : cs) {
      f->comp[f->n] = c.first;
      f->w[f->n] = c.second;
      if (c.first == C_B88 || c.first == C_LYP) f->gga = 1;
      ++f->n;
    }
    f->hfex = hf;
  };
  if (s == "slater") set({{C_LDAX, 1.0}}, 0.0);
  else if (s == "svwn5") set({{C_LDAX, 1.0}, {C_VWN5, 1.0}}, 0.0);
  else if (s == "blyp") set({{C_B88, 1.0}, {C_LYP, 1.0}}, 0.0);
  else if (s == "b3lyp5") set({{C_LDAX, 0.08}, {C_B88, 0.72}, {C_VWN5, 0.19}, {C_LYP, 0.81}}, 0.20);
  else if (s == "b3lyp") set({{C_LDAX, 0.08}, {C_B88, 0.72}, {C_VWNRPA, 0.19}, {C_LYP, 0.81}}, 0.20);
  else if (s == "bhhlyp") set({{C_B88, 0.5}, {C_LYP, 1.0}}, 0.5);
  else return false;
  return true;
}

// forward-mode dual number over (rho_a, rho_b, sigma_aa, sigma_ab, sigma_bb)
struct Dual {
  double v, d[5];
};
__device__ inline Dual mk(double c) {
  Dual r;
  r.v = c;
  for (int k = 0; k < 5; ++k) r.d[k] = 0.0;
  return r;
}
__device__ inline Dual var(double x, int i) {
  Dual r = mk(x);
  r.d[i] = 1.0;
  return r;
}
__device__ inline Dual operator+(const Dual& a, const Dual& b) {
  Dual r;
  r.v = a.v + b.v;
  for (int k = 0; k < 5; ++k) r.d[k] = a.d[k] + b.d[k];
  return r;
}
__device__ inline Dual operator-(const Dual& a, const Dual& b) {
  Dual r;
  r.v = a.v - b.v;
  for (int k = 0; k < 5; ++k) r.d[k] = a.d[k] - b.d[k];
  return r;
}
__device__ inline Dual operator*(const Dual& a, const Dual& b) {
  Dual r;
  r.v = a.v * b.v;
  for (int k = 0; k < 5; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
  return r;
}
__device__ inline Dual operator/(const Dual& a, const Dual& b) {
  Dual r;
  const double ib = 1.0 / b.v;
  r.v = a.v * ib;
  for (int k = 0; k < 5; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) * ib;
  return r;
}
__device__ inline Dual operator*(double c, const Dual& a) {
  Dual r;
  r.v = c * a.v;
  for (int k = 0; k < 5; ++k) r.d[k] = c * a.d[k];
  return r;
}
__device__ inline Dual operator+(double c, const Dual& a) {
  Dual r = a;
  r.v += c;
  return r;
}
__device__ inline Dual operator-(double c, const Dual& a) { return c + (-1.0) * a; }
__device__ inline Dual operator/(double c, const Dual& a) { return mk(c) / a; }
__device__ inline Dual chain(const Dual& a, double f, double fp) {
  Dual r;
  r.v = f;
  for (int k = 0; k < 5; ++k) r.d[k] = fp * a.d[k];
  return r;
}
__device__ inline Dual dpow(const Dual& a, double p) { return chain(a, pow(a.v, p), p * pow(a.v, p - 1.0)); }
__device__ inline Dual dexp(const Dual& a) {
  const double e = exp(a.v);
  return chain(a, e, e);
}
__device__ inline Dual dlog(const Dual& a) { return chain(a, log(a.v), 1.0 / a.v); }
__device__ inline Dual dsqrt(const Dual& a) {
  const double s = sqrt(a.v);
  return chain(a, s, 0.5 / s);
}
__device__ inline Dual datan(const Dual& a) { return chain(a, atan(a.v), 1.0 / (1.0 + a.v * a.v)); }
// g(y) = x asinh(x) with x = sqrt(y); g'(y) = asinh(x) / (2x) + 1 / (2 sqrt(1 + y)), -> 1 at y = 0
__device__ inline Dual dxasinhx(const Dual& y) {
  const double x = sqrt(y.v), as = asinh(x);
  const double g1 = (x > 1e-8 ? as / (2.0 * x) : 0.5) + 0.5 / sqrt(1.0 + y.v);
  return chain(y, x * as, g1);
}

// (3/4) (6/pi)^(1/3)
__device__ inline double cx() { return 0.75 * cbrt(6.0 / PI); }

__device__ Dual lda_x(const Dual& ra, const Dual& rb) { return (-cx()) * (dpow(ra, 4.0 / 3.0) + dpow(rb, 4.0 / 3.0)); }

__device__ Dual b88(const Dual& ra, const Dual& rb, const Dual& saa, const Dual& sbb) {
  const double beta = 0.0042;
  Dual out = mk(0.0);
  for (int s = 0; s < 2; ++s) {
    const Dual& r = s ? rb : ra;
    const Dual& sg = s ? sbb : saa;
    const Dual r43 = dpow(r, 4.0 / 3.0);
    const Dual y = sg / (r43 * r43);  // x^2
    out = out - cx() * r43 - beta * (r43 * y) / (1.0 + (6.0 * beta) * dxasinhx(y));
  }
  return out;
}

__device__ Dual vwn_aux(double A, double b, double c, double x0, const Dual& rs) {
  const Dual x = dsqrt(rs);
  const Dual X = x * x + b * x + mk(c);
  const double X0 = x0 * x0 + b * x0 + c, Q = sqrt(4.0 * c - b * b);
  const Dual at = datan(Q / (2.0 * x + mk(b)));
  const Dual xm = x + mk(-x0);
  return A * (dlog(x * x / X) + (2.0 * b / Q) * at -
              (b * x0 / X0) * (dlog(xm * xm / X) + (2.0 * (b + 2.0 * x0) / Q) * at));
}

__device__ Dual vwn(const Dual& ra, const Dual& rb, bool rpa) {
  const Dual r = ra + rb;
  const Dual rs = cbrt(3.0 / (4.0 * PI)) * dpow(r, -1.0 / 3.0);
  const Dual z = (ra - rb) / r;
  const double den = pow(2.0, 4.0 / 3.0) - 2.0;
  const Dual fz = (1.0 / den) * (dpow(1.0 + z, 4.0 / 3.0) + dpow(1.0 - z, 4.0 / 3.0) + mk(-2.0));
  if (rpa) {
    const Dual eP = vwn_aux(0.0310907, 13.0720, 42.7198, -0.409286, rs);
    const Dual eF = vwn_aux(0.01554535, 20.1231, 101.578, -0.743294, rs);
    return r * (eP * (1.0 - fz) + eF * fz);
  }
  const Dual eP = vwn_aux(0.0310907, 3.72744, 12.9352, -0.10498, rs);
  const Dual eF = vwn_aux(0.01554535, 7.06042, 18.0578, -0.32500, rs);
  const Dual ac = vwn_aux(-1.0 / (6.0 * PI * PI), 1.13107, 13.0045, -0.0047584, rs);
  const double fpp0 = 4.0 / (9.0 * (cbrt(2.0) - 1.0));
  const Dual z2 = z * z, z4 = z2 * z2;
  return r * (eP + (1.0 / fpp0) * (ac * fz * (1.0 - z4)) + (eF - eP) * fz * z4);
}

__device__ Dual lyp(const Dual& ra, const Dual& rb, const Dual& saa, const Dual& sab, const Dual& sbb) {
  const double a = 0.04918, b = 0.132, c = 0.2533, d = 0.349;
  const Dual r = ra + rb;
  const Dual rm = dpow(r, -1.0 / 3.0);
  const Dual den = 1.0 + d * rm;
  const Dual om = dexp((-c) * rm) / den * dpow(r, -11.0 / 3.0);
  const Dual dl = c * rm + d * rm / den;
  const double cf = 0.3 * pow(3.0 * PI * PI, 2.0 / 3.0);
  const Dual s = saa + 2.0 * sab + sbb;
  Dual t = ra * rb *
           ((pow(2.0, 11.0 / 3.0) * cf) * (dpow(ra, 8.0 / 3.0) + dpow(rb, 8.0 / 3.0)) +
            (47.0 / 18.0 - (7.0 / 18.0) * dl) * s - (2.5 - (1.0 / 18.0) * dl) * (saa + sbb) -
            (1.0 / 9.0) * (dl + mk(-11.0)) * (ra / r * saa + rb / r * sbb));
  const Dual r2 = r * r;
  t = t - (2.0 / 3.0) * r2 * s + ((2.0 / 3.0) * r2 - ra * ra) * sbb + ((2.0 / 3.0) * r2 - rb * rb) * saa;
  return (-4.0 * a) * (ra * rb / (r * den)) - (a * b) * om * t;
}

__device__ Dual energy_density(const XCFunc& f, const Dual& ra, const Dual& rb, const Dual& saa, const Dual& sab,
                               const Dual& sbb) {
  Dual e = mk(0.0);
  for (int i = 0; i < f.n; ++i) {
    Dual c;
    switch (f.comp[i]) {
      case C_LDAX: c = lda_x(ra, rb); break;
      case C_B88: c = b88(ra, rb, saa, sbb); break;
      case C_VWN5: c = vwn(ra, rb, false); break;
      case C_VWNRPA: c = vwn(ra, rb, true); break;
      default: c = lyp(ra, rb, saa, sab, sbb); break;
    }
    e = e + f.w[i] * c;
  }
  return e;
}

// ------------------------------------------------------------ grid kernel
struct GridArgs {
  const double* xyz;    // natm x 3
  const double* rinv;   // natm x natm inverse distances
  const double* rad;    // natm x nrad x 2 (r, w r^2 dr)
  const double* ang;    // nang x 4 (unit vector, weight)
  int natm, nrad, nang;
  long long n;          // natm nrad nang
  double* out;          // n x 4
};

__device__ inline double ssf_s(double mu) {
  if (mu <= -SSF_A) return 1.0;
  if (mu >= SSF_A) return 0.0;
  const double x = mu / SSF_A, x2 = x * x;
  const double g = x * (35.0 + x2 * (-35.0 + x2 * (21.0 - 5.0 * x2))) / 16.0;
  return 0.5 * (1.0 - g);
}

__device__ inline double dist(const double* xyz, int a, double x, double y, double z) {
  const double dx = x - xyz[3 * a], dy = y - xyz[3 * a + 1], dz = z - xyz[3 * a + 2];
  return sqrt(dx * dx + dy * dy + dz * dz);
}

// P_B = prod_{C != B} s(mu_BC) at (x, y, z), 0 as soon as a factor is 0
__device__ double cell(const GridArgs& g, int B, double dB, double x, double y, double z) {
  double P = 1.0;
  for (int C = 0; C < g.natm && P != 0.0; ++C) {
    if (C == B) continue;
    P *= ssf_s((dB - dist(g.xyz, C, x, y, z)) * g.rinv[(size_t)B * g.natm + C]);
  }
  return P;
}

__global__ void k_grid(GridArgs g) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.n) return;
  const int j = (int)(t % g.nang);
  const long long ar = t / g.nang;
  const int i = (int)(ar % g.nrad), A = (int)(ar / g.nrad);
  const double r = g.rad[((size_t)A * g.nrad + i) * 2], wr = g.rad[((size_t)A * g.nrad + i) * 2 + 1];
  const double* v = g.ang + 4 * j;
  const double x = g.xyz[3 * A] + r * v[0], y = g.xyz[3 * A + 1] + r * v[1], z = g.xyz[3 * A + 2] + r * v[2];
  double w = wr * v[3];
  if (g.natm > 1) {
    const double dA = dist(g.xyz, A, x, y, z);
    const double PA = cell(g, A, dA, x, y, z);
    double tot = 0.0;
    if (PA != 0.0) {
      for (int B = 0; B < g.natm; ++B) {
        if (B == A) {
          tot += PA;
          continue;
        }
        const double dB = dist(g.xyz, B, x, y, z);
        if (ssf_s((dB - dA) * g.rinv[(size_t)B * g.natm + A]) == 0.0) continue;  // factor C = A vanishes
        tot += cell(g, B, dB, x, y, z);
      }
    }
    w = PA != 0.0 ? w * (PA / tot) : 0.0;
  }
  double* o = g.out + 4 * t;
  o[0] = x;
  o[1] = y;
  o[2] = z;
  o[3] = w;
}

// ------------------------------------------------------------ basis functions
struct AOArgs {
  const double* pts;  // batch points, 4 per point
  int npts, nsh, nb;  // points, significant shells, columns of the batch
  const int* shl;     // significant shells (session indices)
  const int* col;     // their first column
  const int *am, *harm, *atom, *nprim, *poff;
  const double *ex, *cc, *xyz, *xf;
  double *phi, *dphi;  // column-major npts x nb (point fastest), 3 of them
  int grad;
};

__global__ void k_ao(AOArgs a) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)a.npts * a.nsh) return;
  const int p = (int)(t % a.npts);
  const int q = (int)(t / a.npts);
  const int sh = a.shl[q], l = a.am[sh], A = a.atom[sh];
  const double dx = a.pts[4 * (size_t)p] - a.xyz[3 * A], dy = a.pts[4 * (size_t)p + 1] - a.xyz[3 * A + 1],
               dz = a.pts[4 * (size_t)p + 2] - a.xyz[3 * A + 2];
  const double r2 = dx * dx + dy * dy + dz * dz;
  double R = 0.0, dR = 0.0;
  for (int k = 0; k < a.nprim[sh]; ++k) {
    const double al = a.ex[a.poff[sh] + k], e = a.cc[a.poff[sh] + k] * exp(-al * r2);
    R += e;
    dR -= 2.0 * al * e;
  }
  const int nc = (l + 1) * (l + 2) / 2;
  double c[10], gx[10], gy[10], gz[10];
  double px[5], py[5], pz[5];
  px[0] = py[0] = pz[0] = 1.0;
  for (int k = 1; k <= l + 1 && k < 5; ++k) {
    px[k] = px[k - 1] * dx;
    py[k] = py[k - 1] * dy;
    pz[k] = pz[k - 1] * dz;
  }
  // canonical order: lx descending, then ly descending
  int i = 0;
  for (int lx = l; lx >= 0; --lx)
    for (int ly = l - lx; ly >= 0; --ly, ++i) {
      const int lz = l - lx - ly;
      const double m = px[lx] * py[ly] * pz[lz];
      c[i] = m * R;
      gx[i] = (lx ? lx * px[lx - 1] * py[ly] * pz[lz] * R : 0.0) + m * dR * dx;
      gy[i] = (ly ? ly * px[lx] * py[ly - 1] * pz[lz] * R : 0.0) + m * dR * dy;
      gz[i] = (lz ? lz * px[lx] * py[ly] * pz[lz - 1] * R : 0.0) + m * dR * dz;
    }
  const size_t ld = (size_t)a.npts, row = (size_t)a.col[q] * ld + p, plane = (size_t)a.npts * a.nb;
  if (l < 2) {
    for (int o = 0; o < nc; ++o) {
      a.phi[row + o * ld] = c[o];
      if (a.grad) {
        a.dphi[row + o * ld] = gx[o];
        a.dphi[plane + row + o * ld] = gy[o];
        a.dphi[2 * plane + row + o * ld] = gz[o];
      }
    }
    return;
  }
  const int pure = a.harm[sh] ? 1 : 0;
  const int no = pure ? 2 * l + 1 : nc;
  const double* X = a.xf + ((l - 2) * 2 + pure) * 100;
  for (int o = 0; o < no; ++o) {
    double v = 0.0, vx = 0.0, vy = 0.0, vz = 0.0;
    for (int k = 0; k < nc; ++k) {
      const double f = X[o * 10 + k];
      v += f * c[k];
      vx += f * gx[k];
      vy += f * gy[k];
      vz += f * gz[k];
    }
    a.phi[row + o * ld] = v;
    if (a.grad) {
      a.dphi[row + o * ld] = vx;
      a.dphi[plane + row + o * ld] = vy;
      a.dphi[2 * plane + row + o * ld] = vz;
    }
  }
}

// ------------------------------------------------------------ per-point XC
struct PointArgs {
  const double* pts;  // 4 per point (weight in [3])
  int npts, nb, ns, gga;
  const double *phi, *dphi, *X;  // X: ns x npts x nb
  double* Y;                     // ns x npts x nb
  double* acc;                   // [0] E_xc, [1] N_elec
  XCFunc f;
};

__device__ inline double wsum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// 64 points per block (one per lane) x XC_SLICES wavefronts that split the
// basis functions; matrices point-fastest: element (p, m) at m npts + p
#define XC_SLICES 8
__global__ __launch_bounds__(64 * XC_SLICES) void k_xc_point(PointArgs a) {
  const int lane = threadIdx.x & 63, sl = threadIdx.x >> 6;
  const long long p = (long long)blockIdx.x * 64 + lane;
  __shared__ double part[XC_SLICES][2][4][64];  // per slice: rho, grad rho per spin
  __shared__ double coef[2][4][64];             // per spin: c0, v (3)
  __shared__ double red[2][64];
  const bool live = p < a.npts;
  const int n = a.nb;
  const size_t ld = (size_t)a.npts, plane = ld * n;
  const int m0 = (int)((long long)n * sl / XC_SLICES), m1 = (int)((long long)n * (sl + 1) / XC_SLICES);
  for (int s = 0; s < a.ns; ++s) {
    double r = 0.0, g0 = 0.0, g1 = 0.0, g2 = 0.0;
    if (live) {
      const double* X = a.X + s * plane + p;
      const double* phi = a.phi + p;
      const double* gp = a.dphi + p;
      if (a.gga) {
        for (int m = m0; m < m1; ++m) {
          const double x = X[m * ld];
          r += x * phi[m * ld];
          g0 += x * gp[m * ld];
          g1 += x * gp[plane + m * ld];
          g2 += x * gp[2 * plane + m * ld];
        }
      } else {
        for (int m = m0; m < m1; ++m) r += X[m * ld] * phi[m * ld];
      }
    }
    part[sl][s][0][lane] = r;
    part[sl][s][1][lane] = g0;
    part[sl][s][2][lane] = g1;
    part[sl][s][3][lane] = g2;
  }
  __syncthreads();
  if (sl == 0) {
    double rho[2] = {0.0, 0.0}, g[2][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
    for (int s = 0; s < a.ns; ++s)
      for (int j = 0; j < XC_SLICES; ++j) {
        rho[s] += part[j][s][0][lane];
        for (int k = 0; k < 3; ++k) g[s][k] += 2.0 * part[j][s][k + 1][lane];
      }
    double ra, rb, ga[3], gb[3];
    if (a.ns == 1) {
      ra = rb = 0.5 * rho[0];
      for (int k = 0; k < 3; ++k) ga[k] = gb[k] = 0.5 * g[0][k];
    } else {
      ra = rho[0];
      rb = rho[1];
      for (int k = 0; k < 3; ++k) {
        ga[k] = g[0][k];
        gb[k] = g[1][k];
      }
    }
    const double w = live ? a.pts[4 * p + 3] : 0.0, rt = ra + rb;
    double e = 0.0, c0[2] = {0.0, 0.0}, vec[2][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
    if (live && rt >= RHO_CUT) {
      const double saa = ga[0] * ga[0] + ga[1] * ga[1] + ga[2] * ga[2];
      const double sab = ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2];
      const double sbb = gb[0] * gb[0] + gb[1] * gb[1] + gb[2] * gb[2];
      const Dual F = energy_density(a.f, var(fmax(ra, RHO_FLOOR), 0), var(fmax(rb, RHO_FLOOR), 1), var(saa, 2),
                                    var(sab, 3), var(sbb, 4));
      e = w * F.v;
      c0[0] = 0.5 * w * F.d[0];
      c0[1] = 0.5 * w * F.d[1];
      for (int k = 0; k < 3; ++k) {
        vec[0][k] = w * (2.0 * F.d[2] * ga[k] + F.d[3] * gb[k]);
        vec[1][k] = w * (2.0 * F.d[4] * gb[k] + F.d[3] * ga[k]);
      }
    }
    for (int s = 0; s < 2; ++s) {
      coef[s][0][lane] = c0[s];
      for (int k = 0; k < 3; ++k) coef[s][k + 1][lane] = vec[s][k];
    }
    red[0][lane] = e;
    red[1][lane] = w * rt;
  }
  __syncthreads();
  if (live)
    for (int s = 0; s < a.ns; ++s) {
      const double c = coef[s][0][lane], v0 = coef[s][1][lane], v1 = coef[s][2][lane], v2 = coef[s][3][lane];
      double* Y = a.Y + s * plane + p;
      const double* phi = a.phi + p;
      const double* gp = a.dphi + p;
      if (a.gga) {
        for (int m = m0; m < m1; ++m)
          Y[m * ld] = c * phi[m * ld] + v0 * gp[m * ld] + v1 * gp[plane + m * ld] + v2 * gp[2 * plane + m * ld];
      } else {
        for (int m = m0; m < m1; ++m) Y[m * ld] = c * phi[m * ld];
      }
    }
  if (sl == 0) {
    const double e = wsum(red[0][lane]), nel = wsum(red[1][lane]);
    if (lane == 0) {
      atomicAdd(a.acc, e);
      atomicAdd(a.acc + 1, nel);
    }
  }
}

// ------------------------------------------------------------ gathers and packing
__global__ void k_unpack(const double* pk, int nbf, double* full) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nbf * nbf) return;
  const int i = (int)(t / nbf), j = (int)(t % nbf);
  const int hi = i > j ? i : j, lo = i > j ? j : i;
  full[t] = pk[(size_t)hi * (hi + 1) / 2 + lo];
}

__global__ void k_gather(const double* full, int nbf, const int* idx, int nb, double* out) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nb * nb) return;
  const int i = (int)(t / nb), j = (int)(t % nb);
  out[t] = full[(size_t)idx[i] * nbf + idx[j]];
}

// V[idx[m], idx[n]] += sum over the nparts partials of Vb(m, n) + Vb(n, m);
// each partial column-major nb x nb
__global__ void k_scatter(const double* Vb, int nb, int nparts, const int* idx, int nbf, double* V) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nb * nb) return;
  const int m = (int)(t % nb), n = (int)(t / nb);
  const size_t nb2 = (size_t)nb * nb;
  double v = 0.0;
  for (int k = 0; k < nparts; ++k) v += Vb[k * nb2 + (size_t)n * nb + m] + Vb[k * nb2 + (size_t)m * nb + n];
  V[(size_t)idx[m] * nbf + idx[n]] += v;
}

// packed lower triangle of the sum of the lanes' matrices
struct LaneSum {
  const double* v[4];
  int n;
};
__global__ void k_pack(LaneSum ls, int nbf, double* pk) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long ntri = (long long)nbf * (nbf + 1) / 2;
  if (t >= ntri) return;
  int i = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) / 2.0);
  while ((long long)i * (i + 1) / 2 > t) --i;
  while ((long long)(i + 1) * (i + 2) / 2 <= t) ++i;
  const int j = (int)(t - (long long)i * (i + 1) / 2);
  double v = 0.0;
  for (int k = 0; k < ls.n; ++k) v += ls.v[k][(size_t)i * nbf + j];
  pk[t] = v;
}

inline unsigned nblk(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// ------------------------------------------------------------ device buffers
template <typename T>
struct Buf {
  T* p = nullptr;
  size_t n = 0;
  bool ensure(size_t m) {
    if (m <= n && p) return true;
    if (p) hipFree(p);
    p = nullptr;
    n = 0;
    if (hipMalloc((void**)&p, std::max<size_t>(m, 1) * sizeof(T)) != hipSuccess) {
      p = nullptr;
      err("XC: hipMalloc of %zu bytes failed", m * sizeof(T));
      return false;
    }
    n = std::max<size_t>(m, 1);
    return true;
  }
  void release() {
    if (p) hipFree(p);
    p = nullptr;
    n = 0;
  }
};

struct Batch {
  long long p0;
  int npts, sh0, nsh, ao0, nb;
};

// a stream with its own batch buffers, GEMM handle and Vxc accumulator;
// batches go round-robin over the lanes so that the GEMMs of one overlap the
// memory-bound kernels of the others (three measured fastest: two and four
// lanes were 12 % and 8 % slower)
constexpr int XC_LANES = 3;
static_assert(XC_LANES >= 1 && XC_LANES <= 4, "k_pack sums at most four lanes");
struct Lane {
  hipStream_t st = nullptr;
  hipEvent_t done = nullptr;
  rocblas_handle blas = nullptr;
  Buf<double> phi, dphi, X, Y, Db, Vb, V;
  void release() {
    for (auto* b : {&phi, &dphi, &X, &Y, &Db, &Vb, &V}) b->release();
    if (blas) rocblas_destroy_handle(blas);
    if (done) hipEventDestroy(done);
    if (st) hipStreamDestroy(st);
    blas = nullptr;
    done = nullptr;
    st = nullptr;
  }
};

struct XCState {
  XCFunc f{};
  bool configured = false;
  int nrad = 96, ntheta = 15;
  // cache key of the grid
  std::vector<double> key_xyz, key_ex, key_cc;
  std::vector<int> key_am;
  int key_nbf = -1, key_nrad = 0, key_ntheta = 0;
  bool grid_ready = false;
  std::vector<double> grid;  // kept points (x, y, z, w), atom-major
  std::vector<Batch> batches;
  int max_nb = 0, max_pts = 0;
  Buf<double> d_pts, d_xf, d_D, d_pk, d_acc;
  Buf<int> d_shl, d_col, d_idx, d_am, d_harm, d_atom, d_nprim, d_poff;
  Buf<double> d_ex, d_cc, d_xyz;
  Lane lane[XC_LANES];
  hipEvent_t ready = nullptr;
  void release() {
    for (auto* b : {&d_pts, &d_xf, &d_D, &d_pk, &d_acc, &d_ex, &d_cc, &d_xyz}) b->release();
    for (auto* b : {&d_shl, &d_col, &d_idx, &d_am, &d_harm, &d_atom, &d_nprim, &d_poff}) b->release();
    for (auto& l : lane) l.release();
    if (ready) hipEventDestroy(ready);
    ready = nullptr;
    grid_ready = false;
  }
};

std::mutex g_xmtx;
std::map<oqph_session*, XCState*> g_xc;

XCState* state(oqph_session* s, bool create) {
  std::lock_guard<std::mutex> lk(g_xmtx);
  auto it = g_xc.find(s);
  if (it != g_xc.end()) return it->second;
  if (!create) return nullptr;
  XCState* x = new XCState();
  g_xc[s] = x;
  return x;
}

// Treutler-Ahlrichs xi (J. Chem. Phys. 102, 346 (1995)); 1.0 beyond Kr
double ta_xi(int z) {
  static const double t[37] = {1.0, 0.8, 0.9, 1.8, 1.4, 1.3, 1.1, 0.9, 0.9, 0.9, 0.9, 1.4, 1.3,
                               1.3, 1.2, 1.1, 1.0, 1.0, 1.0, 1.5, 1.4, 1.3, 1.2, 1.2, 1.2, 1.2,
                               1.2, 1.2, 1.1, 1.1, 1.1, 1.1, 1.0, 0.9, 0.9, 0.9, 0.9};
  return (z >= 1 && z <= 36) ? t[z] : 1.0;
}

// Gauss-Legendre nodes and weights on [-1, 1] (Newton on P_n)
void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  for (int i = 0; i < n; ++i) {
    double z = cos(PI * (i + 0.75) / (n + 0.5)), dp = 0.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = z;
      for (int k = 2; k <= n; ++k) {
        const double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      if (n == 1) p0 = 1.0;
      dp = n * (z * p1 - p0) / (z * z - 1.0);
      const double dz = p1 / dp;
      z -= dz;
      if (fabs(dz) < 1e-16) break;
    }
    double p0 = 1.0, p1 = z;
    for (int k = 2; k <= n; ++k) {
      const double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
      p0 = p1;
      p1 = p2;
    }
    dp = n * (z * p1 - p0) / (z * z - 1.0);
    x[n - 1 - i] = z;
    w[n - 1 - i] = 2.0 / ((1.0 - z * z) * dp * dp);
  }
}

template <typename T>
bool upload(Buf<T>& b, const T* h, size_t n, hipStream_t st) {
  if (!b.ensure(n)) return false;
  if (n && hipMemcpyAsync(b.p, h, n * sizeof(T), hipMemcpyHostToDevice, st) != hipSuccess) {
    err("XC: host-to-device copy failed");
    return false;
  }
  return true;
}

int build_grid(XCState* x, const oqph::SessionView& v) {
  const int natm = v.natm, nrad = x->nrad, nth = x->ntheta, nphi = 2 * nth, nang = nth * nphi;
  hipStream_t st = v.stream;
  // radial and angular tables
  std::vector<double> rad((size_t)natm * nrad * 2), ang((size_t)nang * 4);
  for (int A = 0; A < natm; ++A) {
    const double xi = ta_xi((int)lround(v.zn[A])), al = 0.6, c = xi / log(2.0);
    for (int i = 1; i <= nrad; ++i) {
      const double th = i * PI / (nrad + 1), xx = cos(th), lg = log(2.0 / (1.0 - xx));
      const double r = c * pow(1.0 + xx, al) * lg;
      const double drdx = c * (al * pow(1.0 + xx, al - 1.0) * lg + pow(1.0 + xx, al) / (1.0 - xx));
      rad[((size_t)A * nrad + i - 1) * 2] = r;
      rad[((size_t)A * nrad + i - 1) * 2 + 1] = PI / (nrad + 1) * sin(th) * drdx * r * r;
    }
  }
  std::vector<double> gt, gw;
  gauss_legendre(nth, gt, gw);
  for (int a = 0; a < nth; ++a)
    for (int b = 0; b < nphi; ++b) {
      const double ph = (b + 0.5) * (2.0 * PI / nphi), s = sqrt(1.0 - gt[a] * gt[a]);
      double* o = &ang[(size_t)(a * nphi + b) * 4];
      o[0] = s * cos(ph);
      o[1] = s * sin(ph);
      o[2] = gt[a];
      o[3] = gw[a] * (2.0 * PI / nphi);
    }
  std::vector<double> rinv((size_t)natm * natm, 0.0);
  for (int a = 0; a < natm; ++a)
    for (int b = 0; b < natm; ++b)
      if (a != b) {
        const double dx = v.xyz[3 * a] - v.xyz[3 * b], dy = v.xyz[3 * a + 1] - v.xyz[3 * b + 1],
                     dz = v.xyz[3 * a + 2] - v.xyz[3 * b + 2];
        rinv[(size_t)a * natm + b] = 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
      }
  const long long nraw = (long long)natm * nrad * nang;
  Buf<double> d_rad, d_ang, d_rinv, d_raw;
  bool ok = upload(x->d_xyz, v.xyz, (size_t)3 * natm, st) && upload(d_rad, rad.data(), rad.size(), st) &&
            upload(d_ang, ang.data(), ang.size(), st) && upload(d_rinv, rinv.data(), rinv.size(), st) &&
            d_raw.ensure((size_t)nraw * 4);
  std::vector<double> raw;
  if (ok) {
    GridArgs g{x->d_xyz.p, d_rinv.p, d_rad.p, d_ang.p, natm, nrad, nang, nraw, d_raw.p};
    hipLaunchKernelGGL(k_grid, dim3(nblk(nraw, 256)), dim3(256), 0, st, g);
    raw.resize((size_t)nraw * 4);
    ok = hipMemcpyAsync(raw.data(), d_raw.p, raw.size() * sizeof(double), hipMemcpyDeviceToHost, st) == hipSuccess &&
         hipStreamSynchronize(st) == hipSuccess;
    if (!ok) err("XC: grid kernel failed: %s", hipGetErrorString(hipGetLastError()));
  }
  d_rad.release();
  d_ang.release();
  d_rinv.release();
  d_raw.release();
  if (!ok) return 2;
  // keep w > 1e-20 (atom-major order: the exported grid)
  x->grid.clear();
  for (long long t = 0; t < nraw; ++t)
    if (raw[4 * t + 3] > 1e-20) x->grid.insert(x->grid.end(), &raw[4 * t], &raw[4 * t] + 4);
  const long long n = (long long)x->grid.size() / 4;
  // boxes
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (long long t = 0; t < n; ++t)
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], x->grid[4 * t + k]);
      hi[k] = std::max(hi[k], x->grid[4 * t + k]);
    }
  long long nbx[3];
  for (int k = 0; k < 3; ++k) nbx[k] = std::max<long long>(1, (long long)((hi[k] - lo[k]) / BOX) + 1);
  std::vector<long long> key(n);
  for (long long t = 0; t < n; ++t) {
    long long b[3];
    for (int k = 0; k < 3; ++k) b[k] = std::min(nbx[k] - 1, (long long)((x->grid[4 * t + k] - lo[k]) / BOX));
    key[t] = (b[0] * nbx[1] + b[1]) * nbx[2] + b[2];
  }
  // counting sort by box (stable: atom-major order within a box)
  const long long nkeys = nbx[0] * nbx[1] * nbx[2];
  std::vector<long long> start(nkeys + 1, 0), order(n);
  for (long long t = 0; t < n; ++t) ++start[key[t] + 1];
  for (long long k = 0; k < nkeys; ++k) start[k + 1] += start[k];
  for (long long t = 0; t < n; ++t) order[start[key[t]]++] = t;
  std::vector<double> sorted((size_t)n * 4);
  for (long long t = 0; t < n; ++t) memcpy(&sorted[4 * t], &x->grid[4 * order[t]], 4 * sizeof(double));
  // batches and their shells
  std::vector<double> rcut(v.nsh);
  for (int s = 0; s < v.nsh; ++s) {
    double rc = 0.0;
    for (int k = 0; k < v.nprim[s]; ++k) {
      const double c = fabs(v.cc[v.poff[s] + k]), al = v.ex[v.poff[s] + k];
      if (c > AO_EPS) rc = std::max(rc, sqrt(log(c / AO_EPS) / al));
    }
    rcut[s] = rc;
  }
  std::vector<int> shl, col, idx;
  x->batches.clear();
  x->max_nb = 0;
  x->max_pts = 0;
  for (long long t0 = 0; t0 < n;) {
    long long t1 = t0;
    while (t1 < n && key[order[t1]] == key[order[t0]] && t1 - t0 < MAX_BATCH) ++t1;
    // the batch's bounding box; a shell is kept when its centre lies within
    // its cutoff radius of the box
    double blo[3] = {1e300, 1e300, 1e300}, bhi[3] = {-1e300, -1e300, -1e300};
    for (long long t = t0; t < t1; ++t)
      for (int k = 0; k < 3; ++k) {
        blo[k] = std::min(blo[k], sorted[4 * t + k]);
        bhi[k] = std::max(bhi[k], sorted[4 * t + k]);
      }
    Batch b{t0, (int)(t1 - t0), (int)shl.size(), 0, (int)idx.size(), 0};
    for (int s = 0; s < v.nsh; ++s) {
      const int A = v.atom[s];
      double d2 = 0.0;
      for (int k = 0; k < 3; ++k) {
        const double a = v.xyz[3 * A + k], o = a < blo[k] ? blo[k] - a : (a > bhi[k] ? a - bhi[k] : 0.0);
        d2 += o * o;
      }
      if (sqrt(d2) > rcut[s]) continue;
      const int l = v.am[s], no = (v.harm[s] && l >= 2) ? 2 * l + 1 : (l + 1) * (l + 2) / 2;
      shl.push_back(s);
      col.push_back(b.nb);
      for (int o = 0; o < no; ++o) idx.push_back(v.aooff[s] + o);
      b.nb += no;
      ++b.nsh;
    }
    if (b.nsh > 0) {
      x->batches.push_back(b);
      x->max_nb = std::max(x->max_nb, b.nb);
      x->max_pts = std::max(x->max_pts, b.npts);
    }
    t0 = t1;
  }
  // device copies
  std::vector<double> xf(400);
  oqph_host::build_xform(reinterpret_cast<double(*)[2][10][10]>(xf.data()));
  std::vector<int> harm(v.harm, v.harm + v.nsh);
  int nprim_tot = 0;
  for (int s = 0; s < v.nsh; ++s) nprim_tot = std::max(nprim_tot, v.poff[s] + v.nprim[s]);
  const size_t mp = (size_t)x->max_pts, mb = (size_t)x->max_nb;
  ok = upload(x->d_pts, sorted.data(), sorted.size(), st) && upload(x->d_xf, xf.data(), xf.size(), st) &&
       upload(x->d_shl, shl.data(), shl.size(), st) && upload(x->d_col, col.data(), col.size(), st) &&
       upload(x->d_idx, idx.data(), idx.size(), st) && upload(x->d_am, v.am, (size_t)v.nsh, st) &&
       upload(x->d_harm, harm.data(), harm.size(), st) && upload(x->d_atom, v.atom, (size_t)v.nsh, st) &&
       upload(x->d_nprim, v.nprim, (size_t)v.nsh, st) && upload(x->d_poff, v.poff, (size_t)v.nsh, st) &&
       upload(x->d_ex, v.ex, (size_t)nprim_tot, st) && upload(x->d_cc, v.cc, (size_t)nprim_tot, st);
  for (auto& l : x->lane)
    ok = ok && l.phi.ensure(mp * mb) && l.dphi.ensure(3 * mp * mb) && l.X.ensure(2 * mp * mb) &&
         l.Y.ensure(2 * mp * mb) && l.Db.ensure(mb * mb) && l.Vb.ensure((mp / KCHUNK + 1) * mb * mb);
  ok = ok && hipStreamSynchronize(st) == hipSuccess;
  if (!ok) return 3;
  x->grid_ready = true;
  return 0;
}

bool grid_current(XCState* x, const oqph::SessionView& v) {
  if (!x->grid_ready || x->key_nbf != v.nbf || x->key_nrad != x->nrad || x->key_ntheta != x->ntheta) return false;
  if (x->key_xyz.size() != (size_t)3 * v.natm || !std::equal(x->key_xyz.begin(), x->key_xyz.end(), v.xyz))
    return false;
  int np = 0;
  for (int s = 0; s < v.nsh; ++s) np = std::max(np, v.poff[s] + v.nprim[s]);
  return x->key_ex.size() == (size_t)np && std::equal(x->key_ex.begin(), x->key_ex.end(), v.ex) &&
         std::equal(x->key_cc.begin(), x->key_cc.end(), v.cc) && x->key_am.size() == (size_t)v.nsh &&
         std::equal(x->key_am.begin(), x->key_am.end(), v.am);
}

int ensure_grid(XCState* x, const oqph::SessionView& v) {
  if (grid_current(x, v)) return 0;
  x->grid_ready = false;
  const int rc = build_grid(x, v);
  if (rc) return rc;
  x->key_nbf = v.nbf;
  x->key_nrad = x->nrad;
  x->key_ntheta = x->ntheta;
  x->key_xyz.assign(v.xyz, v.xyz + 3 * v.natm);
  int np = 0;
  for (int s = 0; s < v.nsh; ++s) np = std::max(np, v.poff[s] + v.nprim[s]);
  x->key_ex.assign(v.ex, v.ex + np);
  x->key_cc.assign(v.cc, v.cc + np);
  x->key_am.assign(v.am, v.am + v.nsh);
  return 0;
}

int vxc_run(oqph_session* s, const double* d, int nfocks, double* fxc, double* exc, double* nel) {
  oqph::SessionView v;
  if (!oqph::session_view(s, &v)) {
    err("oqph_vxc: the session has no basis or geometry");
    return 4;
  }
  XCState* x = state(s, false);
  if (!x || !x->configured) {
    err("oqph_vxc: no functional set (oqph_session_set_xc)");
    return 1;
  }
  if (nfocks != 1 && nfocks != 2) {
    err("oqph_vxc: nfocks %d (1 or 2)", nfocks);
    return 1;
  }
  if (!d || !fxc || !exc || !nel) {
    err("oqph_vxc: null argument");
    return 4;
  }
  hipSetDevice(v.device);
  hipStream_t st = v.stream;
  int rc = ensure_grid(x, v);
  if (rc) return rc;
  const int nbf = v.nbf;
  const size_t ntri = (size_t)nbf * (nbf + 1) / 2, nn = (size_t)nbf * nbf;
  if (!x->ready && hipEventCreateWithFlags(&x->ready, hipEventDisableTiming) != hipSuccess) {
    x->ready = nullptr;
    err("oqph_vxc: hipEventCreate failed");
    return 2;
  }
  for (auto& l : x->lane) {
    if (!l.st && hipStreamCreateWithFlags(&l.st, hipStreamNonBlocking) != hipSuccess) {
      l.st = nullptr;
      err("oqph_vxc: hipStreamCreate failed");
      return 2;
    }
    if (!l.done && hipEventCreateWithFlags(&l.done, hipEventDisableTiming) != hipSuccess) {
      l.done = nullptr;
      err("oqph_vxc: hipEventCreate failed");
      return 2;
    }
    if (!l.blas && rocblas_create_handle(&l.blas) != rocblas_status_success) {
      l.blas = nullptr;
      err("oqph_vxc: rocblas_create_handle failed");
      return 2;
    }
    rocblas_set_stream(l.blas, l.st);
    if (!l.V.ensure(nn * nfocks)) return 3;
  }
  if (!upload(x->d_pk, d, ntri * nfocks, st) || !x->d_D.ensure(nn * nfocks) || !x->d_acc.ensure(2)) return 3;
  for (int sp = 0; sp < nfocks; ++sp)
    hipLaunchKernelGGL(k_unpack, dim3(nblk((long long)nn, 256)), dim3(256), 0, st, x->d_pk.p + sp * ntri, nbf,
                       x->d_D.p + sp * nn);
  hipMemsetAsync(x->d_acc.p, 0, 2 * sizeof(double), st);
  hipEventRecord(x->ready, st);
  for (auto& l : x->lane) {
    hipStreamWaitEvent(l.st, x->ready, 0);
    hipMemsetAsync(l.V.p, 0, nn * nfocks * sizeof(double), l.st);
  }
  // on an error, let the work already queued on the lanes finish before the
  // buffers can be reused
  auto drain = [&](int code) {
    for (auto& l : x->lane) hipStreamSynchronize(l.st);
    hipStreamSynchronize(st);
    return code;
  };
  const double one = 1.0, zero = 0.0;
  for (size_t ib = 0; ib < x->batches.size(); ++ib) {
    const Batch& b = x->batches[ib];
    Lane& L = x->lane[ib % XC_LANES];
    hipStream_t ls = L.st;
    const int np = b.npts, nb = b.nb;
    AOArgs aa{x->d_pts.p + 4 * b.p0, np, b.nsh, nb, x->d_shl.p + b.sh0, x->d_col.p + b.sh0, x->d_am.p,
              x->d_harm.p, x->d_atom.p, x->d_nprim.p, x->d_poff.p, x->d_ex.p, x->d_cc.p, x->d_xyz.p, x->d_xf.p,
              L.phi.p, L.dphi.p, x->f.gga};
    hipLaunchKernelGGL(k_ao, dim3(nblk((long long)np * b.nsh, 256)), dim3(256), 0, ls, aa);
    const size_t plane = (size_t)np * nb;
    for (int sp = 0; sp < nfocks; ++sp) {
      hipLaunchKernelGGL(k_gather, dim3(nblk((long long)nb * nb, 256)), dim3(256), 0, ls, x->d_D.p + sp * nn, nbf,
                         x->d_idx.p + b.ao0, nb, L.Db.p);
      // X (np x nb, column-major) = phi (np x nb) D_b (nb x nb)
      if (rocblas_dgemm(L.blas, rocblas_operation_none, rocblas_operation_none, np, nb, nb, &one, L.phi.p, np, L.Db.p,
                        nb, &zero, L.X.p + sp * plane, np) != rocblas_status_success) {
        err("oqph_vxc: rocblas_dgemm failed");
        return drain(2);
      }
    }
    PointArgs pa{x->d_pts.p + 4 * b.p0, np, nb, nfocks, x->f.gga, L.phi.p, L.dphi.p, L.X.p, L.Y.p, x->d_acc.p, x->f};
    hipLaunchKernelGGL(k_xc_point, dim3(nblk(np, 64)), dim3(64 * XC_SLICES), 0, ls, pa);
    // V_b = phi^T Y, V_b(m, n) = sum_p phi(p, m) Y(p, n), as partial sums over
    // KCHUNK-point slices (one strided-batched GEMM plus the remainder): the
    // nb x nb result alone gives too few tiles to fill the device
    const int nfull = np / KCHUNK, rem = np % KCHUNK, nparts = nfull + (rem > 0 ? 1 : 0);
    const long long nb2 = (long long)nb * nb;
    for (int sp = 0; sp < nfocks; ++sp) {
      const double* Ys = L.Y.p + sp * plane;
      if (nfull > 0 &&
          rocblas_dgemm_strided_batched(L.blas, rocblas_operation_transpose, rocblas_operation_none, nb, nb, KCHUNK,
                                        &one, L.phi.p, np, KCHUNK, Ys, np, KCHUNK, &zero, L.Vb.p, nb, nb2,
                                        nfull) != rocblas_status_success) {
        err("oqph_vxc: rocblas_dgemm_strided_batched failed");
        return drain(2);
      }
      if (rem > 0 && rocblas_dgemm(L.blas, rocblas_operation_transpose, rocblas_operation_none, nb, nb, rem, &one,
                                   L.phi.p + (size_t)nfull * KCHUNK, np, Ys + (size_t)nfull * KCHUNK, np, &zero,
                                   L.Vb.p + nfull * nb2, nb) != rocblas_status_success) {
        err("oqph_vxc: rocblas_dgemm failed");
        return drain(2);
      }
      hipLaunchKernelGGL(k_scatter, dim3(nblk(nb2, 256)), dim3(256), 0, ls, L.Vb.p, nb, nparts, x->d_idx.p + b.ao0,
                         nbf, L.V.p + sp * nn);
    }
  }
  for (auto& l : x->lane) {
    hipEventRecord(l.done, l.st);
    hipStreamWaitEvent(st, l.done, 0);
  }
  for (int sp = 0; sp < nfocks; ++sp) {
    LaneSum lsum{};
    lsum.n = XC_LANES;
    for (int k = 0; k < XC_LANES; ++k) lsum.v[k] = x->lane[k].V.p + sp * nn;
    hipLaunchKernelGGL(k_pack, dim3(nblk((long long)ntri, 256)), dim3(256), 0, st, lsum, nbf, x->d_pk.p + sp * ntri);
  }
  double acc[2];
  if (hipMemcpyAsync(fxc, x->d_pk.p, ntri * nfocks * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(acc, x->d_acc.p, 2 * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess) {
    err("oqph_vxc: device error: %s", hipGetErrorString(hipGetLastError()));
    return 2;
  }
  *exc = acc[0];
  *nel = acc[1];
  return 0;
}

}  // namespace

void oqph::xc_release(oqph_session* s) {
  XCState* x = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_xmtx);
    auto it = g_xc.find(s);
    if (it == g_xc.end()) return;
    x = it->second;
    g_xc.erase(it);
  }
  x->release();
  delete x;
}

extern "C" int oqph_xc_functional_info(const char* name, double* hfex, int* gga) {
  XCFunc f;
  if (!lookup(name, &f)) return 1;
  if (hfex) *hfex = f.hfex;
  if (gga) *gga = f.gga;
  return 0;
}

extern "C" int oqph_session_set_xc(oqph_session* s, const char* name, int nrad, int ntheta) {
  if (!s) return 4;
  XCFunc f;
  if (!lookup(name, &f)) {
    err("oqph_session_set_xc: unknown functional '%s'", name ? name : "");
    return 1;
  }
  if (nrad < 2 || nrad > 1000 || ntheta < 2 || ntheta > 200) {
    err("oqph_session_set_xc: grid %d x %d out of range", nrad, ntheta);
    return 4;
  }
  XCState* x = state(s, true);
  x->f = f;
  x->configured = true;
  x->nrad = nrad;
  x->ntheta = ntheta;
  return 0;
}

extern "C" int oqph_vxc(oqph_session* s, const double* d, int nfocks, double* fxc, double* exc, double* nel) {
  if (!s) return 4;
  return vxc_run(s, d, nfocks, fxc, exc, nel);
}

extern "C" long long oqph_xc_grid(oqph_session* s, double* pts, long long cap) {
  oqph::SessionView v;
  if (!s || !oqph::session_view(s, &v)) return -4;
  XCState* x = state(s, false);
  if (!x || !x->configured) return -1;
  hipSetDevice(v.device);
  if (ensure_grid(x, v)) return -2;
  const long long n = (long long)x->grid.size() / 4;
  if (pts) memcpy(pts, x->grid.data(), (size_t)std::min(n, cap) * 4 * sizeof(double));
  return n;
}

extern "C" int oqph_xc_grid_stats(oqph_session* s, long long* out) {
  oqph::SessionView v;
  if (!s || !out || !oqph::session_view(s, &v)) return 4;
  XCState* x = state(s, false);
  if (!x || !x->configured) return 1;
  hipSetDevice(v.device);
  if (ensure_grid(x, v)) return 2;
  long long a = 0, b2 = 0;
  for (const Batch& b : x->batches) {
    a += (long long)b.npts * b.nb;
    b2 += (long long)b.npts * b.nb * b.nb;
  }
  out[0] = (long long)x->grid.size() / 4;
  out[1] = (long long)x->batches.size();
  out[2] = a;
  out[3] = b2;
  out[4] = x->max_nb;
  return 0;
}

extern "C" void routec_vxc(const double* d, double* fxc, const int* nbf, const int* nfocks, double* eexc,
                           double* totele, int* info) {
  int rc = 1;
  oqph_session* s = oqph::default_session();
  oqph::SessionView v;
  if (!info) return;
  if (!s || !nbf || !nfocks || !oqph::session_view(s, &v)) {
    err("routec_vxc: no default session");
  } else if (*nbf != v.nbf) {
    err("routec_vxc: nbf %d does not match the session basis (%d)", *nbf, v.nbf);
  } else {
    rc = vxc_run(s, d, *nfocks, fxc, eexc, totele);
    if (rc != 0) rc = 1;
  }
  *info = rc;
}
