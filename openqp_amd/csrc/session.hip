// session.hip -- host side of liboqph: session state, pair data, screening
// lists, launches, and the C ABI declared in include/oqph.h.
//
// Per geometry: primitive-pair data screened as OpenQP screens it
// (source/integrals/int2_pairs.F90:74-266, cutoffs from int2.F90:266-272),
// the Schwarz table on the device with the tight cutoffs of ints_exchange
// (int2.F90:1582-1737, :1600-1604), and per-class shell-pair lists sorted by
// Schwarz bound.  Per build: the shell-pair density maxima on the device,
// the coarse per-bra ket counts (a prefix of each sorted ket list survives
// Q_bra Q_ket 4 maxD >= c), and one launch per class pair; the exact OpenQP
// screens run inside the kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/oqph.h"
#include "oqph_common.h"
#include "grad_params.h"
#include "session_view.h"

namespace oqph {
bool launch_jk(int la, int lb, int lc, int ld, bool pure, const JKParams& p, const WaveWork& cw, hipStream_t s);
bool launch_jkb(int la, int lb, int lc, int ld, bool pure, const JKParams& p, const BraWork& bw, const ColMap& cm,
                size_t lds, hipStream_t s);
bool jkb_info(int la, int lb, int lc, int ld, bool pure, int* nrow, size_t* static_lds, size_t lds_max);
bool launch_jkg(int la, int lb, int lc, int ld, bool pure, const JKParams& p, const BraWork& bw, const ColMap& cm,
                const ChunkView& cv, size_t lds, hipStream_t s);
int jkg_info(int la, int lb, int lc, int ld, bool pure, size_t* static_lds, size_t lds_max);
bool launch_jkt(int la, int lb, int lc, int ld, bool pure, const JKParams& p, const BraWork& bw, const ColMap& cm,
                const ChunkView& cv, size_t lds, hipStream_t s);
int jkt_info(int la, int lb, int lc, int ld, bool pure, size_t* static_lds, size_t lds_max);
bool launch_schwarz(int la, int lb, bool pure, const ShellPair* pairs, const PrimJK* prims, const int* list, int n,
                    const double* tab, const double* xf, double qcut, double imu2, double* Q, hipStream_t s);
bool kernel_resources(int kind, int la, int lb, int lc, int ld, bool pure, KernelRes* r);
void launch_shell_density(const double* D, const double* D2, const int* ao_off, const int* nao, int nsh, double* dsh,
                          hipStream_t s);
void launch_sumdiff(const double* a, const double* b, double* p, double* m, double c, long long n, hipStream_t s);
void launch_jscatter(const ShellPair* pairs, const int* joff, int npair, bool pure, const double* Jk, double* F,
                     hipStream_t s);
void launch_max_reduce(const double* x, long long n, unsigned long long* out, hipStream_t s);
bool kernels_support(int lmax);
bool jk_template(int la, int lb, int lc, int ld);
bool grad_template(int la, int lb, int lc, int ld);
HLDims hl_dims(int la, int lb, int lc, int ld, int pure, int grad, int natm);
void launch_jkh(const HLDims& h, const JKParams& p, const WaveWork& ww, hipStream_t s);
void launch_jkh_multi(const JKParams& p, const HLTaskList& tl, long long nblocks, int lds, hipStream_t s);
void launch_schwarzh(const HLDims& h, const ShellPair* pairs, const PrimJK* prims, const int* list, int n,
                     const double* rys, const double* xf, double qcut, double imu2, double* Q, hipStream_t s);
struct GradParams;
void launch_gradh(const HLDims& h, const GradParams& p, const WaveWork& ww, hipStream_t s);
void launch_gradh_multi(const GradParams& p, const HLTaskList& tl, long long nblocks, int lds, hipStream_t s);
bool launch_grad(int la, int lb, int lc, int ld, const GradParams& p, const WaveWork& ww, hipStream_t s, bool f32,
                 bool mrsf);
void launch_cart_matrix(const double* M, int nbf, int nsh, const int* am, const int* pure, const int* ao_off,
                        const int* coff, const double* xf, int ncf, double* Mc, hipStream_t s);
void launch_cart_density(const double* D, int nsh, const int* am, const int* pure, const int* ao_off, const int* coff,
                         const double* xf, int ncf, double* Dc, hipStream_t s);
const double* rys_table_host(size_t* n);
const float* rys32_table_host(size_t* n);
}  // namespace oqph

using namespace oqph;

// ------------------------------------------------------------------ errors
static thread_local std::string g_err;
static void set_err(const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
}
#define HIPCHK(x)                                                              \
  do {                                                                         \
    hipError_t e_ = (x);                                                       \
    if (e_ != hipSuccess) {                                                    \
      set_err("%s:%d %s: %s", __FILE__, __LINE__, #x, hipGetErrorString(e_)); \
      return e_ == hipErrorOutOfMemory ? 3 : 2;                                \
    }                                                                          \
  } while (0)

// fault injection for host-side tests of the caller's fallback:
// OQPH_FAULT=decline -> every build declines with 1; =oom -> with 3.
static int fault_code() {
  const char* f = getenv("OQPH_FAULT");
  if (!f || !*f) return 0;
  if (!strcmp(f, "decline")) return 1;
  if (!strcmp(f, "oom")) return 3;
  return 0;
}

// ------------------------------------------------------------------ device buffers
template <typename T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  int ensure(size_t m) {
    if (m <= n && p) return 0;
    if (p) hipFree(p);
    p = nullptr;
    n = 0;
    if (m == 0) m = 1;
    hipError_t e = hipMalloc((void**)&p, m * sizeof(T));
    if (e != hipSuccess) {
      set_err("hipMalloc(%zu bytes): %s", m * sizeof(T), hipGetErrorString(e));
      p = nullptr;
      return e == hipErrorOutOfMemory ? 3 : 2;
    }
    n = m;
    return 0;
  }
  void release() {
    if (p) hipFree(p);
    p = nullptr;
    n = 0;
  }
};

static inline int ncart(int l) { return (l + 1) * (l + 2) / 2; }
static inline int class_of(int la, int lb) { return la * (la + 1) / 2 + lb; }  // la >= lb
// Rys model FLOPs of one primitive quartet of a class (eri_device.h's
// flops_prim_rt), for ordering work by cost on the host
static inline long long flops_prim_host(int la, int lb, int lc, int ld) {
  const int lab = la + lb, lcd = lc + ld, nr = (lab + lcd) / 2 + 1;
  const int nc = ncart(la) * ncart(lb) * ncart(lc) * ncart(ld);
  return (long long)nr * (56 + 12 * (lab + 1) * (lcd + 1) + 6 * (la + 1) * (lb + 1) * (lcd + 1) +
                          6 * (la + 1) * (lb + 1) * (lc + 1) * (ld + 1) + 3 * nc);
}
static inline int nao_of(int l, bool pure) { return (pure && l >= 2) ? 2 * l + 1 : (l + 1) * (l + 2) / 2; }
// primitive-pair-count buckets of the class lists
static constexpr int NBKT = OQPH_NBKT;
// Primitive-pair-count buckets: the 64 kets of a wavefront come from one
// bucket, and its lanes wait for the most contracted of them (14 buckets
// growing by 1.3x measured no faster overall on (H2O)64/cc-pVDZ: the s/p
// classes gained 5-8 %, the d classes lost as much to partial wavefronts).
static inline int bucket_of(int nprim) {
  static const int hi[OQPH_NBKT - 1] = {1, 4, 9, 20, 40};
  int b = 0;
  while (b < OQPH_NBKT - 1 && nprim > hi[b]) ++b;
  return b;
}

// Extra streams of a session: the class launches of one build (or gradient)
// are independent (each adds into F or de with atomics), so they are spread
// over the session stream and these, which lets the small launches of a
// small molecule fill CUs that one launch alone leaves idle.
#define OQPH_SIDE_STREAMS 3

struct oqph_session {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int nstreams = 1 + OQPH_SIDE_STREAMS;  // OQPH_STREAMS (1..4): streams a build's launches use
  hipStream_t side[OQPH_SIDE_STREAMS] = {};
  hipEvent_t ev_fork = nullptr, ev_join[OQPH_SIDE_STREAMS] = {};
  // basis
  int nsh = 0, nbf = 0, natm = 0, lmax = 0;
  bool pure = false;
  std::vector<int> am, harm, atom, nprim, poff, aooff, nao;
  std::vector<double> ex, cc, xyz, zn;
  double cutoff = 5.0e-11, grad_cutoff = 1.0e-10;
  int policy = 0;
  double fp32_cut = 0.0;
  // geometry-derived state
  bool ready = false;
  std::vector<ShellPair> pairs;              // regular cutoffs, indexed by pair id i(i+1)/2+j
  std::vector<double> Q;                     // Schwarz per pair id
  std::vector<std::vector<int>> cls;         // per class: pair ids sorted by Q desc
  std::vector<std::vector<double>> clsQ;     // matching Q values
  DBuf<ShellPair> d_pairs;
  DBuf<PrimJK> d_prims;
  DBuf<PrimF32> d_prims32;                   // FP32 copies of d_prims (mixed mode)
  std::vector<ShellPair> h_dpairs;           // host copy of d_pairs
  std::vector<long long> dev_pid;            // pair id i(i+1)/2+j of each device pair
  // erf-attenuated pass (CAM): pair records with the attenuated Schwarz bounds
  DBuf<ShellPair> d_pairs_att;
  double att_mu = -1.0;
  // OpenQP's MPI share of the bra shell pairs (int2.F90:759-761): rank of size
  int share_rank = 0, share_size = 1;
  bool share_ready = false;
  DBuf<unsigned char> d_own;                 // per device pair: 1 when the share owns it
  DBuf<double> d_rys, d_xform;
  DBuf<float> d_rys32;
  DBuf<int> d_cls;                           // concatenated class lists
  // ket-side Coulomb buffer (JKParams::Jk): device pair i owns the entries
  // [joff[i], joff[i + 1]) (its AO block); class c's block starts at jbase[c]
  std::vector<int> h_joff, jbase;
  DBuf<int> d_joff;
  DBuf<double> d_jk;
  DBuf<int> d_kbs;                           // per class: bucket starts (NBKT+1)
  std::vector<std::vector<int>> kbs;
  std::vector<size_t> cls_off;
  DBuf<int> d_aooff, d_nao;
  // per build
  DBuf<double> d_D, d_F, d_dsh, d_tmp;
  DBuf<double> d_gD, d_gF;  // response-pass densities and outputs (nd x nbf x nbf)
  DBuf<int> d_gcoul;
  DBuf<long long> d_off;
  DBuf<unsigned long long> d_ctr;
  double cached_maxden = -1.0;
  bool cached_two = false;  // lists prepared for the open-shell pass (K rows of two densities in LDS)
  // response pass over non-symmetric densities (oqph_mrsf_int2): every class
  // on the run-time-class kernel
  bool general = false, cached_general = false;
  std::vector<long long> h_off;
  struct GLaunch {
    int la, lb, lc, ld;
    WaveWork ww;
    bool hl;     // run-time-class kernel (a class without a template), layout hd
    HLDims hd;
    bool multi = false;  // several run-time classes in one launch (tl, nblocks, lds)
    HLTaskList tl{};
    long long nblocks = 0;
    int lds = 0;
    bool f32 = false;  // the FP32 form of the class template (mixed-precision gradient)
  };
  // Grouped bra-row kernel (k_jkg): per shell, the shell whose primitive set
  // it is contracted from (fam[i] = i for a family's master); per bra class
  // and member cap, the chunks of pairs that share primitive pairs
  std::vector<int> fam;
  struct ChunkSet {
    std::vector<int> hdr, pool;          // host copies of the device tables
    std::vector<std::vector<int>> pos;   // per chunk: members' class-list positions
    std::vector<int> ncp, nrow, nshl, stride;
    DBuf<int> d_hdr, d_pool;
    DBuf<double> d_cp;
    DBuf<float> d_cp32;
    ChunkView view() const { return {d_hdr.p, d_pool.p, d_cp.p, d_cp32.p}; }
  };
  std::map<int, std::unique_ptr<ChunkSet>> chunksets;  // key: bra class * 16 + cap
  int jkg_mode = 1;  // OQPH_JKG: 0 = segmented bra-row kernel only, 1 = default class pairs, 2 = all
  int jkg_cap = 0;   // OQPH_JKG_CAP: at most this many members per chunk (0: the kernel's)
  int jkt_mode = 0;  // OQPH_JKT: 0 (default) = no tiles, 1 = (ps|ps), 2 = every class pair with a tiled kernel
  struct WLaunch {
    int la, lb, lc, ld;
    bool grouped = false;  // grouped bra-row kernel (k_jkg; bw, cm, cv, lds)
    bool tiled = false;    // its tiled form (tiles in the work lists; bw, cm, lds)
    ChunkView cv{};
    bool rows;   // bra-row kernel (bw, cm, lds) instead of wavefront tasks (ww)
    bool hl;     // run-time-class kernel on the wavefront tasks (f shells), layout hd
    HLDims hd;
    WaveWork ww;   // the list pointers are set per build (bind_lists)
    BraWork bw;
    ColMap cm;
    size_t lds;
    size_t seg0, ch0, item0;  // offsets of this launch in the work lists
    long long slots = 0;      // ket lanes the launch's wavefronts visit (64 per ket iteration)
    // multi-class launch of the run-time-class kernel: tasks [task0, task0 +
    // ntask) of the task list, nblocks blocks, lds bytes (the largest class)
    bool multi = false;
    int task0 = 0, ntask = 0;
    long long nblocks = 0;
    HLTaskList tl{};
  };
  // bra-row kernel: LDS column spaces per angular momentum and its work items
  std::vector<int> h_colidx, h_colao, colbase, ncolL;
  DBuf<int> d_colidx, d_colao, d_items;
  std::vector<int> h_items;
  size_t lds_max = 0;
  // OQPH_JK_ROWS: 0 = wavefront-task kernel only, 2 = bra-row kernel wherever its
  // K rows fit in LDS, 3 = as 2 without the LDS density rows, otherwise
  // (default) the bra-row kernel where it fits and also saves atomics
  int rows_mode = 1;
  // OQPH_LAUNCH_ORDER: 1 (default) = the launches of a build or gradient in
  // decreasing per-quartet cost (run-time classes first), 0 = class order
  int launch_order = 1;
  // OQPH_F_TEMPLATES=0: every class with an f shell on the run-time-class
  // kernels (k_jkh, k_gradh) instead of the class templates where they exist
  bool f_templates = true;
  // OQPH_GRAD_HL_NG=n: gradient classes with more than n Cartesian components
  // per quartet run on the run-time-class kernel k_gradh (0: templates wherever
  // they exist and fit).  Default 300, as for the J/K build: (H2O)16/cc-pVTZ
  // gradient 548 -> 507 ms, C20H42 unchanged (150: 526 and 24.6 ms, 80: slower)
  int grad_hl_ng = 300;
  // OQPH_JK_HL_NG=n: the same for the J/K build (k_jkh).  Default 300: the
  // d classes above 300 Cartesian components ((dp|dp), (dd|pp), (dd|dp))
  // run on k_jkh, whose multi-class launches now beat their spilling
  // templates ((H2O)64/cc-pVDZ 366 -> 359 ms mixed, (H2O)16/cc-pVTZ 450 -> 435)
  int jk_hl_ng = 300;
  // OQPH_HL_KPB=64|32|16|8: kets per block of every run-time-class launch
  // (0, default: chosen per launch by hl_kpb)
  int hl_kpb = 0;
  int ncu = 0;  // compute units of the device
  std::vector<WLaunch> launches;
  unsigned long long list_gen = 0;       // bumped whenever the host work lists change
  unsigned long long list_uploaded = 0;  // list_gen of the copy in d_seg / d_choff / d_items
  unsigned long long serial = 0;         // process-unique session number
  DBuf<int> d_seg;                 // seg_bra | seg_k0 | seg_n of all launches
  DBuf<long long> d_choff;
  DBuf<HLTask> d_tasks;            // classes of the multi-class run-time-class launches
  DBuf<HLTask> d_gtasks;           // the same for the gradient
  std::vector<int> h_seg_bra, h_seg_k0, h_seg_n;
  std::vector<long long> h_choff;
  std::vector<HLTask> h_tasks;
  int hl_multi = 1;                // OQPH_HL_MULTI=0: one launch per run-time class
  // gradient state (pair data with the grd2 cutoffs, coarse-screened lists)
  DBuf<ShellPair> d_gpairs;
  DBuf<PrimPair> d_gprims;
  DBuf<int> d_gcls, d_coff, d_satom, d_am, d_pure;
  DBuf<int> d_gseg;                // gradient seg_bra | seg_k0 | seg_n
  DBuf<long long> d_gchoff;
  DBuf<double> d_Dc, d_de;
  DBuf<double> d_Dt;  // total density of a UHF build
  int ncf = 0;
  double gcut_ready = -1.0;
  // gradient precision (oqph_session_set_grad_precision): 1 = quartets with
  // Q_ij Q_kl below grad_fp32_cut run in the FP32 form of their template
  int grad_policy = 0;
  // MRSF gradient (oqph_grad2_mrsf): the MRSF form of the class templates
  // where it exists (up to 27 Cartesian components), else the run-time-class kernel
  bool grad_all_hl = false, ghl_ready = false;
  DBuf<double> d_gM, d_gMc;  // the eleven MRSF matrices, AO (row-major) and Cartesian
  double grad_fp32_cut = 1.0e-4;
  int gpolicy_ready = -1;
  double gf32_ready = -1.0;
  std::vector<GLaunch> glaunches;  // gradient wavefront tasks
  // optional per-launch profile of the last build (HIP events on the build stream)
  bool profile = false;
  std::vector<hipEvent_t> pev;
  std::vector<float> prof_ms;
  std::vector<unsigned long long> prof_ctr;   // 8 per launch
};

static oqph_session* g_default = nullptr;
namespace oqph {
// the default session (sigma.hip) and the error text its entry points set
oqph_session* default_session();
void set_error_text(const char* msg);
}  // namespace oqph
static std::mutex g_mtx;
static unsigned long long g_serial = 0;  // last session serial handed out (under g_mtx)

// ------------------------------------------------ no-allocation scratch
// Device memory of oqph_fock_jk_noalloc: the shell-pair density maxima, the
// statistics counters and the work lists of one build, in a static array
// that every GPU the library runs on holds until the process ends.  Host code
// never uses the array's name as an address: each device's copy is looked up
// once with hipGetSymbolAddress (oqph_session_create, device current) and
// that pointer is what kernels, memsets and copies receive.
static __device__ __align__(256) unsigned char scratch[OQPH_NOALLOC_SCRATCH_BYTES];
#define OQPH_MAX_DEVICES 64
struct ScratchSlot {
  unsigned char* ptr = nullptr;     // this device's copy of scratch
  int looked_up = 0;                // 1 after hipGetSymbolAddress ran (ptr null if it failed)
  unsigned long long serial = 0;    // session and list generation whose work
  unsigned long long gen = 0;       // lists the copy holds
};
static ScratchSlot g_scratch[OQPH_MAX_DEVICES];

// once per device, with that device current (called under g_mtx)
static void scratch_lookup(int device) {
  if (device < 0 || device >= OQPH_MAX_DEVICES || g_scratch[device].looked_up) return;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  g_scratch[device].looked_up = 1;
  g_scratch[device].ptr = (e == hipSuccess) ? (unsigned char*)ptr : nullptr;
  if (e != hipSuccess) (void)hipGetLastError();  // leave no sticky error for the next call
}

#include "ao_xform.h"
using oqph_host::build_xform;

// ------------------------------------------------------------------ pair data
struct Cut {
  double quartet, expo;
};

static void build_pairs(const oqph_session& s, const Cut& cut, std::vector<ShellPair>& pairs,
                        std::vector<PrimPair>& prims) {
  const double sqrtpi52 = std::sqrt(2.0) * std::pow(M_PI, 1.25);
  const size_t np = (size_t)s.nsh * (s.nsh + 1) / 2;
  pairs.assign(np, ShellPair{});
  prims.clear();
  for (int i = 0; i < s.nsh; ++i)
    for (int j = 0; j <= i; ++j) {
      const size_t pid = (size_t)i * (i + 1) / 2 + j;
      int a = i, b = j;
      if (s.am[j] > s.am[i]) std::swap(a, b);
      ShellPair sp{};
      const double* A = &s.xyz[3 * s.atom[a]];
      const double* B = &s.xyz[3 * s.atom[b]];
      double ab2 = 0.0;
      for (int d = 0; d < 3; ++d) {
        sp.A[d] = A[d];
        sp.B[d] = B[d];
        ab2 += (A[d] - B[d]) * (A[d] - B[d]);
      }
      sp.sa = a; sp.sb = b;
      sp.ao_a = s.aooff[a]; sp.ao_b = s.aooff[b];
      sp.si = i; sp.sj = j;
      sp.la = s.am[a]; sp.lb = s.am[b];
      sp.prim0 = (int)prims.size();
      for (int p = 0; p < s.nprim[a]; ++p)
        for (int q = 0; q < s.nprim[b]; ++q) {
          const double al = s.ex[s.poff[a] + p], be = s.ex[s.poff[b] + q];
          const double g = al + be;
          double e12 = al * be * ab2;
          if (e12 > g * cut.expo) continue;          // int2_pairs.F90:248
          e12 /= g;
          const double k1 = s.cc[s.poff[a] + p] * s.cc[s.poff[b] + q] * std::exp(-e12);
          if (std::fabs(k1) < cut.quartet) continue;  // int2_pairs.F90:252
          PrimPair pp{};
          pp.g = g;
          pp.Kg = sqrtpi52 * k1 / g;
          for (int d = 0; d < 3; ++d) pp.P[d] = (al * A[d] + be * B[d]) / g;
          pp.a = al;
          pp.b = be;
          pp.h = 0.5 / g;
          prims.push_back(pp);
        }
      sp.nprim = (int)prims.size() - sp.prim0;
      pairs[pid] = sp;
    }
}

static int upload(void* dst, const void* src, size_t bytes, hipStream_t st) {
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st));
  return 0;
}
// upload from a host buffer that may go out of scope right after: the copy
// has completed when this returns (set-up paths only)
static int upload_sync(void* dst, const void* src, size_t bytes, hipStream_t st) {
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
  return 0;
}

// J/K primitive-pair records of one shell pair: product centre relative to
// the pair's first centre A, formed in FP64 (and the FP32 copy)
static PrimJK jk_record(const PrimPair& q, const ShellPair& sp) {
  return {q.g, q.Kg, {q.P[0] - sp.A[0], q.P[1] - sp.A[1], q.P[2] - sp.A[2]}, q.h};
}
static PrimF32 f32_record(const PrimJK& q) {
  return {(float)q.g, (float)q.Kg, {(float)q.X[0], (float)q.X[1], (float)q.X[2]}, (float)q.h};
}

// Schwarz bounds Q = sqrt(max |(ab|ab)|) of every shell pair (pair id
// i(i+1)/2 + j) with the tight cutoffs of ints_exchange (exponent 50,
// quartet 1e-17; int2.F90:1600-1604); imu2 = 1/mu^2 gives the bounds of the
// erf-attenuated integrals (0: 1/r).
static int schwarz_bounds(oqph_session* s, double imu2, std::vector<double>& Qout) {
  int rc;
  std::vector<ShellPair> tp;
  std::vector<PrimPair> tpr;
  build_pairs(*s, Cut{1.0e-17, 50.0}, tp, tpr);
  const size_t np = tp.size();
  std::vector<PrimJK> tjk;
  tjk.reserve(tpr.size());
  for (size_t pid = 0; pid < np; ++pid)
    for (int k = tp[pid].prim0; k < tp[pid].prim0 + tp[pid].nprim; ++k) tjk.push_back(jk_record(tpr[k], tp[pid]));
  std::vector<std::vector<int>> byc(10);
  for (size_t pid = 0; pid < np; ++pid) byc[class_of(tp[pid].la, tp[pid].lb)].push_back((int)pid);
  std::vector<int> flat;
  std::vector<size_t> foff(11, 0);
  for (int c = 0; c < 10; ++c) {
    foff[c] = flat.size();
    flat.insert(flat.end(), byc[c].begin(), byc[c].end());
  }
  foff[10] = flat.size();
  DBuf<ShellPair> dtp;
  DBuf<PrimJK> dtpr;
  DBuf<int> dlist;
  DBuf<double> dQ;
  if ((rc = dtp.ensure(np)) || (rc = dtpr.ensure(tjk.size())) || (rc = dlist.ensure(np)) || (rc = dQ.ensure(np)))
    return rc;
  if ((rc = upload_sync(dtp.p, tp.data(), np * sizeof(ShellPair), s->stream))) return rc;
  if (!tjk.empty() && (rc = upload_sync(dtpr.p, tjk.data(), tjk.size() * sizeof(PrimJK), s->stream))) return rc;
  if ((rc = upload_sync(dlist.p, flat.data(), flat.size() * sizeof(int), s->stream))) return rc;
  HIPCHK(hipMemsetAsync(dQ.p, 0, np * sizeof(double), s->stream));
  for (int la = 0; la <= s->lmax; ++la)
    for (int lb = 0; lb <= la; ++lb) {
      const int c = class_of(la, lb);
      if (byc[c].empty()) continue;
      KernelRes kr{};
      const bool tmpl = la < 3 && kernel_resources(3, la, lb, la, lb, s->pure, &kr) &&
                        kr.private_bytes <= OQPH_PRIVATE_BUDGET;
      if (!tmpl) {
        const HLDims h = hl_dims(la, lb, la, lb, s->pure, 0, 0);
        if ((size_t)h.lds > s->lds_max) {
          set_err("Schwarz class (%d %d): %d bytes of LDS exceed the device's %zu", la, lb, h.lds, s->lds_max);
          HIPCHK(hipStreamSynchronize(s->stream));
          return 1;
        }
        launch_schwarzh(h, dtp.p, dtpr.p, dlist.p + foff[c], (int)byc[c].size(), s->d_rys.p, s->d_xform.p, 1.0e-17,
                        imu2, dQ.p, s->stream);
        continue;
      }
      if (!launch_schwarz(la, lb, s->pure, dtp.p, dtpr.p, dlist.p + foff[c], (int)byc[c].size(), s->d_rys.p,
                          s->d_xform.p, 1.0e-17, imu2, dQ.p, s->stream)) {
        set_err("no Schwarz kernel for class (%d %d)", la, lb);
        HIPCHK(hipStreamSynchronize(s->stream));
        return 1;
      }
    }
  Qout.assign(np, 0.0);
  const hipError_t e1 = hipGetLastError();
  const hipError_t e2 = hipMemcpyAsync(Qout.data(), dQ.p, np * sizeof(double), hipMemcpyDeviceToHost, s->stream);
  const hipError_t e3 = hipStreamSynchronize(s->stream);  // before the temporaries are freed
  HIPCHK(e1);
  HIPCHK(e2);
  HIPCHK(e3);
  return 0;
}

// OpenQP's shell-pair work list order (int2_build_shell_pair_map,
// source/integrals/int2.F90:864-921): pairs (i, j), i >= j (1-based), in
// classes of log2 of the cost Q_ij nbf_i nbf_j i (Cartesian counts), most
// expensive class first, i descending then j ascending inside a class.
// Returns the 1-based position ij_pair of every pair id i(i+1)/2 + j (0-based
// i, j); rank r of size n owns the pairs with mod(ij_pair, n) == r
// (:759-761).
static std::vector<long long> reference_pair_order(const oqph_session& s) {
  const int nsh = s.nsh, NCLASS = 64, OFFS = 42;
  auto cls = [&](int i, int j) {  // 0-based shells
    const int nbi = ncart(s.am[i]), nbj = ncart(s.am[j]);
    const double sw = std::max(s.Q[(size_t)i * (i + 1) / 2 + j], 1.0e-30);
    const double cost = sw * (double)(nbi * nbj) * (double)(i + 1);
    return NCLASS - std::max(0, std::min(NCLASS, (int)(std::log(cost) / std::log(2.0)) + OFFS));
  };
  std::vector<long long> off(NCLASS + 1, 0), cnt(NCLASS + 1, 0);
  for (int i = nsh - 1; i >= 0; --i)
    for (int j = 0; j <= i; ++j) ++cnt[cls(i, j)];
  for (int c = 1; c <= NCLASS; ++c) off[c] = off[c - 1] + cnt[c - 1];
  std::vector<long long> pos((size_t)nsh * (nsh + 1) / 2);
  for (int i = nsh - 1; i >= 0; --i)
    for (int j = 0; j <= i; ++j) pos[(size_t)i * (i + 1) / 2 + j] = ++off[cls(i, j)];
  return pos;
}

// Per-device-pair ownership flags of the current (rank, size) share.
static int prepare_share(oqph_session* s) {
  if (s->share_size <= 1) return 0;
  if (s->share_ready) return 0;
  const std::vector<long long> pos = reference_pair_order(*s);
  std::vector<unsigned char> own(s->dev_pid.size());
  for (size_t k = 0; k < s->dev_pid.size(); ++k)
    own[k] = (pos[s->dev_pid[k]] % s->share_size) == s->share_rank ? 1 : 0;
  int rc;
  if ((rc = s->d_own.ensure(std::max<size_t>(own.size(), 1)))) return rc;
  if (!own.empty() && (rc = upload_sync(s->d_own.p, own.data(), own.size(), s->stream))) return rc;
  s->share_ready = true;
  return 0;
}

// Families of the grouped bra-row kernel: shell i is evaluated with the
// primitive set of fam[i], the first shell on the same atom with the same
// angular momentum and the same exponents (cc-pVDZ oxygen: 1s and 2s over one
// set of nine s exponents).  Shells whose exponents are a proper subset (the
// single-primitive 3s) stay apart: with them a chunk would evaluate its whole
// primitive grid for kets that only the small member keeps.
static void build_families(oqph_session* s) {
  const int n = s->nsh;
  s->fam.assign(n, 0);
  auto subset = [&](int i, int j) {  // exponents of i all among those of j
    for (int p = 0; p < s->nprim[i]; ++p) {
      const double e = s->ex[s->poff[i] + p];
      bool found = false;
      for (int q = 0; q < s->nprim[j] && !found; ++q) found = s->ex[s->poff[j] + q] == e;
      if (!found) return false;
    }
    return true;
  };
  for (int i = 0; i < n; ++i) {
    int m = i;
    for (int j = 0; j < n; ++j)
      if (j < m && s->atom[j] == s->atom[i] && s->am[j] == s->am[i] && s->nprim[j] == s->nprim[i] && subset(i, j))
        m = j;
    s->fam[i] = m;
  }
}

// Chunks of bra class cb with at most cap members: the class's pairs grouped
// by the families of their two shells (same two atoms, same primitive
// exponents), members in decreasing primitive-pair count, cut into chunks of
// cap.  Each chunk's primitive records cover the union of its members'
// primitive pairs (each member keeps its own K/zeta, 0 where its pair
// screens dropped the pair); its K rows are the AOs of the members' distinct
// shells.
static int build_chunkset(oqph_session* s, int cb, int cap, const std::vector<PrimPair>& prims,
                          oqph_session::ChunkSet& cs) {
  const std::vector<int>& list = s->cls[cb];
  std::map<std::pair<int, int>, std::vector<int>> groups;  // (fam a, fam b) -> positions
  std::vector<std::pair<int, int>> order;
  for (int p = 0; p < (int)list.size(); ++p) {
    const ShellPair& sp = s->pairs[list[p]];
    const std::pair<int, int> key{s->fam[sp.sa], s->fam[sp.sb]};
    auto it = groups.find(key);
    if (it == groups.end()) {
      order.push_back(key);
      groups[key] = {p};
    } else {
      it->second.push_back(p);
    }
  }
  std::vector<double> cp;
  std::vector<float> cp32;
  for (const auto& key : order) {
    std::vector<int> mem = groups[key];
    std::stable_sort(mem.begin(), mem.end(),
                     [&](int x, int y) { return s->pairs[list[x]].nprim > s->pairs[list[y]].nprim; });
    for (size_t c0 = 0; c0 < mem.size(); c0 += cap) {
      const int nm = (int)std::min<size_t>(cap, mem.size() - c0);
      const ShellPair& p0 = s->pairs[list[mem[c0]]];
      // union of the members' primitive pairs, keyed by the two exponents
      std::map<std::pair<double, double>, int> gi;
      std::vector<const PrimPair*> base;
      std::vector<std::vector<double>> kg;
      for (int m = 0; m < nm; ++m) {
        const ShellPair& sp = s->pairs[list[mem[c0 + m]]];
        for (int k = sp.prim0; k < sp.prim0 + sp.nprim; ++k) {
          const PrimPair& q = prims[k];
          auto r = gi.emplace(std::make_pair(q.a, q.b), (int)base.size());
          if (r.second) {
            base.push_back(&q);
            kg.emplace_back(nm, 0.0);
          }
          kg[r.first->second][m] = q.Kg;
        }
      }
      // distinct shells of the members, their K rows and dsh slots
      std::vector<int> shl, row0;
      int nrow = 0;
      auto slot = [&](int sh) {
        for (size_t t = 0; t < shl.size(); ++t)
          if (shl[t] == sh) return (int)t;
        shl.push_back(sh);
        row0.push_back(nrow);
        nrow += s->nao[sh];
        return (int)shl.size() - 1;
      };
      std::vector<int> ent;
      for (int m = 0; m < nm; ++m) {
        const ShellPair& sp = s->pairs[list[mem[c0 + m]]];
        const int ta = slot(sp.sa), tb = slot(sp.sb);
        const int ti = slot(sp.si), tj = slot(sp.sj);
        ent.insert(ent.end(), {(int)s->cls_off[cb] + mem[c0 + m], row0[ta], row0[tb], ti, tj});
      }
      const int stride = 6 + nm;
      const int c = (int)cs.hdr.size();
      cs.hdr.push_back((int)cs.pool.size());
      const int o = (int)cs.pool.size();
      const int hdr_len = OQPH_JKG_HDR + OQPH_JKG_MEM * nm;
      cs.pool.insert(cs.pool.end(), {nm, (int)cp.size(), (int)base.size(), nrow, (int)shl.size(), stride,
                                     hdr_len, hdr_len + nrow});
      cs.pool.insert(cs.pool.end(), ent.begin(), ent.end());
      for (size_t t = 0; t < shl.size(); ++t)
        for (int k = 0; k < s->nao[shl[t]]; ++k) cs.pool.push_back(s->aooff[shl[t]] + k);
      cs.pool.insert(cs.pool.end(), shl.begin(), shl.end());
      (void)o;
      for (size_t g = 0; g < base.size(); ++g) {
        const PrimPair& q = *base[g];
        double km = 0.0;
        for (int m = 0; m < nm; ++m) km = std::max(km, std::fabs(kg[g][m]));
        const double rec[6] = {q.g, q.P[0] - p0.A[0], q.P[1] - p0.A[1], q.P[2] - p0.A[2], q.h, km};
        for (double v : rec) {
          cp.push_back(v);
          cp32.push_back((float)v);
        }
        for (int m = 0; m < nm; ++m) {
          cp.push_back(kg[g][m]);
          cp32.push_back((float)kg[g][m]);
        }
      }
      std::vector<int> pos(mem.begin() + c0, mem.begin() + c0 + nm);
      cs.pos.push_back(pos);
      cs.ncp.push_back((int)base.size());
      cs.nrow.push_back(nrow);
      cs.nshl.push_back((int)shl.size());
      cs.stride.push_back(stride);
      (void)c;
    }
  }
  int rc;
  if ((rc = cs.d_hdr.ensure(std::max<size_t>(cs.hdr.size(), 1))) ||
      (rc = cs.d_pool.ensure(std::max<size_t>(cs.pool.size(), 1))) ||
      (rc = cs.d_cp.ensure(std::max<size_t>(cp.size(), 1))) || (rc = cs.d_cp32.ensure(std::max<size_t>(cp32.size(), 1))))
    return rc;
  if (!cs.hdr.empty() &&
      ((rc = upload_sync(cs.d_hdr.p, cs.hdr.data(), cs.hdr.size() * sizeof(int), s->stream)) ||
       (rc = upload_sync(cs.d_pool.p, cs.pool.data(), cs.pool.size() * sizeof(int), s->stream)) ||
       (rc = upload_sync(cs.d_cp.p, cp.data(), cp.size() * sizeof(double), s->stream)) ||
       (rc = upload_sync(cs.d_cp32.p, cp32.data(), cp32.size() * sizeof(float), s->stream))))
    return rc;
  return 0;
}

// The member cap of a (bra class, ket class) launch of the grouped kernel;
// 0 when the launch keeps the segmented kernels.  By default the grouped
// kernel runs (ss|ss) and (ps|ps), the class pairs where it measured faster
// than the segmented bra-row kernel on (H2O)64/cc-pVDZ ((ss|ss) FP64 72 -> 61
// ms, (ps|ps) mixed 78 -> 74 ms); for (ps|ss), (ds|ss) and (ds|ps) its
// larger register file (three waves per SIMD instead of four) cost more than
// the shared quadrature saved.  OQPH_JKG=2 sends every class pair that has
// a grouped kernel to it.
static int jkg_cap_for(const oqph_session* s, int la, int lb, int lc, int ld) {
  if (!s->jkg_mode) return 0;
  if (s->jkg_mode == 1 && !((la == 0 && lb == 0 && lc == 0 && ld == 0) || (la == 1 && lb == 0 && lc == 1 && ld == 0)))
    return 0;
  size_t st = 0;
  int nm = jkg_info(la, lb, lc, ld, s->pure, &st, s->lds_max);
  if (s->jkg_cap > 0) nm = std::min(nm, s->jkg_cap);
  return nm >= 2 ? nm : 0;
}

// Chunk sets of every (bra class, cap) a launch can use, built with the
// geometry so that builds never allocate them.
// Members per tile of a (bra class, ket class) launch of the tiled kernel; 0
// when the launch keeps the segmented bra-row kernel for its single bras.
// Tiles are off by default: they were written to share each ket's J_cd atomic
// among the bras of a tile ((ps|ps) on (H2O)64/cc-pVDZ 58 -> 38 ms mixed), but
// the ket-side Coulomb buffer removed that cost for every class ((ps|ps) 28
// ms untiled, 29 ms tiled; FP64 36 vs 47 ms).  OQPH_JKT=1 tiles (ps|ps),
// OQPH_JKT=2 every class pair that has a tiled kernel.
static int jkt_cap_for(const oqph_session* s, int la, int lb, int lc, int ld) {
  if (!s->jkt_mode) return 0;
  if (s->jkt_mode == 1 && !(la == 1 && lb == 0 && lc == 1 && ld == 0)) return 0;
  size_t st = 0;
  const int nm = jkt_info(la, lb, lc, ld, s->pure, &st, s->lds_max);
  return nm >= 2 ? nm : 0;
}

static int prepare_chunks(oqph_session* s, const std::vector<PrimPair>& prims) {
  s->chunksets.clear();
  if (!s->jkg_mode) return 0;
  const int ncls = class_of(s->lmax, s->lmax) + 1;
  for (int cb = 0; cb < ncls; ++cb) {
    if (s->cls[cb].empty()) continue;
    // does the class have any pair that shares its primitives with another?
    std::map<std::pair<int, int>, int> cnt;
    bool shared = false;
    for (int pid : s->cls[cb]) {
      const ShellPair& sp = s->pairs[pid];
      shared |= ++cnt[{s->fam[sp.sa], s->fam[sp.sb]}] > 1;
    }
    if (!shared) continue;
    int la = 0, lb = 0;
    for (int x = 0; x <= 3; ++x)
      for (int y = 0; y <= x; ++y)
        if (class_of(x, y) == cb) { la = x; lb = y; }
    for (int ck = 0; ck <= cb; ++ck) {
      int lc = 0, ld = 0;
      for (int x = 0; x <= 3; ++x)
        for (int y = 0; y <= x; ++y)
          if (class_of(x, y) == ck) { lc = x; ld = y; }
      const int cap = jkg_cap_for(s, la, lb, lc, ld);
      if (cap < 2 || s->chunksets.count(cb * 16 + cap)) continue;
      auto cs = std::make_unique<oqph_session::ChunkSet>();
      int rc = build_chunkset(s, cb, cap, prims, *cs);
      if (rc) return rc;
      s->chunksets[cb * 16 + cap] = std::move(cs);
    }
  }
  return 0;
}

static int prepare_geometry(oqph_session* s) {
  if (s->ready) return 0;
  if (s->nsh == 0 || s->natm == 0) {
    set_err("basis or geometry not set");
    return 4;
  }
  HIPCHK(hipSetDevice(s->device));
  int rc;
  {
    int v = 0;
    HIPCHK(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, s->device));
    s->lds_max = (size_t)v;
    HIPCHK(hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, s->device));
    s->ncu = std::max(v, 1);
  }
  // Rys tables and AO transforms
  size_t nt;
  const double* tab = rys_table_host(&nt);
  if ((rc = s->d_rys.ensure(nt))) return rc;
  if ((rc = upload_sync(s->d_rys.p, tab, nt * sizeof(double), s->stream))) return rc;
  size_t n32;
  const float* tab32 = rys32_table_host(&n32);
  if ((rc = s->d_rys32.ensure(n32))) return rc;
  if ((rc = upload_sync(s->d_rys32.p, tab32, n32 * sizeof(float), s->stream))) return rc;
  double xf[2][2][10][10];
  build_xform(xf);
  if ((rc = s->d_xform.ensure(400))) return rc;
  if ((rc = upload_sync(s->d_xform.p, &xf[0][0][0][0], 400 * sizeof(double), s->stream))) return rc;
  {
    const char* ft = getenv("OQPH_F_TEMPLATES");
    s->f_templates = !(ft && ft[0] == '0');
    const char* gh = getenv("OQPH_GRAD_HL_NG");
    s->grad_hl_ng = gh ? atoi(gh) : 300;
    const char* jh = getenv("OQPH_JK_HL_NG");
    s->jk_hl_ng = jh ? atoi(jh) : 300;
    const char* lo = getenv("OQPH_LAUNCH_ORDER");
    s->launch_order = lo ? atoi(lo) : 1;
    const char* kp = getenv("OQPH_HL_KPB");
    const int kpv = kp ? atoi(kp) : 0;
    s->hl_kpb = (kpv == 64 || kpv == 32 || kpv == 16 || kpv == 8) ? kpv : 0;
    const char* e = getenv("OQPH_JK_ROWS");
    s->rows_mode = (e && (e[0] == '0' || e[0] == '2' || e[0] == '3')) ? e[0] - '0' : 1;
    const char* hm = getenv("OQPH_HL_MULTI");
    s->hl_multi = (hm && hm[0] == '0') ? 0 : 1;
    const char* g = getenv("OQPH_JKG");
    s->jkg_mode = (g && (g[0] == '0' || g[0] == '2')) ? g[0] - '0' : 1;
    const char* jt = getenv("OQPH_JKT");
    s->jkt_mode = (jt && (jt[0] == '1' || jt[0] == '2')) ? jt[0] - '0' : 0;
    const char* gc = getenv("OQPH_JKG_CAP");
    s->jkg_cap = gc ? atoi(gc) : 0;
  }
  if ((rc = schwarz_bounds(s, 0.0, s->Q))) return rc;
  const size_t np = s->Q.size();
  // regular pair data (int2.F90:266-272): quartet 1e-4 c, exponent 25 ln 10
  std::vector<PrimPair> prims;
  build_pairs(*s, Cut{1.0e-4 * s->cutoff, 25.0 * std::log(10.0)}, s->pairs, prims);
  for (size_t pid = 0; pid < np; ++pid) s->pairs[pid].Q = s->Q[pid];
  // class lists of significant pairs, Q descending
  s->cls.assign(10, {});
  s->clsQ.assign(10, {});
  for (size_t pid = 0; pid < np; ++pid) {
    const ShellPair& sp = s->pairs[pid];
    if (sp.nprim == 0 || sp.Q <= 0.0) continue;
    s->cls[class_of(sp.la, sp.lb)].push_back((int)pid);
  }
  std::vector<int> all, kbs_all;
  s->cls_off.assign(11, 0);
  s->kbs.assign(10, std::vector<int>(NBKT + 1, 0));
  for (int c = 0; c < 10; ++c) {
    auto& v = s->cls[c];
    // primitive-count bucket first (uniform wavefronts), then Q descending
    std::stable_sort(v.begin(), v.end(), [&](int x, int y) {
      const int bx = bucket_of(s->pairs[x].nprim), by = bucket_of(s->pairs[y].nprim);
      return bx != by ? bx < by : s->Q[x] > s->Q[y];
    });
    for (int x : v) s->clsQ[c].push_back(s->Q[x]);
    for (int b = 0; b <= NBKT; ++b) {
      int k = 0;
      while (k < (int)v.size() && bucket_of(s->pairs[v[k]].nprim) < b) ++k;
      s->kbs[c][b] = (b == NBKT) ? (int)v.size() : k;
    }
    kbs_all.insert(kbs_all.end(), s->kbs[c].begin(), s->kbs[c].end());
    s->cls_off[c] = all.size();
    all.insert(all.end(), v.begin(), v.end());
  }
  s->cls_off[10] = all.size();
  s->dev_pid.assign(all.begin(), all.end());
  // Device pair and primitive-pair records in class-list order: device pair
  // id i is s->pairs[all[i]], so the 64 kets of a wavefront (consecutive
  // list entries) read consecutive records; the class lists on the device
  // are then the identity.
  {
    std::vector<ShellPair> dp(all.size());
    std::vector<PrimJK> dpr;
    std::vector<PrimF32> dpr32;
    dpr.reserve(prims.size());
    dpr32.reserve(prims.size());
    for (size_t i = 0; i < all.size(); ++i) {
      dp[i] = s->pairs[all[i]];
      dp[i].prim0 = (int)dpr.size();
      const ShellPair& o = s->pairs[all[i]];
      for (int k = o.prim0; k < o.prim0 + o.nprim; ++k) {
        dpr.push_back(jk_record(prims[k], o));
        dpr32.push_back(f32_record(dpr.back()));
      }
    }
    s->h_dpairs = dp;
    s->h_joff.assign(dp.size() + 1, 0);
    for (size_t i = 0; i < dp.size(); ++i)
      s->h_joff[i + 1] = s->h_joff[i] + nao_of(dp[i].la, s->pure) * nao_of(dp[i].lb, s->pure);
    s->jbase.assign(11, 0);
    for (int c = 0; c <= 10; ++c) s->jbase[c] = s->h_joff[s->cls_off[c]];
    if ((rc = s->d_joff.ensure(s->h_joff.size())) ||
        (rc = upload_sync(s->d_joff.p, s->h_joff.data(), s->h_joff.size() * sizeof(int), s->stream)))
      return rc;
    std::vector<int> ident(all.size());
    for (size_t i = 0; i < all.size(); ++i) ident[i] = (int)i;
    if ((rc = s->d_pairs.ensure(std::max<size_t>(dp.size(), 1))) ||
        (rc = s->d_prims.ensure(std::max<size_t>(dpr.size(), 1))) ||
        (rc = s->d_prims32.ensure(std::max<size_t>(dpr.size(), 1))) ||
        (rc = s->d_cls.ensure(std::max<size_t>(all.size(), 1))))
      return rc;
    if (!dp.empty() && ((rc = upload_sync(s->d_pairs.p, dp.data(), dp.size() * sizeof(ShellPair), s->stream)) ||
                        (rc = upload_sync(s->d_prims.p, dpr.data(), dpr.size() * sizeof(PrimJK), s->stream)) ||
                        (rc = upload_sync(s->d_prims32.p, dpr32.data(), dpr32.size() * sizeof(PrimF32), s->stream)) ||
                        (rc = upload_sync(s->d_cls.p, ident.data(), ident.size() * sizeof(int), s->stream))))
      return rc;
  }
  build_families(s);
  if ((rc = prepare_chunks(s, prims))) return rc;
  if ((rc = s->d_kbs.ensure(kbs_all.size() + 2))) return rc;
  if ((rc = upload_sync(s->d_kbs.p, kbs_all.data(), kbs_all.size() * sizeof(int), s->stream))) return rc;
  if ((rc = s->d_aooff.ensure(s->nsh)) || (rc = s->d_nao.ensure(s->nsh))) return rc;
  if ((rc = upload_sync(s->d_aooff.p, s->aooff.data(), s->nsh * sizeof(int), s->stream))) return rc;
  if ((rc = upload_sync(s->d_nao.p, s->nao.data(), s->nsh * sizeof(int), s->stream))) return rc;
  // LDS column spaces of the bra-row kernel: the AOs of the shells of each
  // angular momentum, in shell order
  s->h_colidx.assign(s->nsh, 0);
  s->h_colao.clear();
  s->colbase.assign(s->lmax + 1, 0);
  s->ncolL.assign(s->lmax + 1, 0);
  for (int L = 0; L <= s->lmax; ++L) {
    s->colbase[L] = (int)s->h_colao.size();
    int off = 0;
    for (int i = 0; i < s->nsh; ++i)
      if (s->am[i] == L) {
        s->h_colidx[i] = off;
        for (int k = 0; k < s->nao[i]; ++k) s->h_colao.push_back(s->aooff[i] + k);
        off += s->nao[i];
      }
    s->ncolL[L] = off;
  }
  if ((rc = s->d_colidx.ensure(s->nsh)) || (rc = s->d_colao.ensure(s->h_colao.size()))) return rc;
  if ((rc = upload_sync(s->d_colidx.p, s->h_colidx.data(), s->nsh * sizeof(int), s->stream)) ||
      (rc = upload_sync(s->d_colao.p, s->h_colao.data(), s->h_colao.size() * sizeof(int), s->stream)))
    return rc;
  const size_t ntri = (size_t)s->nbf * (s->nbf + 1) / 2;
  if ((rc = s->d_D.ensure(ntri)) || (rc = s->d_F.ensure(ntri)) || (rc = s->d_dsh.ensure((size_t)s->nsh * s->nsh)) ||
      (rc = s->d_ctr.ensure(8)))
    return rc;
  HIPCHK(hipStreamSynchronize(s->stream));
  s->cached_maxden = -1.0;
  s->gcut_ready = -1.0;
  s->att_mu = -1.0;
  s->share_ready = false;
  s->ready = true;
  return 0;
}

// Device pair records of the erf-attenuated pass (mu): the regular records
// with Q replaced by the attenuated Schwarz bound, which the in-kernel screens
// read (int2.F90:675-681).  The attenuated bound never exceeds the regular
// one (erfc(mu r)/r is a positive kernel too), so the regular coarse lists
// stay a superset and the kernels' exact screens decide.
static int prepare_attenuated(oqph_session* s, double mu) {
  if (s->att_mu == mu) return 0;
  std::vector<double> Qa;
  int rc = schwarz_bounds(s, 1.0 / (mu * mu), Qa);
  if (rc) return rc;
  std::vector<ShellPair> dp = s->h_dpairs;
  for (size_t k = 0; k < dp.size(); ++k) dp[k].Q = Qa[s->dev_pid[k]];
  if ((rc = s->d_pairs_att.ensure(std::max<size_t>(dp.size(), 1)))) return rc;
  if (!dp.empty() && (rc = upload_sync(s->d_pairs_att.p, dp.data(), dp.size() * sizeof(ShellPair), s->stream)))
    return rc;
  s->att_mu = mu;
  return 0;
}

// Kets per block of a run-time-class launch (k_jkh, k_gradh and their
// multi-class forms): their blocks are one wavefront, LDS-limited to per_cu
// per CU, and walk a chunk's surviving quartets one after another, so a
// launch of few 64-ket chunks (a small molecule's f or (dd|dd) classes) would
// leave most of the device idle; such launches split each chunk over 2, 4 or
// 8 blocks until they have two rounds of resident blocks.
static int hl_kpb(const oqph_session* s, long long chunks, int lds) {
  if (s->hl_kpb > 0) return s->hl_kpb;
  const long long per_cu = std::min<long long>(16, (long long)(s->lds_max / (size_t)std::max(lds, 1)));
  const long long target = 2 * std::max<long long>(per_cu, 1) * s->ncu;
  int kpb = 64;
  while (kpb > 8 && chunks * (64 / kpb) < target) kpb /= 2;
  return kpb;
}

// Per-quartet cost of a class (Cartesian block x roots) for the launch order:
// launches whose wavefronts are long and few start first, so the many short
// ones fill the device behind them instead of leaving them as the tail.
static long long class_weight(int la, int lb, int lc, int ld, bool runtime) {
  const long long w = (long long)ncart(la) * ncart(lb) * ncart(lc) * ncart(ld) * ((la + lb + lc + ld) / 2 + 1);
  return runtime ? w + (1ll << 40) : w;
}

// gradient pair data and coarse-screened quartet lists (grd2.F90:219-328):
// pair data with exponent cutoff 20 ln 10 and prefactor cutoff 1e-20; a
// quartet survives the coarse screen when Q_ij Q_kl >= grad_cutoff.
static int prepare_grad(oqph_session* s) {
  int rc = prepare_geometry(s);
  if (rc) return rc;
  if (s->gcut_ready == s->grad_cutoff && s->gpolicy_ready == s->grad_policy &&
      (s->grad_policy == 0 || s->gf32_ready == s->grad_fp32_cut) && s->ghl_ready == s->grad_all_hl)
    return 0;
  // the per-atom gradient sums are staged in LDS (3 natm doubles per block)
  if (24 * (size_t)s->natm + 16384 > s->lds_max) {
    set_err("gradient: %d atoms need %zu bytes of LDS per block for the per-atom sums (device: %zu)", s->natm,
            24 * (size_t)s->natm, s->lds_max);
    return 1;
  }
  std::vector<ShellPair> gp;
  std::vector<PrimPair> gpr;
  build_pairs(*s, Cut{1.0e-20, 20.0 * std::log(10.0)}, gp, gpr);
  const size_t np = gp.size();
  for (size_t pid = 0; pid < np; ++pid) gp[pid].Q = s->Q[pid];
  // class lists ordered by primitive-pair bucket, then Q descending (as the
  // J/K lists), so a wavefront's kets have similar contraction lengths
  std::vector<std::vector<int>> cls(10);
  std::vector<std::vector<double>> clsQ(10);
  std::vector<std::array<int, NBKT + 1>> kbs(10);
  for (size_t pid = 0; pid < np; ++pid)
    if (gp[pid].nprim > 0 && gp[pid].Q > 0.0) cls[class_of(gp[pid].la, gp[pid].lb)].push_back((int)pid);
  std::vector<int> all;
  std::vector<size_t> coff(11, 0);
  for (int c = 0; c < 10; ++c) {
    auto& v = cls[c];
    std::stable_sort(v.begin(), v.end(), [&](int x, int y) {
      const int bx = bucket_of(gp[x].nprim), by = bucket_of(gp[y].nprim);
      return bx != by ? bx < by : s->Q[x] > s->Q[y];
    });
    for (int x : v) clsQ[c].push_back(s->Q[x]);
    for (int b = 0; b <= NBKT; ++b) {
      int k = 0;
      while (k < (int)v.size() && bucket_of(gp[v[k]].nprim) < b) ++k;
      kbs[c][b] = (b == NBKT) ? (int)v.size() : k;
    }
    coff[c] = all.size();
    all.insert(all.end(), v.begin(), v.end());
  }
  // device records in class-list order (device pair id i = gp[all[i]]), as
  // for the J/K build; the device class lists are the identity
  std::vector<ShellPair> dgp(all.size());
  std::vector<PrimPair> dgpr;
  dgpr.reserve(gpr.size());
  std::vector<int> ident(all.size());
  for (size_t i = 0; i < all.size(); ++i) {
    dgp[i] = gp[all[i]];
    dgp[i].prim0 = (int)dgpr.size();
    dgpr.insert(dgpr.end(), gpr.begin() + gp[all[i]].prim0, gpr.begin() + gp[all[i]].prim0 + gp[all[i]].nprim);
    ident[i] = (int)i;
  }
  if ((rc = s->d_gpairs.ensure(std::max<size_t>(dgp.size(), 1))) ||
      (rc = s->d_gprims.ensure(std::max<size_t>(dgpr.size(), 1))) ||
      (rc = s->d_gcls.ensure(std::max<size_t>(all.size(), 1))))
    return rc;
  if (!all.empty() && ((rc = upload_sync(s->d_gpairs.p, dgp.data(), dgp.size() * sizeof(ShellPair), s->stream)) ||
                       (rc = upload_sync(s->d_gprims.p, dgpr.data(), dgpr.size() * sizeof(PrimPair), s->stream)) ||
                       (rc = upload_sync(s->d_gcls.p, ident.data(), ident.size() * sizeof(int), s->stream))))
    return rc;
  // segments (bra p, ket bucket b): the kets with Q_p Q_k >= grad_cutoff form
  // a prefix of the bucket; within one class each unordered pair is taken
  // once, with the bra the later entry of the (bucket, Q) order
  std::vector<int> sbra, sk0, sn;
  std::vector<long long> choff;
  struct Pend { int la, lb, lc, ld, ck; size_t seg0, ch0; int nseg; long long nwaves; bool f32; };
  std::vector<Pend> pend;
  const int ncls = class_of(s->lmax, s->lmax) + 1;
  for (int cb = 0; cb < ncls; ++cb)
    for (int ck = 0; ck <= cb; ++ck) {
      const auto& QB = clsQ[cb];
      const auto& QK = clsQ[ck];
      if (QB.empty() || QK.empty()) continue;
      const auto& kb = kbs[ck];
      const auto& bb = kbs[cb];
      int kmax[NBKT] = {0};
      for (int b = 0; b < NBKT; ++b)
        for (int k = kb[b]; k < kb[b + 1]; ++k) kmax[b] = std::max(kmax[b], gp[cls[ck][k]].nprim);
      struct Seg { int bra, k0, n; long long cost; };
      std::vector<Seg> segs, segs32;  // FP64 and, in the mixed-precision gradient, FP32 kets
      int bp = 0;
      for (int p = 0; p < (int)QB.size(); ++p) {
        while (bp + 1 < NBKT && p >= bb[bp + 1]) ++bp;
        const double thr = s->grad_cutoff / QB[p];
        const int bra = cls[cb][p];
        for (int b = 0; b < NBKT; ++b) {
          if (kb[b + 1] <= kb[b]) continue;
          int k = (int)(std::lower_bound(QK.begin() + kb[b], QK.begin() + kb[b + 1], thr,
                                         [](double a, double t) { return a >= t; }) - (QK.begin() + kb[b]));
          if (cb == ck) {
            if (b > bp) k = 0;
            else if (b == bp) k = std::min(k, p - bb[bp] + 1);
          }
          if (k <= 0) continue;
          // mixed precision: the kets with Q_p Q_k below grad_fp32_cut (a
          // suffix of the prefix) go to the FP32 form; the FP64 part is
          // rounded up to whole wavefronts, so the split adds no partly
          // filled wavefront
          int khi = k;
          if (s->grad_policy == 1 && !s->grad_all_hl) {
            khi = (int)(std::lower_bound(QK.begin() + kb[b], QK.begin() + kb[b] + k, s->grad_fp32_cut / QB[p],
                                         [](double a, double t) { return a >= t; }) - (QK.begin() + kb[b]));
            khi = std::min(k, (khi + 63) / 64 * 64);
          }
          const long long cost = (long long)gp[bra].nprim * kmax[b];
          if (khi > 0) segs.push_back({(int)coff[cb] + p, kb[b], khi, cost});
          if (k > khi) segs32.push_back({(int)coff[cb] + p, kb[b] + khi, k - khi, cost});
        }
      }
      if (segs.empty() && segs32.empty()) continue;
      int la = 0, lb = 0, lc = 0, ld = 0;
      for (int x = 0; x <= 3; ++x)
        for (int y = 0; y <= x; ++y) {
          if (class_of(x, y) == cb) { la = x; lb = y; }
          if (class_of(x, y) == ck) { lc = x; ld = y; }
        }
      for (int f32 = 0; f32 < 2; ++f32) {
      std::vector<Seg>& segs_ = f32 ? segs32 : segs;
      if (segs_.empty()) continue;
      // longest wavefronts first
      std::stable_sort(segs_.begin(), segs_.end(), [](const Seg& x, const Seg& y) { return x.cost > y.cost; });
      const size_t seg0 = sbra.size(), ch0 = choff.size();
      long long waves = 0;
      for (const Seg& g : segs_) {
        sbra.push_back(g.bra);
        sk0.push_back(g.k0);
        sn.push_back(g.n);
        choff.push_back(waves);
        waves += (g.n + 63) / 64;
      }
      choff.push_back(waves);
      pend.push_back({la, lb, lc, ld, ck, seg0, ch0, (int)segs_.size(), waves, f32 != 0});
      }
    }
  const size_t nseg = sbra.size();
  if ((rc = s->d_gseg.ensure(3 * nseg + 1)) || (rc = s->d_gchoff.ensure(choff.size() + 1))) return rc;
  if (nseg) {
    if ((rc = upload_sync(s->d_gseg.p, sbra.data(), nseg * sizeof(int), s->stream)) ||
        (rc = upload_sync(s->d_gseg.p + nseg, sk0.data(), nseg * sizeof(int), s->stream)) ||
        (rc = upload_sync(s->d_gseg.p + 2 * nseg, sn.data(), nseg * sizeof(int), s->stream)) ||
        (rc = upload_sync(s->d_gchoff.p, choff.data(), choff.size() * sizeof(long long), s->stream)))
      return rc;
  }
  s->glaunches.clear();
  for (auto& p : pend) {
    oqph_session::GLaunch L{};
    L.la = p.la; L.lb = p.lb; L.lc = p.lc; L.ld = p.ld;
    L.ww.seg_bra = s->d_gseg.p + p.seg0;
    L.ww.seg_k0 = s->d_gseg.p + nseg + p.seg0;
    L.ww.seg_n = s->d_gseg.p + 2 * nseg + p.seg0;
    L.ww.choff = s->d_gchoff.p + p.ch0;
    L.ww.ket = s->d_gcls.p + coff[p.ck];
    L.ww.nseg = p.nseg;
    L.ww.nwaves = p.nwaves;
    L.hl = !grad_template(p.la, p.lb, p.lc, p.ld) || (!s->f_templates && (p.la == 3 || p.lc == 3)) ||
           (s->grad_hl_ng > 0 && ncart(p.la) * ncart(p.lb) * ncart(p.lc) * ncart(p.ld) > s->grad_hl_ng);
    if (!L.hl) {  // templates whose private segment exceeds the budget run on k_gradh
      KernelRes kr{};
      const int kind = s->grad_all_hl ? (2 | 32) : 2;  // the MRSF form in the MRSF gradient
      L.hl = !kernel_resources(kind, p.la, p.lb, p.lc, p.ld, s->pure, &kr) || kr.private_bytes > OQPH_PRIVATE_BUDGET ||
             (size_t)kr.lds_bytes + 24 * (size_t)s->natm > s->lds_max;
    }
    if (!L.hl && p.f32) {  // the FP32 form where its resources fit as well (else the FP64 form)
      KernelRes kr{};
      L.f32 = kernel_resources(2 | 16, p.la, p.lb, p.lc, p.ld, s->pure, &kr) &&
              kr.private_bytes <= OQPH_PRIVATE_BUDGET && (size_t)kr.lds_bytes + 24 * (size_t)s->natm <= s->lds_max;
    }
    if (L.hl) {
      L.hd = hl_dims(p.la, p.lb, p.lc, p.ld, s->pure, 1, s->natm);
      if ((size_t)L.hd.lds > s->lds_max) {
        set_err("gradient class (%d %d|%d %d): %d bytes of LDS exceed the device's %zu", p.la, p.lb, p.lc, p.ld,
                L.hd.lds, s->lds_max);
        return 1;
      }
      L.ww.kpb = hl_kpb(s, L.ww.nwaves, L.hd.lds);
    }
    s->glaunches.push_back(L);
  }
  // run-time classes merged into one launch per occupancy tier (as the J/K
  // build's k_jkh_multi)
  if (s->hl_multi) {
    std::map<int, std::vector<size_t>> tiers;
    for (size_t i = 0; i < s->glaunches.size(); ++i) {
      const auto& L = s->glaunches[i];
      if (!L.hl || L.ww.nwaves == 0) continue;
      tiers[(int)std::min<size_t>(16, s->lds_max / std::max(L.hd.lds, 1))].push_back(i);
    }
    std::vector<HLTask> tasks;
    std::vector<oqph_session::GLaunch> merged;
    std::vector<char> drop(s->glaunches.size(), 0);
    for (auto& kv : tiers) {
      auto& t = kv.second;
      if (t.size() < 2) continue;
      std::stable_sort(t.begin(), t.end(), [&](size_t x, size_t y) {
        const auto &a = s->glaunches[x], &b = s->glaunches[y];
        return flops_prim_host(a.la, a.lb, a.lc, a.ld) > flops_prim_host(b.la, b.lb, b.lc, b.ld);
      });
      oqph_session::GLaunch M{};
      const auto& L0 = s->glaunches[t[0]];
      M.la = L0.la; M.lb = L0.lb; M.lc = L0.lc; M.ld = L0.ld;
      M.hl = true;
      M.multi = true;
      M.tl.ntask = (int)t.size();
      M.tl.task = (const HLTask*)(uintptr_t)tasks.size();  // offset until the table is on the device
      M.tl.seg = s->d_gseg.p;
      M.tl.nseg = (int)nseg;
      M.tl.choff = s->d_gchoff.p;
      M.tl.ket = s->d_gcls.p;
      long long chunks = 0;
      for (size_t i : t) {
        chunks += s->glaunches[i].ww.nwaves;
        M.lds = std::max(M.lds, s->glaunches[i].hd.lds);
      }
      const int kpb = hl_kpb(s, chunks, M.lds);
      M.ww.kpb = kpb;
      long long w0 = 0;
      for (size_t i : t) {
        const auto& L = s->glaunches[i];
        HLTask k{};
        k.h = L.hd;
        k.wave0 = w0;
        k.nwaves = L.ww.nwaves;
        k.seg0 = (int)(L.ww.seg_bra - s->d_gseg.p);
        k.ch0 = (int)(L.ww.choff - s->d_gchoff.p);
        k.nseg = L.ww.nseg;
        k.ket0 = (int)(L.ww.ket - s->d_gcls.p);
        k.kpb = kpb;
        tasks.push_back(k);
        w0 += L.ww.nwaves * (64 / kpb);
        drop[i] = 1;
      }
      M.nblocks = w0;
      merged.push_back(M);
    }
    if (!merged.empty()) {
      if ((rc = s->d_gtasks.ensure(tasks.size())) ||
          (rc = upload_sync(s->d_gtasks.p, tasks.data(), tasks.size() * sizeof(HLTask), s->stream)))
        return rc;
      std::vector<oqph_session::GLaunch> keep;
      for (size_t i = 0; i < s->glaunches.size(); ++i)
        if (!drop[i]) keep.push_back(s->glaunches[i]);
      for (auto& M : merged) {
        M.tl.task = s->d_gtasks.p + (uintptr_t)M.tl.task;
        keep.push_back(M);
      }
      s->glaunches.swap(keep);
    }
  }
  // longest launches first, so that the launches spread over the streams
  // start with the expensive ones (estimated by chunks x model FLOPs)
  std::stable_sort(s->glaunches.begin(), s->glaunches.end(),
                   [](const oqph_session::GLaunch& x, const oqph_session::GLaunch& y) {
                     auto cost = [](const oqph_session::GLaunch& L) {
                       // (64-ket chunks of the launch)
                       return (double)(L.multi ? L.nblocks * (L.ww.kpb > 0 ? L.ww.kpb : 64) / 64 : L.ww.nwaves) *
                              (double)flops_prim_host(L.la, L.lb, L.lc, L.ld) * (L.hl ? 4.0 : 1.0);
                     };
                     return cost(x) > cost(y);
                   });
  // Cartesian offsets, shell atoms and flags
  std::vector<int> cof(s->nsh), pu(s->nsh);
  int ncf = 0;
  for (int i = 0; i < s->nsh; ++i) {
    cof[i] = ncf;
    ncf += ncart(s->am[i]);
    pu[i] = (s->harm[i] && s->am[i] >= 2) ? 1 : 0;
  }
  s->ncf = ncf;
  if ((rc = s->d_coff.ensure(s->nsh)) || (rc = s->d_satom.ensure(s->nsh)) || (rc = s->d_am.ensure(s->nsh)) ||
      (rc = s->d_pure.ensure(s->nsh)) || (rc = s->d_Dc.ensure((size_t)ncf * ncf)) ||
      (rc = s->d_de.ensure(3 * (size_t)s->natm * OQPH_DE_SLOTS)))
    return rc;
  if ((rc = upload_sync(s->d_coff.p, cof.data(), s->nsh * sizeof(int), s->stream)) ||
      (rc = upload_sync(s->d_satom.p, s->atom.data(), s->nsh * sizeof(int), s->stream)) ||
      (rc = upload_sync(s->d_am.p, s->am.data(), s->nsh * sizeof(int), s->stream)) ||
      (rc = upload_sync(s->d_pure.p, pu.data(), s->nsh * sizeof(int), s->stream)))
    return rc;
  HIPCHK(hipStreamSynchronize(s->stream));
  if (s->launch_order == 1)
    std::stable_sort(s->glaunches.begin(), s->glaunches.end(),
                     [](const oqph_session::GLaunch& x, const oqph_session::GLaunch& y) {
                       return class_weight(x.la, x.lb, x.lc, x.ld, x.hl || x.multi) >
                              class_weight(y.la, y.lb, y.lc, y.ld, y.hl || y.multi);
                     });
  s->gcut_ready = s->grad_cutoff;
  s->ghl_ready = s->grad_all_hl;
  s->gpolicy_ready = s->grad_policy;
  s->gf32_ready = s->grad_fp32_cut;
  return 0;
}

// Wavefront tasks for a given maxden.  For bra pair p and ket bucket b the
// kets that survive Q_bra Q_ket 4 maxD >= c form a prefix of the bucket; for a
// class with itself each unordered pair is taken once (the bra is the later
// entry of the (bucket, Q) order).  Non-empty (p, b) segments are stored with
// a prefix over their 64-ket chunks.
static int prepare_lists(oqph_session* s, double maxden, bool two) {
  if (maxden == s->cached_maxden && two == s->cached_two && s->general == s->cached_general) return 0;
  s->launches.clear();
  s->h_seg_bra.clear();
  s->h_seg_k0.clear();
  s->h_seg_n.clear();
  s->h_choff.clear();
  struct Pend {
    int la, lb, lc, ld, ck;
    size_t seg0, ch0;
    int nseg;
    long long nwaves;
    bool rows = false;
    size_t item0 = 0;
    int nitem = 0, ncolC = 0, ncol = 0;
    size_t lds = 0;
    bool drows = false;
    bool hl = false;  // run-time-class kernel (k_jkh)
    bool grouped = false;  // grouped bra-row kernel (k_jkg) on chunk set cv
    bool tiled = false;    // tiled form of k_jkg (tiles in h_items)
    ChunkView cv{};
    long long slots = 0;
  };
  s->h_items.clear();
  std::vector<Pend> pend;
  const double c = s->cutoff;
  const int ncls = class_of(s->lmax, s->lmax) + 1;
  for (int cb = 0; cb < ncls; ++cb)
    for (int ck = 0; ck <= cb; ++ck) {
      const auto& QB = s->clsQ[cb];
      const auto& QK = s->clsQ[ck];
      if (QB.empty() || QK.empty()) continue;
      const auto& kb = s->kbs[ck];
      const auto& bb = s->kbs[cb];
      int la = 0, lb = 0, lc = 0, ld = 0;
      for (int x = 0; x <= 3; ++x)
        for (int y = 0; y <= x; ++y) {
          if (class_of(x, y) == cb) { la = x; lb = y; }
          if (class_of(x, y) == ck) { lc = x; ld = y; }
        }
      // largest ket primitive-pair count of each bucket
      int kmax[NBKT] = {0};
      for (int b = 0; b < NBKT; ++b)
        for (int k = kb[b]; k < kb[b + 1]; ++k) kmax[b] = std::max(kmax[b], s->pairs[s->cls[ck][k]].nprim);
      // surviving kets of every bra: a prefix of each ket bucket
      const int nb = (int)QB.size();
      std::vector<int> cnt((size_t)nb * NBKT, 0);
      long long nket = 0;
      int nbra = 0, bp = 0;
      for (int p = 0; p < nb; ++p) {
        while (bp + 1 < NBKT && p >= bb[bp + 1]) ++bp;
        const double thr = c / (4.0 * maxden * QB[p]);
        long long any = 0;
        for (int b = 0; b < NBKT; ++b) {
          if (kb[b + 1] <= kb[b]) continue;
          long long k = std::lower_bound(QK.begin() + kb[b], QK.begin() + kb[b + 1], thr,
                                         [](double a, double t) { return a >= t; }) - (QK.begin() + kb[b]);
          if (cb == ck) {
            if (b > bp) k = 0;
            else if (b == bp) k = std::min<long long>(k, (long long)p - bb[bp] + 1);
          }
          cnt[(size_t)p * NBKT + b] = (int)k;
          any += k;
        }
        nket += any;
        nbra += any > 0;
      }
      if (nket == 0) continue;
      // bra-row kernel when its K rows fit in LDS and it saves global atomics:
      // per quartet it leaves OC*OD (J_cd) of OC*OD + (OA+OB)(OC+OD), plus one
      // flush of the K rows per bra
      const int oa = nao_of(la, s->pure), obn = nao_of(lb, s->pure), oc = nao_of(lc, s->pure),
                od = nao_of(ld, s->pure);
      const int ncC = s->ncolL[lc], ncol = (lc == ld) ? ncC : ncC + s->ncolL[ld];
      int nrow = 0;
      size_t slds = 0;
      // run-time-class kernel for the f classes without a class template
      // templates whose private segment exceeds the budget run on k_jkh instead
      KernelRes kr0{}, kr1{};
      const bool tmpl = !s->general && jk_template(la, lb, lc, ld) && !(!s->f_templates && (la == 3 || lc == 3)) &&
                        !(s->jk_hl_ng > 0 && ncart(la) * ncart(lb) * ncart(lc) * ncart(ld) > s->jk_hl_ng);
      const int variant = two ? 16 : 0;  // the open-shell kernels when two
      const bool fit0 = tmpl && kernel_resources(0 | variant, la, lb, lc, ld, s->pure, &kr0) &&
                        kr0.private_bytes <= OQPH_PRIVATE_BUDGET;
      const bool fit1 = tmpl && kernel_resources(1 | variant, la, lb, lc, ld, s->pure, &kr1) &&
                        kr1.private_bytes <= OQPH_PRIVATE_BUDGET;
      const bool hl = !fit0 && !fit1;
      bool rows = fit1 && s->rows_mode > 0 && jkb_info(la, lb, lc, ld, s->pure, &nrow, &slds, s->lds_max);
      // dynamic LDS: K rows (nrow x ncol), plus D rows (nrow x ncol) and the
      // bra rows of dsh (2 x nsh) when those fit as well; the open-shell
      // pass holds the K rows (and D rows) of both densities
      const size_t nd = two ? 2 : 1;
      const size_t dynk = nd * nrow * ncol * sizeof(double);
      const size_t dynd = (2 * nd * nrow * ncol + 2 * (size_t)s->nsh) * sizeof(double);
      const bool drows = rows && s->rows_mode != 3 && slds + dynd <= s->lds_max;
      const size_t dynb = drows ? dynd : dynk;
      rows = rows && slds + dynb <= s->lds_max;
      if (rows && s->rows_mode == 1 && fit0) {
        const double old_at = (double)nket * (oc * od + (oa + obn) * (oc + od));
        const double new_at = (double)nket * oc * od + (double)nbra * nrow * ncol;
        rows = new_at < 0.5 * old_at;
      }
      if (rows) {
        // grouped bra-row kernel: one block per chunk of pairs sharing primitives
        const int cap = jkg_cap_for(s, la, lb, lc, ld);
        const auto csit = cap >= 2 ? s->chunksets.find(cb * 16 + cap) : s->chunksets.end();
        // bras the grouped launch takes (members of chunks of two or more);
        // the single pairs stay on the segmented bra-row kernel below
        std::vector<char> grouped_bra;
        if (csit != s->chunksets.end()) {
          const oqph_session::ChunkSet& cs = *csit->second;
          size_t st = 0;
          jkg_info(la, lb, lc, ld, s->pure, &st, s->lds_max);
          struct Itg {
            int c;
            long long cost;
            std::array<int, NBKT> n;
          };
          std::vector<Itg> its;
          size_t need_d = 0, need_n = 0;
          for (int ci = 0; ci < (int)cs.pos.size(); ++ci) {
            if (cs.pos[ci].size() < 2) continue;
            Itg t{ci, 0, {}};
            long long any = 0;
            for (int b = 0; b < NBKT; ++b) {
              int m = 0;
              for (int p : cs.pos[ci]) m = std::max(m, cnt[(size_t)p * NBKT + b]);
              t.n[b] = m;
              any += m;
              t.cost += (long long)m * kmax[b];
            }
            if (any == 0) continue;
            t.cost *= cs.ncp[ci];
            its.push_back(t);
            const size_t rows_b = (size_t)cs.nrow[ci] * ncol * sizeof(double);
            const size_t prims_b = (size_t)cs.ncp[ci] * cs.stride[ci] * (sizeof(double) + sizeof(float));
            need_n = std::max(need_n, nd * rows_b + prims_b);
            need_d = std::max(need_d, 2 * nd * rows_b + (size_t)cs.nshl[ci] * s->nsh * sizeof(double) + prims_b);
          }
          const bool gd = s->rows_mode != 3 && st + need_d <= s->lds_max;
          if (!its.empty() && (gd || st + need_n <= s->lds_max)) {
            grouped_bra.assign(nb, 0);
            for (const Itg& t : its)
              for (int p : cs.pos[t.c]) grouped_bra[p] = 1;
            std::stable_sort(its.begin(), its.end(), [](const Itg& x, const Itg& y) { return x.cost > y.cost; });
            const size_t i0 = s->h_items.size();
            long long slots = 0;
            for (const Itg& t : its) {
              s->h_items.push_back(t.c);
              for (int b = 0; b < NBKT; ++b) {
                s->h_items.push_back(t.n[b]);
                slots += 64ll * ((t.n[b] + 63) / 64);
              }
            }
            Pend q{la, lb, lc, ld, ck, 0, 0, 0, 0};
            q.rows = true;
            q.grouped = true;
            q.cv = cs.view();
            q.item0 = i0;
            q.nitem = (int)its.size();
            q.ncolC = ncC;
            q.ncol = ncol;
            q.lds = gd ? need_d : need_n;
            q.drows = gd;
            q.slots = slots;
            pend.push_back(q);
          }
        }
        struct It {
          int p;
          long long cost;
        };
        std::vector<It> its;
        for (int p = 0; p < nb; ++p) {
          if (!grouped_bra.empty() && grouped_bra[p]) continue;
          long long cost = 0;
          for (int b = 0; b < NBKT; ++b) cost += (long long)cnt[(size_t)p * NBKT + b] * kmax[b];
          if (cost > 0) its.push_back({p, cost * s->pairs[s->cls[cb][p]].nprim});
        }
        // longest blocks first
        std::stable_sort(its.begin(), its.end(), [](const It& x, const It& y) { return x.cost > y.cost; });
        // tiled form: the bras that share their first shell, in tiles of up
        // to tcap (consecutive in cost order, so their ket prefixes are alike)
        const int tcap = jkt_cap_for(s, la, lb, lc, ld);
        if (tcap >= 2 && !its.empty()) {
          size_t st = 0;
          jkt_info(la, lb, lc, ld, s->pure, &st, s->lds_max);
          std::map<int, std::vector<int>> bysa;  // first shell -> positions in its (cost order)
          std::vector<int> saorder;
          for (size_t i = 0; i < its.size(); ++i) {
            const int sa = s->pairs[s->cls[cb][its[i].p]].sa;
            auto it = bysa.find(sa);
            if (it == bysa.end()) {
              saorder.push_back(sa);
              bysa[sa] = {(int)i};
            } else {
              it->second.push_back((int)i);
            }
          }
          struct Tile {
            std::vector<int> mem;  // positions in its
            long long cost = 0;
            int nprim = 0;
          };
          std::vector<Tile> tiles;
          bool ok = true;
          for (int sa : saorder) {
            Tile t;
            for (int i : bysa[sa]) {
              const int np = s->pairs[s->cls[cb][its[i].p]].nprim;
              if (np > OQPH_JKT_MAXP) ok = false;
              if (!t.mem.empty() && ((int)t.mem.size() == tcap || t.nprim + np > OQPH_JKT_MAXP)) {
                tiles.push_back(t);
                t = Tile{};
              }
              t.mem.push_back(i);
              t.cost += its[i].cost;
              t.nprim += np;
            }
            if (!t.mem.empty()) tiles.push_back(t);
          }
          std::stable_sort(tiles.begin(), tiles.end(), [](const Tile& x, const Tile& y) { return x.cost > y.cost; });
          // headers (chunk layout) into the work lists, then the items
          std::vector<int> hoff;
          const size_t h0 = s->h_items.size();
          size_t need_d = 0, need_n = 0;
          const size_t nd2 = two ? 2 : 1;
          for (const Tile& t : tiles) {
            std::vector<int> shl, row0;
            int nrow = 0;
            auto slot = [&](int sh) {
              for (size_t u = 0; u < shl.size(); ++u)
                if (shl[u] == sh) return (int)u;
              shl.push_back(sh);
              row0.push_back(nrow);
              nrow += s->nao[sh];
              return (int)shl.size() - 1;
            };
            std::vector<int> ent;
            for (int i : t.mem) {
              const int p = its[i].p;
              const ShellPair& sp = s->pairs[s->cls[cb][p]];
              const int ta = slot(sp.sa), tb = slot(sp.sb);
              const int ti = slot(sp.si), tj = slot(sp.sj);
              ent.insert(ent.end(), {(int)s->cls_off[cb] + p, row0[ta], row0[tb], ti, tj});
            }
            const int nm = (int)t.mem.size();
            const int hdr_len = OQPH_JKG_HDR + OQPH_JKG_MEM * nm;
            hoff.push_back((int)s->h_items.size());
            s->h_items.insert(s->h_items.end(), {nm, 0, t.nprim, nrow, (int)shl.size(), 0, hdr_len, hdr_len + nrow});
            s->h_items.insert(s->h_items.end(), ent.begin(), ent.end());
            for (size_t u = 0; u < shl.size(); ++u)
              for (int k = 0; k < s->nao[shl[u]]; ++k) s->h_items.push_back(s->aooff[shl[u]] + k);
            s->h_items.insert(s->h_items.end(), shl.begin(), shl.end());
            const size_t rows_b = (size_t)nrow * ncol * sizeof(double);
            const size_t prims_b = (size_t)t.nprim * 6 * (sizeof(double) + sizeof(float));  // PrimV<double> + PrimV<float>
            need_n = std::max(need_n, nd2 * rows_b + prims_b);
            need_d = std::max(need_d, 2 * nd2 * rows_b + shl.size() * (size_t)s->nsh * sizeof(double) + prims_b);
          }
          const bool gd = s->rows_mode != 3 && st + need_d <= s->lds_max;
          if (ok && (gd || st + need_n <= s->lds_max)) {
            const size_t i0 = s->h_items.size();
            long long slots = 0;
            for (size_t ti = 0; ti < tiles.size(); ++ti) {
              s->h_items.push_back(hoff[ti]);
              for (int b = 0; b < NBKT; ++b) {
                int m = 0;
                for (int i : tiles[ti].mem) m = std::max(m, cnt[(size_t)its[i].p * NBKT + b]);
                s->h_items.push_back(m);
                slots += 64ll * ((m + 63) / 64);
              }
            }
            Pend q{la, lb, lc, ld, ck, 0, 0, 0, 0};
            q.rows = true;
            q.tiled = true;
            q.item0 = i0;
            q.nitem = (int)tiles.size();
            q.ncolC = ncC;
            q.ncol = ncol;
            q.lds = gd ? need_d : need_n;
            q.drows = gd;
            q.slots = slots;
            pend.push_back(q);
            continue;
          }
          s->h_items.resize(h0);  // does not fit: the segmented bra-row kernel below
        }
        const size_t i0 = s->h_items.size();
        long long slots = 0;
        for (const It& t : its) {
          s->h_items.push_back((int)s->cls_off[cb] + t.p);  // device pair id
          for (int b = 0; b < NBKT; ++b) {
            s->h_items.push_back(cnt[(size_t)t.p * NBKT + b]);
            slots += 64ll * ((cnt[(size_t)t.p * NBKT + b] + 63) / 64);
          }
        }
        Pend q{la, lb, lc, ld, ck, 0, 0, 0, 0};
        q.rows = true;
        q.item0 = i0;
        q.nitem = (int)its.size();
        q.ncolC = ncC;
        q.ncol = ncol;
        q.lds = dynb;
        q.drows = drows;
        q.slots = slots;
        pend.push_back(q);
        continue;
      }
      if (!fit0 && !hl) {
        set_err("class (%d %d|%d %d): its bra-row kernel does not fit in LDS and its wavefront kernel exceeds the "
                "private-segment budget", la, lb, lc, ld);
        return 1;
      }
      const size_t seg0 = s->h_seg_bra.size(), ch0 = s->h_choff.size();
      struct Seg {
        int bra, k0, n;
        long long cost;  // primitive quartets one wavefront of the segment loops over
      };
      std::vector<Seg> segs;
      for (int p = 0; p < nb; ++p) {
        const int bra = s->cls[cb][p];
        for (int b = 0; b < NBKT; ++b) {
          const int k = cnt[(size_t)p * NBKT + b];
          if (k == 0) continue;
          segs.push_back({(int)s->cls_off[cb] + p, kb[b], k, (long long)s->pairs[bra].nprim * kmax[b]});
        }
      }
      // longest wavefronts first, so the launch does not end on a tail of
      // heavily contracted quartets
      std::stable_sort(segs.begin(), segs.end(), [](const Seg& x, const Seg& y) { return x.cost > y.cost; });
      long long waves = 0;
      for (const Seg& g : segs) {
        s->h_seg_bra.push_back(g.bra);
        s->h_seg_k0.push_back(g.k0);
        s->h_seg_n.push_back(g.n);
        s->h_choff.push_back(waves);
        waves += (g.n + 63) / 64;
      }
      const int nseg = (int)(s->h_seg_bra.size() - seg0);
      s->h_choff.push_back(waves);  // terminator of this launch's prefix
      pend.push_back({la, lb, lc, ld, ck, seg0, ch0, nseg, waves});
      pend.back().hl = hl;
      pend.back().slots = 64ll * waves;
    }
  for (auto& p : pend) {
    oqph_session::WLaunch L{};
    L.la = p.la; L.lb = p.lb; L.lc = p.lc; L.ld = p.ld;
    L.rows = p.rows;
    L.grouped = p.grouped;
    L.tiled = p.tiled;
    L.cv = p.cv;
    L.slots = p.slots;
    L.seg0 = p.seg0;
    L.ch0 = p.ch0;
    L.item0 = p.item0;
    if (p.rows) {
      L.bw.ket = s->d_cls.p + s->cls_off[p.ck];
      L.bw.jbase = s->jbase[p.ck];
      L.bw.kofs = (int)s->cls_off[p.ck];
      for (int b = 0; b <= NBKT; ++b) L.bw.kb[b] = s->kbs[p.ck][b];
      L.bw.nitem = p.nitem;
      L.cm.colidx = s->d_colidx.p;
      L.cm.aoC = s->d_colao.p + s->colbase[p.lc];
      L.cm.aoD = s->d_colao.p + s->colbase[p.ld];
      L.cm.ncolC = p.ncolC;
      L.cm.ncol = p.ncol;
      L.cm.dofs = (p.lc == p.ld) ? 0 : p.ncolC;
      L.cm.drows = p.drows ? 1 : 0;
      L.lds = p.lds;
      s->launches.push_back(L);
      continue;
    }
    L.ww.ket = s->d_cls.p + s->cls_off[p.ck];
    L.ww.jbase = s->jbase[p.ck];
    L.ww.kofs = (int)s->cls_off[p.ck];
    L.ww.nseg = p.nseg;
    L.ww.nwaves = p.nwaves;
    L.hl = p.hl;
    if (L.hl) {
      L.hd = hl_dims(p.la, p.lb, p.lc, p.ld, s->pure, 0, 0);
      if ((size_t)L.hd.lds > s->lds_max) {
        set_err("class (%d %d|%d %d): %d bytes of LDS exceed the device's %zu", p.la, p.lb, p.lc, p.ld, L.hd.lds,
                s->lds_max);
        return 1;
      }
      L.ww.kpb = hl_kpb(s, L.ww.nwaves, L.hd.lds);
    }
    s->launches.push_back(L);
  }
  // the run-time-class launches (classes with an f shell, (dd|dd)) merged
  // into one launch per occupancy tier: a launch is sized for the LDS of its
  // largest class, so only classes that fit the same number of blocks per CU
  // (LDS per CU over the class's LDS, at most 4 waves per SIMD of one-wave
  // blocks) share a launch, and no class loses occupancy by the merge
  s->h_tasks.clear();
  if (s->hl_multi) {
    std::vector<oqph_session::WLaunch> merged;
    std::map<int, std::vector<size_t>> tiers;
    for (size_t i = 0; i < s->launches.size(); ++i) {
      const auto& L = s->launches[i];
      if (!L.hl || L.ww.nwaves == 0) continue;
      const int per_cu = (int)std::min<size_t>(16, s->lds_max / std::max(L.hd.lds, 1));
      tiers[per_cu].push_back(i);
    }
    std::vector<std::vector<size_t>> tier;
    for (auto& kv : tiers) tier.push_back(kv.second);
    std::vector<char> drop(s->launches.size(), 0);
    for (auto& t : tier) {
      if (t.size() < 2) continue;
      // most expensive chunks first: model FLOPs of one primitive quartet
      std::stable_sort(t.begin(), t.end(), [&](size_t x, size_t y) {
        const auto &a = s->launches[x], &b = s->launches[y];
        return flops_prim_host(a.la, a.lb, a.lc, a.ld) > flops_prim_host(b.la, b.lb, b.lc, b.ld);
      });
      oqph_session::WLaunch M{};
      const auto& L0 = s->launches[t[0]];
      M.la = L0.la; M.lb = L0.lb; M.lc = L0.lc; M.ld = L0.ld;
      M.hl = true;
      M.multi = true;
      M.task0 = (int)s->h_tasks.size();
      M.ntask = (int)t.size();
      size_t lds = 0;
      long long chunks = 0;
      for (size_t i : t) {
        chunks += s->launches[i].ww.nwaves;
        lds = std::max(lds, (size_t)s->launches[i].hd.lds);
      }
      const int kpb = hl_kpb(s, chunks, (int)lds);
      long long w0 = 0;
      for (size_t i : t) {
        const auto& L = s->launches[i];
        HLTask k{};
        k.h = L.hd;
        k.wave0 = w0;
        k.nwaves = L.ww.nwaves;
        k.seg0 = (int)L.seg0;
        k.ch0 = (int)L.ch0;
        k.nseg = L.ww.nseg;
        k.ket0 = (int)s->cls_off[class_of(L.lc, L.ld)];
        k.kpb = kpb;
        s->h_tasks.push_back(k);
        w0 += L.ww.nwaves * (64 / kpb);
        M.slots += L.slots;
        drop[i] = 1;
      }
      M.nblocks = w0;
      M.lds = lds;
      merged.push_back(M);
    }
    if (!merged.empty()) {
      std::vector<oqph_session::WLaunch> keep;
      for (size_t i = 0; i < s->launches.size(); ++i)
        if (!drop[i]) keep.push_back(s->launches[i]);
      keep.insert(keep.end(), merged.begin(), merged.end());
      s->launches.swap(keep);
    }
  }
  if (s->launch_order == 1)
    std::stable_sort(s->launches.begin(), s->launches.end(),
                     [](const oqph_session::WLaunch& x, const oqph_session::WLaunch& y) {
                       return class_weight(x.la, x.lb, x.lc, x.ld, x.hl || x.multi) >
                              class_weight(y.la, y.lb, y.lc, y.ld, y.hl || y.multi);
                     });
  ++s->list_gen;
  s->cached_maxden = maxden;
  s->cached_two = two;
  s->cached_general = s->general;
  return 0;
}

// Device work lists of the current launches: element counts of seg (three
// int arrays of nseg), choff (long long) and items (int).
struct ListSizes {
  size_t seg, choff, items, tasks;
};
static ListSizes list_sizes(const oqph_session* s) {
  return {3 * s->h_seg_bra.size() + 1, s->h_choff.size() + 1, std::max<size_t>(s->h_items.size(), 1),
          std::max<size_t>(s->h_tasks.size(), 1)};
}

// Copy the host work lists into device buffers of at least list_sizes();
// synchronises the stream, so the host vectors may change afterwards.
static int upload_lists(oqph_session* s, int* seg, long long* choff, int* items, HLTask* tasks, hipStream_t stream) {
  const size_t nseg = s->h_seg_bra.size();
  int rc;
  if (nseg) {
    if ((rc = upload(seg, s->h_seg_bra.data(), nseg * sizeof(int), stream)) ||
        (rc = upload(seg + nseg, s->h_seg_k0.data(), nseg * sizeof(int), stream)) ||
        (rc = upload(seg + 2 * nseg, s->h_seg_n.data(), nseg * sizeof(int), stream)) ||
        (rc = upload(choff, s->h_choff.data(), s->h_choff.size() * sizeof(long long), stream)))
      return rc;
  }
  if (!s->h_items.empty() &&
      (rc = upload(items, s->h_items.data(), s->h_items.size() * sizeof(int), stream)))
    return rc;
  if (!s->h_tasks.empty() &&
      (rc = upload(tasks, s->h_tasks.data(), s->h_tasks.size() * sizeof(HLTask), stream)))
    return rc;
  HIPCHK(hipStreamSynchronize(stream));
  return 0;
}

// The session's own list buffers (grown on demand), refreshed when the host
// lists changed since the last upload.
static int session_lists(oqph_session* s, hipStream_t stream) {
  if (s->list_uploaded == s->list_gen) return 0;
  const ListSizes n = list_sizes(s);
  int rc;
  if ((rc = s->d_seg.ensure(n.seg)) || (rc = s->d_choff.ensure(n.choff)) || (rc = s->d_items.ensure(n.items)) ||
      (rc = s->d_tasks.ensure(n.tasks)))
    return rc;
  if ((rc = upload_lists(s, s->d_seg.p, s->d_choff.p, s->d_items.p, s->d_tasks.p, stream))) return rc;
  s->list_uploaded = s->list_gen;
  return 0;
}

// Point one launch's work descriptors at a device copy of the lists.
static void bind_lists(const oqph_session* s, oqph_session::WLaunch& L, const int* seg, const long long* choff,
                       const int* items, const HLTask* tasks) {
  const size_t nseg = s->h_seg_bra.size();
  if (L.multi) {
    L.tl.task = tasks + L.task0;
    L.tl.ntask = L.ntask;
    L.tl.seg = seg;
    L.tl.nseg = (int)nseg;
    L.tl.choff = choff;
    L.tl.ket = s->d_cls.p;
    return;
  }
  if (L.rows) {
    L.bw.item = items + L.item0;
    if (L.tiled) L.cv.pool = items;  // the tiles' headers live in the same lists
    return;
  }
  L.ww.seg_bra = seg + L.seg0;
  L.ww.seg_k0 = seg + nseg + L.seg0;
  L.ww.seg_n = seg + 2 * nseg + L.seg0;
  L.ww.choff = choff + L.ch0;
}

// The no-allocation entry point's scratch: the calling device's copy of the
// static array, and a host record of which session's work lists it holds.
struct StaticScratch {
  unsigned char* base;        // device address from hipGetSymbolAddress
  size_t bytes;
  unsigned long long* serial;  // session serial and list generation of the
  unsigned long long* gen;     // lists last uploaded into it
};

// Fork the side streams off `main` (they wait for everything enqueued on it
// so far), pick the stream of launch i, and join them back into `main`.
// n: the streams used (1: everything on main; per-launch profiles use 1,
// so each launch's events time it alone on the device).
static int streams_fork(oqph_session* s, hipStream_t main, int n) {
  if (n < 2) return 0;
  HIPCHK(hipEventRecord(s->ev_fork, main));
  for (int j = 0; j + 1 < n; ++j) HIPCHK(hipStreamWaitEvent(s->side[j], s->ev_fork, 0));
  return 0;
}
static hipStream_t stream_of(const oqph_session* s, hipStream_t main, size_t i, int n) {
  const int k = (int)(i % (size_t)n);
  return k == 0 ? main : s->side[k - 1];
}
static int streams_join(oqph_session* s, hipStream_t main, int n) {
  if (n < 2) return 0;
  for (int j = 0; j + 1 < n; ++j) {
    HIPCHK(hipEventRecord(s->ev_join[j], s->side[j]));
    HIPCHK(hipStreamWaitEvent(main, s->ev_join[j], 0));
  }
  return 0;
}

// G[D] on device buffers; D already on the device.  The screening densities
// scr_a (and scr_b) default to D; zero_f = false accumulates into dF.  mu > 0:
// the erf(mu r)/r attenuated pass (pair records with the attenuated bounds).
// ss: run in the static scratch of the no-allocation entry point instead of
// the session's buffers (no device allocation; profiling is off there).
static int build_device(oqph_session* s, const double* dD, double* dF, double se, double sc, oqph_stats* st,
                        hipStream_t stream, const double* scr_a = nullptr, const double* scr_b = nullptr,
                        bool zero_f = true, double mu = 0.0, const StaticScratch* ss = nullptr,
                        const double* dD2 = nullptr, double* dF2 = nullptr) {
  int rc;
  if (mu > 0.0 && (rc = prepare_attenuated(s, mu))) return rc;
  if ((rc = prepare_share(s))) return rc;
  const size_t ntri = (size_t)s->nbf * (s->nbf + 1) / 2;
  const size_t nsh2 = (size_t)s->nsh * s->nsh;
  // scratch layout: counters 0..7 (density maximum), dsh, per-launch
  // counters, work lists; 256-byte aligned slices
  size_t off = 0;
  auto carve = [&](size_t bytes) -> unsigned char* {
    unsigned char* q = ss->base + off;
    off += (bytes + 255) & ~(size_t)255;
    return q;
  };
  unsigned long long* ctr0 = s->d_ctr.p;
  double* dsh = s->d_dsh.p;
  if (ss) {
    ctr0 = (unsigned long long*)carve(8 * sizeof(unsigned long long));
    dsh = (double*)carve(nsh2 * sizeof(double));
    if (off > ss->bytes) {
      set_err("no-allocation build: %zu shells need %zu bytes of scratch for the shell-pair maxima alone; the "
              "static scratch holds %zu", (size_t)s->nsh, off, ss->bytes);
      return 1;
    }
  }
  launch_shell_density(scr_a ? scr_a : dD, scr_b, s->d_aooff.p, s->d_nao.p, s->nsh, dsh, stream);
  HIPCHK(hipMemsetAsync(ctr0, 0, 8 * sizeof(unsigned long long), stream));
  launch_max_reduce(dsh, (long long)nsh2, ctr0 + 4, stream);
  unsigned long long mbits = 0;
  HIPCHK(hipMemcpyAsync(&mbits, ctr0 + 4, sizeof(mbits), hipMemcpyDeviceToHost, stream));
  HIPCHK(hipStreamSynchronize(stream));
  double maxden;
  memcpy(&maxden, &mbits, sizeof(double));
  if (zero_f) HIPCHK(hipMemsetAsync(dF, 0, ntri * sizeof(double), stream));
  if (zero_f && dF2) HIPCHK(hipMemsetAsync(dF2, 0, ntri * sizeof(double), stream));
  if (maxden <= 0.0) {
    if (st) memset(st, 0, sizeof(*st));
    return 0;
  }
  if ((rc = prepare_lists(s, maxden, dF2 != nullptr))) return rc;
  const size_t nl = s->launches.size();
  const size_t nc = 8 * (size_t)OQPH_CTR_SLOTS;  // counters per launch
  unsigned long long* ctr = nullptr;  // per-launch counters
  int* lseg;
  long long* lchoff;
  int* litems;
  HLTask* ltasks;
  const size_t njk = std::max<size_t>(s->h_joff.empty() ? 0 : (size_t)s->h_joff.back(), 1);
  double* Jk = nullptr;
  if (ss) {
    const ListSizes n = list_sizes(s);
    ctr = (unsigned long long*)carve(nc * std::max<size_t>(nl, 1) * sizeof(unsigned long long));
    lseg = (int*)carve(n.seg * sizeof(int));
    lchoff = (long long*)carve(n.choff * sizeof(long long));
    litems = (int*)carve(n.items * sizeof(int));
    ltasks = (HLTask*)carve(n.tasks * sizeof(HLTask));
    Jk = (double*)carve(njk * sizeof(double));
    if (off > ss->bytes) {
      set_err("no-allocation build: this basis and density need %zu bytes of scratch (shell-pair maxima, "
              "counters, work lists and the ket-side Coulomb buffer); the static scratch holds %zu", off,
              ss->bytes);
      return 1;
    }
    if (*ss->serial != s->serial || *ss->gen != s->list_gen) {
      if ((rc = upload_lists(s, lseg, lchoff, litems, ltasks, stream))) return rc;
      *ss->serial = s->serial;
      *ss->gen = s->list_gen;
    }
  } else {
    if ((rc = session_lists(s, stream))) return rc;
    if ((rc = s->d_ctr.ensure(8 + nc * nl))) return rc;
    if ((rc = s->d_jk.ensure(njk))) return rc;
    Jk = s->d_jk.p;
    ctr = s->d_ctr.p + 8;
    lseg = s->d_seg.p;
    lchoff = s->d_choff.p;
    litems = s->d_items.p;
    ltasks = s->d_tasks.p;
  }
  for (auto& L : s->launches) bind_lists(s, L, lseg, lchoff, litems, ltasks);
  HIPCHK(hipMemsetAsync(ctr, 0, nc * nl * sizeof(unsigned long long), stream));
  const bool profile = s->profile && !ss;
  if (profile && s->pev.size() < 2 * nl) {
    for (size_t i = s->pev.size(); i < 2 * nl; ++i) {
      hipEvent_t e;
      HIPCHK(hipEventCreate(&e));
      s->pev.push_back(e);
    }
  }
  JKParams p{};
  p.pairs = mu > 0.0 ? s->d_pairs_att.p : s->d_pairs.p;
  p.prims = s->d_prims.p;
  p.prims32 = s->d_prims32.p;
  p.inv_mu2 = mu > 0.0 ? 1.0 / (mu * mu) : 0.0;
  p.own = s->share_size > 1 ? s->d_own.p : nullptr;
  p.rys = s->d_rys.p;
  p.rys32 = s->d_rys32.p;
  p.xform = s->d_xform.p;
  HIPCHK(hipMemsetAsync(Jk, 0, njk * sizeof(double), stream));
  p.Jk = Jk;
  p.D = dD;
  p.F = dF;
  p.D2 = dF2 ? dD2 : nullptr;
  p.F2 = dF2;
  p.dsh = dsh;
  p.counters = ctr;
  p.nsh = s->nsh;
  p.cutoff = s->cutoff;
  p.maxden = maxden;
  p.sc = sc;
  p.se = se;
  p.mixed = s->policy;
  p.fp32_cut = s->fp32_cut;
  HIPCHK(hipEventRecord(s->ev0, stream));
  const int nst = profile ? 1 : s->nstreams;
  if ((rc = streams_fork(s, stream, nst))) return rc;
  for (size_t i = 0; i < nl; ++i) {
    auto& L = s->launches[i];
    const hipStream_t ls = stream_of(s, stream, i, nst);
    p.counters = ctr + nc * i;
    if (profile) HIPCHK(hipEventRecord(s->pev[2 * i], ls));
    bool ok = true;
    if (L.multi) launch_jkh_multi(p, L.tl, L.nblocks, (int)L.lds, ls);
    else if (L.hl) launch_jkh(L.hd, p, L.ww, ls);
    else if (L.grouped) ok = launch_jkg(L.la, L.lb, L.lc, L.ld, s->pure, p, L.bw, L.cm, L.cv, L.lds, ls);
    else if (L.tiled) ok = launch_jkt(L.la, L.lb, L.lc, L.ld, s->pure, p, L.bw, L.cm, L.cv, L.lds, ls);
    else if (L.rows) ok = launch_jkb(L.la, L.lb, L.lc, L.ld, s->pure, p, L.bw, L.cm, L.lds, ls);
    else ok = launch_jk(L.la, L.lb, L.lc, L.ld, s->pure, p, L.ww, ls);
    if (!ok) {
      set_err("no J/K kernel for class (%d %d|%d %d)", L.la, L.lb, L.lc, L.ld);
      return 1;
    }
    if (profile) HIPCHK(hipEventRecord(s->pev[2 * i + 1], ls));
  }
  HIPCHK(hipGetLastError());
  if ((rc = streams_join(s, stream, nst))) return rc;
  // the ket-side Coulomb blocks into F
  launch_jscatter(s->d_pairs.p, s->d_joff.p, (int)(s->h_joff.size() - 1), s->pure != 0, Jk, dF, stream);
  HIPCHK(hipEventRecord(s->ev1, stream));
  if (st || profile) {
    std::vector<unsigned long long> raw(nc * nl);
    if (nl) HIPCHK(hipMemcpyAsync(raw.data(), ctr, nc * nl * sizeof(unsigned long long),
                                  hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    s->prof_ctr.assign(8 * nl, 0ull);
    for (size_t i = 0; i < nl; ++i) {
      for (size_t k = 0; k < nc; ++k) s->prof_ctr[8 * i + k % 8] += raw[nc * i + k];
      s->prof_ctr[8 * i + 6] = (unsigned long long)s->launches[i].slots;
    }
    s->prof_ms.assign(nl, 0.f);
    if (profile)
      for (size_t i = 0; i < nl; ++i) HIPCHK(hipEventElapsedTime(&s->prof_ms[i], s->pev[2 * i], s->pev[2 * i + 1]));
  }
  if (st) {
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, s->ev0, s->ev1));
    memset(st, 0, sizeof(*st));
    for (size_t i = 0; i < nl; ++i) {
      const unsigned long long* c = &s->prof_ctr[8 * i];
      st->quartets_computed += (long long)c[0];
      st->quartets_screened += (long long)c[1];
      st->quartets_fp32 += (long long)c[2];
      st->model_flops += (double)c[3];
      st->prim_quartets += (long long)c[4];
    }
    st->kernel_ms = ms;
  }
  return 0;
}

// Response pass over nd full nbf x nbf densities (device slabs dD, outputs
// dF): every class on k_jkh / k_jkh_multi with the general digestion; dsh and
// maxden come from the caller (the reference's screening density).
static int build_general(oqph_session* s, int nd, const double* dD, double* dF, const int* dcoul, double maxden,
                         double se, double sc, oqph_stats* st, hipStream_t stream) {
  int rc;
  if ((rc = prepare_share(s))) return rc;
  const size_t n2 = (size_t)s->nbf * s->nbf;
  HIPCHK(hipMemsetAsync(dF, 0, (size_t)nd * n2 * sizeof(double), stream));
  if (st) memset(st, 0, sizeof(*st));
  if (maxden <= 0.0) return 0;
  s->general = true;
  rc = prepare_lists(s, maxden, false);
  s->general = false;
  if (rc) return rc;
  const size_t nl = s->launches.size();
  const size_t nc = 8 * (size_t)OQPH_CTR_SLOTS;
  if ((rc = session_lists(s, stream))) return rc;
  if ((rc = s->d_ctr.ensure(8 + nc * nl))) return rc;
  unsigned long long* ctr = s->d_ctr.p + 8;
  for (auto& L : s->launches) bind_lists(s, L, s->d_seg.p, s->d_choff.p, s->d_items.p, s->d_tasks.p);
  HIPCHK(hipMemsetAsync(ctr, 0, nc * nl * sizeof(unsigned long long), stream));
  JKParams p{};
  p.pairs = s->d_pairs.p;
  p.prims = s->d_prims.p;
  p.prims32 = s->d_prims32.p;
  p.own = s->share_size > 1 ? s->d_own.p : nullptr;
  p.rys = s->d_rys.p;
  p.rys32 = s->d_rys32.p;
  p.xform = s->d_xform.p;
  p.dsh = s->d_dsh.p;
  p.nsh = s->nsh;
  p.cutoff = s->cutoff;
  p.maxden = maxden;
  p.sc = sc;
  p.se = se;
  p.Dg = dD;
  p.Fg = dF;
  p.gcoul = dcoul;
  p.gnd = nd;
  p.nbf = s->nbf;
  HIPCHK(hipEventRecord(s->ev0, stream));
  const int nst = s->nstreams;
  if ((rc = streams_fork(s, stream, nst))) return rc;
  for (size_t i = 0; i < nl; ++i) {
    auto& L = s->launches[i];
    const hipStream_t ls = stream_of(s, stream, i, nst);
    p.counters = ctr + nc * i;
    if (L.multi) launch_jkh_multi(p, L.tl, L.nblocks, (int)L.lds, ls);
    else if (L.hl) launch_jkh(L.hd, p, L.ww, ls);
    else {
      set_err("response pass: class (%d %d|%d %d) is not on the run-time-class kernel", L.la, L.lb, L.lc, L.ld);
      return 1;
    }
  }
  HIPCHK(hipGetLastError());
  if ((rc = streams_join(s, stream, nst))) return rc;
  HIPCHK(hipEventRecord(s->ev1, stream));
  std::vector<unsigned long long> raw(nc * nl);
  if (nl) HIPCHK(hipMemcpyAsync(raw.data(), ctr, nc * nl * sizeof(unsigned long long), hipMemcpyDeviceToHost, stream));
  HIPCHK(hipStreamSynchronize(stream));
  if (st) {
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, s->ev0, s->ev1));
    for (size_t k = 0; k < raw.size(); ++k) {
      const unsigned long long c = raw[k];
      switch (k % 8) {
        case 0: st->quartets_computed += (long long)c; break;
        case 1: st->quartets_screened += (long long)c; break;
        case 3: st->model_flops += (double)c; break;
        case 4: st->prim_quartets += (long long)c; break;
        default: break;
      }
    }
    st->kernel_ms = ms;
  }
  return 0;
}

bool oqph::session_view(oqph_session* s, SessionView* v) {
  if (!s || s->nsh == 0 || s->natm == 0 || (int)s->xyz.size() != 3 * s->natm) return false;
  v->device = s->device;
  v->stream = s->stream;
  v->nsh = s->nsh;
  v->nbf = s->nbf;
  v->natm = s->natm;
  v->am = s->am.data();
  v->harm = s->harm.data();
  v->atom = s->atom.data();
  v->nprim = s->nprim.data();
  v->poff = s->poff.data();
  v->aooff = s->aooff.data();
  v->ex = s->ex.data();
  v->cc = s->cc.data();
  v->xyz = s->xyz.data();
  v->zn = s->zn.data();
  return true;
}

oqph_session* oqph::default_session() {
  std::lock_guard<std::mutex> lk(g_mtx);
  return g_default;
}
void oqph::set_error_text(const char* msg) { set_err("%s", msg); }

// ------------------------------------------------------------------ C ABI
extern "C" {

const char* oqph_last_error(void) { return g_err.c_str(); }

int oqph_session_create(oqph_session** out, int device) {
  if (!out) return 4;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    set_err("no HIP device");
    return 2;
  }
  if (device < 0 || device >= ndev) {
    set_err("device %d out of range (%d devices)", device, ndev);
    return 4;
  }
  HIPCHK(hipSetDevice(device));
  oqph_session* s = new oqph_session();
  s->device = device;
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&s->ev0) != hipSuccess || hipEventCreate(&s->ev1) != hipSuccess) {
    set_err("stream/event creation failed");
    delete s;
    return 2;
  }
  {
    const char* ns = getenv("OQPH_STREAMS");
    if (ns) s->nstreams = std::max(1, std::min(1 + OQPH_SIDE_STREAMS, atoi(ns)));
    // side streams at the default priority, like the session stream
    for (int j = 0; j + 1 < s->nstreams; ++j)
      if (hipStreamCreateWithFlags(&s->side[j], hipStreamNonBlocking) != hipSuccess ||
          hipEventCreateWithFlags(&s->ev_join[j], hipEventDisableTiming) != hipSuccess) {
        set_err("stream/event creation failed");
        return 2;
      }
    if (s->nstreams > 1 && hipEventCreateWithFlags(&s->ev_fork, hipEventDisableTiming) != hipSuccess) {
      set_err("stream/event creation failed");
      return 2;
    }
  }
  *out = s;
  std::lock_guard<std::mutex> lk(g_mtx);
  s->serial = ++g_serial;
  scratch_lookup(device);
  if (!g_default) g_default = s;
  return 0;
}

int oqph_session_destroy(oqph_session* s) {
  if (!s) return 0;
  {
    std::lock_guard<std::mutex> lk(g_mtx);
    if (g_default == s) g_default = nullptr;
  }
  hipSetDevice(s->device);
  oqph::xc_release(s);
  s->d_pairs.release(); s->d_prims.release(); s->d_rys.release(); s->d_cls.release();
  s->d_xform.release(); s->d_rys32.release(); s->d_kbs.release(); s->d_seg.release(); s->d_choff.release();
  s->d_gpairs.release(); s->d_gprims.release(); s->d_gcls.release(); s->d_coff.release(); s->d_satom.release();
  s->d_am.release(); s->d_pure.release(); s->d_gseg.release(); s->d_gchoff.release(); s->d_Dc.release(); s->d_de.release();
  s->d_aooff.release(); s->d_nao.release(); s->d_colidx.release(); s->d_colao.release(); s->d_items.release(); s->d_Dt.release(); s->d_D.release(); s->d_F.release(); s->d_dsh.release();
  s->d_tmp.release(); s->d_off.release(); s->d_ctr.release();
  s->d_gD.release(); s->d_gF.release(); s->d_gcoul.release(); s->d_gM.release(); s->d_gMc.release();
  for (auto e : s->pev) hipEventDestroy(e);
  if (s->ev0) hipEventDestroy(s->ev0);
  if (s->ev1) hipEventDestroy(s->ev1);
  if (s->ev_fork) hipEventDestroy(s->ev_fork);
  for (int j = 0; j < OQPH_SIDE_STREAMS; ++j) {
    if (s->ev_join[j]) hipEventDestroy(s->ev_join[j]);
    if (s->side[j]) hipStreamDestroy(s->side[j]);
  }
  if (s->stream) hipStreamDestroy(s->stream);
  delete s;
  return 0;
}

int oqph_set_default_session(oqph_session* s) {
  std::lock_guard<std::mutex> lk(g_mtx);
  g_default = s;
  return 0;
}

int oqph_session_set_basis(oqph_session* s, int nsh, const int* am, const int* harmonic, const int* atom,
                           const int* nprim, const double* ex, const double* cc) {
  if (!s || nsh <= 0 || !am || !harmonic || !atom || !nprim || !ex || !cc) {
    set_err("set_basis: invalid argument");
    return 4;
  }
  s->nsh = nsh;
  s->am.assign(am, am + nsh);
  s->harm.assign(harmonic, harmonic + nsh);
  s->atom.assign(atom, atom + nsh);
  s->nprim.assign(nprim, nprim + nsh);
  s->poff.resize(nsh);
  s->aooff.resize(nsh);
  s->nao.resize(nsh);
  int p = 0, a = 0, lmax = 0, npure = 0, nhigh = 0;
  for (int i = 0; i < nsh; ++i) {
    if (am[i] < 0 || nprim[i] <= 0) {
      set_err("set_basis: shell %d invalid", i);
      return 4;
    }
    s->poff[i] = p;
    s->aooff[i] = a;
    const bool pu = harmonic[i] && am[i] >= 2;
    s->nao[i] = pu ? 2 * am[i] + 1 : ncart(am[i]);
    if (am[i] >= 2) {
      ++nhigh;
      if (pu) ++npure;
    }
    p += nprim[i];
    a += s->nao[i];
    lmax = std::max(lmax, am[i]);
  }
  s->nbf = a;
  s->lmax = lmax;
  s->ex.assign(ex, ex + p);
  s->cc.assign(cc, cc + p);
  if (npure != 0 && npure != nhigh) {
    set_err("set_basis: mixed pure/Cartesian shells of l >= 2 are not supported");
    return 1;
  }
  s->pure = npure > 0;
  if (!kernels_support(lmax)) {
    set_err("set_basis: angular momentum %d not supported by the device kernels", lmax);
    return 1;
  }
  s->ready = false;
  return 0;
}

int oqph_session_set_geometry(oqph_session* s, int natm, const double* xyz, const double* zn) {
  if (!s || natm <= 0 || !xyz) {
    set_err("set_geometry: invalid argument");
    return 4;
  }
  for (int i = 0; i < s->nsh; ++i)
    if (s->atom[i] < 0 || s->atom[i] >= natm) {
      set_err("set_geometry: shell %d refers to atom %d of %d", i, s->atom[i], natm);
      return 4;
    }
  s->natm = natm;
  s->xyz.assign(xyz, xyz + 3 * natm);
  s->zn.assign(natm, 0.0);
  if (zn) s->zn.assign(zn, zn + natm);
  s->ready = false;
  return 0;
}

int oqph_session_set_thresholds(oqph_session* s, double c, double gc) {
  if (!s || !(c > 0.0) || !(gc > 0.0)) {
    set_err("set_thresholds: invalid argument");
    return 4;
  }
  if (c != s->cutoff) s->ready = false;
  s->cutoff = c;
  s->grad_cutoff = gc;
  return 0;
}

int oqph_session_set_precision(oqph_session* s, int policy, double fp32_cut) {
  if (!s || policy < 0 || policy > 1 || fp32_cut < 0.0) {
    set_err("set_precision: invalid argument");
    return 4;
  }
  s->policy = policy;
  s->fp32_cut = fp32_cut;
  return 0;
}

int oqph_session_set_grad_precision(oqph_session* s, int policy, double fp32_cut) {
  if (!s || policy < 0 || policy > 1 || fp32_cut < 0.0) {
    set_err("set_grad_precision: invalid argument");
    return 4;
  }
  s->grad_policy = policy;
  s->grad_fp32_cut = fp32_cut;
  return 0;
}

int oqph_session_nbf(oqph_session* s) { return s ? s->nbf : 0; }

int oqph_session_set_profiling(oqph_session* s, int on) {
  if (!s) return 4;
  s->profile = on != 0;
  return 0;
}

int oqph_last_profile(oqph_session* s, int maxn, int* cls, double* ms, double* flops, long long* quartets,
                      long long* prims) {
  if (!s) return -4;
  const int n = (int)std::min<size_t>((size_t)maxn, s->launches.size());
  for (int i = 0; i < n; ++i) {
    const auto& L = s->launches[i];
    const unsigned long long* c = i * 8 + 8 <= (int)s->prof_ctr.size() ? &s->prof_ctr[8 * i] : nullptr;
    // a multi-class launch reports 40000 + its number of classes
    if (cls) cls[i] = L.multi ? 40000 + L.ntask
                              : (L.tiled ? 60000 : L.hl ? 20000 : L.grouped ? 30000 : L.rows ? 10000 : 0) +
                                    L.la * 1000 + L.lb * 100 +
                                    L.lc * 10 + L.ld;
    if (ms) ms[i] = i < (int)s->prof_ms.size() ? s->prof_ms[i] : 0.0;
    if (flops) flops[i] = c ? (double)c[3] : 0.0;
    if (quartets) quartets[i] = c ? (long long)c[0] : 0;
    if (prims) prims[i] = c ? (long long)c[4] : 0;
  }
  return n;
}

int oqph_last_counters(oqph_session* s, int maxn, unsigned long long* ctr) {
  if (!s || !ctr) return -4;
  const int n = (int)std::min<size_t>((size_t)maxn, s->prof_ctr.size() / 8);
  for (int i = 0; i < 8 * n; ++i) ctr[i] = s->prof_ctr[i];
  return n;
}

int oqph_schwarz(oqph_session* s, double* q) {
  if (!s || !q) return 4;
  int rc = prepare_geometry(s);
  if (rc) return rc;
  for (int i = 0; i < s->nsh; ++i)
    for (int j = 0; j <= i; ++j) q[i * s->nsh + j] = q[j * s->nsh + i] = s->Q[(size_t)i * (i + 1) / 2 + j];
  return 0;
}

// UHF/ROHF (nfocks = 2, int2_urohf_data_t, int2.F90:1488-1578): F_s =
// sc J[Da + Db] - se K[Ds] (packed, halved and diagonal-doubled as for RHF),
// in one pass over the integrals.  With D+ = Da + Db and D- = Da - Db,
// K[Da] = (K[D+] + K[D-]) / 2 and K[Db] = (K[D+] - K[D-]) / 2, so
//   F_a = G(D+; se, sc) + X,  F_b = G(D+; se, sc) - X,  X = -se K[D-] / 2:
// the build digests every quartet for D+ (Coulomb and exchange, into fa) and
// for D- (exchange only, into fb), then forms fa + fb and fa - fb.  The
// quartet screens use the shell maxima over both spin densities, as shlden
// does for nfocks = 2 (int2.F90:999-1047).  d and f hold alpha then beta;
// zero_f = false adds to what f holds (its alpha/beta pair is first taken to
// the same half-sum / half-difference form).
static int build_uhf(oqph_session* s, const double* d_dev, double* f_dev, double se, double sc, oqph_stats* st,
                     hipStream_t stream, bool zero_f = true, double mu = 0.0) {
  const size_t ntri = (size_t)s->nbf * (s->nbf + 1) / 2;
  int rc;
  if ((rc = s->d_Dt.ensure(2 * ntri))) return rc;
  const double *da = d_dev, *db = d_dev + ntri;
  double *fa = f_dev, *fb = f_dev + ntri;
  double *dp = s->d_Dt.p, *dm = s->d_Dt.p + ntri;
  launch_sumdiff(da, db, dp, dm, 1.0, (long long)ntri, stream);
  if (!zero_f) launch_sumdiff(fa, fb, fa, fb, 0.5, (long long)ntri, stream);
  if ((rc = build_device(s, dp, fa, se, sc, st, stream, da, db, zero_f, mu, nullptr, dm, fb))) return rc;
  launch_sumdiff(fa, fb, fa, fb, 1.0, (long long)ntri, stream);
  HIPCHK(hipGetLastError());
  return 0;
}

static void add_stats(oqph_stats* a, const oqph_stats& b) {
  a->quartets_computed += b.quartets_computed;
  a->quartets_screened += b.quartets_screened;
  a->quartets_fp32 += b.quartets_fp32;
  a->kernel_ms += b.kernel_ms;
  a->model_flops += b.model_flops;
  a->prim_quartets += b.prim_quartets;
}

int oqph_fock_jk_device(oqph_session* s, const double* d_dev, double* f_dev, int nfocks, double se, double sc,
                        oqph_stats* st, void* stream) {
  if (!s || !d_dev || !f_dev) return 4;
  if (int fc = fault_code()) {
    set_err("OQPH_FAULT injected decline");
    return fc;
  }
  if (nfocks != 1 && nfocks != 2) {
    set_err("nfocks=%d: only nfocks = 1 (RHF) and 2 (UHF/ROHF) are supported", nfocks);
    return 1;
  }
  HIPCHK(hipSetDevice(s->device));
  int rc = prepare_geometry(s);
  if (rc) return rc;
  hipStream_t st_ = stream ? (hipStream_t)stream : s->stream;
  if (nfocks == 2) return build_uhf(s, d_dev, f_dev, se, sc, st, st_);
  return build_device(s, d_dev, f_dev, se, sc, st, st_);
}

int oqph_fock_jk_noalloc(oqph_session* s, const double* d_dev, double* f_dev, double se, double sc, oqph_stats* st,
                         void* stream) {
  if (!s || !d_dev || !f_dev) return 4;
  if (int fc = fault_code()) {
    set_err("OQPH_FAULT injected decline");
    return fc;
  }
  if (!s->ready) {
    set_err("no-allocation build: the session is not prepared (call oqph_session_prepare after setting the basis, "
            "geometry and thresholds; preparing allocates)");
    return 4;
  }
  if (s->share_size > 1 && !s->share_ready) {
    set_err("no-allocation build: the rank share is not prepared (call oqph_session_prepare after "
            "oqph_session_set_share)");
    return 4;
  }
  HIPCHK(hipSetDevice(s->device));
  std::lock_guard<std::mutex> lk(g_mtx);  // the scratch tag of the device
  ScratchSlot* slot = (s->device < OQPH_MAX_DEVICES) ? &g_scratch[s->device] : nullptr;
  if (!slot || !slot->ptr) {
    set_err("no-allocation build: the static scratch of device %d is unavailable", s->device);
    return 2;
  }
  const StaticScratch ss{slot->ptr, (size_t)OQPH_NOALLOC_SCRATCH_BYTES, &slot->serial, &slot->gen};
  hipStream_t st_ = stream ? (hipStream_t)stream : s->stream;
  return build_device(s, d_dev, f_dev, se, sc, st, st_, nullptr, nullptr, true, 0.0, &ss);
}

int oqph_fock_jk(oqph_session* s, const double* d, double* f, int nfocks, double se, double sc, oqph_stats* st) {
  if (!s || !d || !f) return 4;
  if (int fc = fault_code()) {
    set_err("OQPH_FAULT injected decline");
    return fc;
  }
  if (nfocks != 1 && nfocks != 2) {
    set_err("nfocks=%d: only nfocks = 1 (RHF) and 2 (UHF/ROHF) are supported", nfocks);
    return 1;
  }
  HIPCHK(hipSetDevice(s->device));
  int rc = prepare_geometry(s);
  if (rc) return rc;
  const size_t ntri = (size_t)s->nbf * (s->nbf + 1) / 2;
  const size_t nt = ntri * (size_t)nfocks;
  if ((rc = s->d_D.ensure(nt)) || (rc = s->d_F.ensure(nt))) return rc;
  if ((rc = upload(s->d_D.p, d, nt * sizeof(double), s->stream))) return rc;
  if (nfocks == 2) rc = build_uhf(s, s->d_D.p, s->d_F.p, se, sc, st, s->stream);
  else rc = build_device(s, s->d_D.p, s->d_F.p, se, sc, st, s->stream);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(f, s->d_F.p, nt * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return 0;
}

int oqph_fock_jk_inc(oqph_session* s, const double* d, double* d_old, double* f, double* f_old, int nfocks, double se,
                     double sc, oqph_stats* st) {
  if (!s || !d || !d_old || !f || !f_old) return 4;
  const size_t ntri = (size_t)s->nbf * (s->nbf + 1) / 2;
  std::vector<double> dd(ntri * (size_t)std::max(nfocks, 1));
  for (size_t i = 0; i < dd.size(); ++i) dd[i] = d[i] - d_old[i];
  int rc = oqph_fock_jk(s, dd.data(), f, nfocks, se, sc, st);
  if (rc) return rc;
  for (size_t i = 0; i < dd.size(); ++i) {
    f[i] += f_old[i];
    f_old[i] = f[i];
    d_old[i] = d[i];
  }
  return 0;
}

int oqph_mrsf_int2(oqph_session* s, int nf, int nmat, int ncoul, const double* d3, double* f3, double se,
                   double sc, oqph_stats* st) {
  if (!s || !d3 || !f3 || nf < 1 || nmat < 1 || ncoul < 0 || ncoul > nmat) {
    set_err("mrsf_int2: invalid argument");
    return 4;
  }
  if (int fc = fault_code()) {
    set_err("OQPH_FAULT injected decline");
    return fc;
  }
  HIPCHK(hipSetDevice(s->device));
  int rc = prepare_geometry(s);
  if (rc) return rc;
  const int nbf = s->nbf, nsh = s->nsh;
  const size_t n2 = (size_t)nbf * nbf, nd = (size_t)nf * nmat;
  // slabs q = v nmat + c, each in the reference's (mu, nu) column order
  std::vector<double> h(nd * n2);
  size_t x = 0;
  for (int nu = 0; nu < nbf; ++nu)
    for (int mu = 0; mu < nbf; ++mu)
      for (int c = 0; c < nmat; ++c)
        for (int v = 0; v < nf; ++v) h[((size_t)v * nmat + c) * n2 + (size_t)nu * nbf + mu] = d3[x++];
  std::vector<int> coul(nd);
  for (size_t q = 0; q < nd; ++q) coul[q] = (int)(q % nmat) < ncoul;
  // shell_den_screen_mrsf (tdhf_mrsf_lib.F90:176-204): the last component
  // only, dsh(i, j) = dsh(j, i) = max |M(mu, nu)|, mu in shell j, nu in
  // shell i >= j
  std::vector<double> dsh((size_t)nsh * nsh, 0.0);
  double maxden = 0.0;
  for (int si = 0; si < nsh; ++si)
    for (int sj = 0; sj <= si; ++sj) {
      double dm = 0.0;
      for (int v = 0; v < nf; ++v) {
        const double* M = h.data() + ((size_t)v * nmat + nmat - 1) * n2;
        for (int nu = s->aooff[si]; nu < s->aooff[si] + s->nao[si]; ++nu)
          for (int mu = s->aooff[sj]; mu < s->aooff[sj] + s->nao[sj]; ++mu) dm = std::max(dm, fabs(M[(size_t)nu * nbf + mu]));
      }
      dsh[(size_t)si * nsh + sj] = dsh[(size_t)sj * nsh + si] = dm;
      maxden = std::max(maxden, dm);
    }
  if ((rc = s->d_dsh.ensure((size_t)nsh * nsh)) || (rc = s->d_gD.ensure(nd * n2)) ||
      (rc = s->d_gF.ensure(nd * n2)) || (rc = s->d_gcoul.ensure(nd)))
    return rc;
  // synchronous: h, dsh and coul are host locals that an early return frees
  if ((rc = upload_sync(s->d_gD.p, h.data(), nd * n2 * sizeof(double), s->stream)) ||
      (rc = upload_sync(s->d_dsh.p, dsh.data(), dsh.size() * sizeof(double), s->stream)) ||
      (rc = upload_sync(s->d_gcoul.p, coul.data(), nd * sizeof(int), s->stream)))
    return rc;
  if ((rc = build_general(s, (int)nd, s->d_gD.p, s->d_gF.p, s->d_gcoul.p, maxden, se, sc, st, s->stream))) return rc;
  HIPCHK(hipMemcpyAsync(h.data(), s->d_gF.p, nd * n2 * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  x = 0;
  for (int nu = 0; nu < nbf; ++nu)
    for (int mu = 0; mu < nbf; ++mu)
      for (int c = 0; c < nmat; ++c)
        for (int v = 0; v < nf; ++v) f3[x++] = h[((size_t)v * nmat + c) * n2 + (size_t)nu * nbf + mu];
  return 0;
}

// The gradient launches over the Cartesian density (p.Dc) or the MRSF
// matrices (p.Mc) the caller set in p; de receives the natm x 3 sums.
static int grad_run(oqph_session* s, GradParams& p, double* de, oqph_stats* st) {
  int rc;
  hipStream_t stream = s->stream;
  const size_t nde = 3 * (size_t)s->natm;
  HIPCHK(hipMemsetAsync(s->d_de.p, 0, nde * OQPH_DE_SLOTS * sizeof(double), stream));
  const size_t nl = s->glaunches.size();
  const size_t nc = 8 * (size_t)OQPH_CTR_SLOTS;  // counters per launch
  if ((rc = s->d_ctr.ensure(8 + nc * std::max(nl, s->launches.size())))) return rc;
  HIPCHK(hipMemsetAsync(s->d_ctr.p + 8, 0, nc * nl * sizeof(unsigned long long), stream));
  double zbig = 0.0;
  for (double e : s->ex) zbig = std::max(zbig, e);
  double dabcut = 1.0e-11;                 // grd2.F90:224-226
  if (zbig > 1.0e6) dabcut /= 10;
  if (zbig > 1.0e7) dabcut /= 10;
  p.pairs = s->d_gpairs.p;
  p.prims = s->d_gprims.p;
  p.rys = s->d_rys.p;
  p.rys32 = s->d_rys32.p;
  p.coff = s->d_coff.p;
  p.shell_atom = s->d_satom.p;
  p.de = s->d_de.p;
  p.ncf = s->ncf;
  p.natm = s->natm;
  p.gcut = s->grad_cutoff;
  p.dabcut2 = dabcut * dabcut;
  HIPCHK(hipEventRecord(s->ev0, stream));
  const int nst = s->profile ? 1 : s->nstreams;
  if ((rc = streams_fork(s, stream, nst))) return rc;
  for (size_t i = 0; i < nl; ++i) {
    auto& L = s->glaunches[i];
    const hipStream_t ls = stream_of(s, stream, i, nst);
    p.counters = s->d_ctr.p + 8 + nc * i;
    if (L.multi) {
      launch_gradh_multi(p, L.tl, L.nblocks, L.lds, ls);
    } else if (L.hl) {
      launch_gradh(L.hd, p, L.ww, ls);
    } else if (!launch_grad(L.la, L.lb, L.lc, L.ld, p, L.ww, ls, L.f32, p.Mc != nullptr)) {
      set_err("no gradient kernel for class (%d %d|%d %d)", L.la, L.lb, L.lc, L.ld);
      return 1;
    }
  }
  HIPCHK(hipGetLastError());
  if ((rc = streams_join(s, stream, nst))) return rc;
  HIPCHK(hipEventRecord(s->ev1, stream));
  std::vector<double> deraw(nde * OQPH_DE_SLOTS);
  HIPCHK(hipMemcpyAsync(deraw.data(), s->d_de.p, deraw.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
  std::vector<unsigned long long> raw(nc * nl);
  if (nl) HIPCHK(hipMemcpyAsync(raw.data(), s->d_ctr.p + 8, nc * nl * sizeof(unsigned long long),
                                hipMemcpyDeviceToHost, stream));
  HIPCHK(hipStreamSynchronize(stream));
  for (size_t i = 0; i < nde; ++i) {
    double v = 0.0;
    for (int k = 0; k < OQPH_DE_SLOTS; ++k) v += deraw[k * nde + i];
    de[i] = v;
  }
  std::vector<unsigned long long> ctr(8 * nl, 0ull);
  for (size_t i = 0; i < nl; ++i)
    for (size_t k = 0; k < nc; ++k) ctr[8 * i + k % 8] += raw[nc * i + k];
  if (st) {
    memset(st, 0, sizeof(*st));
    for (size_t i = 0; i < nl; ++i) {
      st->quartets_computed += (long long)ctr[8 * i];
      st->quartets_screened += (long long)ctr[8 * i + 1];
      st->model_flops += (double)ctr[8 * i + 3];
      st->prim_quartets += (long long)ctr[8 * i + 4];
    }
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, s->ev0, s->ev1));
    st->kernel_ms = ms;
  }
  return 0;
}

int oqph_grad2(oqph_session* s, const double* d, double* de, double hfscale, double coulscale, oqph_stats* st) {
  if (!s || !d || !de) return 4;
  if (int fc = fault_code()) {
    set_err("OQPH_FAULT injected decline");
    return fc;
  }
  HIPCHK(hipSetDevice(s->device));
  int rc = prepare_grad(s);
  if (rc) return rc;
  const size_t ntri = (size_t)s->nbf * (s->nbf + 1) / 2;
  if ((rc = upload(s->d_D.p, d, ntri * sizeof(double), s->stream))) return rc;
  launch_cart_density(s->d_D.p, s->nsh, s->d_am.p, s->d_pure.p, s->d_aooff.p, s->d_coff.p, s->d_xform.p, s->ncf,
                      s->d_Dc.p, s->stream);
  GradParams p{};
  p.Dc = s->d_Dc.p;
  p.cs = coulscale;
  p.hs = hfscale;
  return grad_run(s, p, de, st);
}

int oqph_grad2_mrsf(oqph_session* s, const double* d, const double* pd, const double* spc, double* de, double hfscale,
                    double hfscale2, double coulscale, const double* spcscale, int mrst, oqph_stats* st) {
  if (!s || !d || !pd || !spc || !de || !spcscale) return 4;
  if (int fc = fault_code()) {
    set_err("OQPH_FAULT injected decline");
    return fc;
  }
  HIPCHK(hipSetDevice(s->device));
  s->grad_all_hl = true;
  int rc = prepare_grad(s);
  s->grad_all_hl = false;
  if (rc) return rc;
  const int nbf = s->nbf;
  const size_t n2 = (size_t)nbf * nbf;
  // the sums and differences of the raw spin densities (grd2_mrsf_compute_data_t_init,
  // tdhf_mrsf_gradient.F90:341-353) and the seven response densities, row-major
  std::vector<double> h(11 * n2);
  static const int comp[7] = {6, 0, 1, 2, 3, 4, 5};  // ball, bo2v, bo1v, bco1, bco2, o21v, co12 in spc(1:7)
  for (int i = 0; i < nbf; ++i)
    for (int j = 0; j < nbf; ++j) {
      const size_t f = (size_t)i + (size_t)nbf * j, r = (size_t)i * nbf + j;
      h[0 * n2 + r] = d[f] + d[f + n2];
      h[1 * n2 + r] = d[f] - d[f + n2];
      h[2 * n2 + r] = pd[f] + pd[f + n2];
      h[3 * n2 + r] = pd[f] - pd[f + n2];
      for (int x = 0; x < 7; ++x) h[(4 + x) * n2 + r] = spc[comp[x] + 7 * f];
    }
  const size_t nc2 = (size_t)s->ncf * s->ncf;
  if ((rc = s->d_gM.ensure(11 * n2)) || (rc = s->d_gMc.ensure(11 * nc2))) return rc;
  if ((rc = upload_sync(s->d_gM.p, h.data(), 11 * n2 * sizeof(double), s->stream))) return rc;
  for (int x = 0; x < 11; ++x)
    launch_cart_matrix(s->d_gM.p + x * n2, nbf, s->nsh, s->d_am.p, s->d_pure.p, s->d_aooff.p, s->d_coff.p,
                       s->d_xform.p, s->ncf, s->d_gMc.p + x * nc2, s->stream);
  GradParams p{};
  p.Mc = s->d_gMc.p;
  p.cs = coulscale;
  p.hs = hfscale;
  p.hs2 = hfscale2;
  p.qf0 = spcscale[0];
  p.qf1 = spcscale[1];
  p.qf2 = spcscale[2];
  p.sgnk = mrst == 3 ? -1.0 : 1.0;
  rc = grad_run(s, p, de, st);
  // the next ground-state gradient rebuilds its lists with the class templates
  return rc;
}

// CAM build (int2_run_cam, source/integrals/int2.F90:538-584): pass 1 with
// 1/r, Coulomb scale alpha_coulomb and exchange scale alpha; pass 2 with
// erf(mu r)/r, Coulomb scale beta_coulomb and exchange scale beta; both
// accumulate into f, which then carries fock_jk's scaling.
int oqph_fock_jk_cam(oqph_session* s, const double* d, double* f, int nfocks, double alpha, double beta, double mu,
                     double alpha_coulomb, double beta_coulomb, oqph_stats* st) {
  if (!s || !d || !f) return 4;
  if (int fc = fault_code()) {
    set_err("OQPH_FAULT injected decline");
    return fc;
  }
  if (nfocks != 1 && nfocks != 2) {
    set_err("nfocks=%d: only nfocks = 1 (RHF) and 2 (UHF/ROHF) are supported", nfocks);
    return 1;
  }
  if (!(mu > 0.0)) {
    set_err("fock_jk_cam: mu must be positive");
    return 4;
  }
  HIPCHK(hipSetDevice(s->device));
  int rc = prepare_geometry(s);
  if (rc) return rc;
  const size_t ntri = (size_t)s->nbf * (s->nbf + 1) / 2;
  const size_t nt = ntri * (size_t)nfocks;
  if ((rc = s->d_D.ensure(nt)) || (rc = s->d_F.ensure(nt))) return rc;
  if ((rc = upload(s->d_D.p, d, nt * sizeof(double), s->stream))) return rc;
  oqph_stats s1{}, s2{};
  if (nfocks == 2) {
    if ((rc = build_uhf(s, s->d_D.p, s->d_F.p, alpha, alpha_coulomb, &s1, s->stream, true, 0.0))) return rc;
    if ((rc = build_uhf(s, s->d_D.p, s->d_F.p, beta, beta_coulomb, &s2, s->stream, false, mu))) return rc;
  } else {
    if ((rc = build_device(s, s->d_D.p, s->d_F.p, alpha, alpha_coulomb, &s1, s->stream, nullptr, nullptr, true, 0.0)))
      return rc;
    if ((rc = build_device(s, s->d_D.p, s->d_F.p, beta, beta_coulomb, &s2, s->stream, nullptr, nullptr, false, mu)))
      return rc;
  }
  HIPCHK(hipMemcpyAsync(f, s->d_F.p, nt * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  if (st) {
    *st = s1;
    add_stats(st, s2);
  }
  return 0;
}

int oqph_session_set_share(oqph_session* s, int rank, int size) {
  if (!s || size < 1 || rank < 0 || rank >= size) {
    set_err("set_share: need 0 <= rank < size");
    return 4;
  }
  s->share_rank = rank;
  s->share_size = size;
  s->share_ready = false;
  return 0;
}

int oqph_session_prepare(oqph_session* s) {
  if (!s) return 4;
  HIPCHK(hipSetDevice(s->device));
  int rc = prepare_geometry(s);
  if (rc) return rc;
  return prepare_share(s);
}

int oqph_kernel_resources(int kind, int la, int lb, int lc, int ld, int pure, int* out) {
  if (!out) return 4;
  KernelRes r{};
  if (!kernel_resources(kind, la, lb, lc, ld, pure != 0, &r)) {
    set_err("no class-templated kernel of kind %d for (%d %d|%d %d)", kind, la, lb, lc, ld);
    return 1;
  }
  out[0] = r.private_bytes;
  out[1] = r.lds_bytes;
  out[2] = r.regs;
  out[3] = r.max_threads;
  return 0;
}

void routec_fock_jk(const double* d, double* f, const int* nbf, const int* nfocks, const double* se,
                    const double* sc, int* info) {
  if (!info) return;
  *info = 1;
  oqph_session* s;
  {
    std::lock_guard<std::mutex> lk(g_mtx);
    s = g_default;
  }
  if (!s || !d || !f || !nbf || !nfocks || !se || !sc) {
    set_err("routec_fock_jk: no default session");
    return;
  }
  if (*nbf != s->nbf) {
    set_err("routec_fock_jk: nbf %d does not match the session basis (%d)", *nbf, s->nbf);
    *info = 4;
    return;
  }
  *info = oqph_fock_jk(s, d, f, *nfocks, *se, *sc, nullptr);
}

void routec_grad2(const double* d, const double* xyz, double* de, const int* nbf, const int* natm,
                  const double* hfscale, const double* coulscale, int* info) {
  if (!info) return;
  *info = 1;
  oqph_session* s;
  {
    std::lock_guard<std::mutex> lk(g_mtx);
    s = g_default;
  }
  if (!s || !d || !xyz || !de || !nbf || !natm || !hfscale || !coulscale) {
    set_err("routec_grad2: no default session");
    return;
  }
  if (*nbf != s->nbf || *natm != s->natm) {
    set_err("routec_grad2: dimensions do not match the session");
    *info = 4;
    return;
  }
  // the caller passes the current geometry: refresh the session's copy
  bool moved = false;
  for (int i = 0; i < 3 * s->natm; ++i)
    if (xyz[i] != s->xyz[i]) moved = true;
  if (moved) {
    s->xyz.assign(xyz, xyz + 3 * s->natm);
    s->ready = false;
  }
  *info = oqph_grad2(s, d, de, *hfscale, *coulscale, nullptr);
}

void routec_grad2_mrsf(const double* d, const double* pd, const double* spc, const double* xyz, double* de,
                       const int* nbf, const int* natm, const double* hfscale, const double* hfscale2,
                       const double* coulscale, const double* spcscale, const int* mrst, int* info) {
  if (!info) return;
  *info = 1;
  oqph_session* s;
  {
    std::lock_guard<std::mutex> lk(g_mtx);
    s = g_default;
  }
  if (!s || !d || !pd || !spc || !xyz || !de || !nbf || !natm || !hfscale || !hfscale2 || !coulscale ||
      !spcscale || !mrst) {
    set_err("routec_grad2_mrsf: no default session");
    return;
  }
  if (*nbf != s->nbf || *natm != s->natm) {
    set_err("routec_grad2_mrsf: dimensions do not match the session");
    *info = 4;
    return;
  }
  bool moved = false;
  for (int i = 0; i < 3 * s->natm; ++i)
    if (xyz[i] != s->xyz[i]) moved = true;
  if (moved) {
    s->xyz.assign(xyz, xyz + 3 * s->natm);
    s->ready = false;
  }
  *info = oqph_grad2_mrsf(s, d, pd, spc, de, *hfscale, *hfscale2, *coulscale, spcscale, *mrst, nullptr);
}

}  // extern "C"
