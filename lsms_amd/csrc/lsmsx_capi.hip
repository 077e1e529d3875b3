// lsmsx_capi.hip -- context, device residency and the C ABI of the MI355X
// tau00 path (declared in include/lsmsx.h).
//
// Device residency mirrors the reference's DeviceStorage
// (Accelerator/DeviceStorage.hpp:56-184): t-matrix store, LIZ geometry and
// angular-momentum tables are uploaded once and stay on the device; the
// batched call assembles and factors every (atom, energy) matrix of a chunk
// without host round trips, and only the kkrsz_ns^2 tau00 blocks come back.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <complex>
#include <chrono>
#include <cmath>
#include <tuple>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "../../include/lsmsx.h"
#include "lsmsx_device.hpp"

static_assert(lsmsx::kInfoSweepStall == LSMSX_INFO_SWEEP_STALL, "info code");
static_assert(lsmsx::kInfoRefineUnconverged == LSMSX_INFO_REFINE_UNCONVERGED, "info code");

namespace lsmsx {
template <int NSC, int MODE>
__global__ void k_assemble(const SolveDesc *, const double2 *, long, const double *,
                           const double2 *, int);
template <int MODE>
__global__ void k_assemble_from_g(const SolveDesc *, const double2 *, long);
__global__ void k_pad_identity(const SolveDesc *);
template <class T>
__global__ void k_lu_diag(const SolveDesc *, int);
template <class T>
__global__ void k_lu_panel(const SolveDesc *, int);
template <class T>
__global__ void k_lu_update(const SolveDesc *, int, int);
template <class T>
__global__ void k_lu_cross(const SolveDesc *, int, int);
template <class T>
__global__ void k_lu_trail(const SolveDesc *, int, int);
template <class T>
__global__ void k_lu_last(const SolveDesc *);
template <class T>
__global__ void k_ir_init(const SolveDesc *);
template <class T, int NB>
__global__ void k_ir_residual(const SolveDesc *);
template <class T, bool FWD>
__global__ void k_ir_step(const SolveDesc *, int);
template <class T>
__global__ void k_ir_accum(const SolveDesc *, int);
__global__ void k_ir_decide(const SolveDesc *, int, double);
template <bool FIRST>
__global__ void k_ir_sweep(const SolveDesc *, double);
__global__ void k_ir_residual_g(const SolveDesc *);
__global__ void k_ir_residual_g_finish(const SolveDesc *);
__global__ void k_copy_m(const SolveDesc *);
__global__ void k_fitpot(const double *, const double *, const int *, int, int, double *);
__global__ void k_single_site(const double *, const double *, const SingleSitePot *,
                              const double2 *, const double2 *, int, int, int, int, int,
                              double2 *, double2 *, double2 *, int *);
__global__ void k_green_mt_canted(const double2 *, const double2 *, const double2 *,
                                  const double2 *, const double *, const GreenPot *, const int *,
                                  int, int, int, double2 *, double2 *);
__global__ void k_tmat_fill_canted(const double2 *, const int *, const double2 *, int, int, int,
                                   int, long, long, double2 *);
__global__ void k_jlr_scale(const double *, const SingleSitePot *, const double2 *,
                            const double2 *, int, int, int, int, double2 *);
__global__ void k_tmat_fill(const double2 *, const int *, const int *, int, int, int, long, long,
                            double2 *);
__global__ void k_green_mt(const double2 *, const double2 *, const double2 *, const double *,
                           const GreenPot *, const int *, const int *, int, int, int, double,
                           double2 *, double2 *);
__global__ void k_ir_extract(const SolveDesc *);
__global__ void k_count_conv(const SolveDesc *, int, int *);
__global__ void k_lloyd(const SolveDesc *, const double2 *, const double2 *, long, const double *,
                        const double2 *, int);
__global__ void k_inv_zero(const SolveDesc *);
__global__ void k_inv_fwd_panel(const SolveDesc *, int);
__global__ void k_inv_fwd_update(const SolveDesc *, int);
__global__ void k_inv_bwd_panel(const SolveDesc *, int);
__global__ void k_inv_bwd_update(const SolveDesc *, int);
__global__ void k_inv_x(const SolveDesc *);
__global__ void k_tau_full_out(const SolveDesc *, const double2 *, long, const int *, const int *,
                               double2 *, long);
__global__ void k_lloyd_reduce(const SolveDesc *, int, double2 *);
__global__ void k_tau_grad(const SolveDesc *, const double2 *, double2 *, int);
__global__ void k_grad_reduce(const double2 *, int, const int *, const long *, const int *,
                              const int *, double2 *, long, int);
__global__ void k_zero_c(double2 *, long);
__global__ void k_lu_trail_h(const SolveDesc *, int, int);
__global__ void k_interstitial(int, int, const double2 *, const double2 *, const double2 *,
                               const double *, const int *, const int *, int, int, const double *,
                               const double *, const double2 *, const double *, int,
                               const double2 *, double, double2 *, double2 *, double2 *);
__global__ void k_rel_tstore(double2 *, long, long, int, const double *, const double2 *);
__global__ void k_rel_tau(double2 *, int, const double *, const double2 *);
__global__ void k_rel_blocks(double2 *, long, int, const double *);
__global__ void k_lu_cross_diag_h(const SolveDesc *, int, int);
__global__ void k_lu_panel_h(const SolveDesc *, int, int, int);

// Copy a caller matrix (central block first, n x n) into the solver layout;
// rowperm (or null) reorders the n - k non-central rows: solver row xr takes
// non-central row rowperm[xr].  A row permutation that leaves the central
// rows in place leaves tau00 = [M^-1]_00 t0 unchanged.
__global__ __launch_bounds__(256) void k_permute_in(const SolveDesc *__restrict__ descs,
                                                    const double2 *__restrict__ src, long lds,
                                                    int n, int k, const int *__restrict__ rowperm) {
  const SolveDesc &d = descs[0];
  const int Ns = d.Ns, pad = d.pad;
  const int c = blockIdx.x;  // solver-layout column
  if (c >= Ns) return;
  double2 *col = d.M + d.ld * (long)c;
  for (int r = threadIdx.x; r < Ns; r += blockDim.x) {
    double2 v;
    if (r < pad || c < pad) {
      v = make_double2(r == c ? 1.0 : 0.0, 0.0);
    } else {
      int xr = r - pad, xc = c - pad;
      int orr = xr < n - k ? (rowperm ? rowperm[xr] : xr) + k : xr - (n - k);
      int oc = xc < n - k ? xc + k : xc - (n - k);
      v = src[orr + lds * oc];
    }
    col[r] = v;
  }
}

// Right-hand side [t0; 0] of the column-pivoted fallback: B (N x kc, column
// stride N) from the central site's t block t0 (column stride ldt).
__global__ void k_fallback_rhs(double2 *__restrict__ B, int N, int kc,
                               const double2 *__restrict__ t0, int ldt) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= (long)N * kc) return;
  const int r = (int)(e % N), col = (int)(e / N);
  B[e] = r < kc ? t0[r + (long)ldt * col] : make_double2(0.0, 0.0);
}

__global__ void k_init_status(const SolveDesc *__restrict__ descs, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    *descs[i].info = 0;
    *descs[i].minpiv = 1e300;
    if (descs[i].conv) *descs[i].conv = 0;
  }
}

// Scratch of lsmsx_solve_tau00_matrix_noalloc, one copy per GPU for the life
// of the process: the solver-layout copy of the caller's matrix (at most
// kNoAllocMaxNs^2 complex), the one reused diagonal-tile inverse, the row
// permutations, the smallest pivot and the solve descriptor.  Host code
// reaches it only through the address hipGetSymbolAddress returns for the
// current device (noalloc_scratch_for_device), never through the name.
constexpr int kNoAllocMaxNs = 1024;
constexpr size_t kNoAllocInvOff = (size_t)kNoAllocMaxNs * kNoAllocMaxNs * sizeof(double2);
constexpr size_t kNoAllocPermOff = kNoAllocInvOff + (size_t)kTile * kTile * sizeof(double2);
constexpr size_t kNoAllocMinOff = kNoAllocPermOff + (size_t)kNoAllocMaxNs * sizeof(int);
constexpr size_t kNoAllocDescOff = kNoAllocMinOff + 256;
constexpr size_t kNoAllocBytes = kNoAllocDescOff + ((sizeof(SolveDesc) + 255) / 256) * 256;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

// The descriptor travels as a kernel argument and is stored by the device,
// so the no-allocation entry point needs no staging buffer for it.
__global__ void k_put_desc(SolveDesc d, SolveDesc *__restrict__ dst) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *dst = d;
}
}  // namespace lsmsx

using namespace lsmsx;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(call)                                                              \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess)                                                         \
      return fail(-100 - (int)e_, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

struct AtomDev {
  int numLIZ = 0;
  int nrmat = 0;
  int kkrsz = 0;
  std::vector<int> lmax, store;
  std::vector<double> pos;  // host copy of LIZPos
  int geo = -1;             // atoms with bitwise equal LIZPos and LIZlmax share this id
  double *d_pos = nullptr;
  int *d_lmax = nullptr, *d_store = nullptr;
  int *d_off[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};  // [solver layout?][nsc-1]
};

template <class T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  int ensure(size_t want) {
    if (want <= n) return 0;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    hipError_t e = hipMalloc((void **)&p, want * sizeof(T));
    if (e != hipSuccess) {
      p = nullptr;
      return fail(-100 - (int)e, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    n = want;
    return 0;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};

int roundup(int x, int m) { return ((x + m - 1) / m) * m; }

// checks of one potential's mesh for the muffin-tin Green function and the
// cinterp.f window at sqrt(r_sph) on x_0 = 0, x_k = sqrt(r_k), k = 1..jmt
const char *make_green_pot(const double *r, int jmt, int jws, double r_sph, int iprpts,
                           GreenPot &P) {
  P.jmt = jmt;
  P.jws = jws;
  if (jmt < 4 || jmt > iprpts || jws < 0 || jws > iprpts) return "jmt or jws outside 4..iprpts";
  if (!(r[0] > 0.0)) return "the radial mesh must start above r = 0";
  for (int j = 0; j + 1 < jmt; j++)
    if (!(r[j + 1] > r[j])) return "the radial mesh must increase";
  if (!(r_sph > 0.0)) return "r_sph must be > 0";
  P.xs = std::sqrt(r_sph);
  const int nr = jmt + 1;
  int ip1 = 1;
  for (int i = 1; i <= nr - 2; i++)
    if (P.xs > std::sqrt(r[i - 1])) ip1 = i + 1;
  const int i = ip1 - 1;
  P.ip1 = ip1;
  P.i1 = (i > 0) ? i - 1 : std::min(4, nr) - 1;
  P.i2 = (ip1 < nr - 1) ? ip1 + 1 : std::max(nr - 3, 1) - 1;
  return nullptr;
}

}  // namespace

struct lsmsx_ctx {
  int device = 0;
  unsigned char *noalloc = nullptr;  // this device's copy of `scratch`
  hipStream_t stream = nullptr;
  hipStream_t stream2 = nullptr;  // second half of a chunk (LSMSX_SCHED_TWO_STREAMS)
  hipEvent_t ev_asm = nullptr, ev_join = nullptr;  // second-half start and join
  int lmax_tab = -1;
  DevBuf<double> cgnt;
  DevBuf<double2> illp;
  DevBuf<double2> tmat, dtmat;
  long ldT = 0, blk = 0;
  int nstore = 0, group = 0;
  std::map<int, AtomDev> atoms;
  // batch pool
  DevBuf<double2> M, inv, tau, tmpM, t0buf, X, A64, R64;
  DevBuf<float2> W, R;
  DevBuf<int> perm, info, conv, iters, count;
  DevBuf<double> minpiv, dxn;
  DevBuf<unsigned short> hp;  // split-FP16 panel planes
  DevBuf<int> he;
  DevBuf<SolveDesc> desc, desc2, desc_own;
  // explicit inverse / adjoint
  DevBuf<double2> Minv, gx, lpart, lloyd, vbar, gblk, gstore, tfull;
  DevBuf<int> keyi, colmap;
  // single-site scattering batch
  DevBuf<double> ss_r, ss_v, ss_coef;
  DevBuf<SingleSitePot> ss_pot;
  DevBuf<int> ss_iend, ss_status;
  DevBuf<double2> ss_e, ss_prel, ss_matom, ss_zlr, ss_jlr;
  DevBuf<GreenPot> gf_pot;
  DevBuf<int> gf_potof, gf_src;
  // radial functions of the last single-site call kept in ss_zlr/ss_jlr
  int ss_resident = 0, ss_valid = 0, ss_ne = 0, ss_npot = 0, ss_lmax = 0, ss_iprpts = 0;
  int ss_have_matom = 0;  // ss_matom holds the last call's ss_ne x ss_npot x (ss_lmax+1)
  DevBuf<int> ss_storeof;
  DevBuf<double2> ss_uu;
  std::vector<GreenPot> ss_gpot;
  std::vector<double> ss_hr, ss_hrsph;  // host copies of the last call's meshes and r_sph
  DevBuf<double2> gf_tau, gf_dos, gf_green, gf_w;
  DevBuf<long> keyl;
  std::vector<int> last_iters;
  std::vector<double> last_minpiv;
  long ws_limit = 0;
  int ir_max_sweeps = 12;
  // residual form of the mixed path (LSMSX_RESIDUAL_*) and the G-half buffers
  int residual_form = LSMSX_RESIDUAL_AUTO;
  // structure constants once per group of solves with the same LIZ geometry
  // and energy (lsmsx_set_shared_structure_constants)
  int share_g = 1;
  int next_geo = 0;
  DevBuf<double2> GS;  // their storage on the paths that keep no G otherwise
  DevBuf<int> rowperm;  // retry order of lsmsx_solve_tau00_matrix
  // column-pivoted fallback of the batched calls (rocSOLVER getrf/getrs on
  // the FP64 matrix in the reference layout)
  rocblas_handle blas = nullptr;
  DevBuf<double2> fbM, fbB;
  DevBuf<int> fbpiv, fbinfo;
  DevBuf<SolveDesc> fbdesc;
  DevBuf<double2> GH, YB, GP;
  // relativistic store (lsmsx_upload_tmats_rel): the solver store holds
  // t' = f U^T t U, f = prel^2/E per energy slot; rel_finv[slot] = 1/f
  int rel = 0, rel_kk2 = 0;
  std::vector<double2> rel_finv;
  DevBuf<double> relU;
  DevBuf<double2> relf;
  // interstitial cell integrals (lsmsx_green_interstitial)
  DevBuf<double2> it_c2;  // prel, matom left/right, wylm, tau00, pzz, pzj, dos (one buffer)
  DevBuf<double> it_d;    // r_sph, grwylm, gwwylm
  DevBuf<int> it_i;       // cell_of, ncrit
  double ir_tol = 1e-12;
  double pivot_guard = 1e-2;  // plain FP64 solves with a smaller tile pivot are refined
  // schedule (LSMSX_SCHED_* flags) and hipGraph replay of the batched path
  int sched = 0;
  int split_group = 8;                     // steps per trailing update, split-FP16 path
  double group_ms[3] = {-1.0, -1.0, -1.0};  // last autotune: split groups 4, 8, 16
  bool use_graph = false;
  std::map<std::vector<long>, hipGraphExec_t> graphs;
  // profiling
  bool profiling = false;
  double prof_ms[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  long prof_n[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;
  std::vector<hipEvent_t> event_pool;
};

namespace {

hipEvent_t get_event(lsmsx_ctx *c) {
  if (!c->event_pool.empty()) {
    hipEvent_t e = c->event_pool.back();
    c->event_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

struct ProfScope {
  lsmsx_ctx *c;
  int kind;
  hipEvent_t a = nullptr, b = nullptr;
  ProfScope(lsmsx_ctx *c_, int k) : c(c_), kind(k) {
    if (c->profiling) {
      a = get_event(c);
      b = get_event(c);
      if (a) (void)hipEventRecord(a, c->stream);
    }
  }
  ~ProfScope() {
    if (c->profiling && a && b) {
      (void)hipEventRecord(b, c->stream);
      c->pending.push_back({kind, {a, b}});
    }
  }
};

// KKR assembly of a batch (descriptors dd) with n_spin_cant = nsc, writing
// the outputs `mode` names (kAsmM: M; kAsmMW: M and W; kAsmGW: the
// structure constants once per site pair and W, collinear).
void launch_assemble(lsmsx_ctx *c, int nsc, dim3 grid, hipStream_t st, const SolveDesc *dd,
                     int mode = kAsmM) {
#define LSMSX_ASM_LAUNCH(NSC, MODE)                                                         \
  hipLaunchKernelGGL((k_assemble<NSC, MODE>), grid, dim3(256), 0, st, dd, c->tmat.p, c->ldT, \
                     c->cgnt.p, c->illp.p, c->lmax_tab)
  if (nsc == 2) {
    if (mode == kAsmMW) LSMSX_ASM_LAUNCH(2, kAsmMW);
    else LSMSX_ASM_LAUNCH(2, kAsmM);
  } else if (mode == kAsmG) {
    LSMSX_ASM_LAUNCH(1, kAsmG);
  } else if (mode == kAsmGW) {
    LSMSX_ASM_LAUNCH(1, kAsmGW);
  } else if (mode == kAsmMW) {
    LSMSX_ASM_LAUNCH(1, kAsmMW);
  } else {
    LSMSX_ASM_LAUNCH(1, kAsmM);
  }
#undef LSMSX_ASM_LAUNCH
}

// Blocks of every solve of a batch from structure constants stored once per
// group of solves (d.gh), collinear: mode kAsmM, kAsmMW or kAsmGW.
void launch_assemble_from_g(lsmsx_ctx *c, dim3 grid, hipStream_t st, const SolveDesc *dd,
                            int mode) {
  if (mode == kAsmGW)
    hipLaunchKernelGGL(k_assemble_from_g<kAsmGW>, grid, dim3(256), 0, st, dd, c->tmat.p, c->ldT);
  else if (mode == kAsmMW)
    hipLaunchKernelGGL(k_assemble_from_g<kAsmMW>, grid, dim3(256), 0, st, dd, c->tmat.p, c->ldT);
  else
    hipLaunchKernelGGL(k_assemble_from_g<kAsmM>, grid, dim3(256), 0, st, dd, c->tmat.p, c->ldT);
}

int drain_profile(lsmsx_ctx *c) {
  for (auto &p : c->pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, p.second.first, p.second.second) == hipSuccess) {
      c->prof_ms[p.first] += ms;
      c->prof_n[p.first] += 1;
    }
    c->event_pool.push_back(p.second.first);
    c->event_pool.push_back(p.second.second);
  }
  c->pending.clear();
  return 0;
}

int check_launch(const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(-100 - (int)e, std::string(what) + ": " + hipGetErrorString(e));
  return 0;
}

// Offsets of the site blocks: reference order (site 0 first, no padding) or
// solver order (padding, sites 1..n-1, site 0 last).
std::vector<int> site_offsets(const AtomDev &a, int nsc, bool solver) {
  std::vector<int> off(a.numLIZ);
  int N = nsc * a.nrmat;
  if (!solver) {
    int o = 0;
    for (int s = 0; s < a.numLIZ; s++) {
      off[s] = o;
      o += nsc * (a.lmax[s] + 1) * (a.lmax[s] + 1);
    }
  } else {
    int pad = roundup(N, kTile) - N;
    int o = pad;
    for (int s = 1; s < a.numLIZ; s++) {
      off[s] = o;
      o += nsc * (a.lmax[s] + 1) * (a.lmax[s] + 1);
    }
    off[0] = o;
  }
  return off;
}

// Steps per trailing update: with the split-FP16 update the context's
// split_group (4, 8 or 16, default 8: the trailing matrix is read and
// written once per group, the fused panels of a group's later steps apply
// up to group - 1 earlier steps), two otherwise.
inline int lu_group(const lsmsx_ctx *c, bool split_f16) { return split_f16 ? c->split_group : 2; }

// Launch the blocked LU over a chunk of nb solves already in solver layout
// (each descriptor's group = lu_group(c, split_f16)): steps taken in groups
// g .. g+G-1 -- panels of g, then for each later step t of the group a cross
// update of tile row/column t with the group's earlier steps and the panels
// of t (on the split-FP16 path the cross update is fused into the panel
// kernel, and only the diagonal tile's has a kernel of its own) -- then one
// K = 64 G trailing update, so the trailing matrix is read and written once
// per group.
template <class T>
int launch_lu(lsmsx_ctx *c, const SolveDesc *dd, int nb, int nt_max, bool split_f16 = false,
              hipStream_t st = nullptr) {
  if (!st) st = c->stream;
  const int swz = (c->sched & LSMSX_SCHED_XCD_SWIZZLE) ? 1 : 0;
  const int G = lu_group(c, split_f16);
  auto diag_panel = [&](int s) -> int {
    const int nrest = nt_max - 1 - s;
    if (nrest <= 0) return 0;
    {
      ProfScope p(c, 1);
      hipLaunchKernelGGL(k_lu_diag<T>, dim3(nb), dim3(256), 0, st, dd, s);
    }
    {
      ProfScope p(c, 2);
      hipLaunchKernelGGL(k_lu_panel<T>, dim3(2 * nrest, nb), dim3(256), 0, st, dd, s);
    }
    return check_launch("k_lu_diag/k_lu_panel");
  };
  // the steps s .. s+n-1 of a group: n <= H, panels and crosses step by
  // step; otherwise the first half, the strip update of the second half's
  // columns and rows with the first half, then the second half
  if (split_f16) {
    // split-FP16 path: for each step t of a group, the cross update of the
    // diagonal tile (steps g .. t-1), its factorization, and the panel with
    // the cross update of its column and row fused in; then the trailing
    // update with the whole group
    for (int s = 0; s < nt_max - 1; s += G) {
      const int b = std::min(s + G - 1, nt_max - 2);
      for (int t = s; t <= b; t++) {
        const int u = t - s;
        if (u > 0) {
          ProfScope p(c, 9);
          hipLaunchKernelGGL(k_lu_cross_diag_h, dim3(nb), dim3(128), 0, st, dd, s, u);
        }
        if (int r = check_launch("k_lu_cross_diag_h")) return r;
        {
          ProfScope p(c, 1);
          hipLaunchKernelGGL(k_lu_diag<T>, dim3(nb), dim3(256), 0, st, dd, t);
        }
        {
          ProfScope p(c, 2);
          hipLaunchKernelGGL(k_lu_panel_h, dim3(2 * (nt_max - 1 - t), nb), dim3(256), 0, st, dd,
                             t, u, swz);
        }
        if (int r = check_launch("k_lu_diag/k_lu_panel_h")) return r;
      }
      const int nrest = nt_max - 1 - b;
      if (nrest > 0) {
        ProfScope p(c, 3);
        hipLaunchKernelGGL(k_lu_trail_h, dim3(((nrest + 1) / 2) * ((nrest + 1) / 2), nb),
                           dim3(512), 0, st, dd, s, swz);
      }
      if (int r = check_launch("k_lu_trail_h")) return r;
    }
    {
      ProfScope p(c, 4);
      hipLaunchKernelGGL(k_lu_last<T>, dim3(nb), dim3(256), 0, st, dd);
    }
    return check_launch("k_lu_last");
  }
  // steps s .. s+n-1 of a group: panels and crosses step by step
  auto block = [&](int s, int n) -> int {
    if (int r = diag_panel(s)) return r;
    for (int u = 1; u < n && s + u < nt_max - 1; u++) {
      {
        ProfScope p(c, 9);
        const int n1 = nt_max - s - u;
        hipLaunchKernelGGL(k_lu_cross<T>, dim3(2 * n1 - 1, nb), dim3(256), 0, st, dd, s, u);
      }
      if (int r = check_launch("k_lu_cross")) return r;
      if (int r = diag_panel(s + u)) return r;
    }
    return 0;
  };
  for (int s = 0; s < nt_max - 1; s += G) {
    if (int r = block(s, G)) return r;
    const int last = std::min(s + G - 1, nt_max - 2);  // the group's last step, largest solve
    const int nrest = nt_max - 1 - last;
    if (nrest > 0) {
      ProfScope p(c, 3);
      hipLaunchKernelGGL(k_lu_trail<T>, dim3(nrest * nrest, nb), dim3(256), 0, st, dd, s,
                         swz);
    }
    if (int r = check_launch("k_lu_trail")) return r;
  }
  {
    ProfScope p(c, 4);
    hipLaunchKernelGGL(k_lu_last<T>, dim3(nb), dim3(256), 0, st, dd);
  }
  return check_launch("k_lu_last");
}

// Refinement sweeps, one launch per tile column (k_ir_step), on a chunk whose
// factors (T = float: mixed path with 32 right-hand sides; T = double:
// FP64-refined path) are in place.  Returns the number of
// unconverged solves left in *left.
template <class T>
int refine(lsmsx_ctx *c, const SolveDesc *dd, int nb, int nt_max, int kc, int max_sweeps,
           double tol, int *left) {
  hipLaunchKernelGGL(k_ir_init<T>, dim3(nb), dim3(256), 0, c->stream, dd);
  if (int r = check_launch("k_ir_init")) return r;
  int done = 0;
  for (int it = 0; it < max_sweeps; it++) {
    if (it > 0) {  // k_ir_init set the first residual, B
      ProfScope p(c, 5);
      if (kc == 32)
        hipLaunchKernelGGL((k_ir_residual<T, 2>), dim3(nt_max, nb), dim3(256), 0, c->stream, dd);
      else
        hipLaunchKernelGGL((k_ir_residual<T, 1>), dim3(nt_max, nb), dim3(256), 0, c->stream, dd);
    }
    if (int r = check_launch("k_ir_residual")) return r;
    {
      ProfScope p(c, 6);
      if (it == 0) {  // R = B: only the last tile of each solve (kIrFirstSweep)
        hipLaunchKernelGGL((k_ir_step<T, true>), dim3(1, nb), dim3(256), 0, c->stream, dd,
                           kIrFirstSweep);
      } else {
        for (int s = -1; s <= nt_max - 2; s++)
          hipLaunchKernelGGL((k_ir_step<T, true>), dim3(s < 0 ? 1 : nt_max - 1 - s, nb), dim3(256),
                             0, c->stream, dd, s);
      }
      // the central rows are final after the forward steps: test convergence
      // before the backward steps, which converged solves then skip
      hipLaunchKernelGGL(k_ir_accum<T>, dim3(1, nb), dim3(256), 0, c->stream, dd, 1);
      hipLaunchKernelGGL(k_ir_decide, dim3((nb + 255) / 256), dim3(256), 0, c->stream, dd, nb, tol);
      for (int j = 0; j <= nt_max - 2; j++)
        hipLaunchKernelGGL((k_ir_step<T, false>), dim3(nt_max - 1 - j, nb), dim3(256), 0, c->stream,
                           dd, j);
      if (nt_max > 1)
        hipLaunchKernelGGL(k_ir_accum<T>, dim3(nt_max - 1, nb), dim3(256), 0, c->stream, dd, 0);
    }
    if (int r = check_launch("k_ir_step")) return r;
    if (it >= 1) {  // the first correction never meets the tolerance
      hipLaunchKernelGGL(k_count_conv, dim3(1), dim3(256), 0, c->stream, dd, nb, c->count.p);
      HIPCHK(hipMemcpyAsync(&done, c->count.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
      if (done == nb) break;
    }
  }
  hipLaunchKernelGGL(k_ir_extract, dim3(nb), dim3(256), 0, c->stream, dd);
  if (int r = check_launch("k_ir_extract")) return r;
  hipLaunchKernelGGL(k_count_conv, dim3(1), dim3(256), 0, c->stream, dd, nb, c->count.p);
  HIPCHK(hipMemcpyAsync(&done, c->count.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  *left = nb - done;
  return 0;
}

// Refinement of a chunk part on stream st with no host round trip (mixed
// path, kc = 16, k_ir_sweep): max_sweeps sweeps are enqueued up front; the
// residual and sweep kernels return at once for solves already converged,
// so a solve's arithmetic and sweep count are those of refine<float>.  The
// part's refinement can then run beside another part's factorization.
int refine_async(lsmsx_ctx *c, const SolveDesc *dd, int nb, int nt_max, int max_sweeps,
                 double tol, hipStream_t st, int nbands_max) {
  hipLaunchKernelGGL(k_ir_init<float>, dim3(nb), dim3(256), 0, st, dd);
  for (int it = 0; it < max_sweeps; it++) {
    if (it > 0) {
      ProfScope p(c, 5);
      if (nbands_max > 0) {  // structure constants once per site pair
        hipLaunchKernelGGL(k_ir_residual_g, dim3((nbands_max + 1) / 2, nb), dim3(512), 0, st, dd);
        hipLaunchKernelGGL(k_ir_residual_g_finish, dim3(nbands_max, nb), dim3(512), 0, st, dd);
      } else {
        hipLaunchKernelGGL((k_ir_residual<float, 1>), dim3(nt_max, nb), dim3(256), 0, st, dd);
      }
    }
    ProfScope p(c, 6);
    if (it == 0)
      hipLaunchKernelGGL(k_ir_sweep<true>, dim3(nb), dim3(kSweepWaves * 64), 0, st, dd, tol);
    else
      hipLaunchKernelGGL(k_ir_sweep<false>, dim3(nb), dim3(kSweepWaves * 64), 0, st, dd, tol);
  }
  hipLaunchKernelGGL(k_ir_extract, dim3(nb), dim3(256), 0, st, dd);
  return check_launch("refine_async");
}

// The (l m s) -> (kappa mu) Clebsch-Gordan matrix of clebsch.f as a dense
// kk2 x kk2 column-major U, kk2 = 2 (lmax+1)^2: kappa-mu row ir (0-based
// 2 kappa^2 + kappa + m - 1, and the next row for spin up), column s kkr + lm
// with s = 0 for the spin-down coefficient u1 and s = 1 for the spin-up u2,
// so that relmtrx(G) = U diag_spin(G) U^T.
std::vector<double> clebsch_u(int lmax) {
  const int kkr = (lmax + 1) * (lmax + 1), kk2 = 2 * kkr;
  std::vector<double> U((size_t)kk2 * kk2, 0.0);
  auto at = [&](int r, int col) -> double & { return U[r + (size_t)kk2 * col]; };
  for (int l = 0; l <= lmax; l++) {
    const double d = 2.0 * l + 1.0;
    for (int m = -l; m <= l; m++) {
      const int lm = l * l + l + m;
      if (l > 0) {  // j = l - 1/2, kappa = l
        const int r = 2 * l * l + l + m - 1;
        at(r, lm) = std::sqrt((l + m) / d);
        at(r + 1, kkr + lm) = -std::sqrt((l - m) / d);
      }
      const int kap = -l - 1;  // j = l + 1/2
      const int r = 2 * kap * kap + kap + m - 1;
      at(r, lm) = std::sqrt((l - m + 1) / d);
      at(r + 1, kkr + lm) = std::sqrt((l + m + 1) / d);
    }
  }
  return U;
}

int validate_common(lsmsx_ctx *c, int nsc, int nspola) {
  if (!c) return fail(-1, "null context");
  if (c->lmax_tab < 0) return fail(-2, "tables not set");
  if (!c->tmat.p) return fail(-3, "t-matrix store not uploaded");
  if (nsc != 1 && nsc != 2) return fail(-4, "n_spin_cant must be 1 or 2");
  if (nspola != 1 && nspola != 2) return fail(-4, "n_spin_pola must be 1 or 2");
  if (nspola < nsc) return fail(-4, "n_spin_pola < n_spin_cant");
  return 0;
}

// Batched solve of nsolve (atom, spin, slot) problems; results to host.
//   FP64            tile-LU in place, tau00 from the Schur complement; solves
//                   whose smallest tile pivot is below the pivot guard are
//                   redone with FP64_REFINED;
//   MIXED_FP32/F16  FP32 factors of an FP32 copy + refinement against M;
//                   unconverged solves are redone with FP64_REFINED;
//   FP64_REFINED    FP64 factors of an FP64 copy + refinement against M;
//                   unconverged solves fall back to the plain FP64 result.
int assemble_reference_layout(lsmsx_ctx *c, AtomDev &a, int ispin, int iie, const double *prel,
                              int nsc, int nspola, double2 *dst, SolveDesc *ddev);
int column_pivoted_fallback(lsmsx_ctx *c, AtomDev &a, int ispin, int iie, const double *prel,
                            int nsc, int nspola, double *tau, int *info);
// last_iters value of a solve whose result comes from the column-pivoted fallback
constexpr int kItersColumnPivoted = -2;

int run_batched(lsmsx_ctx *c, int nsolve, const int *atom_ids, const int *ispins,
                const int *iies, const double *prel, int nsc, int nspola, int precision,
                double *tau_out, int *info_out) {
  if (int r = validate_common(c, nsc, nspola)) return r;
  if (precision != LSMSX_PREC_FP64 && precision != LSMSX_PREC_MIXED_FP32 &&
      precision != LSMSX_PREC_MIXED_F16 && precision != LSMSX_PREC_FP64_REFINED)
    return fail(-8, "unsupported precision mode");
  const bool plain = precision == LSMSX_PREC_FP64;
  const bool refined64 = precision == LSMSX_PREC_FP64_REFINED;
  const bool split = precision == LSMSX_PREC_MIXED_F16;
  const bool mixed = precision == LSMSX_PREC_MIXED_FP32 || split;
  const bool refining = mixed || refined64;
  if (nsolve <= 0) return 0;
  if (c->rel && (nsc != 2 || nspola != 2))
    return fail(-4, "the relativistic store needs n_spin_cant = n_spin_pola = 2");
  HIPCHK(hipSetDevice(c->device));
  c->last_iters.assign(nsolve, 0);
  c->last_minpiv.assign(nsolve, 0.0);
  // per-solve sizes (bytes of all per-solve device buffers)
  std::vector<long> bytes(nsolve);
  int kc = -1;
  // mixed path with 16 right-hand sides, collinear, every site at lmax 3 and
  // at most kSweepMaxTiles tiles: no FP64 M, the structure constants once per
  // site pair (the residual forms B - X + t G X)
  bool gmode = mixed && nsc == 1 && c->residual_form != LSMSX_RESIDUAL_MATRIX;
  for (int i = 0; i < nsolve && gmode; i++) {
    auto it = c->atoms.find(atom_ids[i]);
    if (it == c->atoms.end()) break;  // reported below
    const AtomDev &a = it->second;
    if (a.kkrsz != 16 || roundup(a.nrmat, kTile) / kTile > kSweepMaxTiles) gmode = false;
    for (int s2 = 0; s2 < a.numLIZ && gmode; s2++)
      if (a.lmax[s2] != 3) gmode = false;
  }
  auto g_blocks = [](int n) { return (long)n * (n - 1) / 2; };
  auto g_slots = [](int n) {
    const long nb = (n + kGBand - 1) / kGBand;
    return nb * (nb + 1) / 2 * kGBand;
  };
  for (int i = 0; i < nsolve; i++) {
    auto it = c->atoms.find(atom_ids[i]);
    if (it == c->atoms.end()) return fail(-5, "unknown atom id " + std::to_string(atom_ids[i]));
    const AtomDev &a = it->second;
    int kci = nsc * a.kkrsz;
    if (kc < 0) kc = kci;
    if (kci != kc) return fail(-6, "all atoms of a batch must share kkrsz");
    if (kc > kTile) return fail(-6, "kkrsz_ns > 64 not supported");
    if (refining && kc != 16 && kc != 32)
      return fail(-6, "refined precision modes need kkrsz_ns = 16 or 32");
    if (a.kkrsz > (c->lmax_tab + 1) * (c->lmax_tab + 1))
      return fail(-6, "atom lmax exceeds the lmax of the uploaded tables");
    if (c->rel) {
      if (2 * a.kkrsz != c->rel_kk2) return fail(-6, "atom lmax differs from the relativistic store's");
      for (int s = 0; s < a.numLIZ; s++)
        if (a.lmax[s] != a.lmax[0])
          return fail(-6, "relativistic blocks need every LIZ site at the central lmax");
    }
    if (iies[i] < 0 || iies[i] >= c->group) return fail(-7, "energy slot out of range");
    if (ispins[i] < 0 || ispins[i] >= nspola) return fail(-7, "spin index out of range");
    for (int s = 0; s < a.numLIZ; s++)
      if (a.store[s] < 0 || a.store[s] >= c->nstore) return fail(-7, "LIZStoreIdx out of range");
    long Ns = roundup(nsc * a.nrmat, kTile), nt = Ns / kTile;
    bytes[i] = gmode ? (g_blocks(a.numLIZ) + a.numLIZ + g_slots(a.numLIZ)) * 256 * 16
                     : Ns * Ns * 16;
    if (plain) bytes[i] += kTile * kTile * 16;
    if (mixed) bytes[i] += Ns * Ns * 8 + nt * kTile * kTile * 8 + Ns * kc * 24;
    if (refined64) bytes[i] += Ns * Ns * 16 + nt * kTile * kTile * 16 + Ns * kc * 32;
    if (split) bytes[i] += (2 * lu_group(c, true) * nt + 2) * ((long)kPlaneHalves * 2 + 4);
  }
  // the explicit-inverse route's buffers are not used here: give them back
  // before sizing the chunks by the free memory
  c->Minv.release();
  c->gblk.release();
  long cap = c->ws_limit;
  if (cap <= 0) {
    size_t fr = 0, tot = 0;
    HIPCHK(hipMemGetInfo(&fr, &tot));
    long held = (long)((c->M.n + c->inv.n + c->A64.n + c->GH.n + c->YB.n + c->GP.n) *
                           sizeof(double2) +
                       c->W.n * sizeof(float2));
    cap = (long)((double)(fr + held) * 0.7);
  }
  int done = 0;
  while (done < nsolve) {
    // chunk [done, end)
    long acc = 0;
    int end = done;
    while (end < nsolve && (end == done || acc + bytes[end] <= cap)) acc += bytes[end++];
    int nb = end - done;
    long mtot = 0, xtot = 0, ivtot = 0, htot = 0, ghtot = 0, ybtot = 0, gptot = 0;
    int maxliz = 1, nt_max = 1, nbands_max = 0;
    for (int b = 0; b < nb; b++) {
      const AtomDev &a = c->atoms[atom_ids[done + b]];
      long Ns = roundup(nsc * a.nrmat, kTile);
      mtot += Ns * Ns;
      xtot += Ns * kc;
      ivtot += refining ? (Ns / kTile) * kTile * kTile : kTile * kTile;
      htot += 2 * lu_group(c, split) * (Ns / kTile) + 2;
      maxliz = std::max(maxliz, a.numLIZ);
      nt_max = std::max(nt_max, (int)(Ns / kTile));
      if (gmode) {
        ghtot += g_blocks(a.numLIZ) * 256;
        ybtot += (long)a.numLIZ * 256;
        gptot += g_slots(a.numLIZ) * 256;
        nbands_max = std::max(nbands_max, (a.numLIZ + kGBand - 1) / kGBand);
      }
    }
    if (gmode) {
      if (int r = c->GH.ensure(std::max(ghtot, 1L))) return r;
      if (int r = c->YB.ensure(ybtot)) return r;
      if (int r = c->GP.ensure(gptot)) return r;
    } else if (int r = c->M.ensure(mtot)) {
      return r;
    }
    // inv holds double2 elements; FP32 inverses take half as many
    if (int r = c->inv.ensure(mixed ? (ivtot + 1) / 2 : ivtot)) return r;
    if (int r = c->perm.ensure((size_t)nb * nt_max * kTile)) return r;
    if (int r = c->info.ensure((size_t)nb)) return r;
    if (int r = c->minpiv.ensure((size_t)nb)) return r;
    if (int r = c->tau.ensure((size_t)nb * kc * kc)) return r;
    if (int r = c->desc.ensure((size_t)nb)) return r;
    if (mixed) {
      if (int r = c->W.ensure(mtot)) return r;
      if (int r = c->R.ensure(xtot)) return r;
    }
    if (refined64) {
      if (int r = c->A64.ensure(mtot)) return r;
      if (int r = c->R64.ensure(xtot)) return r;
    }
    if (refining) {
      if (int r = c->X.ensure(xtot)) return r;
      if (int r = c->conv.ensure((size_t)nb)) return r;
      if (int r = c->iters.ensure((size_t)nb)) return r;
      if (int r = c->dxn.ensure((size_t)4 * nb)) return r;
      if (int r = c->count.ensure(1)) return r;
    }
    if (split) {
      if (int r = c->hp.ensure((size_t)htot * kPlaneHalves)) return r;
      if (int r = c->he.ensure((size_t)htot)) return r;
    }
    std::vector<SolveDesc> hd(nb);
    long moff = 0, xoff = 0, ioff = 0, hoff = 0, ghoff = 0, yboff = 0, gpoff = 0;
    for (int b = 0; b < nb; b++) {
      int i = done + b;
      AtomDev &a = c->atoms[atom_ids[i]];
      SolveDesc &d = hd[b];
      d = SolveDesc{};
      int N = nsc * a.nrmat, Ns = roundup(N, kTile);
      if (split) {
        d.hp = c->hp.p + hoff * kPlaneHalves;
        d.he = c->he.p + hoff;
        hoff += 2 * lu_group(c, split) * (Ns / kTile);
        d.ihp = c->hp.p + hoff * kPlaneHalves;
        d.ihe = c->he.p + hoff;
        hoff += 2;
      }
      d.pos = a.d_pos;
      d.lmax = a.d_lmax;
      d.store = a.d_store;
      d.site_off = a.d_off[1][nsc - 1];
      d.numLIZ = a.numLIZ;
      d.n_spin_cant = nsc;
      d.kkrsz = a.kkrsz;
      d.kc = kc;
      d.prel_re = prel[2 * i];
      d.prel_im = prel[2 * i + 1];
      long jsm = (nspola != nsc) ? (long)kc * kc * ispins[i] : 0;
      d.t_off = (long)iies[i] * c->blk + jsm;
      d.ldt = kc;
      d.M = gmode ? nullptr : c->M.p + moff;
      if (gmode) {
        d.gh = c->GH.p + ghoff;
        d.ybuf = c->YB.p + yboff;
        d.gpart = c->GP.p + gpoff;
        ghoff += g_blocks(a.numLIZ) * 256;
        yboff += (long)a.numLIZ * 256;
        gpoff += g_slots(a.numLIZ) * 256;
      }
      d.tmat = c->tmat.p;
      d.ldT = c->ldT;
      d.ld = Ns;
      d.Ns = Ns;
      d.pad = Ns - N;
      d.nt = Ns / kTile;
      // right alignment starts each solve late enough to end with the
      // largest; the start is a multiple of the group size so that every
      // solve groups its own steps (0..G-1), (G..2G-1), ... whatever its
      // size: the arithmetic of a solve never depends on the schedule or the
      // batch
      d.group = lu_group(c, split);
      d.step0 = (c->sched & LSMSX_SCHED_RIGHT_ALIGN) ? ((nt_max - d.nt) & ~(d.group - 1)) : 0;
      d.perm = c->perm.p + (long)b * nt_max * kTile;
      d.info = c->info.p + b;
      d.minpiv = c->minpiv.p + b;
      d.tau00 = c->tau.p + (long)b * kc * kc;
      d.t0 = c->tmat.p + c->ldT * a.store[0] + d.t_off;
      if (refining) {
        d.inv_stride = kTile * kTile;
        d.X = c->X.p + xoff;
        d.conv = c->conv.p + b;
        d.iters = c->iters.p + b;
        d.dxn = c->dxn.p + b;
        d.dxn_rho2 = c->dxn.p + nb + b;
        d.nrm = c->dxn.p + 2 * nb + 2 * b;
      }
      if (mixed) {
        d.A = c->W.p + moff;
        d.tiled = 1;  // the assembly writes the FP32 copy tile-major
        d.w32 = c->W.p + moff;
        d.R = c->R.p + xoff;
        d.inv = (float2 *)c->inv.p + ioff;
      } else if (refined64) {
        d.A = c->A64.p + moff;
        d.R = c->R64.p + xoff;
        d.inv = c->inv.p + ioff;
      } else {
        d.A = d.M;
        d.inv = c->inv.p + (long)b * kTile * kTile;
        d.inv_stride = 0;
      }
      if (refining) ioff += (long)d.nt * kTile * kTile;
      xoff += (long)Ns * kc;
      moff += (long)Ns * Ns;
    }
    // Structure constants shared by the solves of the chunk that have the same
    // LIZ geometry (bitwise equal positions and lmax), energy slot and prel:
    // the first solve of each such group evaluates them once
    // (k_assemble<1, kAsmG>) and every solve builds its blocks from them
    // (k_assemble_from_g) -- the same FP64 values, so the same blocks.
    // Collinear, non-relativistic, and only when some group has two solves.
    std::vector<SolveDesc> own;
    int own_liz = 1;
    if (c->share_g && nsc == 1 && !c->rel) {
      std::map<std::tuple<int, int, long, long>, int> first;
      std::vector<int> owner(nb);
      bool any = false;
      for (int b = 0; b < nb; b++) {
        const int i = done + b;
        long pr = 0, pi = 0;
        std::memcpy(&pr, &prel[2 * i], sizeof(pr));
        std::memcpy(&pi, &prel[2 * i + 1], sizeof(pi));
        auto key = std::make_tuple(c->atoms[atom_ids[i]].geo, iies[i], pr, pi);
        auto it = first.find(key);
        if (it == first.end()) {
          first.emplace(key, b);
          owner[b] = b;
        } else {
          owner[b] = it->second;
          any = true;
        }
      }
      if (any) {
        if (!gmode) {  // G storage for the owners (the mixed path with G keeps its own)
          long gs = 0;
          for (int b = 0; b < nb; b++)
            if (owner[b] == b) gs += g_blocks(hd[b].numLIZ) * 256;
          if (int r = c->GS.ensure(std::max(gs, 1L))) return r;
          long go = 0;
          for (int b = 0; b < nb; b++)
            if (owner[b] == b) {
              hd[b].gh = c->GS.p + go;
              go += g_blocks(hd[b].numLIZ) * 256;
            }
        }
        for (int b = 0; b < nb; b++) {
          if (owner[b] == b) {
            own.push_back(hd[b]);
            own_liz = std::max(own_liz, hd[b].numLIZ);
          } else {
            hd[b].gh = hd[owner[b]].gh;
          }
        }
        if (int r = c->desc_own.ensure(own.size())) return r;
        HIPCHK(hipMemcpyAsync(c->desc_own.p, own.data(), own.size() * sizeof(SolveDesc),
                              hipMemcpyHostToDevice, c->stream));
      }
    }
    const bool share = !own.empty();
    const int asm_mode = gmode ? kAsmGW : (mixed ? kAsmMW : kAsmM);
    // two halves on two streams (see enqueue below)
    const bool two = (c->sched & LSMSX_SCHED_TWO_STREAMS) && nb >= 2 && !c->profiling;
    if (share) {
      // the chunk's solves run with the solves of a group next to each other
      // (each descriptor keeps its own output slots), so that a group's
      // stored structure constants -- read by the block builder and by every
      // refinement residual -- come from the caches once its first solve has
      // read them
      std::stable_sort(hd.begin(), hd.end(),
                       [](const SolveDesc &x, const SolveDesc &y) { return x.gh < y.gh; });
    }
    // position in hd of the chunk's solve b (its output slots are index b)
    std::vector<int> pos_of(nb);
    for (int k = 0; k < nb; k++) pos_of[hd[k].info - c->info.p] = k;
    HIPCHK(hipMemcpyAsync(c->desc.p, hd.data(), nb * sizeof(SolveDesc), hipMemcpyHostToDevice,
                          c->stream));
    // assembly + factorization of the chunk: launched eagerly, or captured
    // once per chunk shape into a hipGraph and replayed (the kernels read
    // the freshly uploaded descriptors, so only the shape and the buffer
    // addresses passed as arguments key the graph)
    // solves [b0, b0 + n) of the chunk on stream st; after_asm, if set, is
    // recorded once the assembly is enqueued
    // the mixed path with 16 right-hand sides refines each part on its
    // stream right after its factorization (refine_async)
    const bool ir_inline = mixed && kc == 16 && nt_max <= kSweepMaxTiles;
    auto enqueue_part = [&](hipStream_t st, int b0, int n, hipEvent_t after_asm) -> int {
      const SolveDesc *dd = c->desc.p + b0;
      // the step range is the chunk's: right-aligned starts are relative to it
      const int ntm = nt_max;
      int mliz = 1;
      for (int b = b0; b < b0 + n; b++) mliz = std::max(mliz, hd[b].numLIZ);
      hipLaunchKernelGGL(k_init_status, dim3((n + 255) / 256), dim3(256), 0, st, dd, n);
      if (int r = check_launch("k_init_status")) return r;
      {
        ProfScope p(c, 0);
        hipLaunchKernelGGL(k_pad_identity, dim3(n), dim3(256), 0, st, dd);
        if (share)
          launch_assemble_from_g(c, dim3(mliz, n), st, dd, asm_mode);
        else
          launch_assemble(c, nsc, dim3(mliz, n), st, dd, asm_mode);
      }
      if (int r = check_launch("k_assemble")) return r;
      if (after_asm) HIPCHK(hipEventRecord(after_asm, st));
      if (refined64) {
        hipLaunchKernelGGL(k_copy_m, dim3(ntm, n), dim3(256), 0, st, dd);
        if (int r = check_launch("k_copy_m")) return r;
      }
      if (!mixed) return launch_lu<double>(c, dd, n, ntm, false, st);
      if (int r = launch_lu<float>(c, dd, n, ntm, split, st)) return r;
      if (ir_inline)
        return refine_async(c, dd, n, ntm, c->ir_max_sweeps, c->ir_tol, st, gmode ? nbands_max : 0);
      return 0;
    };
    // two halves on two streams, the second started once the first half's
    // assembly is done: its assembly (FP64 VALU work) runs beside the first
    // half's factorization (MFMA and HBM work), and one half's latency-bound
    // steps (diagonal tiles, panels) overlap the other half's trailing updates
    auto enqueue = [&]() -> int {
      if (share) {  // the shared structure constants, before either part
        ProfScope p(c, 0);
        launch_assemble(c, 1, dim3(own_liz, (int)own.size()), c->stream, c->desc_own.p, kAsmG);
        if (int r = check_launch("k_assemble (structure constants)")) return r;
      }
      if (!two) return enqueue_part(c->stream, 0, nb, nullptr);
      const int h = nb / 2;
      if (int r = enqueue_part(c->stream, 0, h, c->ev_asm)) return r;
      HIPCHK(hipStreamWaitEvent(c->stream2, c->ev_asm, 0));
      if (int r = enqueue_part(c->stream2, h, nb - h, nullptr)) return r;
      HIPCHK(hipEventRecord(c->ev_join, c->stream2));
      HIPCHK(hipStreamWaitEvent(c->stream, c->ev_join, 0));
      return 0;
    };
    if (c->use_graph) {
      long tol_bits = 0;
      std::memcpy(&tol_bits, &c->ir_tol, sizeof(tol_bits));
      std::vector<long> key = {nb, nt_max, maxliz, precision, c->sched, (long)c->desc.p,
                               (long)c->tmat.p, c->ldT, (long)c->cgnt.p, (long)c->illp.p,
                               c->lmax_tab, c->ir_max_sweeps, tol_bits, gmode ? nbands_max : -1,
                               lu_group(c, split), share ? (long)own.size() : -1,
                               (long)c->desc_own.p};
      auto it = c->graphs.find(key);
      if (it == c->graphs.end()) {
        const bool prof = c->profiling;
        c->profiling = false;  // no event records inside the captured sequence
        HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
        int r = enqueue();
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        c->profiling = prof;
        if (r) {
          if (g) (void)hipGraphDestroy(g);
          return r;
        }
        if (e != hipSuccess) return fail(-100 - (int)e, std::string("graph capture: ") + hipGetErrorString(e));
        hipGraphExec_t ex = nullptr;
        e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return fail(-100 - (int)e, std::string("graph instantiate: ") + hipGetErrorString(e));
        it = c->graphs.emplace(key, ex).first;
      }
      HIPCHK(hipGraphLaunch(it->second, c->stream));
    } else {
      if (int r = enqueue()) return r;
    }
    std::vector<int> hiters(nb, 0);
    std::vector<char> unconverged64(nb, 0);
    if (refining) {
      int left = 0;
      int r = ir_inline ? 0
              : mixed   ? refine<float>(c, c->desc.p, nb, nt_max, kc, c->ir_max_sweeps, c->ir_tol,
                                        &left)
                        : refine<double>(c, c->desc.p, nb, nt_max, kc, c->ir_max_sweeps,
                                         c->ir_tol, &left);
      if (r) return r;
      std::vector<int> hconv(nb);
      HIPCHK(hipMemcpyAsync(hiters.data(), c->iters.p, nb * sizeof(int), hipMemcpyDeviceToHost,
                            c->stream));
      HIPCHK(hipMemcpyAsync(hconv.data(), c->conv.p, nb * sizeof(int), hipMemcpyDeviceToHost,
                            c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
      for (int b = 0; b < nb; b++)
        if (!hconv[b]) {
          hiters[b] = -1;
          if (refined64) unconverged64[b] = 1;
        }
      if (refined64 && left > 0) {
        // last resort for a refinement that did not converge: the plain
        // FP64 result from the untouched M
        std::vector<SolveDesc> fb;
        for (int b = 0; b < nb; b++)
          if (!hconv[b]) {
            SolveDesc d = hd[pos_of[b]];
            d.step0 = 0;  // the fallback chunk has its own step count
            d.A = d.M;
            d.tiled = 0;
            d.X = nullptr;
            d.conv = nullptr;
            d.inv = c->inv.p + (long)fb.size() * kTile * kTile;
            d.inv_stride = 0;
            fb.push_back(d);
          }
        int nf = (int)fb.size();
        int fnt = 1;
        for (auto &d : fb) fnt = std::max(fnt, d.nt);
        if (int r2 = c->desc2.ensure((size_t)nf)) return r2;
        HIPCHK(hipMemcpyAsync(c->desc2.p, fb.data(), nf * sizeof(SolveDesc),
                              hipMemcpyHostToDevice, c->stream));
        if (int r2 = launch_lu<double>(c, c->desc2.p, nf, fnt)) return r2;
      }
    }
    if (c->rel) {  // kappa-mu tau00 = (E/prel^2) U tau00' U^T
      std::vector<double2> fi(nb);
      for (int b = 0; b < nb; b++) fi[b] = c->rel_finv[iies[done + b]];
      if (int r = c->relf.ensure((size_t)nb)) return r;
      HIPCHK(hipMemcpyAsync(c->relf.p, fi.data(), nb * sizeof(double2), hipMemcpyHostToDevice,
                            c->stream));
      hipLaunchKernelGGL(k_rel_tau, dim3(nb), dim3(256), 0, c->stream, c->tau.p, kc, c->relU.p,
                         c->relf.p);
      if (int r = check_launch("k_rel_tau")) return r;
      HIPCHK(hipStreamSynchronize(c->stream));  // fi is read by the copy above
    }
    HIPCHK(hipMemcpyAsync(tau_out + (long)done * kc * kc * 2, c->tau.p,
                          (size_t)nb * kc * kc * sizeof(double2), hipMemcpyDeviceToHost, c->stream));
    std::vector<int> hinfo(nb);
    std::vector<double> hmin(nb);
    HIPCHK(hipMemcpyAsync(hinfo.data(), c->info.p, nb * sizeof(int), hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipMemcpyAsync(hmin.data(), c->minpiv.p, nb * sizeof(double), hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    drain_profile(c);
    for (int b = 0; b < nb; b++) {
      // a refinement that did not converge returns the plain FP64 result,
      // flagged so that callers can tell it from a refined one
      if (unconverged64[b] && hinfo[b] == 0) hinfo[b] = kInfoRefineUnconverged;
      info_out[done + b] = hinfo[b];
      c->last_iters[done + b] = hiters[b];
      c->last_minpiv[done + b] = hmin[b];
    }
    done = end;
  }
  // Redo with FP64 factors + FP64 refinement: plain solves whose smallest
  // tile-local pivot fell under the guard (element growth can cost digits),
  // and mixed solves whose refinement did not converge.
  std::vector<int> redo;
  for (int i = 0; i < nsolve; i++) {
    if (plain && c->pivot_guard > 0.0 && info_out[i] == 0 && c->last_minpiv[i] < c->pivot_guard)
      redo.push_back(i);
    if (mixed && c->last_iters[i] < 0) redo.push_back(i);
  }
  if (!redo.empty()) {
    const int nr = (int)redo.size();
    std::vector<int> ra(nr), rs(nr), re(nr), rinfo(nr);
    std::vector<double> rp(2 * nr), rtau((size_t)nr * kc * kc * 2);
    for (int j = 0; j < nr; j++) {
      const int i = redo[j];
      ra[j] = atom_ids[i];
      rs[j] = ispins[i];
      re[j] = iies[i];
      rp[2 * j] = prel[2 * i];
      rp[2 * j + 1] = prel[2 * i + 1];
    }
    std::vector<int> outer_iters = c->last_iters;
    std::vector<double> outer_minpiv = c->last_minpiv;
    if (int r = run_batched(c, nr, ra.data(), rs.data(), re.data(), rp.data(), nsc, nspola,
                            LSMSX_PREC_FP64_REFINED, rtau.data(), rinfo.data()))
      return r;
    for (int j = 0; j < nr; j++) {
      const int i = redo[j];
      std::memcpy(tau_out + (long)i * kc * kc * 2, rtau.data() + (long)j * kc * kc * 2,
                  sizeof(double) * 2 * kc * kc);
      info_out[i] = rinfo[j];
      // sweeps of the FP64 refinement (plain solves); -1 marks a mixed
      // solve redone in FP64, -2 one taken by the column-pivoted fallback
      outer_iters[i] = c->last_iters[j] == kItersColumnPivoted ? kItersColumnPivoted
                       : plain                                 ? c->last_iters[j]
                                                               : -1;
      outer_minpiv[i] = c->last_minpiv[j];
    }
    c->last_iters = outer_iters;
    c->last_minpiv = outer_minpiv;
  }
  // Column-pivoted fallback (rocSOLVER on the reference-layout FP64 M) for
  // the solves that are still singular under tile-local pivoting or whose
  // FP64 refinement did not converge (relativistic stores excluded: their
  // tau00 is post-processed in the kappa-mu basis)
  if (!c->rel) {
    for (int i = 0; i < nsolve; i++) {
      if (!(info_out[i] > 0 || info_out[i] == kInfoRefineUnconverged)) continue;
      AtomDev &a = c->atoms[atom_ids[i]];
      if (nsc * a.kkrsz != kc) continue;
      int finfo = 0;
      if (int r = column_pivoted_fallback(c, a, ispins[i], iies[i], prel + 2 * i, nsc, nspola,
                                          tau_out + (long)i * kc * kc * 2, &finfo))
        return r;
      info_out[i] = finfo;
      c->last_iters[i] = kItersColumnPivoted;
    }
  }
  return 0;
}

// LU of every tile (last one included, all diagonal-tile inverses kept) and
// the in-place explicit inverse M^-1 in d.Minv, for nb solves in solver layout.
int launch_inverse(lsmsx_ctx *c, const SolveDesc *dd, int nb, int nt_max) {
  for (int s = 0; s < nt_max; s++) {
    int nrest = nt_max - 1 - s;
    {
      ProfScope p(c, 1);
      hipLaunchKernelGGL(k_lu_diag<double>, dim3(nb), dim3(256), 0, c->stream, dd, s);
    }
    if (nrest > 0) {
      {
        ProfScope p(c, 2);
        hipLaunchKernelGGL(k_lu_panel<double>, dim3(2 * nrest, nb), dim3(256), 0, c->stream, dd, s);
      }
      {
        ProfScope p(c, 3);
        hipLaunchKernelGGL(k_lu_update<double>, dim3(nrest * nrest, nb), dim3(256), 0, c->stream,
                           dd, s, 0);
      }
    }
  }
  if (int r = check_launch("full LU")) return r;
  ProfScope p(c, 7);
  hipLaunchKernelGGL(k_inv_zero, dim3(nt_max, nb), dim3(256), 0, c->stream, dd);
  for (int s = 0; s < nt_max; s++) {
    hipLaunchKernelGGL(k_inv_fwd_panel, dim3(s + 1, nb), dim3(256), 0, c->stream, dd, s);
    int nrest = nt_max - 1 - s;
    if (nrest > 0)
      hipLaunchKernelGGL(k_inv_fwd_update, dim3(nrest * (s + 1), nb), dim3(256), 0, c->stream, dd,
                         s);
  }
  for (int s = nt_max - 1; s >= 0; s--) {
    hipLaunchKernelGGL(k_inv_bwd_panel, dim3(nt_max, nb), dim3(256), 0, c->stream, dd, s);
    if (s > 0)
      hipLaunchKernelGGL(k_inv_bwd_update, dim3(s * nt_max, nb), dim3(256), 0, c->stream, dd, s);
  }
  return check_launch("explicit inverse");
}

// d prel / dE for prel = sqrt(E (1 + E/c^2)), c = 2 * 137.035999206
// (calculateTauMatrix.cpp:846): (1 + 2E/c^2) / (2 prel).
void dprel_de(const double *E, const double *prel, double *out) {
  const double cl = 2.0 * 137.035999206, ic2 = 1.0 / (cl * cl);
  const double nr = 1.0 + 2.0 * E[0] * ic2, ni = 2.0 * E[1] * ic2;
  const double dr = 2.0 * prel[0], di = 2.0 * prel[1];
  const double s = 1.0 / (dr * dr + di * di);
  out[0] = (nr * dr + ni * di) * s;
  out[1] = (ni * dr - nr * di) * s;
}

// Adjoint route over a batch: tau00, Tr(M^-1 dM/dE) and, given the cotangent
// of tau00, the t-store gradient (see lsmsx_tau00_adjoint).
int run_adjoint(lsmsx_ctx *c, int nsolve, const int *atom_ids, const int *ispins, const int *iies,
                const double *energies, const double *prel, int nsc, int nspola, double *tau_out,
                double *lloyd_out, const double *tau_bar, double *tmat_grad, int *info_out) {
  if (int r = validate_common(c, nsc, nspola)) return r;
  if (c->rel) return fail(-6, "the adjoint route does not take a relativistic store");
  if (nsolve <= 0) return 0;
  if (lloyd_out && !c->dtmat.p) return fail(-3, "Lloyd term requested but no dt/dE store uploaded");
  if (lloyd_out && !energies) return fail(-2, "Lloyd term needs the energies");
  HIPCHK(hipSetDevice(c->device));
  std::vector<long> bytes(nsolve);
  int kc = -1;
  for (int i = 0; i < nsolve; i++) {
    auto it = c->atoms.find(atom_ids[i]);
    if (it == c->atoms.end()) return fail(-5, "unknown atom id " + std::to_string(atom_ids[i]));
    const AtomDev &a = it->second;
    int kci = nsc * a.kkrsz;
    if (kc < 0) kc = kci;
    if (kci != kc) return fail(-6, "all atoms of a batch must share kkrsz");
    if (kc != 16 && kc != 32) return fail(-6, "the adjoint route needs kkrsz_ns = 16 or 32");
    if (a.kkrsz > (c->lmax_tab + 1) * (c->lmax_tab + 1))
      return fail(-6, "atom lmax exceeds the lmax of the uploaded tables");
    if (iies[i] < 0 || iies[i] >= c->group) return fail(-7, "energy slot out of range");
    if (ispins[i] < 0 || ispins[i] >= nspola) return fail(-7, "spin index out of range");
    if (prel[2 * i] == 0.0 && prel[2 * i + 1] == 0.0)
      return fail(-7, "the adjoint route needs prel != 0");
    for (int s = 0; s < a.numLIZ; s++)
      if (a.store[s] < 0 || a.store[s] >= c->nstore) return fail(-7, "LIZStoreIdx out of range");
    long Ns = roundup(nsc * a.nrmat, kTile), nt = Ns / kTile;
    bytes[i] = 2 * Ns * Ns * 16 + nt * kTile * kTile * 16 + 2 * Ns * kc * 16 +
               (tau_bar ? (long)a.numLIZ * kc * kc * 16 : 0);
  }
  // the batched mixed and refined paths' buffers are not used here: give
  // them back before sizing the chunks by the free memory
  c->W.release();
  c->R.release();
  c->GH.release();
  c->YB.release();
  c->GP.release();
  c->GS.release();
  c->hp.release();
  c->he.release();
  c->A64.release();
  c->R64.release();
  long cap = c->ws_limit;
  if (cap <= 0) {
    size_t fr = 0, tot = 0;
    HIPCHK(hipMemGetInfo(&fr, &tot));
    long held = (long)((c->M.n + c->Minv.n + c->inv.n + c->gblk.n) * sizeof(double2));
    cap = (long)((double)(fr + held) * 0.7);
  }
  if (tau_bar) {
    if (int r = c->gstore.ensure((size_t)c->ldT * c->nstore)) return r;
    hipLaunchKernelGGL(k_zero_c, dim3(1024), dim3(256), 0, c->stream, c->gstore.p,
                       (long)c->ldT * c->nstore);
  }
  std::vector<double> minpiv_all(nsolve, 0.0);
  int done = 0;
  while (done < nsolve) {
    long acc = 0;
    int end = done;
    while (end < nsolve && (end == done || acc + bytes[end] <= cap)) acc += bytes[end++];
    int nb = end - done;
    long mtot = 0, xtot = 0, ivtot = 0;
    int maxliz = 1, nt_max = 1;
    for (int b = 0; b < nb; b++) {
      const AtomDev &a = c->atoms[atom_ids[done + b]];
      long Ns = roundup(nsc * a.nrmat, kTile);
      mtot += Ns * Ns;
      xtot += Ns * kc;
      ivtot += (Ns / kTile) * kTile * kTile;
      maxliz = std::max(maxliz, a.numLIZ);
      nt_max = std::max(nt_max, (int)(Ns / kTile));
    }
    if (int r = c->M.ensure(mtot)) return r;
    if (int r = c->Minv.ensure(mtot)) return r;
    if (int r = c->inv.ensure(ivtot)) return r;
    if (int r = c->perm.ensure((size_t)nb * nt_max * kTile)) return r;
    if (int r = c->info.ensure((size_t)nb)) return r;
    if (int r = c->minpiv.ensure((size_t)nb)) return r;
    if (int r = c->tau.ensure((size_t)nb * kc * kc)) return r;
    if (int r = c->X.ensure(xtot)) return r;
    if (int r = c->gx.ensure(xtot)) return r;
    if (int r = c->lpart.ensure((size_t)nb * maxliz)) return r;
    if (int r = c->lloyd.ensure((size_t)nb)) return r;
    if (int r = c->desc.ensure((size_t)nb)) return r;
    std::vector<SolveDesc> hd(nb);
    long moff = 0, xoff = 0, ioff = 0;
    for (int b = 0; b < nb; b++) {
      int i = done + b;
      AtomDev &a = c->atoms[atom_ids[i]];
      SolveDesc &d = hd[b];
      d = SolveDesc{};
      int N = nsc * a.nrmat, Ns = roundup(N, kTile);
      d.pos = a.d_pos;
      d.lmax = a.d_lmax;
      d.store = a.d_store;
      d.site_off = a.d_off[1][nsc - 1];
      d.numLIZ = a.numLIZ;
      d.n_spin_cant = nsc;
      d.kkrsz = a.kkrsz;
      d.kc = kc;
      d.prel_re = prel[2 * i];
      d.prel_im = prel[2 * i + 1];
      if (energies) {
        double dp[2];
        dprel_de(energies + 2 * i, prel + 2 * i, dp);
        d.dprel_re = dp[0];
        d.dprel_im = dp[1];
      }
      long jsm = (nspola != nsc) ? (long)kc * kc * ispins[i] : 0;
      d.t_off = (long)iies[i] * c->blk + jsm;
      d.ldt = kc;
      d.M = c->M.p + moff;
      d.A = d.M;
      d.Minv = c->Minv.p + moff;
      d.ld = Ns;
      d.Ns = Ns;
      d.pad = Ns - N;
      d.nt = Ns / kTile;
      d.group = lu_group(c, false);
      d.full_lu = 1;
      d.inv = c->inv.p + ioff;
      d.inv_stride = kTile * kTile;
      ioff += (long)d.nt * kTile * kTile;
      d.perm = c->perm.p + (long)b * nt_max * kTile;
      d.info = c->info.p + b;
      d.minpiv = c->minpiv.p + b;
      d.tau00 = c->tau.p + (long)b * kc * kc;
      d.t0 = c->tmat.p + c->ldT * a.store[0] + d.t_off;
      d.X = c->X.p + xoff;
      d.gx = c->gx.p + xoff;
      d.lpart = c->lpart.p + (long)b * maxliz;
      xoff += (long)Ns * kc;
      moff += (long)Ns * Ns;
    }
    HIPCHK(hipMemcpyAsync(c->desc.p, hd.data(), nb * sizeof(SolveDesc), hipMemcpyHostToDevice,
                          c->stream));
    hipLaunchKernelGGL(k_init_status, dim3((nb + 255) / 256), dim3(256), 0, c->stream, c->desc.p, nb);
    {
      ProfScope p(c, 0);
      hipLaunchKernelGGL(k_pad_identity, dim3(nb), dim3(256), 0, c->stream, c->desc.p);
      launch_assemble(c, nsc, dim3(maxliz, nb), c->stream, c->desc.p);
    }
    if (int r = check_launch("k_assemble")) return r;
    if (int r = launch_inverse(c, c->desc.p, nb, nt_max)) return r;
    {
      ProfScope p(c, 8);
      hipLaunchKernelGGL(k_inv_x, dim3(nt_max, nb), dim3(256), 0, c->stream, c->desc.p);
      hipLaunchKernelGGL(k_lloyd, dim3(maxliz, nb), dim3(256), 0, c->stream, c->desc.p, c->tmat.p,
                         lloyd_out ? c->dtmat.p : nullptr, c->ldT, c->cgnt.p, c->illp.p,
                         c->lmax_tab);
      hipLaunchKernelGGL(k_lloyd_reduce, dim3((nb + 63) / 64), dim3(64), 0, c->stream, c->desc.p,
                         nb, c->lloyd.p);
    }
    if (int r = check_launch("k_lloyd")) return r;
    if (tau_bar) {
      // gradient blocks of every (solve, site), then a keyed scatter-add
      if (int r = c->vbar.ensure((size_t)nb * kc * kc)) return r;
      if (int r = c->gblk.ensure((size_t)nb * maxliz * kc * kc)) return r;
      HIPCHK(hipMemcpyAsync(c->vbar.p, tau_bar + (long)done * kc * kc * 2,
                            (size_t)nb * kc * kc * sizeof(double2), hipMemcpyHostToDevice,
                            c->stream));
      hipLaunchKernelGGL(k_tau_grad, dim3(maxliz, nb), dim3(256), 0, c->stream, c->desc.p,
                         c->vbar.p, c->gblk.p, maxliz);
      std::map<std::pair<int, long>, std::vector<int>> keys;
      for (int b = 0; b < nb; b++) {
        const AtomDev &a = c->atoms[atom_ids[done + b]];
        for (int s = 0; s < a.numLIZ; s++)
          keys[{a.store[s], hd[b].t_off}].push_back(b * maxliz + s);
      }
      std::vector<int> kst, kptr(1, 0), kent;
      std::vector<long> ktoff;
      for (auto &kv : keys) {
        kst.push_back(kv.first.first);
        ktoff.push_back(kv.first.second);
        kent.insert(kent.end(), kv.second.begin(), kv.second.end());
        kptr.push_back((int)kent.size());
      }
      int nk = (int)kst.size();
      if (int r = c->keyi.ensure((size_t)nk + nk + 1 + kent.size())) return r;
      if (int r = c->keyl.ensure((size_t)nk)) return r;
      std::vector<int> packed;
      packed.insert(packed.end(), kst.begin(), kst.end());
      packed.insert(packed.end(), kptr.begin(), kptr.end());
      packed.insert(packed.end(), kent.begin(), kent.end());
      HIPCHK(hipMemcpyAsync(c->keyi.p, packed.data(), packed.size() * sizeof(int),
                            hipMemcpyHostToDevice, c->stream));
      HIPCHK(hipMemcpyAsync(c->keyl.p, ktoff.data(), nk * sizeof(long), hipMemcpyHostToDevice,
                            c->stream));
      hipLaunchKernelGGL(k_grad_reduce, dim3(nk), dim3(256), 0, c->stream, c->gblk.p, kc,
                         c->keyi.p, c->keyl.p, c->keyi.p + nk, c->keyi.p + 2 * nk + 1, c->gstore.p,
                         c->ldT, kc);
      if (int r = check_launch("k_tau_grad")) return r;
      // the host vectors above must outlive the copies
      HIPCHK(hipStreamSynchronize(c->stream));
    }
    if (tau_out)
      HIPCHK(hipMemcpyAsync(tau_out + (long)done * kc * kc * 2, c->tau.p,
                            (size_t)nb * kc * kc * sizeof(double2), hipMemcpyDeviceToHost,
                            c->stream));
    if (lloyd_out)
      HIPCHK(hipMemcpyAsync(lloyd_out + 2L * done, c->lloyd.p, nb * sizeof(double2),
                            hipMemcpyDeviceToHost, c->stream));
    std::vector<int> hinfo(nb);
    std::vector<double> hmin(nb);
    HIPCHK(hipMemcpyAsync(hinfo.data(), c->info.p, nb * sizeof(int), hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipMemcpyAsync(hmin.data(), c->minpiv.p, nb * sizeof(double), hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    drain_profile(c);
    for (int b = 0; b < nb; b++) {
      if (info_out) info_out[done + b] = hinfo[b];
      minpiv_all[done + b] = hmin[b];
    }
    done = end;
  }
  // tau00 of solves with a small tile-local pivot: from the FP64-refined
  // route, as on the plain path (the Lloyd term and the gradient keep the
  // explicit inverse, well inside their tolerances)
  if (tau_out && c->pivot_guard > 0.0) {
    std::vector<int> redo;
    for (int i = 0; i < nsolve; i++)
      if (minpiv_all[i] < c->pivot_guard) redo.push_back(i);
    if (!redo.empty()) {
      const int nr = (int)redo.size();
      std::vector<int> ra(nr), rs(nr), re(nr), rinfo(nr);
      std::vector<double> rp(2 * nr), rtau((size_t)nr * kc * kc * 2);
      for (int j = 0; j < nr; j++) {
        const int i = redo[j];
        ra[j] = atom_ids[i];
        rs[j] = ispins[i];
        re[j] = iies[i];
        rp[2 * j] = prel[2 * i];
        rp[2 * j + 1] = prel[2 * i + 1];
      }
      if (int r = run_batched(c, nr, ra.data(), rs.data(), re.data(), rp.data(), nsc, nspola,
                              LSMSX_PREC_FP64_REFINED, rtau.data(), rinfo.data()))
        return r;
      for (int j = 0; j < nr; j++)
        std::memcpy(tau_out + (long)redo[j] * kc * kc * 2, rtau.data() + (long)j * kc * kc * 2,
                    sizeof(double) * 2 * kc * kc);
    }
  }
  if (tau_bar && tmat_grad) {
    HIPCHK(hipMemcpyAsync(tmat_grad, c->gstore.p, sizeof(double2) * c->ldT * c->nstore,
                          hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  return 0;
}

// The FP64 KKR matrix M = 1 - t G of one (atom, spin, energy) in the
// reference layout (central site first, column-major N x N, N = nsc nrmat)
// into dst, on the context's stream; ddev is a device descriptor slot.  The
// relativistic store's kappa-mu blocks U M'_ij U^T are formed as well.
int assemble_reference_layout(lsmsx_ctx *c, AtomDev &a, int ispin, int iie, const double *prel,
                              int nsc, int nspola, double2 *dst, SolveDesc *ddev) {
  const int N = nsc * a.nrmat;
  SolveDesc d{};
  int kc = nsc * a.kkrsz;
  d.pos = a.d_pos;
  d.lmax = a.d_lmax;
  d.store = a.d_store;
  d.site_off = a.d_off[0][nsc - 1];
  d.numLIZ = a.numLIZ;
  d.n_spin_cant = nsc;
  d.kkrsz = a.kkrsz;
  d.kc = kc;
  d.prel_re = prel[0];
  d.prel_im = prel[1];
  d.t_off = (long)iie * c->blk + ((nspola != nsc) ? (long)kc * kc * ispin : 0);
  d.ldt = kc;
  d.M = dst;
  d.ld = N;
  d.Ns = N;
  d.pad = 0;
  d.nt = 0;
  d.A = d.M;
  HIPCHK(hipMemcpyAsync(ddev, &d, sizeof(d), hipMemcpyHostToDevice, c->stream));
  {
    ProfScope p(c, 0);
    launch_assemble(c, nsc, dim3(a.numLIZ, 1), c->stream, ddev);
  }
  if (int r = check_launch("k_assemble")) return r;
  if (c->rel) {  // the reference's kappa-mu blocks U M'_ij U^T
    if (nsc != 2 || 2 * a.kkrsz != c->rel_kk2)
      return fail(-6, "the relativistic store needs n_spin_cant = 2 and the store's lmax");
    for (int s = 0; s < a.numLIZ; s++)
      if (a.lmax[s] != a.lmax[0])
        return fail(-6, "relativistic blocks need every LIZ site at the central lmax");
    hipLaunchKernelGGL(k_rel_blocks, dim3(a.numLIZ, a.numLIZ), dim3(256), 0, c->stream, dst,
                       (long)N, c->rel_kk2, c->relU.p);
    if (int r = check_launch("k_rel_blocks")) return r;
  }
  return 0;
}

// Column-pivoted fallback of the batched calls: the reference's own route
// (solveTau00zgetrf, linearSolvers_CPU.cpp:98-123: LU of M with partial
// pivoting over whole columns, M X = [t0; 0], tau00 = the top kc x kc block
// of X), with rocSOLVER's zgetrf/zgetrs on the FP64 matrix in the reference
// layout.  For the solves that pivoting inside 64 x 64 diagonal tiles cannot
// factor (a singular diagonal tile, info > 0) or whose refinement against M
// does not converge.  tau (host, kc x kc column-major) and *info receive the
// result; *info is getrf's: 0, or the first exactly singular U(i,i) (1-based,
// reference order), when M itself is singular.
int column_pivoted_fallback(lsmsx_ctx *c, AtomDev &a, int ispin, int iie, const double *prel,
                            int nsc, int nspola, double *tau, int *info) {
  const int N = nsc * a.nrmat, kc = nsc * a.kkrsz;
  if (!c->blas && rocblas_create_handle(&c->blas) != rocblas_status_success)
    return fail(-4, "rocblas_create_handle failed");
  if (rocblas_set_stream(c->blas, c->stream) != rocblas_status_success)
    return fail(-4, "rocblas_set_stream failed");
  if (int r = c->fbM.ensure((size_t)N * N)) return r;
  if (int r = c->fbB.ensure((size_t)N * kc)) return r;
  if (int r = c->fbpiv.ensure((size_t)N)) return r;
  if (int r = c->fbinfo.ensure(1)) return r;
  if (int r = c->fbdesc.ensure(1)) return r;
  if (int r = assemble_reference_layout(c, a, ispin, iie, prel, nsc, nspola, c->fbM.p,
                                        c->fbdesc.p))
    return r;
  const long t_off = (long)iie * c->blk + ((nspola != nsc) ? (long)kc * kc * ispin : 0);
  const double2 *t0 = c->tmat.p + t_off + c->ldT * a.store[0];
  const long nb = (long)N * kc;
  hipLaunchKernelGGL(k_fallback_rhs, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, c->stream,
                     c->fbB.p, N, kc, t0, kc);
  if (int r = check_launch("k_fallback_rhs")) return r;
  auto *A = reinterpret_cast<rocblas_double_complex *>(c->fbM.p);
  auto *B = reinterpret_cast<rocblas_double_complex *>(c->fbB.p);
  if (rocsolver_zgetrf(c->blas, N, N, A, N, c->fbpiv.p, c->fbinfo.p) != rocblas_status_success)
    return fail(-4, "rocsolver_zgetrf failed");
  if (rocsolver_zgetrs(c->blas, rocblas_operation_none, N, kc, A, N, c->fbpiv.p, B, N) !=
      rocblas_status_success)
    return fail(-4, "rocsolver_zgetrs failed");
  int hinfo = 0;
  HIPCHK(hipMemcpy2DAsync(tau, sizeof(double2) * kc, c->fbB.p, sizeof(double2) * N,
                          sizeof(double2) * kc, kc, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(&hinfo, c->fbinfo.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  drain_profile(c);
  *info = hinfo;
  return 0;
}

}  // namespace

namespace {
// Device address of `scratch` on device dev, looked up once per device with
// that device current (each GPU holds its own copy at its own address).
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
unsigned char *g_noalloc_ptr[kMaxDevices] = {};

int noalloc_scratch_for_device(int dev, unsigned char **out) {
  if (dev < 0 || dev >= kMaxDevices) return fail(-1, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  if (!g_noalloc_ptr[dev]) {
    void *p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(-1, "hipGetSymbolAddress returned a null scratch address");
    g_noalloc_ptr[dev] = (unsigned char *)p;
  }
  *out = g_noalloc_ptr[dev];
  return 0;
}
}  // namespace

extern "C" {

const char *lsmsx_last_error(void) { return g_err.c_str(); }

int lsmsx_ctx_create(int device, lsmsx_ctx **out) {
  if (!out) return fail(-1, "null out pointer");
  *out = nullptr;
  int ndev = 0;
  HIPCHK(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return fail(-1, "invalid device index");
  HIPCHK(hipSetDevice(device));
  lsmsx_ctx *c = new lsmsx_ctx();
  c->device = device;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_asm, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming);
  if (e != hipSuccess) {
    if (c->stream) (void)hipStreamDestroy(c->stream);
    if (c->stream2) (void)hipStreamDestroy(c->stream2);
    if (c->ev_asm) (void)hipEventDestroy(c->ev_asm);
    delete c;
    return fail(-100 - (int)e, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  if (int r = noalloc_scratch_for_device(device, &c->noalloc)) {
    (void)hipEventDestroy(c->ev_join);
    (void)hipEventDestroy(c->ev_asm);
    (void)hipStreamDestroy(c->stream2);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return r;
  }
  *out = c;
  return 0;
}

int lsmsx_ctx_destroy(lsmsx_ctx *c) {
  if (!c) return 0;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  drain_profile(c);
  for (auto e : c->event_pool) (void)hipEventDestroy(e);
  for (auto &kv : c->atoms) {
    AtomDev &a = kv.second;
    (void)hipFree(a.d_pos);
    (void)hipFree(a.d_lmax);
    (void)hipFree(a.d_store);
    for (int l = 0; l < 2; l++)
      for (int s = 0; s < 2; s++) (void)hipFree(a.d_off[l][s]);
  }
  c->cgnt.release();
  c->illp.release();
  c->tmat.release();
  c->relU.release();
  c->relf.release();
  c->it_c2.release();
  c->it_d.release();
  c->it_i.release();
  c->ss_r.release();
  c->ss_v.release();
  c->ss_coef.release();
  c->ss_pot.release();
  c->ss_iend.release();
  c->ss_status.release();
  c->ss_e.release();
  c->ss_prel.release();
  c->ss_matom.release();
  c->ss_zlr.release();
  c->ss_jlr.release();
  c->gf_pot.release();
  c->gf_potof.release();
  c->gf_src.release();
  c->ss_storeof.release();
  c->ss_uu.release();
  c->gf_tau.release();
  c->gf_dos.release();
  c->gf_green.release();
  c->gf_w.release();
  c->M.release();
  c->GH.release();
  c->rowperm.release();
  c->fbM.release();
  c->fbB.release();
  c->fbpiv.release();
  c->fbinfo.release();
  c->fbdesc.release();
  if (c->blas) (void)rocblas_destroy_handle(c->blas);
  c->blas = nullptr;
  c->YB.release();
  c->GP.release();
  c->inv.release();
  c->tau.release();
  c->tmpM.release();
  c->t0buf.release();
  c->perm.release();
  c->info.release();
  c->minpiv.release();
  c->desc.release();
  c->desc2.release();
  c->desc_own.release();
  c->GS.release();
  c->X.release();
  c->A64.release();
  c->R64.release();
  c->W.release();
  c->R.release();
  c->conv.release();
  c->iters.release();
  c->count.release();
  c->dxn.release();
  for (auto &g : c->graphs) (void)hipGraphExecDestroy(g.second);
  c->graphs.clear();
  c->dtmat.release();
  c->hp.release();
  c->he.release();
  c->Minv.release();
  c->gx.release();
  c->lpart.release();
  c->lloyd.release();
  c->vbar.release();
  c->gblk.release();
  c->gstore.release();
  c->tfull.release();
  c->keyi.release();
  c->colmap.release();
  c->keyl.release();
  (void)hipStreamSynchronize(c->stream2);
  (void)hipEventDestroy(c->ev_join);
  (void)hipEventDestroy(c->ev_asm);
  (void)hipStreamDestroy(c->stream2);
  (void)hipStreamDestroy(c->stream);
  delete c;
  return 0;
}

int lsmsx_set_tables(lsmsx_ctx *c, int maxlmax, const double *cgnt, const double *illp,
                     const double *ilp1) {
  (void)ilp1;
  if (!c) return fail(-1, "null context");
  if (maxlmax < 0 || maxlmax > kMaxLmax) return fail(-2, "maxlmax must be in 0..3");
  HIPCHK(hipSetDevice(c->device));
  int kkr = (maxlmax + 1) * (maxlmax + 1);
  if (int r = c->cgnt.ensure((size_t)(maxlmax + 1) * kkr * kkr)) return r;
  if (int r = c->illp.ensure((size_t)kkr * kkr)) return r;
  HIPCHK(hipMemcpy(c->cgnt.p, cgnt, sizeof(double) * (maxlmax + 1) * kkr * kkr,
                   hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->illp.p, illp, sizeof(double2) * kkr * kkr, hipMemcpyHostToDevice));
  c->lmax_tab = maxlmax;
  return 0;
}

int lsmsx_upload_tmats(lsmsx_ctx *c, const double *tmat, long ldT, int numStore, long blk,
                       int groupSize) {
  if (!c) return fail(-1, "null context");
  if (ldT < blk * groupSize || numStore <= 0 || blk <= 0 || groupSize <= 0)
    return fail(-2, "inconsistent t-matrix store dimensions");
  HIPCHK(hipSetDevice(c->device));
  if (int r = c->tmat.ensure((size_t)ldT * numStore)) return r;
  HIPCHK(hipMemcpy(c->tmat.p, tmat, sizeof(double2) * ldT * numStore, hipMemcpyHostToDevice));
  c->ldT = ldT;
  c->blk = blk;
  c->nstore = numStore;
  c->group = groupSize;
  c->rel = 0;
  return 0;
}

int lsmsx_upload_tmats_rel(lsmsx_ctx *c, const double *tmat, long ldT, int numStore, long blk,
                           int groupSize, const double *energies) {
  if (!c) return fail(-1, "null context");
  if (c->lmax_tab < 0) return fail(-2, "tables not set");
  if (!energies) return fail(-2, "null energies");
  const int kk2 = 2 * (c->lmax_tab + 1) * (c->lmax_tab + 1);
  if (blk < (long)kk2 * kk2) return fail(-2, "blkSizeTmatStore smaller than (2 kkrsz)^2");
  if (int r = lsmsx_upload_tmats(c, tmat, ldT, numStore, blk, groupSize)) return r;
  const double c2inv = 1.0 / ((2.0 * 137.035999206) * (2.0 * 137.035999206));
  std::vector<double2> fac(groupSize);
  c->rel_finv.assign(groupSize, make_double2(0.0, 0.0));
  for (int e = 0; e < groupSize; e++) {
    const std::complex<double> E(energies[2 * e], energies[2 * e + 1]);
    if (E == 0.0) return fail(-2, "relativistic store at E = 0");
    // setgij.f:52: fac = psq / ce, psq = prel^2 (calculateTauMatrix.cpp:219)
    const std::complex<double> prel = std::sqrt(E * (1.0 + E * c2inv));
    const std::complex<double> f = prel * prel / E;
    fac[e] = make_double2(f.real(), f.imag());
    const std::complex<double> g = 1.0 / f;
    c->rel_finv[e] = make_double2(g.real(), g.imag());
  }
  const std::vector<double> U = clebsch_u(c->lmax_tab);
  if (int r = c->relU.ensure(U.size())) return r;
  if (int r = c->relf.ensure((size_t)groupSize)) return r;
  HIPCHK(hipMemcpy(c->relU.p, U.data(), U.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->relf.p, fac.data(), groupSize * sizeof(double2), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_rel_tstore, dim3(numStore, groupSize), dim3(256), 0, c->stream, c->tmat.p,
                     ldT, blk, kk2, c->relU.p, c->relf.p);
  if (int r = check_launch("k_rel_tstore")) return r;
  HIPCHK(hipStreamSynchronize(c->stream));
  for (auto &g : c->graphs) (void)hipGraphExecDestroy(g.second);
  c->graphs.clear();
  c->rel = 1;
  c->rel_kk2 = kk2;
  return 0;
}

int lsmsx_set_liz(lsmsx_ctx *c, int atom_id, int numLIZ, const double *pos, const int *lmax,
                  const int *store) {
  if (!c) return fail(-1, "null context");
  if (numLIZ <= 0) return fail(-2, "numLIZ must be positive");
  HIPCHK(hipSetDevice(c->device));
  AtomDev &a = c->atoms[atom_id];
  (void)hipFree(a.d_pos);
  (void)hipFree(a.d_lmax);
  (void)hipFree(a.d_store);
  for (int l = 0; l < 2; l++)
    for (int s = 0; s < 2; s++) {
      (void)hipFree(a.d_off[l][s]);
      a.d_off[l][s] = nullptr;
    }
  a.numLIZ = numLIZ;
  a.lmax.assign(lmax, lmax + numLIZ);
  a.store.assign(store, store + numLIZ);
  a.pos.assign(pos, pos + 3 * numLIZ);
  a.geo = -1;
  for (auto &kv : c->atoms) {
    const AtomDev &b = kv.second;
    if (kv.first == atom_id || b.geo < 0 || b.numLIZ != numLIZ || b.lmax != a.lmax) continue;
    if (std::memcmp(b.pos.data(), a.pos.data(), sizeof(double) * 3 * numLIZ) == 0) {
      a.geo = b.geo;
      break;
    }
  }
  if (a.geo < 0) a.geo = c->next_geo++;
  a.nrmat = 0;
  for (int s = 0; s < numLIZ; s++) {
    if (lmax[s] < 0 || lmax[s] > kMaxLmax) return fail(-2, "LIZlmax must be in 0..3");
    if (lmax[s] > lmax[0]) return fail(-2, "LIZlmax exceeds the central site's lmax");
    a.nrmat += (lmax[s] + 1) * (lmax[s] + 1);
  }
  a.kkrsz = (lmax[0] + 1) * (lmax[0] + 1);
  HIPCHK(hipMalloc((void **)&a.d_pos, sizeof(double) * 3 * numLIZ));
  HIPCHK(hipMalloc((void **)&a.d_lmax, sizeof(int) * numLIZ));
  HIPCHK(hipMalloc((void **)&a.d_store, sizeof(int) * numLIZ));
  HIPCHK(hipMemcpy(a.d_pos, pos, sizeof(double) * 3 * numLIZ, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(a.d_lmax, lmax, sizeof(int) * numLIZ, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(a.d_store, store, sizeof(int) * numLIZ, hipMemcpyHostToDevice));
  for (int l = 0; l < 2; l++)
    for (int s = 0; s < 2; s++) {
      std::vector<int> off = site_offsets(a, s + 1, l == 1);
      HIPCHK(hipMalloc((void **)&a.d_off[l][s], sizeof(int) * numLIZ));
      HIPCHK(hipMemcpy(a.d_off[l][s], off.data(), sizeof(int) * numLIZ, hipMemcpyHostToDevice));
    }
  return 0;
}

int lsmsx_build_kkr(lsmsx_ctx *c, int atom_id, int ispin, int iie, const double *prel, int nsc,
                    int nspola, double *m, int ldm) {
  if (int r = validate_common(c, nsc, nspola)) return r;
  auto it = c->atoms.find(atom_id);
  if (it == c->atoms.end()) return fail(-5, "unknown atom id");
  AtomDev &a = it->second;
  if (iie < 0 || iie >= c->group) return fail(-7, "energy slot out of range");
  if (ispin < 0 || ispin >= nspola) return fail(-7, "spin index out of range");
  int N = nsc * a.nrmat;
  if (ldm < N) return fail(-2, "ldm < nrmat_ns");
  if (a.kkrsz > (c->lmax_tab + 1) * (c->lmax_tab + 1))
    return fail(-6, "atom lmax exceeds the lmax of the uploaded tables");
  for (int s = 0; s < a.numLIZ; s++)
    if (a.store[s] < 0 || a.store[s] >= c->nstore) return fail(-7, "LIZStoreIdx out of range");
  HIPCHK(hipSetDevice(c->device));
  if (int r = c->M.ensure((size_t)N * N)) return r;
  if (int r = c->desc.ensure(1)) return r;
  if (int r = assemble_reference_layout(c, a, ispin, iie, prel, nsc, nspola, c->M.p, c->desc.p))
    return r;
  HIPCHK(hipMemcpy2DAsync(m, sizeof(double2) * ldm, c->M.p, sizeof(double2) * N,
                          sizeof(double2) * N, N, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  drain_profile(c);
  return 0;
}

int lsmsx_tau00(lsmsx_ctx *c, int atom_id, int ispin, int iie, const double *prel, int nsc,
                int nspola, int precision, double *tau00, int *info) {
  int inf = 0;
  int r = run_batched(c, 1, &atom_id, &ispin, &iie, prel, nsc, nspola, precision, tau00, &inf);
  if (info) *info = inf;
  return r;
}

int lsmsx_tau00_batched(lsmsx_ctx *c, int nsolve, const int *atom_ids, const int *ispins,
                        const int *iies, const double *prel, int nsc, int nspola, int precision,
                        double *tau00, int *info) {
  return run_batched(c, nsolve, atom_ids, ispins, iies, prel, nsc, nspola, precision, tau00, info);
}

int lsmsx_solve_tau00_matrix(lsmsx_ctx *c, int n, int k, const double *m, int ldm,
                             const double *t0, double *tau00, int *info) {
  if (!c) return fail(-1, "null context");
  if (k <= 0 || k > kTile || n < k || ldm < n) return fail(-2, "bad matrix dimensions");
  HIPCHK(hipSetDevice(c->device));
  int Ns = roundup(n, kTile);
  if (int r = c->tmpM.ensure((size_t)n * n)) return r;
  if (int r = c->M.ensure((size_t)Ns * Ns)) return r;
  if (int r = c->inv.ensure((size_t)kTile * kTile)) return r;
  if (int r = c->perm.ensure((size_t)(Ns / kTile) * kTile)) return r;
  if (int r = c->info.ensure(1)) return r;
  if (int r = c->minpiv.ensure(1)) return r;
  if (int r = c->tau.ensure((size_t)k * k)) return r;
  if (int r = c->t0buf.ensure((size_t)k * k)) return r;
  if (int r = c->desc.ensure(1)) return r;
  HIPCHK(hipMemcpy2DAsync(c->tmpM.p, sizeof(double2) * n, m, sizeof(double2) * ldm,
                          sizeof(double2) * n, n, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->t0buf.p, t0, sizeof(double2) * k * k, hipMemcpyHostToDevice,
                        c->stream));
  SolveDesc d{};
  d.kc = k;
  d.ldt = k;
  d.M = c->M.p;
  d.A = d.M;
  d.ld = Ns;
  d.Ns = Ns;
  d.pad = Ns - n;
  d.nt = Ns / kTile;
  d.group = lu_group(c, false);
  d.inv = c->inv.p;
  d.perm = c->perm.p;
  d.info = c->info.p;
  d.minpiv = c->minpiv.p;
  d.tau00 = c->tau.p;
  d.t0 = c->t0buf.p;
  HIPCHK(hipMemcpyAsync(c->desc.p, &d, sizeof(d), hipMemcpyHostToDevice, c->stream));
  // attempt 0: the caller's row order; attempt 1, only if a diagonal tile
  // met a zero pivot: the non-central rows reordered so that each column in
  // turn gets the largest remaining entry on the diagonal (a greedy static
  // pivot order).  Pivoting inside 64-row tiles cannot exchange rows across
  // tiles, so a non-singular M can still give a singular tile; a row
  // permutation that keeps the central rows leaves tau00 unchanged and brings
  // a nonzero into every diagonal position.  The caller sees info of the
  // first attempt only when the second fails too.
  int hinfo = 0, first_info = 0;
  for (int attempt = 0; attempt < 2; attempt++) {
    const int *rp = nullptr;
    if (attempt == 1) {
      const int nn = n - k;
      std::vector<int> perm(nn);
      std::vector<char> used(nn, 0);
      for (int j = 0; j < nn; j++) {  // column j of the non-central part: caller column j + k
        const double *col = m + 2 * ((long)ldm * (j + k) + k);
        int best = -1;
        double bv = -1.0;
        for (int r = 0; r < nn; r++)
          if (!used[r]) {
            const double v = std::fabs(col[2 * r]) + std::fabs(col[2 * r + 1]);
            if (v > bv) {
              bv = v;
              best = r;
            }
          }
        used[best] = 1;
        perm[j] = best;
      }
      if (int r = c->rowperm.ensure((size_t)nn)) return r;
      HIPCHK(hipMemcpy(c->rowperm.p, perm.data(), nn * sizeof(int), hipMemcpyHostToDevice));
      rp = c->rowperm.p;
    }
    hipLaunchKernelGGL(k_init_status, dim3(1), dim3(64), 0, c->stream, c->desc.p, 1);
    hipLaunchKernelGGL(k_permute_in, dim3(Ns), dim3(256), 0, c->stream, c->desc.p, c->tmpM.p,
                       (long)n, n, k, rp);
    if (int r = check_launch("k_permute_in")) return r;
    if (int r = launch_lu<double>(c, c->desc.p, 1, d.nt)) return r;
    HIPCHK(hipMemcpyAsync(&hinfo, c->info.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (attempt == 0) first_info = hinfo;
    if (hinfo <= 0 || n == k) break;
  }
  HIPCHK(hipMemcpyAsync(tau00, c->tau.p, sizeof(double2) * k * k, hipMemcpyDeviceToHost,
                        c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  drain_profile(c);
  if (info) *info = hinfo > 0 ? first_info : hinfo;
  return 0;
}

int lsmsx_solve_tau00_matrix_noalloc(lsmsx_ctx *c, int n, int k, const double *d_m, int ldm,
                                     const double *d_t0, double *d_tau00, int *d_info) {
  if (!c) return fail(-1, "null context");
  if (k <= 0 || k > kTile || n < k || ldm < n) return fail(-2, "bad matrix dimensions");
  if (!d_m || !d_t0 || !d_tau00 || !d_info) return fail(-2, "null device pointer");
  const int Ns = roundup(n, kTile);
  if (Ns > kNoAllocMaxNs)
    return fail(-9, "n = " + std::to_string(n) + " exceeds the no-allocation limit (padded size " +
                        std::to_string(kNoAllocMaxNs) + ")");
  if (!c->noalloc) return fail(-1, "context has no scratch address");
  HIPCHK(hipSetDevice(c->device));
  unsigned char *base = c->noalloc;
  SolveDesc d{};
  d.kc = k;
  d.ldt = k;
  d.M = (double2 *)base;
  d.A = d.M;
  d.ld = Ns;
  d.Ns = Ns;
  d.pad = Ns - n;
  d.nt = Ns / kTile;
  d.group = lu_group(c, false);
  d.inv = base + kNoAllocInvOff;
  d.inv_stride = 0;
  d.perm = (int *)(base + kNoAllocPermOff);
  d.minpiv = (double *)(base + kNoAllocMinOff);
  d.info = d_info;
  d.tau00 = (double2 *)d_tau00;
  d.t0 = (const double2 *)d_t0;
  SolveDesc *dd = (SolveDesc *)(base + kNoAllocDescOff);
  hipLaunchKernelGGL(k_put_desc, dim3(1), dim3(64), 0, c->stream, d, dd);
  hipLaunchKernelGGL(k_init_status, dim3(1), dim3(64), 0, c->stream, dd, 1);
  hipLaunchKernelGGL(k_permute_in, dim3(Ns), dim3(256), 0, c->stream, dd, (const double2 *)d_m,
                     (long)ldm, n, k, nullptr);
  if (int r = check_launch("k_permute_in")) return r;
  // no profiling events here: recording them may create events
  const bool prof = c->profiling;
  c->profiling = false;
  int r = launch_lu<double>(c, dd, 1, d.nt);
  c->profiling = prof;
  if (r) return r;
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int lsmsx_set_refinement(lsmsx_ctx *c, int max_sweeps, double tol) {
  if (!c) return fail(-1, "null context");
  if (max_sweeps < 1 || max_sweeps > 100 || !(tol > 0.0)) return fail(-2, "bad refinement settings");
  c->ir_max_sweeps = max_sweeps;
  c->ir_tol = tol;
  return 0;
}

int lsmsx_set_residual_form(lsmsx_ctx *c, int form) {
  if (!c) return fail(-1, "null context");
  if (form != LSMSX_RESIDUAL_AUTO && form != LSMSX_RESIDUAL_MATRIX)
    return fail(-2, "unknown residual form");
  c->residual_form = form;
  return 0;
}

int lsmsx_last_batch_stats(lsmsx_ctx *c, int nsolve, int *iters, double *minpiv) {
  if (!c) return fail(-1, "null context");
  if (nsolve < 0 || nsolve > (int)c->last_iters.size()) return fail(-2, "nsolve exceeds the last batch");
  for (int i = 0; i < nsolve; i++) {
    if (iters) iters[i] = c->last_iters[i];
    if (minpiv) minpiv[i] = c->last_minpiv[i];
  }
  return 0;
}

int lsmsx_set_workspace_limit(lsmsx_ctx *c, long bytes) {
  if (!c) return fail(-1, "null context");
  c->ws_limit = bytes;
  return 0;
}

int lsmsx_set_profiling(lsmsx_ctx *c, int on) {
  if (!c) return fail(-1, "null context");
  c->profiling = on != 0;
  return 0;
}

int lsmsx_reset_profile(lsmsx_ctx *c) {
  if (!c) return fail(-1, "null context");
  for (int i = 0; i < 10; i++) {
    c->prof_ms[i] = 0;
    c->prof_n[i] = 0;
  }
  return 0;
}

int lsmsx_get_profile(lsmsx_ctx *c, int kind, double *ms, long *launches) {
  if (!c) return fail(-1, "null context");
  if (kind < 0 || kind > 9) return fail(-2, "bad kernel kind");
  if (ms) *ms = c->prof_ms[kind];
  if (launches) *launches = c->prof_n[kind];
  return 0;
}

int lsmsx_upload_dtmats(lsmsx_ctx *c, const double *dtmat, long ldT, int numStore, long blk,
                        int groupSize) {
  if (!c) return fail(-1, "null context");
  if (!c->tmat.p) return fail(-3, "upload the t-matrix store first");
  if (ldT != c->ldT || numStore != c->nstore || blk != c->blk || groupSize != c->group)
    return fail(-2, "dt/dE store dimensions differ from the t-matrix store");
  HIPCHK(hipSetDevice(c->device));
  if (int r = c->dtmat.ensure((size_t)ldT * numStore)) return r;
  HIPCHK(hipMemcpy(c->dtmat.p, dtmat, sizeof(double2) * ldT * numStore, hipMemcpyHostToDevice));
  return 0;
}

int lsmsx_tau00_adjoint(lsmsx_ctx *c, int nsolve, const int *atom_ids, const int *ispins,
                        const int *iies, const double *energies, const double *prel, int nsc,
                        int nspola, double *tau00, double *lloyd, const double *tau00_bar,
                        double *tmat_grad, int *info) {
  if (tau00_bar && !tmat_grad) return fail(-2, "tau00_bar given without tmat_grad");
  return run_adjoint(c, nsolve, atom_ids, ispins, iies, energies, prel, nsc, nspola, tau00, lloyd,
                     tau00_bar, tmat_grad, info);
}

int lsmsx_tau_full(lsmsx_ctx *c, int atom_id, int ispin, int iie, const double *prel, int nsc,
                   int nspola, double *tau, int ldtau, int *info) {
  if (int r = validate_common(c, nsc, nspola)) return r;
  if (c->rel) return fail(-6, "full tau does not take a relativistic store");
  auto it = c->atoms.find(atom_id);
  if (it == c->atoms.end()) return fail(-5, "unknown atom id");
  AtomDev &a = it->second;
  if (iie < 0 || iie >= c->group) return fail(-7, "energy slot out of range");
  if (ispin < 0 || ispin >= nspola) return fail(-7, "spin index out of range");
  if (a.kkrsz > (c->lmax_tab + 1) * (c->lmax_tab + 1))
    return fail(-6, "atom lmax exceeds the lmax of the uploaded tables");
  for (int s = 0; s < a.numLIZ; s++)
    if (a.store[s] < 0 || a.store[s] >= c->nstore) return fail(-7, "LIZStoreIdx out of range");
  const int N = nsc * a.nrmat, Ns = roundup(N, kTile), nt = Ns / kTile, kc = nsc * a.kkrsz;
  if (ldtau < N) return fail(-2, "ldtau < nrmat_ns");
  if (kc > kTile) return fail(-6, "kkrsz_ns > 64 not supported");
  HIPCHK(hipSetDevice(c->device));
  if (int r = c->M.ensure((size_t)Ns * Ns)) return r;
  if (int r = c->Minv.ensure((size_t)Ns * Ns)) return r;
  if (int r = c->inv.ensure((size_t)nt * kTile * kTile)) return r;
  if (int r = c->perm.ensure((size_t)nt * kTile)) return r;
  if (int r = c->info.ensure(1)) return r;
  if (int r = c->minpiv.ensure(1)) return r;
  if (int r = c->tfull.ensure((size_t)N * N)) return r;
  if (int r = c->colmap.ensure((size_t)2 * Ns)) return r;
  if (int r = c->desc.ensure(1)) return r;
  // solver index -> (site, reference index)
  std::vector<int> soff = site_offsets(a, nsc, true), roff = site_offsets(a, nsc, false);
  std::vector<int> cmap(2 * Ns, -1);
  for (int s = 0; s < a.numLIZ; s++) {
    int w = nsc * (a.lmax[s] + 1) * (a.lmax[s] + 1);
    for (int k = 0; k < w; k++) {
      cmap[soff[s] + k] = s;
      cmap[Ns + soff[s] + k] = roff[s] + k;
    }
  }
  HIPCHK(hipMemcpyAsync(c->colmap.p, cmap.data(), sizeof(int) * 2 * Ns, hipMemcpyHostToDevice,
                        c->stream));
  SolveDesc d{};
  d.pos = a.d_pos;
  d.lmax = a.d_lmax;
  d.store = a.d_store;
  d.site_off = a.d_off[1][nsc - 1];
  d.numLIZ = a.numLIZ;
  d.n_spin_cant = nsc;
  d.kkrsz = a.kkrsz;
  d.kc = kc;
  d.prel_re = prel[0];
  d.prel_im = prel[1];
  d.t_off = (long)iie * c->blk + ((nspola != nsc) ? (long)kc * kc * ispin : 0);
  d.ldt = kc;
  d.M = c->M.p;
  d.A = d.M;
  d.Minv = c->Minv.p;
  d.ld = Ns;
  d.Ns = Ns;
  d.pad = Ns - N;
  d.nt = nt;
  d.group = lu_group(c, false);
  d.full_lu = 1;
  d.inv = c->inv.p;
  d.inv_stride = kTile * kTile;
  d.perm = c->perm.p;
  d.info = c->info.p;
  d.minpiv = c->minpiv.p;
  HIPCHK(hipMemcpyAsync(c->desc.p, &d, sizeof(d), hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_init_status, dim3(1), dim3(64), 0, c->stream, c->desc.p, 1);
  {
    ProfScope p(c, 0);
    hipLaunchKernelGGL(k_pad_identity, dim3(1), dim3(256), 0, c->stream, c->desc.p);
    launch_assemble(c, nsc, dim3(a.numLIZ, 1), c->stream, c->desc.p);
  }
  if (int r = check_launch("k_assemble")) return r;
  if (int r = launch_inverse(c, c->desc.p, 1, nt)) return r;
  hipLaunchKernelGGL(k_tau_full_out, dim3(Ns), dim3(256), 0, c->stream, c->desc.p, c->tmat.p,
                     c->ldT, c->colmap.p, c->colmap.p + Ns, c->tfull.p, (long)N);
  if (int r = check_launch("k_tau_full_out")) return r;
  int hinfo = 0;
  HIPCHK(hipMemcpy2DAsync(tau, sizeof(double2) * ldtau, c->tfull.p, sizeof(double2) * N,
                          sizeof(double2) * N, N, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(&hinfo, c->info.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  drain_profile(c);
  if (info) *info = hinfo;
  return 0;
}

int lsmsx_set_schedule(lsmsx_ctx *c, int flags) {
  if (!c) return fail(-1, "null context");
  if (flags & ~(LSMSX_SCHED_RIGHT_ALIGN | LSMSX_SCHED_XCD_SWIZZLE | LSMSX_SCHED_TWO_STREAMS))
    return fail(-2, "unknown schedule flags");
  c->sched = flags;
  return 0;
}

int lsmsx_get_schedule(lsmsx_ctx *c, int *flags) {
  if (!c || !flags) return fail(-1, "null argument");
  *flags = c->sched;
  return 0;
}

int lsmsx_set_shared_structure_constants(lsmsx_ctx *c, int on) {
  if (!c) return fail(-1, "null context");
  c->share_g = on ? 1 : 0;
  return 0;
}

int lsmsx_set_split_group(lsmsx_ctx *c, int group) {
  if (!c) return fail(-1, "null context");
  if (group != 4 && group != 8 && group != 16) return fail(-2, "split group must be 4, 8 or 16");
  c->split_group = group;
  return 0;
}

int lsmsx_get_split_group(lsmsx_ctx *c, int *group, double *ms) {
  if (!c || !group) return fail(-1, "null argument");
  *group = c->split_group;
  if (ms)
    for (int k = 0; k < 3; k++) ms[k] = c->group_ms[k];
  return 0;
}

int lsmsx_set_graph(lsmsx_ctx *c, int on) {
  if (!c) return fail(-1, "null context");
  c->use_graph = on != 0;
  return 0;
}

int lsmsx_autotune(lsmsx_ctx *c, int nsolve, const int *atom_ids, const int *ispins,
                   const int *iies, const double *prel, int nsc, int nspola, int precision,
                   int *best, double *ms) {
  if (!c) return fail(-1, "null context");
  if (nsolve <= 0) return fail(-2, "empty batch");
  std::vector<double> tau((size_t)nsolve * 64 * 64 * 2);
  std::vector<int> info(nsolve);
  const bool prof = c->profiling;
  c->profiling = false;
  int rc = run_batched(c, nsolve, atom_ids, ispins, iies, prel, nsc, nspola, precision,
                       tau.data(), info.data());  // warm-up: allocations, graphs of sched 0
  double tbest = 1e300;
  int sbest = c->sched;
  for (int sch = 0; sch < 8 && rc == 0; sch++) {
    c->sched = sch;
    if (c->use_graph) {  // capture outside the timed run
      rc = run_batched(c, nsolve, atom_ids, ispins, iies, prel, nsc, nspola, precision,
                       tau.data(), info.data());
      if (rc) break;
    }
    auto t0 = std::chrono::steady_clock::now();
    rc = run_batched(c, nsolve, atom_ids, ispins, iies, prel, nsc, nspola, precision, tau.data(),
                     info.data());
    double dt = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (ms) ms[sch] = dt;
    if (dt < tbest) {
      tbest = dt;
      sbest = sch;
    }
  }
  c->sched = sbest;
  // split-FP16 path: the steps per trailing update under the chosen schedule
  if (precision == LSMSX_PREC_MIXED_F16 && rc == 0) {
    // a single timed run per size: another size replaces the default 8 only
    // when it is faster by more than 2%, so that run-to-run noise does not
    // decide
    const int groups[3] = {4, 8, 16};
    int gbest = c->split_group;
    double gt = 1e300;
    for (int k = 0; k < 3 && rc == 0; k++) {
      c->split_group = groups[k];
      // first run: plane buffers of this size (and the graph, when on)
      rc = run_batched(c, nsolve, atom_ids, ispins, iies, prel, nsc, nspola, precision,
                       tau.data(), info.data());
      if (rc) break;
      auto t0 = std::chrono::steady_clock::now();
      rc = run_batched(c, nsolve, atom_ids, ispins, iies, prel, nsc, nspola, precision,
                       tau.data(), info.data());
      const double dt =
          std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
      c->group_ms[k] = dt;
      if (dt < gt) {
        gt = dt;
        gbest = groups[k];
      }
    }
    if (c->group_ms[1] <= 1.02 * gt) gbest = 8;
    c->split_group = gbest;
  }
  c->profiling = prof;
  if (best) *best = sbest;
  return rc;
}

int lsmsx_set_pivot_guard(lsmsx_ctx *c, double threshold) {
  if (!c) return fail(-1, "null context");
  if (!(threshold >= 0.0)) return fail(-2, "pivot guard must be >= 0");
  c->pivot_guard = threshold;
  return 0;
}

int lsmsx_single_site_nonrel(lsmsx_ctx *c, int nrelv, int lmax, int npot, int iprpts,
                             const double *vr, const double *r_mesh, const int *jmt,
                             const int *jws, const double *r_sph, int ne, const double *energy,
                             const double *prel, double *matom, double *tmat, double *zlr,
                             double *jlr) {
  if (!c) return fail(-1, "null context");
  if (lmax < 1 || lmax > 3) return fail(-2, "single-site lmax must be 1..3");
  if (npot < 0 || ne < 0 || iprpts < 1) return fail(-2, "negative batch size");
  if (npot == 0 || ne == 0) return 0;
  if (!vr || !r_mesh || !jmt || !jws || !r_sph || !energy || !prel || !matom)
    return fail(-2, "null argument");
  // per-potential checks and the matching window (semrel_c.f:130-143, cinterp.f:12-15)
  std::vector<SingleSitePot> hp((size_t)npot);
  std::vector<int> hiend((size_t)npot);
  for (int p = 0; p < npot; p++) {
    const double *r = r_mesh + (size_t)p * iprpts;
    SingleSitePot &P = hp[p];
    P.jmt = jmt[p];
    P.jws = jws[p];
    P.iend = std::max(P.jmt + 11, P.jws);
    P.r_sph = r_sph[p];
    if (P.jmt < 5 || P.jws < 1) return fail(-2, "jmt < 5 or jws < 1");
    if (P.iend > iprpts) return fail(-2, "max(jmt+11, jws) exceeds iprpts");
    if (!(r[0] > 0.0)) return fail(-2, "the radial mesh must start above r = 0");
    for (int j = 0; j + 1 < P.iend; j++)
      if (!(r[j + 1] > r[j])) return fail(-2, "the radial mesh must increase");
    if (!(P.r_sph > r[0]) || !(P.r_sph <= r[P.iend - 1])) return fail(-2, "r_sph outside the mesh");
    const int nr = P.jmt;
    int ip1 = 1;
    for (int i = 1; i <= nr - 2; i++)
      if (P.r_sph > r[i]) ip1 = i + 1;
    const int i = ip1 - 1;
    P.ip1 = ip1;
    P.i1 = (i > 0) ? i - 1 : std::min(4, nr) - 1;
    P.i2 = (ip1 < nr - 1) ? ip1 + 1 : std::max(nr - 3, 1) - 1;
    hiend[p] = P.iend;
  }
  const int nl = lmax + 1;
  const size_t ntask = (size_t)ne * npot * nl;
  const size_t npts = (size_t)npot * iprpts;
  const bool want_z = zlr || c->ss_resident, want_j = jlr || c->ss_resident;
  std::vector<GreenPot> gp;
  if (want_z && want_j) {
    gp.resize((size_t)npot);
    for (int p = 0; p < npot; p++)
      if (const char *e = make_green_pot(r_mesh + (size_t)p * iprpts, jmt[p], jws[p], r_sph[p],
                                         iprpts, gp[p]))
        return fail(-2, e);
  }
  c->ss_valid = 0;
  c->ss_have_matom = 0;
  HIPCHK(hipSetDevice(c->device));
  if (int r = c->ss_r.ensure(npts)) return r;
  if (int r = c->ss_v.ensure(npts)) return r;
  if (int r = c->ss_coef.ensure(4 * npts)) return r;
  if (int r = c->ss_pot.ensure(npot)) return r;
  if (int r = c->ss_iend.ensure(npot)) return r;
  if (int r = c->ss_status.ensure(1)) return r;
  if (int r = c->ss_e.ensure(ne)) return r;
  if (int r = c->ss_prel.ensure(ne)) return r;
  if (int r = c->ss_matom.ensure(ntask)) return r;
  if (want_z)
    if (int r = c->ss_zlr.ensure(ntask * iprpts)) return r;
  if (want_j)
    if (int r = c->ss_jlr.ensure(ntask * iprpts)) return r;
  hipStream_t st = c->stream;
  HIPCHK(hipMemcpyAsync(c->ss_r.p, r_mesh, npts * sizeof(double), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(c->ss_v.p, vr, npts * sizeof(double), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(c->ss_pot.p, hp.data(), npot * sizeof(SingleSitePot),
                        hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(c->ss_iend.p, hiend.data(), npot * sizeof(int), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(c->ss_e.p, energy, ne * sizeof(double2), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(c->ss_prel.p, prel, ne * sizeof(double2), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemsetAsync(c->ss_status.p, 0, sizeof(int), st));
  if (want_z) HIPCHK(hipMemsetAsync(c->ss_zlr.p, 0, ntask * iprpts * sizeof(double2), st));
  if (want_j) HIPCHK(hipMemsetAsync(c->ss_jlr.p, 0, ntask * iprpts * sizeof(double2), st));
  hipLaunchKernelGGL(k_fitpot, dim3((unsigned)((npts + 255) / 256)), dim3(256), 0, st, c->ss_r.p,
                     c->ss_v.p, c->ss_iend.p, iprpts, npot, c->ss_coef.p);
  HIPCHK(hipGetLastError());
  const size_t nthr = (want_j ? 2 : 1) * ntask;  // inward integrations in threads of their own
  hipLaunchKernelGGL(k_single_site, dim3((unsigned)((nthr + 63) / 64)), dim3(64), 0, st,
                     c->ss_r.p, c->ss_coef.p, c->ss_pot.p, c->ss_e.p, c->ss_prel.p, npot, ne,
                     lmax, iprpts, nrelv == 0 ? 1 : 0, c->ss_matom.p,
                     want_z ? c->ss_zlr.p : nullptr, want_j ? c->ss_jlr.p : nullptr,
                     c->ss_status.p);
  HIPCHK(hipGetLastError());
  if (want_j) {
    const size_t npt = ntask * iprpts;
    hipLaunchKernelGGL(k_jlr_scale, dim3((unsigned)((npt + 255) / 256)), dim3(256), 0, st,
                       c->ss_r.p, c->ss_pot.p, c->ss_prel.p, c->ss_matom.p, npot, ne, lmax, iprpts,
                       c->ss_jlr.p);
    HIPCHK(hipGetLastError());
  }
  int status = 0;
  HIPCHK(hipMemcpyAsync(matom, c->ss_matom.p, ntask * sizeof(double2), hipMemcpyDeviceToHost, st));
  if (zlr)
    HIPCHK(hipMemcpyAsync(zlr, c->ss_zlr.p, ntask * iprpts * sizeof(double2),
                          hipMemcpyDeviceToHost, st));
  if (jlr)
    HIPCHK(hipMemcpyAsync(jlr, c->ss_jlr.p, ntask * iprpts * sizeof(double2),
                          hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(&status, c->ss_status.p, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  c->ss_ne = ne;
  c->ss_npot = npot;
  c->ss_lmax = lmax;
  c->ss_iprpts = iprpts;
  c->ss_have_matom = 1;
  if (want_z && want_j) {
    c->ss_valid = 1;
    c->ss_gpot = gp;
    c->ss_hr.assign(r_mesh, r_mesh + npts);
    c->ss_hrsph.assign(r_sph, r_sph + npot);
  }
  if (tmat) {  // diagonal t = 1/matom_l (single_scatterer_nonrel.f:100-111)
    const int kk = nl * nl;
    for (size_t q = 0; q < (size_t)ne * npot; q++) {
      double *tq = tmat + 2 * q * kk * kk;
      std::fill(tq, tq + 2 * (size_t)kk * kk, 0.0);
      int lm = 0;
      for (int l = 0; l <= lmax; l++) {
        const std::complex<double> m(matom[2 * (q * nl + l)], matom[2 * (q * nl + l) + 1]);
        const std::complex<double> t = 1.0 / m;
        for (int mm = -l; mm <= l; mm++, lm++) {
          tq[2 * ((size_t)lm + (size_t)lm * kk)] = t.real();
          tq[2 * ((size_t)lm + (size_t)lm * kk) + 1] = t.imag();
        }
      }
    }
  }
  return status ? 1 : 0;
}

int lsmsx_green_mt(lsmsx_ctx *c, int nsolve, int lmax, int iprpts, int n_spin_pola, int npot,
                   const int *pot_of, const double *r_mesh, const int *jmt, const int *jws,
                   const double *r_sph, const double *tau00, const double *zlr,
                   const double *jlr, double *dos, double *green) {
  if (!c) return fail(-1, "null context");
  if (lmax < 0 || lmax > 3) return fail(-2, "lmax must be 0..3");
  if (nsolve < 0 || npot < 1 || iprpts < 1) return fail(-2, "bad batch size");
  if (n_spin_pola != 1 && n_spin_pola != 2) return fail(-2, "n_spin_pola must be 1 or 2");
  if (nsolve == 0) return 0;
  if (!pot_of || !r_mesh || !jmt || !jws || !r_sph || !tau00 || !zlr || !jlr || !dos)
    return fail(-2, "null argument");
  std::vector<GreenPot> hp((size_t)npot);
  for (int p = 0; p < npot; p++)
    if (const char *e = make_green_pot(r_mesh + (size_t)p * iprpts, jmt[p], jws[p], r_sph[p],
                                       iprpts, hp[p]))
      return fail(-2, e);
  for (int b = 0; b < nsolve; b++)
    if (pot_of[b] < 0 || pot_of[b] >= npot) return fail(-2, "pot_of out of range");
  const int nl = lmax + 1, kk = nl * nl;
  c->ss_valid = 0;  // the radial-function buffers are reused below
  HIPCHK(hipSetDevice(c->device));
  if (int r = c->ss_r.ensure((size_t)npot * iprpts)) return r;
  if (int r = c->gf_pot.ensure(npot)) return r;
  if (int r = c->gf_potof.ensure(nsolve)) return r;
  if (int r = c->gf_tau.ensure((size_t)nsolve * kk * kk)) return r;
  if (int r = c->ss_zlr.ensure((size_t)nsolve * nl * iprpts)) return r;
  if (int r = c->ss_jlr.ensure((size_t)nsolve * nl * iprpts)) return r;
  if (int r = c->gf_dos.ensure(nsolve)) return r;
  if (green)
    if (int r = c->gf_green.ensure((size_t)nsolve * iprpts)) return r;
  hipStream_t st = c->stream;
  HIPCHK(hipMemcpyAsync(c->ss_r.p, r_mesh, (size_t)npot * iprpts * sizeof(double),
                        hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(c->gf_pot.p, hp.data(), npot * sizeof(GreenPot), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(c->gf_potof.p, pot_of, nsolve * sizeof(int), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(c->gf_tau.p, tau00, (size_t)nsolve * kk * kk * sizeof(double2),
                        hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(c->ss_zlr.p, zlr, (size_t)nsolve * nl * iprpts * sizeof(double2),
                        hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(c->ss_jlr.p, jlr, (size_t)nsolve * nl * iprpts * sizeof(double2),
                        hipMemcpyHostToDevice, st));
  if (green)
    HIPCHK(hipMemsetAsync(c->gf_green.p, 0, (size_t)nsolve * iprpts * sizeof(double2), st));
  hipLaunchKernelGGL(k_green_mt, dim3((unsigned)((nsolve + kGreenSolves - 1) / kGreenSolves)),
                     dim3(64), 0, st, c->gf_tau.p, c->ss_zlr.p, c->ss_jlr.p, c->ss_r.p,
                     c->gf_pot.p, c->gf_potof.p, nullptr, nsolve, lmax, iprpts,
                     n_spin_pola == 1 ? 2.0 : 1.0, c->gf_dos.p, green ? c->gf_green.p : nullptr);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(dos, c->gf_dos.p, nsolve * sizeof(double2), hipMemcpyDeviceToHost, st));
  if (green)
    HIPCHK(hipMemcpyAsync(green, c->gf_green.p, (size_t)nsolve * iprpts * sizeof(double2),
                          hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return 0;
}

int lsmsx_green_interstitial(lsmsx_ctx *c, int nsolve, int lmax, int zj_flag, int n_spin_pola,
                             const double *prel, const double *matom_left,
                             const double *matom_right, int ncell, const int *cell_of,
                             const double *r_sph, const int *ncrit, int ngaussr, int maxreg,
                             const double *grwylm, const double *gwwylm, const double *wylm,
                             const double *tau00, double *pzz, double *pzj, double *dos) {
  if (!c) return fail(-1, "null context");
  if (c->lmax_tab < 0) return fail(-2, "tables not set");
  if (lmax < 0 || lmax > 3 || lmax > c->lmax_tab) return fail(-2, "lmax must be 0..3 and <= the tables' lmax");
  if (zj_flag != 0 && zj_flag != 1) return fail(-2, "zj_flag must be 0 or 1");
  if (n_spin_pola != 1 && n_spin_pola != 2) return fail(-2, "n_spin_pola must be 1 or 2");
  if (nsolve < 0 || ncell < 1 || ngaussr < 1 || maxreg < 1) return fail(-2, "bad batch size");
  if (nsolve == 0) return 0;
  if (!prel || !matom_left || !matom_right || !cell_of || !r_sph || !ncrit || !grwylm ||
      !gwwylm || !wylm || !tau00 || !dos)
    return fail(-2, "null argument");
  for (int k = 0; k < ncell; k++)
    if (ncrit[k] < 1 || ncrit[k] - 1 > maxreg) return fail(-2, "ncrit outside 1..maxreg+1");
  for (int b = 0; b < nsolve; b++)
    if (cell_of[b] < 0 || cell_of[b] >= ncell) return fail(-2, "cell_of out of range");
  const int nl = lmax + 1, kk = nl * nl, nw = (2 * lmax + 1) * (lmax + 1);
  const size_t npt = (size_t)ncell * maxreg * ngaussr;
  // one complex buffer: prel | ml | mr | wylm | tau | pzz | pzj | dos
  const size_t o_ml = nsolve, o_mr = o_ml + (size_t)nsolve * nl, o_wy = o_mr + (size_t)nsolve * nl,
               o_tau = o_wy + npt * nw, o_pzz = o_tau + (size_t)nsolve * kk * kk,
               o_pzj = o_pzz + (size_t)nsolve * kk * kk, o_dos = o_pzj + nsolve,
               n_c2 = o_dos + nsolve;
  HIPCHK(hipSetDevice(c->device));
  if (int r = c->it_c2.ensure(n_c2)) return r;
  if (int r = c->it_d.ensure(ncell + 2 * npt)) return r;
  if (int r = c->it_i.ensure(nsolve + ncell)) return r;
  double2 *C2 = c->it_c2.p;
  double *D = c->it_d.p;
  int *Iv = c->it_i.p;
  hipStream_t st = c->stream;
  auto up = [&](void *dst, const void *src, size_t bytes) {
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st);
  };
  HIPCHK(up(C2, prel, nsolve * sizeof(double2)));
  HIPCHK(up(C2 + o_ml, matom_left, (size_t)nsolve * nl * sizeof(double2)));
  HIPCHK(up(C2 + o_mr, matom_right, (size_t)nsolve * nl * sizeof(double2)));
  HIPCHK(up(C2 + o_wy, wylm, npt * nw * sizeof(double2)));
  HIPCHK(up(C2 + o_tau, tau00, (size_t)nsolve * kk * kk * sizeof(double2)));
  HIPCHK(up(D, r_sph, ncell * sizeof(double)));
  HIPCHK(up(D + ncell, grwylm, npt * sizeof(double)));
  HIPCHK(up(D + ncell + npt, gwwylm, npt * sizeof(double)));
  HIPCHK(up(Iv, cell_of, nsolve * sizeof(int)));
  HIPCHK(up(Iv + nsolve, ncrit, ncell * sizeof(int)));
  hipLaunchKernelGGL(k_interstitial, dim3(nsolve), dim3(256), 0, st, lmax, zj_flag, C2, C2 + o_ml,
                     C2 + o_mr, D, Iv, Iv + nsolve, ngaussr, maxreg, D + ncell, D + ncell + npt,
                     C2 + o_wy, c->cgnt.p, c->lmax_tab, C2 + o_tau, n_spin_pola == 1 ? 2.0 : 1.0,
                     C2 + o_pzz, C2 + o_pzj, C2 + o_dos);
  if (int r = check_launch("k_interstitial")) return r;
  HIPCHK(hipMemcpyAsync(dos, C2 + o_dos, nsolve * sizeof(double2), hipMemcpyDeviceToHost, st));
  if (pzz)
    HIPCHK(hipMemcpyAsync(pzz, C2 + o_pzz, (size_t)nsolve * kk * kk * sizeof(double2),
                          hipMemcpyDeviceToHost, st));
  if (pzj)
    HIPCHK(hipMemcpyAsync(pzj, C2 + o_pzj, nsolve * sizeof(double2), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return 0;
}

int lsmsx_set_single_site_resident(lsmsx_ctx *c, int on) {
  if (!c) return fail(-1, "null context");
  c->ss_resident = on ? 1 : 0;
  return 0;
}

int lsmsx_green_mt_resident(lsmsx_ctx *c, int nsolve, int lmax, int iprpts, int n_spin_pola,
                            const int *pair_of, const int *jmt_of_pot, const double *tau00,
                            double *dos, double *green) {
  if (!c) return fail(-1, "null context");
  if (nsolve < 0) return fail(-2, "bad batch size");
  if (n_spin_pola != 1 && n_spin_pola != 2) return fail(-2, "n_spin_pola must be 1 or 2");
  if (nsolve == 0) return 0;
  if (!pair_of || !tau00 || !dos) return fail(-2, "null argument");
  if (!c->ss_valid)
    return fail(-3, "no radial functions on the device: call lsmsx_single_site_nonrel with "
                    "lsmsx_set_single_site_resident on (and no lsmsx_green_mt in between)");
  // the caller's buffers are sized by its lmax and iprpts: they must be the
  // resident batch's, or the copies below would over- or under-run them
  if (lmax != c->ss_lmax || iprpts != c->ss_iprpts)
    return fail(-2, "lmax/iprpts differ from those of the resident single-site batch (lmax " +
                        std::to_string(c->ss_lmax) + ", iprpts " + std::to_string(c->ss_iprpts) +
                        ")");
  const int npair = c->ss_ne * c->ss_npot;
  std::vector<GreenPot> gp = c->ss_gpot;
  if (jmt_of_pot)  // e.g. jws for ASA potentials (energyContourIntegration.cpp:519)
    for (int p = 0; p < c->ss_npot; p++)
      if (const char *e = make_green_pot(c->ss_hr.data() + (size_t)p * iprpts, jmt_of_pot[p],
                                         gp[p].jws, c->ss_hrsph[p], iprpts, gp[p]))
        return fail(-2, e);
  std::vector<int> pot_of((size_t)nsolve);
  for (int b = 0; b < nsolve; b++) {
    if (pair_of[b] < 0 || pair_of[b] >= npair) return fail(-2, "pair_of out of range");
    pot_of[b] = pair_of[b] % c->ss_npot;
  }
  const int nl = lmax + 1, kk = nl * nl;
  HIPCHK(hipSetDevice(c->device));
  if (int r = c->gf_pot.ensure(c->ss_npot)) return r;
  if (int r = c->gf_potof.ensure(nsolve)) return r;
  if (int r = c->gf_src.ensure(nsolve)) return r;
  if (int r = c->gf_tau.ensure((size_t)nsolve * kk * kk)) return r;
  if (int r = c->gf_dos.ensure(nsolve)) return r;
  if (green)
    if (int r = c->gf_green.ensure((size_t)nsolve * iprpts)) return r;
  hipStream_t st = c->stream;
  HIPCHK(hipMemcpyAsync(c->gf_pot.p, gp.data(), c->ss_npot * sizeof(GreenPot),
                        hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(c->gf_potof.p, pot_of.data(), nsolve * sizeof(int), hipMemcpyHostToDevice,
                        st));
  HIPCHK(hipMemcpyAsync(c->gf_src.p, pair_of, nsolve * sizeof(int), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(c->gf_tau.p, tau00, (size_t)nsolve * kk * kk * sizeof(double2),
                        hipMemcpyHostToDevice, st));
  if (green)
    HIPCHK(hipMemsetAsync(c->gf_green.p, 0, (size_t)nsolve * iprpts * sizeof(double2), st));
  hipLaunchKernelGGL(k_green_mt, dim3((unsigned)((nsolve + kGreenSolves - 1) / kGreenSolves)),
                     dim3(64), 0, st, c->gf_tau.p, c->ss_zlr.p, c->ss_jlr.p, c->ss_r.p,
                     c->gf_pot.p, c->gf_potof.p, c->gf_src.p, nsolve, lmax, iprpts,
                     n_spin_pola == 1 ? 2.0 : 1.0, c->gf_dos.p, green ? c->gf_green.p : nullptr);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(dos, c->gf_dos.p, nsolve * sizeof(double2), hipMemcpyDeviceToHost, st));
  if (green)
    HIPCHK(hipMemcpyAsync(green, c->gf_green.p, (size_t)nsolve * iprpts * sizeof(double2),
                          hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return 0;
}

int lsmsx_tmats_from_single_site(lsmsx_ctx *c, long ldT, int numStore, long blk, int groupSize,
                                 const int *store_of_pot, const int *spin_of_pot) {
  if (!c) return fail(-1, "null context");
  if (!c->ss_have_matom) return fail(-3, "no single-site batch on the device");
  if (!store_of_pot || !spin_of_pot) return fail(-2, "null argument");
  const int ne = c->ss_ne, npot = c->ss_npot, kkr = (c->ss_lmax + 1) * (c->ss_lmax + 1);
  if (ldT < blk * groupSize || numStore <= 0 || blk <= 0 || groupSize <= 0)
    return fail(-2, "inconsistent t-matrix store dimensions");
  if (ne > groupSize) return fail(-2, "more energies than store slots");
  for (int p = 0; p < npot; p++) {
    if (store_of_pot[p] < 0 || store_of_pot[p] >= numStore) return fail(-2, "store column out of range");
    if (spin_of_pot[p] < 0 || (long)(spin_of_pot[p] + 1) * kkr * kkr > blk)
      return fail(-2, "spin block outside blkSizeTmatStore");
  }
  HIPCHK(hipSetDevice(c->device));
  if (int r = c->tmat.ensure((size_t)ldT * numStore)) return r;
  if (int r = c->ss_storeof.ensure(2 * (size_t)npot)) return r;
  std::vector<int> hs(2 * (size_t)npot);
  std::copy(store_of_pot, store_of_pot + npot, hs.begin());
  std::copy(spin_of_pot, spin_of_pot + npot, hs.begin() + npot);
  hipStream_t st = c->stream;
  // the store starts from zero, as local.tmatStore = 0 before the solve
  // (energyContourIntegration.cpp:420)
  HIPCHK(hipMemsetAsync(c->tmat.p, 0, sizeof(double2) * (size_t)ldT * numStore, st));
  HIPCHK(hipMemcpyAsync(c->ss_storeof.p, hs.data(), 2 * npot * sizeof(int), hipMemcpyHostToDevice,
                        st));
  const long n = (long)ne * npot * kkr * kkr;
  hipLaunchKernelGGL(k_tmat_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                     c->ss_matom.p, c->ss_storeof.p, c->ss_storeof.p + npot, ne, npot, c->ss_lmax,
                     ldT, blk, c->tmat.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(st));
  c->ldT = ldT;
  c->blk = blk;
  c->nstore = numStore;
  c->group = groupSize;
  return 0;
}

int lsmsx_green_mt_canted(lsmsx_ctx *c, int nsolve, int lmax, int iprpts, int npot,
                          const int *pot_of, const double *r_mesh, const int *jmt,
                          const int *jws, const double *r_sph, const double *tau00_l,
                          const double *zlr, const double *jlr, const double *wxyz,
                          double *dos, double *green) {
  if (!c) return fail(-1, "null context");
  if (lmax < 0 || lmax > 3) return fail(-2, "lmax must be 0..3");
  if (nsolve < 0 || npot < 1 || iprpts < 1) return fail(-2, "bad batch size");
  if (nsolve == 0) return 0;
  if (!pot_of || !r_mesh || !jmt || !jws || !r_sph || !tau00_l || !zlr || !jlr || !wxyz || !dos)
    return fail(-2, "null argument");
  std::vector<GreenPot> hp((size_t)npot);
  for (int p = 0; p < npot; p++)
    if (const char *e = make_green_pot(r_mesh + (size_t)p * iprpts, jmt[p], jws[p], r_sph[p],
                                       iprpts, hp[p]))
      return fail(-2, e);
  for (int b = 0; b < nsolve; b++)
    if (pot_of[b] < 0 || pot_of[b] >= npot) return fail(-2, "pot_of out of range");
  const int nl = lmax + 1, kk = nl * nl;
  const size_t nrad = (size_t)nsolve * 2 * nl * iprpts;
  c->ss_valid = 0;  // the radial-function buffers are reused below
  HIPCHK(hipSetDevice(c->device));
  if (int r = c->ss_r.ensure((size_t)npot * iprpts)) return r;
  if (int r = c->gf_pot.ensure(npot)) return r;
  if (int r = c->gf_potof.ensure(nsolve)) return r;
  if (int r = c->gf_tau.ensure((size_t)nsolve * 4 * kk * kk)) return r;
  if (int r = c->ss_zlr.ensure(nrad)) return r;
  if (int r = c->ss_jlr.ensure(nrad)) return r;
  if (int r = c->gf_w.ensure((size_t)nsolve * 12)) return r;
  if (int r = c->gf_dos.ensure((size_t)nsolve * 4)) return r;
  if (green)
    if (int r = c->gf_green.ensure((size_t)nsolve * 4 * iprpts)) return r;
  hipStream_t st = c->stream;
  HIPCHK(hipMemcpyAsync(c->ss_r.p, r_mesh, (size_t)npot * iprpts * sizeof(double),
                        hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(c->gf_pot.p, hp.data(), npot * sizeof(GreenPot), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(c->gf_potof.p, pot_of, nsolve * sizeof(int), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(c->gf_tau.p, tau00_l, (size_t)nsolve * 4 * kk * kk * sizeof(double2),
                        hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(c->ss_zlr.p, zlr, nrad * sizeof(double2), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(c->ss_jlr.p, jlr, nrad * sizeof(double2), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(c->gf_w.p, wxyz, (size_t)nsolve * 12 * sizeof(double2),
                        hipMemcpyHostToDevice, st));
  if (green)
    HIPCHK(hipMemsetAsync(c->gf_green.p, 0, (size_t)nsolve * 4 * iprpts * sizeof(double2), st));
  hipLaunchKernelGGL(k_green_mt_canted, dim3((unsigned)((nsolve + 1) / 2)), dim3(64), 0, st,
                     c->gf_tau.p, c->ss_zlr.p, c->ss_jlr.p, c->gf_w.p, c->ss_r.p, c->gf_pot.p,
                     c->gf_potof.p, nsolve, lmax, iprpts, c->gf_dos.p,
                     green ? c->gf_green.p : nullptr);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(dos, c->gf_dos.p, (size_t)nsolve * 4 * sizeof(double2),
                        hipMemcpyDeviceToHost, st));
  if (green)
    HIPCHK(hipMemcpyAsync(green, c->gf_green.p, (size_t)nsolve * 4 * iprpts * sizeof(double2),
                          hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return 0;
}

int lsmsx_tmats_from_single_site_canted(lsmsx_ctx *c, long ldT, int numStore, long blk,
                                        int groupSize, int natom, const int *store_of_atom,
                                        const int *pot_up, const int *pot_dn, const double *ubr,
                                        const double *ubrd) {
  if (!c) return fail(-1, "null context");
  if (!c->ss_have_matom) return fail(-3, "no single-site batch on the device");
  if (natom <= 0 || !store_of_atom || !pot_up || !pot_dn || !ubr || !ubrd)
    return fail(-2, "null argument or no atoms");
  const int ne = c->ss_ne, npot = c->ss_npot, kkr = (c->ss_lmax + 1) * (c->ss_lmax + 1);
  if (ldT < blk * groupSize || numStore <= 0 || blk <= 0 || groupSize <= 0)
    return fail(-2, "inconsistent t-matrix store dimensions");
  if (ne > groupSize) return fail(-2, "more energies than store slots");
  if ((long)4 * kkr * kkr > blk) return fail(-2, "blkSizeTmatStore below (2 kkrsz)^2");
  std::vector<int> map(3 * (size_t)natom);
  std::vector<double2> uu(8 * (size_t)natom);
  for (int a = 0; a < natom; a++) {
    if (store_of_atom[a] < 0 || store_of_atom[a] >= numStore)
      return fail(-2, "store column out of range");
    if (pot_up[a] < 0 || pot_up[a] >= npot || pot_dn[a] < 0 || pot_dn[a] >= npot)
      return fail(-2, "potential index out of range");
    map[3 * a] = store_of_atom[a];
    map[3 * a + 1] = pot_up[a];
    map[3 * a + 2] = pot_dn[a];
    for (int k = 0; k < 4; k++) {
      uu[8 * a + k] = make_double2(ubr[2 * (4 * a + k)], ubr[2 * (4 * a + k) + 1]);
      uu[8 * a + 4 + k] = make_double2(ubrd[2 * (4 * a + k)], ubrd[2 * (4 * a + k) + 1]);
    }
  }
  HIPCHK(hipSetDevice(c->device));
  if (int r = c->tmat.ensure((size_t)ldT * numStore)) return r;
  if (int r = c->ss_storeof.ensure(3 * (size_t)natom)) return r;
  if (int r = c->ss_uu.ensure(8 * (size_t)natom)) return r;
  hipStream_t st = c->stream;
  HIPCHK(hipMemsetAsync(c->tmat.p, 0, sizeof(double2) * (size_t)ldT * numStore, st));
  HIPCHK(hipMemcpyAsync(c->ss_storeof.p, map.data(), map.size() * sizeof(int),
                        hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(c->ss_uu.p, uu.data(), uu.size() * sizeof(double2), hipMemcpyHostToDevice,
                        st));
  const long n = (long)ne * natom * 4 * kkr * kkr;
  hipLaunchKernelGGL(k_tmat_fill_canted, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                     c->ss_matom.p, c->ss_storeof.p, c->ss_uu.p, ne, npot, natom, c->ss_lmax, ldT,
                     blk, c->tmat.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(st));
  c->ldT = ldT;
  c->blk = blk;
  c->nstore = numStore;
  c->group = groupSize;
  return 0;
}

}  // extern "C"
