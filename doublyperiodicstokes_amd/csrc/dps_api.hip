// dps_api.hip -- C ABI of the doubly periodic Stokes solver (see
// include/dpstokes_hip.h): parameter validation, constant tables, device
// memory, rocFFT plans and the Mdot launch sequence.
//
// Launch sequence of one Mdot (reference FCM.Mdot, FCM.py:334-441):
//   spread F (+ curl of T)  ->  rocFFT r2c (x-y)  ->  z Chebyshev forward
//   ->  per-mode FP64 BVP (+ wall-correction coefficients, k = 0 mode)
//   ->  z Chebyshev inverse (+ correction profiles)  ->  rocFFT c2r
//   ->  interpolate U (+ derivative kernels for Omega)
#include <hip/hip_runtime.h>
#include <rocfft/rocfft.h>
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>
#include "../../include/dpstokes_hip.h"
#include "dps_internal.h"
#include "dps_tables.h"

using namespace dps;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg)
{
  g_err = msg;
  return code;
}

#define HIPCHK(x)                                                                       \
  do {                                                                                  \
    hipError_t e_ = (x);                                                                \
    if (e_ != hipSuccess)                                                               \
      return fail(DPS_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_));         \
  } while (0)

#define FFTCHK(x)                                                                       \
  do {                                                                                  \
    rocfft_status s_ = (x);                                                             \
    if (s_ != rocfft_status_success)                                                    \
      return fail(DPS_ERR_FFT, std::string(#x) + " failed, status " + std::to_string((int)s_)); \
  } while (0)

int g_rocfft_users = 0;

}  // namespace

// Scratch of dps_mdot_adjoint_noalloc: the adjoint's second velocity grid and
// its derivative records, in one array per GPU that lives as long as the
// process.  Host code never uses the array's name as an address: each device's
// copy is reached through the pointer hipGetSymbolAddress gave on that device
// (g_scratch_ptr), which is what kernels receive.  Its use by host code
// (hipGetSymbolAddress) makes the compiler emit it with global binding in the
// device code object, where the loader can find it; no attribute may force it
// to stay local.
static __device__ __align__(256) unsigned char scratch[DPS_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// Device address of this device's scratch array, looked up once per device
// with that device current (called by dps_create).
int scratch_init(int dev)
{
  if (dev < 0 || dev >= kMaxDevices) return fail(DPS_ERR_ARG, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev]) return DPS_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(DPS_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch_ptr[dev] = (unsigned char*)p;
  return DPS_OK;
}

unsigned char* scratch_of(int dev)
{
  if (dev < 0 || dev >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch_ptr[dev];
}

}  // namespace

struct dps_handle {
  dps_params P;
  int device = 0;   // HIP device current at dps_create
  int N = 0;
  int geometry = 0, prec = 0;
  double hx = 0, hy = 0, Lz = 0, H = 0, hz = 0;   // hz: periodic z spacing
  int Nx = 0, Ny = 0, Nz = 0, Nkx = 0, Nm = 0;
  KernelDesc km{}, kd{};
  bool has_dip = false;
  bool reg = false;   // regular torque kernel with spectral curls
  ZExt zm, zd;
  std::vector<double> zpts, zwts;
  ChebTables cheb;
  ChebTransform ct;
  int TX = 16, TY = 4, ZC = 32, ntx = 0, nty = 0, ntiles = 0;
  int zc_min = 8;   // shortest z chunk the spreading lists are sized for
  size_t lds_max = 65536;   // LDS bytes one workgroup may use on this device
  size_t tsz = 8;   // bytes per grid real

  std::vector<void*> allocs;
  double* d_pos = nullptr;
  int* d_err = nullptr;
  int* h_err = nullptr;
  int4 *d_sup_m = nullptr, *d_sup_d = nullptr;
  void *d_rec_m = nullptr, *d_rec_d = nullptr;
  int rs_m = 0, rs_d = 0;
  int4* d_zrange = nullptr;
  float* d_amax = nullptr;   // partial maxima of |F|, |T| for fixed-point spreading
  // spreading entry lists of each pass (rebuilt whenever positions change):
  // 0 forces, 1 forces and torques with the derivative kernels, 2 torques
  // with the regular kernel
  struct SpreadLists {
    bool used = false;
    int nkind = 1, rows0 = 0;
    bool dip0 = false;   // kind-0 entries use the torque kernel's supports
    int *start = nullptr, *counts = nullptr, *fill = nullptr, *entries = nullptr;
    int cap = 0;
  } lists[3];
  int *d_key = nullptr, *d_rank = nullptr, *d_counts = nullptr, *d_start = nullptr, *d_order = nullptr;
  void* d_grid = nullptr;
  void *d_specA = nullptr, *d_specB = nullptr;
  // regular torque kernel: torque / half-vorticity grid and two more spectra
  void *d_gridT = nullptr, *d_specT = nullptr, *d_specC = nullptr;
  // adjoint (allocated on first use): second velocity grid, force-kernel
  // derivative records
  void* d_gridA = nullptr;
  int4* d_sup_md = nullptr;
  void* d_rec_md = nullptr;
  int4* d_zrange_md = nullptr;
  int rs_md = 0;
  void* d_rec_d2 = nullptr;   // torque-kernel second-derivative records
  int rs_d2 = 0;
  double *d_zext_m = nullptr, *d_zw_m = nullptr, *d_zext_d = nullptr, *d_zw_d = nullptr;
  double *d_fA_m = nullptr, *d_fB_m = nullptr, *d_fA_d = nullptr, *d_fB_d = nullptr;
  int *d_kt_m = nullptr, *d_kt_d = nullptr;
  void *d_TfE = nullptr, *d_TfO = nullptr, *d_TiE = nullptr, *d_TiO = nullptr;
  double *d_zpts = nullptr, *d_kxs = nullptr, *d_kys = nullptr;
  double *d_sd0 = nullptr, *d_sm2 = nullptr, *d_sp2 = nullptr, *d_fm1 = nullptr, *d_fp1 = nullptr;
  double *d_R1 = nullptr, *d_R2 = nullptr, *d_L1 = nullptr, *d_L2 = nullptr;
  int* d_special = nullptr;
  int nspecial = 0;
  double2* d_u2 = nullptr;
  double *s_ib = nullptr, *s_ab = nullptr;
  double2 *s_P = nullptr, *s_CP = nullptr, *s_DP = nullptr, *s_U0 = nullptr, *s_U1 = nullptr, *s_U2 = nullptr;
  double2 *d_corr = nullptr, *d_corr2 = nullptr;
  double* d_stage = nullptr;   // FP64 reference-layout staging for stage entry points
  // fused spectral solve (lane groups)
  bool fused = false;
  int sp_lpm = 0, sp_e = 0, sp_G = 0;
  void *d_TfP = nullptr, *d_TiP = nullptr;
  float *d_TfM = nullptr, *d_TiM = nullptr;
  double *d_fib = nullptr, *d_fab = nullptr, *d_fG = nullptr;
  int* d_fidx = nullptr;   // [Nm] factor-table entry of each mode
  int n_fac = 0;           // distinct |k|^2 (factor-table entries)

  rocfft_plan plan_fwd = nullptr, plan_inv = nullptr;
  bool fft_y = false;        // rocFFT runs the x pass only; k_fft_y the y pass
  bool fft_x = false;        // with fft_y: k_fft_x runs the x pass (no rocFFT plan)
  void* d_twx = nullptr;     // x-pass twiddles exp(-2 pi i k / Nx)
  void* d_twy = nullptr;     // y-pass twiddles exp(-2 pi i k / Ny)
  rocfft_execution_info info = nullptr;
  void* fft_work = nullptr;
  size_t fft_work_size = 0;

  bool positions_set = false;
  bool prof = false;
  hipEvent_t ev[9] = {};
  bool have_timings = false;
  bool have_reg_events = false;   // the profiled Mdot took the regular-kernel path

  template <class X>
  int alloc(X** p, size_t count)
  {
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void**)p, count * sizeof(X));
    if (e != hipSuccess) {
      *p = nullptr;
      return fail(DPS_ERR_ALLOC, "hipMalloc of " + std::to_string(count * sizeof(X)) + " bytes failed");
    }
    allocs.push_back((void*)*p);
    return DPS_OK;
  }
  template <class X>
  int upload(X** p, const std::vector<X>& v)
  {
    int rc = alloc(p, v.size());
    if (rc) return rc;
    if (!v.empty()) {
      hipError_t e = hipMemcpy(*p, v.data(), v.size() * sizeof(X), hipMemcpyHostToDevice);
      if (e != hipSuccess) return fail(DPS_ERR_HIP, "upload failed");
    }
    return DPS_OK;
  }
  void release()
  {
    for (void* p : allocs) (void)hipFree(p);
    allocs.clear();
    if (h_err) { (void)hipHostFree(h_err); h_err = nullptr; }
    if (plan_fwd) rocfft_plan_destroy(plan_fwd);
    if (plan_inv) rocfft_plan_destroy(plan_inv);
    if (info) rocfft_execution_info_destroy(info);
    plan_fwd = plan_inv = nullptr;
    info = nullptr;
    for (auto& e : ev)
      if (e) { (void)hipEventDestroy(e); e = nullptr; }
  }
};

// A set of independent solvers run as one: every instance's Mdot captured,
// forked over a few streams, into one HIP graph that is replayed while the
// argument pointers stay the same.
struct dps_batch {
  std::vector<dps_handle*> hs;
  std::vector<hipStream_t> streams;
  std::vector<hipEvent_t> ev_join;
  hipEvent_t ev_fork = nullptr;
  hipStream_t cap = nullptr;   // capture origin (the caller's stream may be the null stream)
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<const void*> key;   // pointers and dtype the graph was captured with
  void drop_graph()
  {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    key.clear();
  }
  void release()
  {
    drop_graph();
    for (auto e : ev_join) (void)hipEventDestroy(e);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    for (auto st : streams) (void)hipStreamDestroy(st);
    if (cap) (void)hipStreamDestroy(cap);
    cap = nullptr;
    ev_join.clear();
    streams.clear();
    ev_fork = nullptr;
  }
};

namespace {

int setup_kernel(dps_handle* h, int w, double alpha, double beta, bool dipole, KernelDesc* K)
{
  if (w < 1) return fail(DPS_ERR_ARG, "kernel width must be >= 1");
  const double betaw = (beta > 0) ? beta : 1.785 * w;
  const double a = (alpha > 0) ? alpha : w * h->hx / 2;
  K->betaw = betaw;
  K->alpha = a;
  const double norm = es_norm(a, betaw);
  K->inv_norm3 = 1.0 / (norm * (norm * norm));
  K->thresh = dipole ? dipole_thresh(a, betaw) : a;
  if (!(K->thresh > 0) || !std::isfinite(K->thresh)) return fail(DPS_ERR_ARG, "invalid kernel cutoff");
  // |phi'| = phi |dfac| is smooth on [0, thresh] for the dipole cutoff
  // (thresh < alpha): a sampled maximum with a margin bounds it
  K->dmax = 0.0;
  if (dipole) {
    const double a2 = a * a;
    for (int i = 0; i <= 4096; ++i) {
      const double x = K->thresh * i / 4096.0;
      K->dmax = std::max(K->dmax, std::fabs(es_phi(x, betaw, a2) * es_dfac(x, betaw, a2)));
    }
    K->dmax *= 1.05;
  }
  K->wx = (int)std::lround(2 * K->thresh / h->hx);
  K->wy = (int)std::lround(2 * K->thresh / h->hy);
  if (K->wx < 1 || K->wy < 1 || K->wx > DPS_MAX_W || K->wy > DPS_MAX_W)
    return fail(DPS_ERR_ARG, "kernel support of " + std::to_string(K->wx) + "x" +
                                 std::to_string(K->wy) + " points is out of range [1, 32]");
  return DPS_OK;
}

int build_tiles(dps_handle* h, int wmax);

// Triply periodic: uniform periodic z nodes k hz, hz = Lz / Nz (FCM.py:182-186,
// GridAndKernelConfig.py:393-417); z supports like x and y; a 3-D FFT and the
// Fourier solve replace the Chebyshev machinery.
int build_periodic(dps_handle* h, int wmax)
{
  h->hz = h->Lz / h->Nz;
  for (int k = 0; k < (h->has_dip ? 2 : 1); ++k) {
    KernelDesc* K = k ? &h->kd : &h->km;
    K->wzmax = (int)std::lround(2 * K->thresh / h->hz);
    K->nzeff = h->Nz;
    K->ext_up = 0;
    if (K->wzmax < 1 || K->wzmax > DPS_MAX_W || K->wzmax > h->Nz)
      return fail(DPS_ERR_ARG, "kernel support in z is out of range for the periodic z grid");
  }
  h->fused = false;
  return build_tiles(h, wmax);
}

int build(dps_handle* h)
{
  const dps_params& P = h->P;
  h->geometry = P.mode;
  h->prec = P.precision;
  h->Nx = P.nx;
  h->Ny = P.ny;
  h->Nz = P.nz;
  h->hx = P.Lx / P.nx;
  h->hy = P.Ly / P.ny;
  h->Lz = P.zmax - P.zmin;
  h->H = h->Lz / 2;
  h->Nkx = P.nx / 2 + 1;
  h->Nm = h->Ny * h->Nkx;
  h->tsz = (h->prec == DPS_PREC_MIXED) ? 4 : 8;
  int rc = setup_kernel(h, P.w, P.alpha, P.beta, false, &h->km);
  if (rc) return rc;
  h->has_dip = P.w_d > 0;
  if (P.reg_kernel != 0 && P.reg_kernel != 1) return fail(DPS_ERR_ARG, "reg_kernel must be 0 or 1");
  h->reg = h->has_dip && P.reg_kernel == 1;
  if (h->has_dip) {
    rc = setup_kernel(h, P.w_d, P.alpha_d, P.beta_d, !h->reg, &h->kd);
    if (rc) return rc;
  }
  const int wmax = std::max(std::max(h->km.wx, h->km.wy), h->has_dip ? std::max(h->kd.wx, h->kd.wy) : 0);
  if (h->Nx < 2 * wmax || h->Ny < 2 * wmax)
    return fail(DPS_ERR_ARG, "nx and ny must be at least twice the kernel support");
  if (h->geometry == DPS_MODE_PERIODIC) return build_periodic(h, wmax);
  // z grid: Chebyshev nodes on [0, Lz] (FCM.py:220)
  clencurt(h->zpts, h->zwts, 0, h->Lz, h->Nz);
  const bool wall_top = (h->geometry == DPS_MODE_SLIT);
  const bool wall_bot = (h->geometry == DPS_MODE_BOTTOM || h->geometry == DPS_MODE_SLIT);
  h->zm = build_zext(h->zpts, h->zwts, h->km.thresh, wall_top, wall_bot);
  h->km.nzeff = h->zm.nzeff;
  h->km.ext_up = h->zm.ext_up;
  h->km.wzmax = h->zm.wzmax;
  if (h->zm.ext_up != h->zm.ext_down || h->zm.ext_up >= h->Nz)
    return fail(DPS_ERR_ARG, "kernel support is too wide for the z grid");
  if (h->has_dip) {
    h->zd = build_zext(h->zpts, h->zwts, h->kd.thresh, wall_top, wall_bot);
    h->kd.nzeff = h->zd.nzeff;
    h->kd.ext_up = h->zd.ext_up;
    h->kd.wzmax = h->zd.wzmax;
    if (h->zd.ext_up != h->zd.ext_down || h->zd.ext_up >= h->Nz)
      return fail(DPS_ERR_ARG, "dipole kernel support is too wide for the z grid");
  }
  h->cheb = build_cheb_tables(h->Nz, h->H);
  h->ct = build_cheb_transform(h->Nz);
  h->fused = (P.solver == DPS_SOLVER_AUTO) && spectral_shape(h->Nz, &h->sp_lpm, &h->sp_e);
  if (h->fused && h->reg) {
    // the regular torque kernel's curls act on Chebyshev coefficients: around
    // the fused solve they need the matrix-core z transforms
    ZArgs Z{};
    Z.N = h->Nz;
    Z.NH = (h->Nz + 1) / 2;
    Z.NE = (h->Nz + 1) / 2;
    Z.NO = h->Nz / 2;
    h->fused = (h->prec == DPS_PREC_MIXED) ? zgemm_supported<float>(Z) : zgemm_supported<double>(Z);
  }
  if (h->fused) {
    // up to 64 KiB per workgroup keeps several workgroups per CU; beyond
    // that, what the device allows, else the unfused chain
    const bool mx = h->prec == DPS_PREC_MIXED;
    h->sp_G = spectral_modes_per_block(h->sp_lpm, h->sp_e, mx);
    const size_t lds = spectral_lds_bytes(h->sp_lpm, h->sp_e, h->sp_G, mx);
    if (lds > 65536) {
      int dev = 0, lds_max = 0;
      h->fused = hipGetDevice(&dev) == hipSuccess &&
                 hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess &&
                 lds <= (size_t)lds_max;
    }
  }
  if (!h->fused) {
    ZArgs Z;
    Z.N = h->Nz;
    Z.NH = h->ct.NH;
    if (z_lds_bytes(Z, h->prec == DPS_PREC_MIXED) > 160 * 1024)
      return fail(DPS_ERR_ARG, "nz is too large for the z-transform kernels");
  }
  return build_tiles(h, wmax);
}

// spreading tiles and particle records (both z grids)
int build_tiles(dps_handle* h, int wmax)
{
  // tiles for spreading
  h->TX = SP_TX;
  h->TY = SP_TY;
  // 16 rows per z chunk: more, smaller workgroups beat longer chunks with
  // fewer entries replicated across chunks (slit, N=16384, 256x256x96 grid:
  // 8 rows 213 us, 12: 203, 16: 199, 32: 247)
  h->ZC = std::min(16, h->Nz);
  h->ntx = (h->Nx + h->TX - 1) / h->TX;
  h->nty = (h->Ny + h->TY - 1) / h->TY;
  h->ntiles = h->ntx * h->nty;
  // particle weight records: mono [wx][wy][z spread][z interp]; dipole
  // [wx][wy][dx][dy][z spread][dz spread][z interp][dz interp], z parts by
  // interior row from sup.z
  h->rs_m = h->km.wx + h->km.wy + 2 * h->km.wzmax;
  h->rs_d = !h->has_dip ? 0
            : h->reg ? h->kd.wx + h->kd.wy + 2 * h->kd.wzmax
                     : 2 * h->kd.wx + 2 * h->kd.wy + 4 * h->kd.wzmax;
  {
    // the spreading workgroup's LDS stays within 64 KB (spread_lds_cap), so
    // that several workgroups share a CU (slit, N = 16384, 256^2 x 96: 8 rows
    // in 62 KB 121 us, 16 rows in 71 KB 150 us)
    int dev = 0, lim = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && lim > 0)
      h->lds_max = (size_t)lim;
    SpreadArgs S{};
    S.TX = h->TX; S.TY = h->TY; S.ZC = h->ZC; S.rs_max = std::max(h->rs_m, h->rs_d);
    if (h->has_dip && !h->reg)   // paired force-torque entries
      S.rs_max = std::max(S.rs_max, h->km.wx + h->km.wy + h->km.wzmax + 2 * (h->kd.wx + h->kd.wy + h->kd.wzmax) + 7);
    if (spread_npf(S.rs_max) == 0) return fail(DPS_ERR_ARG, "kernel z support too wide for the spreading records");
    while (spread_lds_bytes(S, h->tsz) > std::min<size_t>(65536, h->lds_max) && S.ZC > 1) S.ZC /= 2;
    if (spread_lds_bytes(S, h->tsz) > std::min<size_t>(65536, h->lds_max))
      return fail(DPS_ERR_ARG, "spreading tile does not fit in LDS");
    h->ZC = S.ZC;
  }
  // the spreading passes and their entry lists
  h->zc_min = std::min(h->ZC, 8);
  h->lists[0].used = true;
  if (h->has_dip && !h->reg) {
    h->lists[1].used = true;
    const bool pair = h->km.wx == h->kd.wx && h->km.wy == h->kd.wy;
    h->lists[1].nkind = pair ? 1 : 2;
    h->lists[1].rows0 = pair ? 2 : 0;
  }
  if (h->has_dip && h->reg) {
    h->lists[2].used = true;
    h->lists[2].rows0 = 1;
    h->lists[2].dip0 = true;
  }
  for (const auto& L : h->lists) {
    if (!L.used) continue;
    const int S0 = L.rows0 == 0 ? h->km.wzmax : L.rows0 == 1 ? h->kd.wzmax : h->km.wzmax + h->kd.wzmax;
    const int S = std::max(S0, L.nkind > 1 ? h->kd.wzmax : 0);
    if ((S + h->zc_min - 1) / h->zc_min + 2 > DPS_LIST_MAXC)
      return fail(DPS_ERR_ARG, "kernel z support too long for the spreading lists");
  }
  return DPS_OK;
}

template <class T>
std::vector<T> cast_vec(const std::vector<double>& v)
{
  std::vector<T> r(v.size());
  for (size_t i = 0; i < v.size(); ++i) r[i] = (T)v[i];
  return r;
}

int allocate_fused(dps_handle* h);

int allocate(dps_handle* h)
{
  int rc;
  const int N = h->N;
  const size_t ng = (size_t)3 * h->Nz * h->Ny * h->Nx;
  const size_t nspec = (size_t)3 * h->Nz * h->Nm;
  const size_t ts = h->tsz;
  if ((rc = h->alloc(&h->d_pos, (size_t)3 * N))) return rc;
  if ((rc = h->alloc(&h->d_err, 1))) return rc;
  if (hipHostMalloc((void**)&h->h_err, sizeof(int)) != hipSuccess) return fail(DPS_ERR_ALLOC, "pinned alloc failed");
  if ((rc = h->alloc(&h->d_sup_m, N))) return rc;
  if ((rc = h->alloc((unsigned char**)&h->d_rec_m, (size_t)N * h->rs_m * ts))) return rc;
  if (h->has_dip) {
    if ((rc = h->alloc(&h->d_sup_d, N))) return rc;
    if ((rc = h->alloc((unsigned char**)&h->d_rec_d, (size_t)N * h->rs_d * ts))) return rc;
  }
  if ((rc = h->alloc(&h->d_zrange, N))) return rc;
  if ((rc = h->alloc(&h->d_amax, 2 * DPS_AMAX_BLOCKS))) return rc;
  if ((rc = h->alloc(&h->d_key, N))) return rc;
  if ((rc = h->alloc(&h->d_rank, N))) return rc;
  if ((rc = h->alloc(&h->d_order, N))) return rc;
  if ((rc = h->alloc(&h->d_counts, h->ntiles))) return rc;
  if ((rc = h->alloc(&h->d_start, h->ntiles + 1))) return rc;
  {
    // spreading lists, sized for the shortest z chunk
    const size_t ncells = (size_t)h->ntiles * ((h->Nz + h->zc_min - 1) / h->zc_min);
    for (auto& L : h->lists) {
      if (!L.used) continue;
      const KernelDesc& K0 = L.dip0 ? h->kd : h->km;
      const int S0 = L.rows0 == 0 ? h->km.wzmax : L.rows0 == 1 ? h->kd.wzmax : h->km.wzmax + h->kd.wzmax;
      size_t per = list_cells_bound(K0.wx, K0.wy, S0, h->TX, h->TY, h->zc_min);
      if (L.nkind > 1) per += list_cells_bound(h->kd.wx, h->kd.wy, h->kd.wzmax, h->TX, h->TY, h->zc_min);
      const size_t cap = std::max<size_t>(1, (size_t)N * per);
      if (cap > (size_t)INT32_MAX) return fail(DPS_ERR_ARG, "too many spreading entries");
      L.cap = (int)cap;
      if ((rc = h->alloc(&L.start, ncells + 1))) return rc;
      if ((rc = h->alloc(&L.counts, ncells))) return rc;
      if ((rc = h->alloc(&L.fill, ncells))) return rc;
      if ((rc = h->alloc(&L.entries, cap))) return rc;
    }
  }
  if ((rc = h->alloc((unsigned char**)&h->d_grid, ng * ts))) return rc;
  if ((rc = h->alloc((unsigned char**)&h->d_specA, nspec * 2 * ts))) return rc;
  if ((rc = h->alloc((unsigned char**)&h->d_specB, nspec * 2 * ts))) return rc;
  if (h->reg) {
    if ((rc = h->alloc((unsigned char**)&h->d_gridT, ng * ts))) return rc;
    if ((rc = h->alloc((unsigned char**)&h->d_specT, nspec * 2 * ts))) return rc;
    if ((rc = h->alloc((unsigned char**)&h->d_specC, nspec * 2 * ts))) return rc;
  }
  // triply periodic: the r2c spectrum [c][kz][ky][kx] has the same size, and
  // nothing else is needed
  if (h->geometry == DPS_MODE_PERIODIC) return DPS_OK;
  // z tables
  if ((rc = h->upload(&h->d_zext_m, h->zm.z))) return rc;
  if ((rc = h->upload(&h->d_zw_m, h->zm.w))) return rc;
  if ((rc = h->upload(&h->d_kt_m, h->zm.kt))) return rc;
  if ((rc = h->upload(&h->d_fA_m, h->zm.facA))) return rc;
  if ((rc = h->upload(&h->d_fB_m, h->zm.facB))) return rc;
  if (h->has_dip) {
    if ((rc = h->upload(&h->d_zext_d, h->zd.z))) return rc;
    if ((rc = h->upload(&h->d_zw_d, h->zd.w))) return rc;
    if ((rc = h->upload(&h->d_kt_d, h->zd.kt))) return rc;
    if ((rc = h->upload(&h->d_fA_d, h->zd.facA))) return rc;
    if ((rc = h->upload(&h->d_fB_d, h->zd.facB))) return rc;
  }
  // transform matrices in grid precision, inverse scaled by 1/(Nx Ny) (the
  // c2r transform is unnormalised; Transform.cpp:286-300 divides by Nx Ny Nu)
  const double inv_nxy = 1.0 / ((double)h->Nx * (double)h->Ny);
  std::vector<double> tie = h->ct.TiE, tio = h->ct.TiO;
  for (auto& v : tie) v *= inv_nxy;
  for (auto& v : tio) v *= inv_nxy;
  if (h->prec == DPS_PREC_MIXED) {
    if ((rc = h->upload((float**)&h->d_TfE, cast_vec<float>(h->ct.TfE)))) return rc;
    if ((rc = h->upload((float**)&h->d_TfO, cast_vec<float>(h->ct.TfO)))) return rc;
    if ((rc = h->upload((float**)&h->d_TiE, cast_vec<float>(tie)))) return rc;
    if ((rc = h->upload((float**)&h->d_TiO, cast_vec<float>(tio)))) return rc;
  } else {
    if ((rc = h->upload((double**)&h->d_TfE, h->ct.TfE))) return rc;
    if ((rc = h->upload((double**)&h->d_TfO, h->ct.TfO))) return rc;
    if ((rc = h->upload((double**)&h->d_TiE, tie))) return rc;
    if ((rc = h->upload((double**)&h->d_TiO, tio))) return rc;
  }
  if ((rc = h->upload(&h->d_zpts, h->zpts))) return rc;
  std::vector<double> kxs(h->Nkx), kys(h->Ny);
  for (int i = 0; i < h->Nkx; ++i) {
    const int n = (i <= h->Nx / 2) ? i : i - h->Nx;
    kxs[i] = 2 * M_PI * (double)n / h->P.Lx;
  }
  for (int i = 0; i < h->Ny; ++i) {
    const int n = (i <= h->Ny / 2) ? i : i - h->Ny;
    kys[i] = 2 * M_PI * (double)n / h->P.Ly;
  }
  if ((rc = h->upload(&h->d_kxs, kxs))) return rc;
  if ((rc = h->upload(&h->d_kys, kys))) return rc;
  // Chebyshev bands and BC rows
  if ((rc = h->upload(&h->d_sd0, h->cheb.sd0))) return rc;
  if ((rc = h->upload(&h->d_sm2, h->cheb.sm2))) return rc;
  if ((rc = h->upload(&h->d_sp2, h->cheb.sp2))) return rc;
  if ((rc = h->upload(&h->d_fm1, h->cheb.fm1))) return rc;
  if ((rc = h->upload(&h->d_fp1, h->cheb.fp1))) return rc;
  if ((rc = h->upload(&h->d_R1, h->cheb.R1))) return rc;
  if ((rc = h->upload(&h->d_R2, h->cheb.R2))) return rc;
  if ((rc = h->upload(&h->d_L1, h->cheb.L1))) return rc;
  if ((rc = h->upload(&h->d_L2, h->cheb.L2))) return rc;
  if (h->fused) return allocate_fused(h);
  // modes that touch a Nyquist index (their mirror is solved separately)
  std::vector<int> sp;
  for (int iy = 0; iy < h->Ny; ++iy)
    for (int ix = 0; ix < h->Nkx; ++ix) {
      const bool s = ((h->Nx % 2 == 0) && ix == h->Nx / 2) || ((h->Ny % 2 == 0) && iy == h->Ny / 2);
      if (s && !(ix == 0 && iy == 0)) sp.push_back(iy * h->Nkx + ix);
    }
  h->nspecial = (int)sp.size();
  if ((rc = h->upload(&h->d_special, sp))) return rc;
  if ((rc = h->alloc(&h->d_u2, (size_t)std::max(1, h->nspecial) * 3 * h->Nz))) return rc;
  if ((rc = h->alloc(&h->d_corr2, (size_t)std::max(1, h->nspecial) * 12))) return rc;
  // BVP scratch (one slot per lane) and correction coefficients
  const size_t ms = (size_t)h->Nm + h->nspecial;
  if ((rc = h->alloc(&h->s_ib, (size_t)(h->Nz + 1) * ms))) return rc;
  if ((rc = h->alloc(&h->s_ab, (size_t)h->Nz * ms))) return rc;
  if ((rc = h->alloc(&h->s_P, (size_t)h->Nz * ms))) return rc;
  if ((rc = h->alloc(&h->s_CP, (size_t)h->Nz * ms))) return rc;
  if ((rc = h->alloc(&h->s_DP, (size_t)h->Nz * ms))) return rc;
  if ((rc = h->alloc(&h->s_U0, (size_t)h->Nz * ms))) return rc;
  if ((rc = h->alloc(&h->s_U1, (size_t)h->Nz * ms))) return rc;
  if ((rc = h->alloc(&h->s_U2, (size_t)h->Nz * ms))) return rc;
  if ((rc = h->alloc(&h->d_corr, (size_t)h->Nm * 12))) return rc;
  return DPS_OK;
}

BvpArgs bvp_args(dps_handle* h);

// Tables of the fused spectral kernel: transform rows in lane order and the
// static per-mode factorisation (computed once on the device).
int allocate_fused(dps_handle* h)
{
  int rc;
  const int Nz = h->Nz, NH = h->ct.NH, LPM = h->sp_lpm, E = h->sp_e, R = E / 2;
  const int NP = (Nz + 1) / 2;   // coefficient pairs (2p, 2p+1)
  const double inv_nxy = 1.0 / ((double)h->Nx * (double)h->Ny);
  // forward [NH][LPM][E]: coefficient l E + q of node pair j
  // inverse [NP][LPM][E]: node pair l + LPM r of coefficient 2p (q = r) and
  // of coefficient 2p+1 (q = R + r)
  std::vector<double> tf((size_t)NH * LPM * E, 0.0), ti((size_t)NP * LPM * E, 0.0);
  for (int j = 0; j < NH; ++j)
    for (int l = 0; l < LPM; ++l)
      for (int q = 0; q < E; ++q) {
        const int k = l * E + q;
        if (k >= Nz) continue;
        tf[((size_t)j * LPM + l) * E + q] =
            (k & 1) ? h->ct.TfO[(size_t)(k / 2) * NH + j] : h->ct.TfE[(size_t)(k / 2) * NH + j];
      }
  for (int k = 0; k < Nz; ++k)
    for (int l = 0; l < LPM; ++l)
      for (int r = 0; r < R; ++r) {
        const int j = l + LPM * r;
        if (j >= NH) continue;
        ti[((size_t)(k / 2) * LPM + l) * E + (k & 1) * R + r] =
            inv_nxy * ((k & 1) ? h->ct.TiO[(size_t)j * h->ct.NO + k / 2] : h->ct.TiE[(size_t)j * h->ct.NE + k / 2]);
      }

  if (spectral_mfma(LPM, h->prec == DPS_PREC_MIXED)) {
    // matrix-core operands in v_mfma 16x16x4 lane order: lane L of step s
    // of row tile t holds row 16 t + L % 16 and term 4 s + L / 16 of the
    // parity's table, zero outside it and for the steps past its terms
    const int MT = (LPM * E / 2 + 15) / 16, KSM = LPM * E / 8, NE = h->ct.NE, NO = h->ct.NO;
    std::vector<float> fm((size_t)2 * MT * KSM * 64, 0.f), im((size_t)2 * MT * KSM * 64, 0.f);
    for (int par = 0; par < 2; ++par)
      for (int t = 0; t < MT; ++t)
        for (int ks = 0; ks < KSM; ++ks)
          for (int L = 0; L < 64; ++L) {
            const size_t o = (((size_t)par * MT + t) * KSM + ks) * 64 + L;
            const int r = 16 * t + L % 16, q = 4 * ks + L / 16;
            // forward: coefficient 2 r + par of node pair q
            if (2 * r + par < Nz && q < NH)
              fm[o] = (float)(par ? h->ct.TfO[(size_t)r * NH + q] : h->ct.TfE[(size_t)r * NH + q]);
            // inverse: node pair r of coefficient 2 q + par
            if (r < NH && q < (par ? NO : NE))
              im[o] = (float)(inv_nxy * (par ? h->ct.TiO[(size_t)r * NO + q] : h->ct.TiE[(size_t)r * NE + q]));
          }
    if ((rc = h->upload(&h->d_TfM, fm))) return rc;
    if ((rc = h->upload(&h->d_TiM, im))) return rc;
  }
  if (h->prec == DPS_PREC_MIXED) {
    if ((rc = h->upload((float**)&h->d_TfP, cast_vec<float>(tf)))) return rc;
    if ((rc = h->upload((float**)&h->d_TiP, cast_vec<float>(ti)))) return rc;
  } else {
    if ((rc = h->upload((double**)&h->d_TfP, tf))) return rc;
    if ((rc = h->upload((double**)&h->d_TiP, ti))) return rc;
  }
  // The factorisation depends on |k|^2 only: one table entry per distinct
  // value (bit for bit, as ksq_of forms it in the kernels), shared by the
  // modes (+-kx, +-ky) and, on a square box, (ky, kx) -- about a quarter of
  // the modes on a square grid.
  std::vector<int> fidx(h->Nm);
  std::vector<double> kq;
  {
    std::unordered_map<uint64_t, int> seen;
    seen.reserve((size_t)h->Nm / 2);
    for (int m = 0; m < h->Nm; ++m) {
      const int iy = m / h->Nkx, ix = m - iy * h->Nkx;
      const double k2 = ksq_of(kvec(ix, h->Nx, h->P.Lx), kvec(iy, h->Ny, h->P.Ly));
      uint64_t bits;
      std::memcpy(&bits, &k2, sizeof bits);
      auto it = seen.emplace(bits, (int)kq.size());
      if (it.second) kq.push_back(k2);
      fidx[m] = it.first->second;
    }
  }
  const int nk = (int)kq.size();
  h->n_fac = nk;
  if ((rc = h->upload(&h->d_fidx, fidx))) return rc;
  const size_t nf = (size_t)nk * E * LPM;
  if ((rc = h->alloc(&h->d_fib, nf))) return rc;
  if ((rc = h->alloc(&h->d_fab, nf))) return rc;
  if ((rc = h->alloc(&h->d_fG, (size_t)nk * 4))) return rc;
  double *t_ib = nullptr, *t_ab = nullptr, *t_kq = nullptr;
  if (hipMalloc((void**)&t_ib, (size_t)(Nz + 1) * nk * sizeof(double)) != hipSuccess ||
      hipMalloc((void**)&t_ab, (size_t)Nz * nk * sizeof(double)) != hipSuccess ||
      hipMalloc((void**)&t_kq, (size_t)nk * sizeof(double)) != hipSuccess) {
    if (t_ib) (void)hipFree(t_ib);
    if (t_ab) (void)hipFree(t_ab);
    return fail(DPS_ERR_ALLOC, "hipMalloc of the factorisation scratch failed");
  }
  const BvpArgs A = bvp_args(h);
  hipError_t e0 = hipMemcpy(t_kq, kq.data(), (size_t)nk * sizeof(double), hipMemcpyHostToDevice);
  if (e0 == hipSuccess)
    launch_bvp_factor(A.B, t_kq, nk, t_ib, t_ab, LPM, E, h->d_fib, h->d_fab, h->d_fG, 0);
  hipError_t e1 = (e0 != hipSuccess) ? e0 : hipGetLastError();
  hipError_t e2 = hipStreamSynchronize(0);
  (void)hipFree(t_ib);
  (void)hipFree(t_ab);
  (void)hipFree(t_kq);
  if (e1 != hipSuccess || e2 != hipSuccess) return fail(DPS_ERR_HIP, "factorisation kernel failed");
  return DPS_OK;
}

int make_plans(dps_handle* h)
{
  int rc;
  // rocFFT plans: 2-D over (x, y), batched over (component, z); triply
  // periodic: 3-D over (x, y, z), batched over the components
  const rocfft_precision fp = (h->prec == DPS_PREC_MIXED) ? rocfft_precision_single : rocfft_precision_double;
  const bool tp = h->geometry == DPS_MODE_PERIODIC;
  size_t lengths[3] = {(size_t)h->Nx, (size_t)h->Ny, (size_t)h->Nz};
  // doubly periodic with Ny a power of two or 3 * 2^k: rocFFT does the unit-stride x pass
  // (batched over component, z and y) and k_fft_y the strided y pass
  h->fft_y = !tp && fft_y_supported(h->Ny);
  const size_t dim = tp ? 3 : (h->fft_y ? 1 : 2);
  const size_t batch = tp ? 3 : (size_t)3 * h->Nz * (h->fft_y ? h->Ny : 1);
  if (h->fft_y) {
    std::vector<double> tw(2 * (size_t)h->Ny);
    for (int k = 0; k < h->Ny; ++k) {
      tw[2 * k] = std::cos(2 * M_PI * k / h->Ny);
      tw[2 * k + 1] = -std::sin(2 * M_PI * k / h->Ny);
    }
    if (h->prec == DPS_PREC_MIXED) rc = h->upload((float**)&h->d_twy, cast_vec<float>(tw));
    else rc = h->upload((double**)&h->d_twy, tw);
    if (rc) return rc;
  }
  // Nx a power of two or 3 * 2^k as well: k_fft_x does the x pass and no rocFFT plan is made
  h->fft_x = h->fft_y && fft_x_supported(h->Nx);
  if (h->fft_x) {
    std::vector<double> tw(2 * (size_t)h->Nx);
    for (int k = 0; k < h->Nx; ++k) {
      tw[2 * k] = std::cos(2 * M_PI * k / h->Nx);
      tw[2 * k + 1] = -std::sin(2 * M_PI * k / h->Nx);
    }
    if (h->prec == DPS_PREC_MIXED) rc = h->upload((float**)&h->d_twx, cast_vec<float>(tw));
    else rc = h->upload((double**)&h->d_twx, tw);
    return rc;
  }
  FFTCHK(rocfft_plan_create(&h->plan_fwd, rocfft_placement_notinplace, rocfft_transform_type_real_forward,
                            fp, dim, lengths, batch, nullptr));
  FFTCHK(rocfft_plan_create(&h->plan_inv, rocfft_placement_notinplace, rocfft_transform_type_real_inverse,
                            fp, dim, lengths, batch, nullptr));
  size_t w1 = 0, w2 = 0;
  FFTCHK(rocfft_plan_get_work_buffer_size(h->plan_fwd, &w1));
  FFTCHK(rocfft_plan_get_work_buffer_size(h->plan_inv, &w2));
  h->fft_work_size = std::max(w1, w2);
  if (h->fft_work_size)
    if ((rc = h->alloc((unsigned char**)&h->fft_work, h->fft_work_size))) return rc;
  FFTCHK(rocfft_execution_info_create(&h->info));
  if (h->fft_work_size) FFTCHK(rocfft_execution_info_set_work_buffer(h->info, h->fft_work, h->fft_work_size));
  return DPS_OK;
}

bool mixed(const dps_handle* h) { return h->prec == DPS_PREC_MIXED; }

// planar transform grid -> spectrum (forward) or spectrum -> grid (inverse);
// with the split plan the inverse's y pass runs in place on `in`
int run_fft(dps_handle* h, bool forward, void* in, void* out, hipStream_t s)
{
  const int ns = 3 * h->Nz;
  if (h->fft_y && !forward) {
    if (mixed(h)) launch_fft_y<float>(in, ns, h->Nkx, h->Ny, h->d_twy, 1, s);
    else launch_fft_y<double>(in, ns, h->Nkx, h->Ny, h->d_twy, 1, s);
    HIPCHK(hipGetLastError());
  }
  if (h->fft_x) {
    const int rows = ns * h->Ny;
    if (mixed(h)) launch_fft_x<float>(in, out, rows, h->Nx, h->d_twx, forward ? -1 : 1, s);
    else launch_fft_x<double>(in, out, rows, h->Nx, h->d_twx, forward ? -1 : 1, s);
    HIPCHK(hipGetLastError());
  } else {
    FFTCHK(rocfft_execution_info_set_stream(h->info, s));
    void* ib[1] = {in};
    void* ob[1] = {out};
    FFTCHK(rocfft_execute(forward ? h->plan_fwd : h->plan_inv, ib, ob, h->info));
  }
  if (h->fft_y && forward) {
    if (mixed(h)) launch_fft_y<float>(out, ns, h->Nkx, h->Ny, h->d_twy, -1, s);
    else launch_fft_y<double>(out, ns, h->Nkx, h->Ny, h->d_twy, -1, s);
    HIPCHK(hipGetLastError());
  }
  return DPS_OK;
}

StencilArgs stencil_args(dps_handle* h)
{
  StencilArgs A{};
  A.N = h->N;
  A.xmin = 0;
  A.ymin = 0;
  A.zmin = h->P.zmin;
  A.Lx = h->P.Lx;
  A.Ly = h->P.Ly;
  A.Lz = h->Lz;
  A.hx = h->hx;
  A.hy = h->hy;
  A.Nx = h->Nx;
  A.Ny = h->Ny;
  A.has_dip = h->has_dip;
  A.km = h->km;
  A.kd = h->kd;
  A.zext_m = h->d_zext_m;
  A.zw_m = h->d_zw_m;
  A.zext_d = h->d_zext_d;
  A.zw_d = h->d_zw_d;
  A.pos = h->d_pos;
  A.err = h->d_err;
  A.sup_m = h->d_sup_m;
  A.sup_d = h->d_sup_d;
  A.rec_m = h->d_rec_m;
  A.rec_d = h->d_rec_d;
  A.rs_m = h->rs_m;
  A.rs_d = h->rs_d;
  A.key = h->d_key;
  A.TX = h->TX;
  A.TY = h->TY;
  A.ntx = h->ntx;
  A.kt_m = h->d_kt_m;
  A.kt_d = h->d_kt_d;
  A.fA_m = h->d_fA_m;
  A.fA_d = h->d_fA_d;
  A.fB_d = h->d_fB_d;
  A.zrange = h->d_zrange;
  A.periodic_z = h->geometry == DPS_MODE_PERIODIC;
  A.hz = h->hz;
  A.dip_regular = h->reg;
  A.first_kind = 0;
  return A;
}

SpreadArgs spread_args(dps_handle* h)
{
  SpreadArgs A{};
  A.Nx = h->Nx; A.Ny = h->Ny; A.Nz = h->Nz;
  A.TX = h->TX; A.TY = h->TY; A.ZC = h->ZC; A.ntx = h->ntx; A.nty = h->nty;
  A.nch = (h->Nz + h->ZC - 1) / h->ZC;
  A.lstart = h->lists[0].start;
  A.lentries = h->lists[0].entries;
  A.lcap = h->lists[0].cap;
  A.amax = h->d_amax;
  A.namax = DPS_AMAX_BLOCKS;
  A.K0 = 3 * h->km.inv_norm3;
  A.K1 = h->has_dip ? 3 * h->kd.dmax * h->kd.inv_norm3 : 0.0;
  A.zrange = h->d_zrange;
  A.km = h->km; A.kd = h->kd;
  A.sup_m = h->d_sup_m; A.sup_d = h->d_sup_d;
  A.rec_m = h->d_rec_m; A.rec_d = h->d_rec_d; A.rs_m = h->rs_m; A.rs_d = h->rs_d;
  A.rs_max = std::max(h->rs_m, h->rs_d);
  A.grid = h->d_grid;
  A.periodic_z = h->geometry == DPS_MODE_PERIODIC;
  A.zr_dip = 0;
  A.pair = 0;
  A.ho = A.rs_max;
  return A;
}

// force and torque of a particle as one spreading entry: possible when the two
// kernels have the same x/y supports (derivative torque kernels)
bool spread_pairable(const dps_handle* h)
{
  return h->has_dip && !h->reg && h->km.wx == h->kd.wx && h->km.wy == h->kd.wy;
}

// the pass with torques through the derivative kernels (paired or not)
SpreadArgs spread_args_torques(dps_handle* h)
{
  SpreadArgs A = spread_args(h);
  A.lstart = h->lists[1].start;
  A.lentries = h->lists[1].entries;
  A.lcap = h->lists[1].cap;
  return A;
}

SpreadArgs spread_args_pair(dps_handle* h)
{
  SpreadArgs A = spread_args_torques(h);
  A.pair = 1;
  A.pm = h->km.wx + h->km.wy + h->km.wzmax;
  A.pd = 2 * h->kd.wx + 2 * h->kd.wy + 2 * h->kd.wzmax;
  A.ho = A.pm + A.pd;
  A.rs_max = A.ho + 7;   // staged words: ho + 14
  return A;
}

// the torque kernel presented as the only (kind-0) kernel: a plain ES pass
// over the torque records (regular torque kernel)
SpreadArgs spread_args_torque(dps_handle* h)
{
  SpreadArgs A = spread_args(h);
  A.km = h->kd;
  A.sup_m = h->d_sup_d;
  A.rec_m = h->d_rec_d;
  A.rs_m = h->rs_d;
  A.grid = h->d_gridT;
  A.zr_dip = 1;
  A.K0 = 3 * h->kd.inv_norm3;   // the torques are this pass's forces
  A.K1 = 0.0;
  A.lstart = h->lists[2].start;
  A.lentries = h->lists[2].entries;
  A.lcap = h->lists[2].cap;
  return A;
}

// the spreading lists of every pass, for the located positions and the
// current z chunk (kernels only: a captured position update rebuilds them)
int build_lists(dps_handle* h, hipStream_t s)
{
  for (int i = 0; i < 3; ++i) {
    auto& L = h->lists[i];
    if (!L.used) continue;
    ListArgs A{};
    A.N = h->N; A.Nx = h->Nx; A.Ny = h->Ny; A.Nz = h->Nz;
    A.TX = h->TX; A.TY = h->TY; A.ZC = h->ZC; A.ntx = h->ntx; A.nty = h->nty;
    A.nch = (h->Nz + h->ZC - 1) / h->ZC;
    A.periodic_z = h->geometry == DPS_MODE_PERIODIC;
    A.nkind = L.nkind;
    A.rows0 = L.rows0;
    const KernelDesc& K0 = L.dip0 ? h->kd : h->km;
    A.wx0 = K0.wx; A.wy0 = K0.wy; A.wx1 = h->kd.wx; A.wy1 = h->kd.wy;
    A.sup0 = L.dip0 ? h->d_sup_d : h->d_sup_m;
    A.sup1 = h->d_sup_d;
    A.zrange = h->d_zrange;
    A.counts = L.counts; A.fill = L.fill; A.start = L.start; A.entries = L.entries;
    A.cap = L.cap;
    launch_lists(A, s);
  }
  HIPCHK(hipGetLastError());
  return DPS_OK;
}

InterpArgs interp_args(dps_handle* h)
{
  InterpArgs A{};
  A.N = h->N; A.Nx = h->Nx; A.Ny = h->Ny; A.Nz = h->Nz;
  A.order = h->d_order;
  A.km = h->km; A.kd = h->kd;
  A.sup_m = h->d_sup_m; A.sup_d = h->d_sup_d;
  A.rec_m = h->d_rec_m; A.rec_d = h->d_rec_d; A.rs_m = h->rs_m; A.rs_d = h->rs_d;
  A.kt_m = h->d_kt_m; A.kt_d = h->d_kt_d;
  A.fA_m = h->d_fA_m; A.fA_d = h->d_fA_d; A.fB_d = h->d_fB_d;
  A.grid = h->d_grid;
  A.periodic_z = h->geometry == DPS_MODE_PERIODIC;
  A.pair = 0;
  return A;
}

ZArgs z_args(dps_handle* h)
{
  ZArgs A{};
  A.N = h->Nz; A.NH = h->ct.NH; A.NE = h->ct.NE; A.NO = h->ct.NO;
  A.Nm = h->Nm; A.Nkx = h->Nkx;
  A.TfE = h->d_TfE; A.TfO = h->d_TfO; A.TiE = h->d_TiE; A.TiO = h->d_TiO;
  A.nbasis = (h->geometry == DPS_MODE_BOTTOM) ? 2 : (h->geometry == DPS_MODE_SLIT ? 4 : 0);
  A.corr = h->d_corr;
  A.kxs = h->d_kxs; A.kys = h->d_kys; A.zpts = h->d_zpts;
  A.Hs = h->Lz;
  A.inv_nxy = 1.0 / ((double)h->Nx * (double)h->Ny);
  return A;
}

BvpArgs bvp_args(dps_handle* h)
{
  BvpArgs A{};
  BvpConsts& B = A.B;
  B.Nz = h->Nz; B.geometry = h->geometry; B.H = h->H; B.Hs = h->Lz; B.eta = h->P.viscosity;
  B.sd0 = h->d_sd0; B.sm2 = h->d_sm2; B.sp2 = h->d_sp2;
  B.S0N = h->cheb.S0N; B.S1N1 = h->cheb.S1N1;
  B.fm1 = h->d_fm1; B.fp1 = h->d_fp1; B.F0N1 = h->cheb.F0N1;
  B.R1 = h->d_R1; B.R2 = h->d_R2; B.L1 = h->d_L1; B.L2 = h->d_L2;
  A.Nx = h->Nx; A.Ny = h->Ny; A.Nkx = h->Nkx; A.Nm = h->Nm;
  A.Lx = h->P.Lx; A.Ly = h->P.Ly;
  A.Ms = h->Nm + h->nspecial;
  A.special = h->d_special;
  A.nspecial = h->nspecial;
  A.u2 = h->d_u2;
  A.s_ib = h->s_ib; A.s_ab = h->s_ab;
  A.s_P = h->s_P; A.s_CP = h->s_CP; A.s_DP = h->s_DP;
  A.s_U0 = h->s_U0; A.s_U1 = h->s_U1; A.s_U2 = h->s_U2;
  A.corr = h->d_corr;
  A.corr2 = h->d_corr2;
  return A;
}

SpecArgs spec_args(dps_handle* h)
{
  SpecArgs A{};
  A.B = bvp_args(h).B;
  A.Nx = h->Nx; A.Ny = h->Ny; A.Nkx = h->Nkx; A.Nm = h->Nm;
  A.Lx = h->P.Lx; A.Ly = h->P.Ly;
  A.NH = h->ct.NH;
  A.lpm = h->sp_lpm; A.e = h->sp_e; A.G = h->sp_G;
  A.TfP = h->d_TfP; A.TiP = h->d_TiP;
  A.TfM = h->d_TfM; A.TiM = h->d_TiM;
  A.fac_idx = h->d_fidx; A.fac_ib = h->d_fib; A.fac_ab = h->d_fab; A.fac_G = h->d_fG;
  A.kxs = h->d_kxs; A.kys = h->d_kys;
  A.zpts = h->d_zpts;
  A.Hs = h->Lz;
  A.inv_nxy = 1.0 / ((double)h->Nx * (double)h->Ny);
  return A;
}

void mark(dps_handle* h, int i, hipStream_t s)
{
  if (h->prof) (void)hipEventRecord(h->ev[i], s);
}

void spread_pass(const SpreadArgs& A, const void* F, const void* Tq, bool mx, int dtype, hipStream_t s, int n)
{
  if (mx) {
    // fixed-point accumulation: the pass's input maxima first
    if (dtype == DPS_F32) launch_absmax<float>(F, Tq, 3 * n, const_cast<float*>(A.amax), s);
    else launch_absmax<double>(F, Tq, 3 * n, const_cast<float*>(A.amax), s);
    if (dtype == DPS_F32) launch_spread<float, float>(A, F, Tq, s);
    else launch_spread<float, double>(A, F, Tq, s);
  } else {
    if (dtype == DPS_F32) launch_spread<double, float>(A, F, Tq, s);
    else launch_spread<double, double>(A, F, Tq, s);
  }
}

// forces (and the torque curl of the derivative kernels) into d_grid; with
// the regular torque kernel the torques go, as a plain vector field, into
// d_gridT
// LDS budget of a spreading workgroup: 64 KB.  The device allows more per
// workgroup (160 KB on gfx950) and eager launches take it, but a batch
// graph replaying such a launch produced wrong grids, so no spreading
// configuration goes beyond 64 KB.
size_t spread_lds_cap(const dps_handle* h) { return std::min<size_t>(65536, h->lds_max); }

// whether every spreading pass do_spread runs (with or without torques) fits
// the LDS budget with z chunks of zc levels
bool spread_fits(dps_handle* h, bool torques, int zc)
{
  SpreadArgs passes[2];
  int n = 0;
  if (h->reg) {
    passes[n++] = spread_args(h);
    if (torques) passes[n++] = spread_args_torque(h);
  } else if (torques && spread_pairable(h)) {
    passes[n++] = spread_args_pair(h);
  } else {
    passes[n++] = spread_args(h);
  }
  for (int i = 0; i < n; ++i) {
    passes[i].ZC = zc;
    if (spread_npf(passes[i].rs_max) == 0 || spread_lds_bytes(passes[i], h->tsz) > spread_lds_cap(h)) return false;
  }
  return true;
}

int do_spread(dps_handle* h, const void* F, const void* T, int dtype, hipStream_t s)
{
  const void* Tq = (T && h->has_dip) ? T : nullptr;
  if (h->reg) {
    spread_pass(spread_args(h), F, nullptr, mixed(h), dtype, s, h->N);
    if (Tq) spread_pass(spread_args_torque(h), Tq, nullptr, mixed(h), dtype, s, h->N);
  } else if (Tq && spread_pairable(h)) {
    spread_pass(spread_args_pair(h), F, Tq, mixed(h), dtype, s, h->N);
  } else {
    spread_pass(Tq ? spread_args_torques(h) : spread_args(h), F, Tq, mixed(h), dtype, s, h->N);
  }
  HIPCHK(hipGetLastError());
  return DPS_OK;
}

int do_solve(dps_handle* h, hipStream_t s, int ev0)
{
  int rc;
  if ((rc = run_fft(h, true, h->d_grid, h->d_specA, s))) return rc;
  mark(h, ev0 + 1, s);
  if (h->geometry == DPS_MODE_PERIODIC) {
    // Fourier solve (reported as the bvp stage; no z transforms)
    if (mixed(h)) launch_tp_solve<float>(h->d_specA, h->d_specB, h->Nx, h->Ny, h->Nz, h->P.Lx, h->P.Ly, h->Lz,
                                         h->P.viscosity, s);
    else launch_tp_solve<double>(h->d_specA, h->d_specB, h->Nx, h->Ny, h->Nz, h->P.Lx, h->P.Ly, h->Lz,
                                 h->P.viscosity, s);
    HIPCHK(hipGetLastError());
    mark(h, ev0 + 3, s);
    if ((rc = run_fft(h, false, h->d_specB, h->d_grid, s))) return rc;
    mark(h, ev0 + 5, s);
    return DPS_OK;
  }
  if (h->fused) {
    // z transforms are inside the fused launch: no events of their own
    // (dps_get_timings reports them as zero)
    const SpecArgs A = spec_args(h);
    if (mixed(h)) launch_spectral<float>(A, h->d_specA, h->d_specB, s);
    else launch_spectral<double>(A, h->d_specA, h->d_specB, s);
    HIPCHK(hipGetLastError());
    mark(h, ev0 + 3, s);
    if ((rc = run_fft(h, false, h->d_specB, h->d_grid, s))) return rc;
    mark(h, ev0 + 5, s);
    return DPS_OK;
  }
  const ZArgs Z = z_args(h);
  if (mixed(h)) launch_zfwd<float>(Z, h->d_specA, h->d_specB, s);
  else launch_zfwd<double>(Z, h->d_specA, h->d_specB, s);
  HIPCHK(hipGetLastError());
  mark(h, ev0 + 2, s);
  {
    const BvpArgs B = bvp_args(h);
    if (mixed(h)) launch_bvp<float>(B, h->d_specB, h->d_specA, s);
    else launch_bvp<double>(B, h->d_specB, h->d_specA, s);
    HIPCHK(hipGetLastError());
  }
  mark(h, ev0 + 3, s);
  if (mixed(h)) launch_zinv<float>(Z, h->d_specA, h->d_specB, s);
  else launch_zinv<double>(Z, h->d_specA, h->d_specB, s);
  HIPCHK(hipGetLastError());
  mark(h, ev0 + 4, s);
  if ((rc = run_fft(h, false, h->d_specB, h->d_grid, s))) return rc;
  mark(h, ev0 + 5, s);
  return DPS_OK;
}

// Spectral stage of the regular torque kernel (FCM.py:372-396, 427-439;
// Solvers.py:137-179).  d_grid holds the spread forces and d_gridT the spread
// torques; on return d_grid holds the velocity and d_gridT half the vorticity.
//   doubly periodic: z-forward both spectra, add half the Chebyshev curl of
//   the torques to the forcing, solve, z-inverse (wall corrections included),
//   z-forward that velocity again -- the curl acts on the coefficients of the
//   corrected velocity, as the reference's does -- take half its curl
//   (rescaled by Nx Ny, which the inverse tables divide out) and z-inverse it
//   without corrections;
//   triply periodic: the same with Fourier curls around the Fourier solve.
// Events as do_solve: 2 forward FFTs, 3 z forward + torque curl, 4 BVP,
// 5 z inverse + vorticity, 6 inverse FFTs (triply periodic: 2, 4, 6).
int do_solve_reg(dps_handle* h, hipStream_t s)
{
  int rc;
  const bool mx = mixed(h);
  if ((rc = run_fft(h, true, h->d_grid, h->d_specA, s))) return rc;
  if ((rc = run_fft(h, true, h->d_gridT, h->d_specT, s))) return rc;
  mark(h, 2, s);
  if (h->geometry == DPS_MODE_PERIODIC) {
    const double Lx = h->P.Lx, Ly = h->P.Ly, Lz = h->Lz, eta = h->P.viscosity;
    if (mx) {
      launch_tp_curl<float>(h->d_specT, h->d_specA, h->Nx, h->Ny, h->Nz, Lx, Ly, Lz, 1, s);
      launch_tp_solve<float>(h->d_specA, h->d_specB, h->Nx, h->Ny, h->Nz, Lx, Ly, Lz, eta, s);
      launch_tp_curl<float>(h->d_specB, h->d_specT, h->Nx, h->Ny, h->Nz, Lx, Ly, Lz, 0, s);
    } else {
      launch_tp_curl<double>(h->d_specT, h->d_specA, h->Nx, h->Ny, h->Nz, Lx, Ly, Lz, 1, s);
      launch_tp_solve<double>(h->d_specA, h->d_specB, h->Nx, h->Ny, h->Nz, Lx, Ly, Lz, eta, s);
      launch_tp_curl<double>(h->d_specB, h->d_specT, h->Nx, h->Ny, h->Nz, Lx, Ly, Lz, 0, s);
    }
    HIPCHK(hipGetLastError());
    mark(h, 4, s);
    if ((rc = run_fft(h, false, h->d_specB, h->d_grid, s))) return rc;
    if ((rc = run_fft(h, false, h->d_specT, h->d_gridT, s))) return rc;
    mark(h, 6, s);
    return DPS_OK;
  }
  const ZArgs Z = z_args(h);
  ZArgs Z0 = Z;
  Z0.nbasis = 0;   // the vorticity's inverse carries no wall correction
  const double nxy = (double)h->Nx * (double)h->Ny;
  if (h->fused) {
    // around the fused solve, in values at the nodes: the torque curl's
    // coefficients transformed back and added to the forcing, the velocity
    // transformed forward for its curl (the inverse tables divide by Nx Ny,
    // the curls restore it)
    ZArgs Za = Z0;
    Za.accumulate = 1;
    bool ok;
    if (mx) {
      ok = launch_zfwd_mma<float>(Z, h->d_specT, h->d_specC, s);
      launch_curl_cheb<float>(h->d_specC, h->d_specB, h->Nx, h->Ny, h->Nz, h->Nm, h->d_kxs, h->d_kys, h->H, nxy, 0, s);
      ok = ok && launch_zinv_mma<float>(Za, h->d_specB, h->d_specA, s);
    } else {
      ok = launch_zfwd_mma<double>(Z, h->d_specT, h->d_specC, s);
      launch_curl_cheb<double>(h->d_specC, h->d_specB, h->Nx, h->Ny, h->Nz, h->Nm, h->d_kxs, h->d_kys, h->H, nxy, 0, s);
      ok = ok && launch_zinv_mma<double>(Za, h->d_specB, h->d_specA, s);
    }
    if (!ok) return fail(DPS_ERR_STATE, "matrix-core z transform unavailable for this nz");
    HIPCHK(hipGetLastError());
    mark(h, 3, s);
    const SpecArgs SA = spec_args(h);
    if (mx) launch_spectral<float>(SA, h->d_specA, h->d_specB, s);
    else launch_spectral<double>(SA, h->d_specA, h->d_specB, s);
    HIPCHK(hipGetLastError());
    mark(h, 4, s);
    if (mx) {
      ok = launch_zfwd_mma<float>(Z, h->d_specB, h->d_specC, s);
      launch_curl_cheb<float>(h->d_specC, h->d_specT, h->Nx, h->Ny, h->Nz, h->Nm, h->d_kxs, h->d_kys, h->H, nxy, 0, s);
      ok = ok && launch_zinv_mma<float>(Z0, h->d_specT, h->d_specC, s);
    } else {
      ok = launch_zfwd_mma<double>(Z, h->d_specB, h->d_specC, s);
      launch_curl_cheb<double>(h->d_specC, h->d_specT, h->Nx, h->Ny, h->Nz, h->Nm, h->d_kxs, h->d_kys, h->H, nxy, 0, s);
      ok = ok && launch_zinv_mma<double>(Z0, h->d_specT, h->d_specC, s);
    }
    if (!ok) return fail(DPS_ERR_STATE, "matrix-core z transform unavailable for this nz");
    HIPCHK(hipGetLastError());
    mark(h, 5, s);
    if ((rc = run_fft(h, false, h->d_specB, h->d_grid, s))) return rc;
    if ((rc = run_fft(h, false, h->d_specC, h->d_gridT, s))) return rc;
    mark(h, 6, s);
    return DPS_OK;
  }
  const BvpArgs B = bvp_args(h);
  if (mx) {
    launch_zfwd<float>(Z, h->d_specA, h->d_specB, s);
    launch_zfwd<float>(Z, h->d_specT, h->d_specC, s);
    launch_curl_cheb<float>(h->d_specC, h->d_specB, h->Nx, h->Ny, h->Nz, h->Nm, h->d_kxs, h->d_kys, h->H, 1.0, 1, s);
  } else {
    launch_zfwd<double>(Z, h->d_specA, h->d_specB, s);
    launch_zfwd<double>(Z, h->d_specT, h->d_specC, s);
    launch_curl_cheb<double>(h->d_specC, h->d_specB, h->Nx, h->Ny, h->Nz, h->Nm, h->d_kxs, h->d_kys, h->H, 1.0, 1, s);
  }
  HIPCHK(hipGetLastError());
  mark(h, 3, s);
  if (mx) launch_bvp<float>(B, h->d_specB, h->d_specA, s);
  else launch_bvp<double>(B, h->d_specB, h->d_specA, s);
  HIPCHK(hipGetLastError());
  mark(h, 4, s);
  if (mx) {
    launch_zinv<float>(Z, h->d_specA, h->d_specB, s);
    launch_zfwd<float>(Z, h->d_specB, h->d_specC, s);
    launch_curl_cheb<float>(h->d_specC, h->d_specT, h->Nx, h->Ny, h->Nz, h->Nm, h->d_kxs, h->d_kys, h->H, nxy, 0, s);
    launch_zinv<float>(Z0, h->d_specT, h->d_specC, s);
  } else {
    launch_zinv<double>(Z, h->d_specA, h->d_specB, s);
    launch_zfwd<double>(Z, h->d_specB, h->d_specC, s);
    launch_curl_cheb<double>(h->d_specC, h->d_specT, h->Nx, h->Ny, h->Nz, h->Nm, h->d_kxs, h->d_kys, h->H, nxy, 0, s);
    launch_zinv<double>(Z0, h->d_specT, h->d_specC, s);
  }
  HIPCHK(hipGetLastError());
  mark(h, 5, s);
  if ((rc = run_fft(h, false, h->d_specB, h->d_grid, s))) return rc;
  if ((rc = run_fft(h, false, h->d_specC, h->d_gridT, s))) return rc;
  mark(h, 6, s);
  return DPS_OK;
}

void interp_pass(const InterpArgs& A, void* vel, void* ang, bool mx, int dtype, hipStream_t s)
{
  if (mx) {
    if (dtype == DPS_F32) launch_interp<float, float>(A, vel, ang, s);
    else launch_interp<float, double>(A, vel, ang, s);
  } else {
    if (dtype == DPS_F32) launch_interp<double, float>(A, vel, ang, s);
    else launch_interp<double, double>(A, vel, ang, s);
  }
}

// velocities from d_grid; angular velocities from the derivative kernels on
// d_grid, or (regular torque kernel) the torque kernel on the half-vorticity
// grid d_gridT
int do_interp(dps_handle* h, void* vel, void* ang, int dtype, hipStream_t s)
{
  void* a = (ang && h->has_dip) ? ang : nullptr;
  if (h->reg) {
    interp_pass(interp_args(h), vel, nullptr, mixed(h), dtype, s);
    if (a) {
      InterpArgs A = interp_args(h);
      A.km = h->kd;
      A.sup_m = h->d_sup_d;
      A.rec_m = h->d_rec_d;
      A.rs_m = h->rs_d;
      A.grid = h->d_gridT;
      interp_pass(A, a, nullptr, mixed(h), dtype, s);
    }
  } else {
    InterpArgs A = interp_args(h);
    A.pair = a && spread_pairable(h);
    interp_pass(A, vel, a, mixed(h), dtype, s);
  }
  HIPCHK(hipGetLastError());
  return DPS_OK;
}

// record lengths of the position adjoint's force-kernel derivative records
// and torque-kernel second-derivative records
void adjoint_record_sizes(dps_handle* h)
{
  h->rs_md = 2 * h->km.wx + 2 * h->km.wy + 4 * h->km.wzmax;
  h->rs_d2 = (h->has_dip && !h->reg) ? 4 * h->kd.wx + 4 * h->kd.wy + 4 * h->kd.wzmax : 0;
}

// bytes of the position adjoint's buffers: support records, derivative
// records, z ranges, second velocity grid, torque second-derivative records
void adjoint_buffer_bytes(const dps_handle* h, size_t bytes[5])
{
  const size_t ng = (size_t)3 * h->Nz * h->Ny * h->Nx;
  bytes[0] = (size_t)h->N * sizeof(int4);
  bytes[1] = (size_t)h->N * h->rs_md * h->tsz;
  bytes[2] = (size_t)h->N * sizeof(int4);
  bytes[3] = ng * h->tsz;
  bytes[4] = (size_t)h->N * h->rs_d2 * h->tsz;
}

// buffers of the position adjoint, allocated on its first call
int ensure_adjoint(dps_handle* h)
{
  if (h->d_gridA) return DPS_OK;
  int rc;
  adjoint_record_sizes(h);
  size_t b[5];
  adjoint_buffer_bytes(h, b);
  if ((rc = h->alloc((unsigned char**)&h->d_sup_md, b[0]))) return rc;
  if ((rc = h->alloc((unsigned char**)&h->d_rec_md, b[1]))) return rc;
  if ((rc = h->alloc((unsigned char**)&h->d_zrange_md, b[2]))) return rc;
  if ((rc = h->alloc((unsigned char**)&h->d_gridA, b[3]))) return rc;
  if (h->rs_d2 > 0 && (rc = h->alloc((unsigned char**)&h->d_rec_d2, b[4]))) return rc;
  return DPS_OK;
}

// force-kernel derivative records (dipole layout, ghost factors of the force
// kernel in x, y and z, zero derivative where an offset was clamped) from the
// located positions
KernelDesc force_deriv_kernel(const dps_handle* h)
{
  KernelDesc K = h->km;
  K.pos_deriv = 1;
  return K;
}

int locate_force_derivatives(dps_handle* h, hipStream_t s)
{
  StencilArgs A = stencil_args(h);
  A.first_kind = 1;
  A.has_dip = 1;
  A.dip_regular = 0;
  A.kd = force_deriv_kernel(h);
  A.zmin = 0;   // d_pos is already shifted and wrapped
  A.zext_d = h->d_zext_m;
  A.zw_d = h->d_zw_m;
  A.kt_d = h->d_kt_m;
  A.fA_d = h->d_fA_m;
  A.fB_d = h->d_fA_m;
  A.sup_d = h->d_sup_md;
  A.rec_d = h->d_rec_md;
  A.rs_d = h->rs_md;
  A.zrange = h->d_zrange_md;
  if (mixed(h)) launch_stencil<double, float>(A, h->d_pos, s);
  else launch_stencil<double, double>(A, h->d_pos, s);
  HIPCHK(hipGetLastError());
  return DPS_OK;
}

int check_handle(const dps_handle* h, bool need_pos)
{
  if (!h) return fail(DPS_ERR_STATE, "DPStokes is not initialized. Call Initialize first");
  if (need_pos && !h->positions_set) return fail(DPS_ERR_STATE, "positions have not been set");
  return DPS_OK;
}

// argument rules shared by the two adjoint entry points
int adjoint_check(dps_handle* h, const void* forces, const void* torques, const void* grad_velocities,
                  const void* grad_angular, void* grad_forces, void* grad_torques, void* grad_positions,
                  int dtype)
{
  int rc = check_handle(h, true);
  if (rc) return rc;
  if (dtype != DPS_F32 && dtype != DPS_F64) return fail(DPS_ERR_ARG, "dtype must be DPS_F32 or DPS_F64");
  if (h->N > 0 && !grad_velocities) return fail(DPS_ERR_ARG, "null grad_velocities");
  if ((grad_angular || grad_torques || torques) && !h->has_dip)
    return fail(DPS_ERR_ARG, "torque terms given but no dipole kernel (w_d = 0)");
  if (grad_torques && !grad_forces) return fail(DPS_ERR_ARG, "grad_torques needs grad_forces");
  if (!grad_positions) {
    if (grad_torques && !grad_angular)
      return fail(DPS_ERR_ARG, "grad_torques without positions needs grad_angular (zeros if W is unused)");
    return DPS_OK;
  }
  if ((torques || grad_angular || grad_torques) && h->reg)
    return fail(DPS_ERR_ARG, "position gradients with torques need the derivative torque kernels");
  if (h->N > 0 && !forces) return fail(DPS_ERR_ARG, "null forces");
  return DPS_OK;
}

// the position adjoint on the handle's adjoint buffers (d_sup_md, d_rec_md,
// d_zrange_md, d_gridA, d_rec_d2), which the caller has pointed at memory
int adjoint_positions(dps_handle* h, const void* forces, const void* torques, const void* grad_velocities,
                      const void* grad_angular, void* grad_forces, void* grad_torques, void* grad_positions,
                      int dtype, hipStream_t s)
{
  int rc = DPS_OK;
  const bool tterms = torques || grad_angular;
  const bool p = h->prof;
  h->prof = false;
  rc = locate_force_derivatives(h, s);
  // u = velocity grid of (F, T) into d_gridA, v = that of (gU, gW) into d_grid
  if (!rc) rc = do_spread(h, forces, torques, dtype, s);
  if (!rc) rc = do_solve(h, s, 1);
  std::swap(h->d_grid, h->d_gridA);
  if (!rc) rc = do_spread(h, grad_velocities, grad_angular, dtype, s);
  if (!rc) rc = do_solve(h, s, 1);
  if (!rc && grad_forces) rc = do_interp(h, grad_forces, grad_torques, dtype, s);
  if (!rc) {
    const InterpArgs A = interp_args(h);
    const KernelDesc K = force_deriv_kernel(h);
    if (mixed(h)) {
      if (dtype == DPS_F32)
        launch_interp_grad<float, float>(A, h->d_sup_md, h->d_rec_md, h->rs_md, K, h->d_gridA, h->d_grid,
                                         grad_velocities, forces, grad_positions, s);
      else
        launch_interp_grad<float, double>(A, h->d_sup_md, h->d_rec_md, h->rs_md, K, h->d_gridA, h->d_grid,
                                          grad_velocities, forces, grad_positions, s);
    } else {
      if (dtype == DPS_F32)
        launch_interp_grad<double, float>(A, h->d_sup_md, h->d_rec_md, h->rs_md, K, h->d_gridA, h->d_grid,
                                          grad_velocities, forces, grad_positions, s);
      else
        launch_interp_grad<double, double>(A, h->d_sup_md, h->d_rec_md, h->rs_md, K, h->d_gridA, h->d_grid,
                                           grad_velocities, forces, grad_positions, s);
    }
    if (tterms) {
      const StencilArgs S = stencil_args(h);
      const void* mu = grad_angular;
      const void* tq = torques;
      if (mixed(h)) {
        launch_stencil_d2<float>(S, h->d_sup_d, h->d_rec_d2, h->rs_d2, s);
        if (dtype == DPS_F32)
          launch_interp_grad_t<float, float>(A, h->d_sup_d, h->d_rec_d2, h->rs_d2, h->kd, h->d_gridA, h->d_grid,
                                             mu, tq, grad_positions, s);
        else
          launch_interp_grad_t<float, double>(A, h->d_sup_d, h->d_rec_d2, h->rs_d2, h->kd, h->d_gridA, h->d_grid,
                                              mu, tq, grad_positions, s);
      } else {
        launch_stencil_d2<double>(S, h->d_sup_d, h->d_rec_d2, h->rs_d2, s);
        if (dtype == DPS_F32)
          launch_interp_grad_t<double, float>(A, h->d_sup_d, h->d_rec_d2, h->rs_d2, h->kd, h->d_gridA, h->d_grid,
                                              mu, tq, grad_positions, s);
        else
          launch_interp_grad_t<double, double>(A, h->d_sup_d, h->d_rec_d2, h->rs_d2, h->kd, h->d_gridA,
                                               h->d_grid, mu, tq, grad_positions, s);
      }
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) rc = fail(DPS_ERR_HIP, std::string("kernel launch failed: ") + hipGetErrorString(e));
  }
  std::swap(h->d_grid, h->d_gridA);
  h->prof = p;
  return rc;
}

}  // namespace

extern "C" {

int dps_create(const dps_params* params, int number_particles, dps_handle** out)
{
  if (!out) return fail(DPS_ERR_ARG, "null output handle");
  *out = nullptr;
  if (!params) return fail(DPS_ERR_ARG, "null parameters");
  const dps_params& P = *params;
  if (P.mode != DPS_MODE_NOWALL && P.mode != DPS_MODE_BOTTOM && P.mode != DPS_MODE_SLIT &&
      P.mode != DPS_MODE_PERIODIC)
    return fail(DPS_ERR_ARG, "mode must be nowall, bottom, slit or periodic");
  if (P.precision != DPS_PREC_DOUBLE && P.precision != DPS_PREC_MIXED)
    return fail(DPS_ERR_ARG, "precision must be DPS_PREC_DOUBLE or DPS_PREC_MIXED");
  if (P.solver != DPS_SOLVER_AUTO && P.solver != DPS_SOLVER_LANE)
    return fail(DPS_ERR_ARG, "solver must be DPS_SOLVER_AUTO or DPS_SOLVER_LANE");
  if (number_particles < 0) return fail(DPS_ERR_ARG, "number of particles must be >= 0");
  if (!(P.Lx > 0) || !(P.Ly > 0) || !(P.zmax > P.zmin) || !(P.viscosity > 0))
    return fail(DPS_ERR_ARG, "box lengths, z extent and viscosity must be positive");
  if (P.nx < 4 || P.ny < 4 || P.nz < 5) return fail(DPS_ERR_ARG, "need nx, ny >= 4 and nz >= 5");
  if ((long)P.nx * P.ny * P.nz > (1L << 31) / 3) return fail(DPS_ERR_ARG, "grid too large");
  dps_handle* h = new dps_handle();
  h->P = P;
  h->N = number_particles;
  int rc = build(h);
  if (!rc) rc = hipGetDevice(&h->device) == hipSuccess ? DPS_OK : fail(DPS_ERR_HIP, "hipGetDevice failed");
  if (!rc) rc = scratch_init(h->device);
  bool fft_user = false;
  if (!rc) {
    if (g_rocfft_users++ == 0) rocfft_setup();
    fft_user = true;
    rc = allocate(h);
    if (!rc) rc = make_plans(h);
  }
  if (rc) {
    const std::string msg = g_err;
    h->release();
    delete h;
    if (fft_user && --g_rocfft_users == 0) rocfft_cleanup();
    g_err = msg;
    return rc;
  }
  *out = h;
  return DPS_OK;
}

int dps_destroy(dps_handle* h)
{
  if (!h) return DPS_OK;
  h->release();
  delete h;
  if (g_rocfft_users > 0 && --g_rocfft_users == 0) rocfft_cleanup();
  return DPS_OK;
}

int dps_set_positions(dps_handle* h, const void* positions, int dtype, void* stream)
{
  int rc = check_handle(h, false);
  if (rc) return rc;
  if (dtype != DPS_F32 && dtype != DPS_F64) return fail(DPS_ERR_ARG, "dtype must be DPS_F32 or DPS_F64");
  if (!positions && h->N > 0) return fail(DPS_ERR_ARG, "null positions");
  hipStream_t s = (hipStream_t)stream;
  launch_reset_err(h->d_err, s);
  const StencilArgs A = stencil_args(h);
  if (mixed(h)) {
    if (dtype == DPS_F32) launch_stencil<float, float>(A, positions, s);
    else launch_stencil<double, float>(A, positions, s);
  } else {
    if (dtype == DPS_F32) launch_stencil<float, double>(A, positions, s);
    else launch_stencil<double, double>(A, positions, s);
  }
  HIPCHK(hipGetLastError());
  launch_sort(h->d_key, h->d_counts, h->d_rank, h->d_start, h->d_order, h->N, h->ntiles, s);
  HIPCHK(hipGetLastError());
  int rcl = build_lists(h, s);
  if (rcl) return rcl;
  HIPCHK(hipMemcpyAsync(h->h_err, h->d_err, sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (*h->h_err) {
    h->positions_set = false;
    return fail(DPS_ERR_RANGE, "Position is outside of z extent");
  }
  h->positions_set = true;
  return DPS_OK;
}

int dps_update_positions(dps_handle* h, void* positions, const void* velocities, double dt, int dtype,
                         void* stream)
{
  int rc = check_handle(h, false);
  if (rc) return rc;
  if (dtype != DPS_F32 && dtype != DPS_F64) return fail(DPS_ERR_ARG, "dtype must be DPS_F32 or DPS_F64");
  if (!positions && h->N > 0) return fail(DPS_ERR_ARG, "null positions");
  if (!(dt == dt)) return fail(DPS_ERR_ARG, "dt is not a number");
  hipStream_t s = (hipStream_t)stream;
  // the unit cell is [0, Lx] x [0, Ly] in x, y (as for dps_set_positions)
  const double Lx = h->P.Lx, Ly = h->P.Ly;
  if (dtype == DPS_F32) launch_advance<float>(positions, velocities, dt, h->N, 0.0, Lx, 0.0, Ly, s);
  else launch_advance<double>(positions, velocities, dt, h->N, 0.0, Lx, 0.0, Ly, s);
  HIPCHK(hipGetLastError());
  // without valid positions (first call, or after a reported range error)
  // this is a set: synchronous, with a fresh error flag
  if (!h->positions_set) return dps_set_positions(h, positions, dtype, stream);
  // locate as dps_set_positions does, without its synchronisation: a z-range
  // error stays set on the device until dps_check_positions reads it
  const StencilArgs A = stencil_args(h);
  if (mixed(h)) {
    if (dtype == DPS_F32) launch_stencil<float, float>(A, positions, s);
    else launch_stencil<double, float>(A, positions, s);
  } else {
    if (dtype == DPS_F32) launch_stencil<float, double>(A, positions, s);
    else launch_stencil<double, double>(A, positions, s);
  }
  HIPCHK(hipGetLastError());
  launch_sort(h->d_key, h->d_counts, h->d_rank, h->d_start, h->d_order, h->N, h->ntiles, s);
  HIPCHK(hipGetLastError());
  return build_lists(h, s);
}

int dps_check_positions(dps_handle* h, void* stream)
{
  int rc = check_handle(h, false);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  HIPCHK(hipMemcpyAsync(h->h_err, h->d_err, sizeof(int), hipMemcpyDeviceToHost, s));
  launch_reset_err(h->d_err, s);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s));
  if (*h->h_err) {
    h->positions_set = false;
    return fail(DPS_ERR_RANGE, "Position is outside of z extent");
  }
  return DPS_OK;
}

int dps_mdot(dps_handle* h, const void* forces, const void* torques, void* velocities, void* angular,
             int dtype, void* stream)
{
  int rc = check_handle(h, true);
  if (rc) return rc;
  if (dtype != DPS_F32 && dtype != DPS_F64) return fail(DPS_ERR_ARG, "dtype must be DPS_F32 or DPS_F64");
  if (h->N > 0 && (!forces || !velocities)) return fail(DPS_ERR_ARG, "null forces or velocities");
  if (torques && !h->has_dip) return fail(DPS_ERR_ARG, "torques given but no dipole kernel (w_d = 0)");
  hipStream_t s = (hipStream_t)stream;
  if (h->prof) {
    for (auto& e : h->ev)
      if (!e) HIPCHK(hipEventCreate(&e));
  }
  mark(h, 0, s);
  if ((rc = do_spread(h, forces, torques, dtype, s))) return rc;
  mark(h, 1, s);
  if ((rc = (h->reg && torques) ? do_solve_reg(h, s) : do_solve(h, s, 1))) return rc;
  if ((rc = do_interp(h, velocities, torques ? angular : nullptr, dtype, s))) return rc;
  mark(h, 7, s);
  h->have_timings = h->prof;
  h->have_reg_events = h->reg && torques;
  return DPS_OK;
}

int dps_mdot_adjoint(dps_handle* h, const void* forces, const void* torques, const void* grad_velocities,
                     const void* grad_angular, void* grad_forces, void* grad_torques, void* grad_positions,
                     int dtype, void* stream)
{
  int rc = adjoint_check(h, forces, torques, grad_velocities, grad_angular, grad_forces, grad_torques,
                         grad_positions, dtype);
  if (rc) return rc;
  if (!grad_positions) {
    // the solve is taken as self-adjoint: [dL/dF; dL/dT] = M [gU; gW]
    if (!grad_forces) return DPS_OK;
    return dps_mdot(h, grad_velocities, grad_angular, grad_forces, grad_torques, dtype, stream);
  }
  if ((rc = ensure_adjoint(h))) return rc;
  return adjoint_positions(h, forces, torques, grad_velocities, grad_angular, grad_forces, grad_torques,
                           grad_positions, dtype, (hipStream_t)stream);
}

int dps_mdot_adjoint_noalloc(dps_handle* h, const void* forces, const void* torques,
                             const void* grad_velocities, const void* grad_angular, void* grad_forces,
                             void* grad_torques, void* grad_positions, int dtype, void* stream)
{
  int rc = adjoint_check(h, forces, torques, grad_velocities, grad_angular, grad_forces, grad_torques,
                         grad_positions, dtype);
  if (rc) return rc;
  if (!grad_positions) {
    if (!grad_forces) return DPS_OK;
    return dps_mdot(h, grad_velocities, grad_angular, grad_forces, grad_torques, dtype, stream);
  }
  adjoint_record_sizes(h);
  size_t b[5], off[5], need = 0;
  adjoint_buffer_bytes(h, b);
  for (int i = 0; i < 5; ++i) {
    off[i] = need;
    need += (b[i] + 255) / 256 * 256;
  }
  if (need > (size_t)DPS_NOALLOC_SCRATCH_BYTES)
    return fail(DPS_ERR_ARG, "the position adjoint of this solver needs " + std::to_string(need) +
                                 " bytes of scratch, more than the " +
                                 std::to_string((size_t)DPS_NOALLOC_SCRATCH_BYTES) +
                                 " bytes of dps_mdot_adjoint_noalloc's scratch; use dps_mdot_adjoint");
  unsigned char* base = scratch_of(h->device);
  if (!base) return fail(DPS_ERR_STATE, "the no-allocation scratch is not set up on this device");
  // point the adjoint's buffers at the scratch for this call, then restore
  int4* sup_md = h->d_sup_md;
  void* rec_md = h->d_rec_md;
  int4* zrange_md = h->d_zrange_md;
  void* gridA = h->d_gridA;
  void* rec_d2 = h->d_rec_d2;
  h->d_sup_md = (int4*)(base + off[0]);
  h->d_rec_md = base + off[1];
  h->d_zrange_md = (int4*)(base + off[2]);
  h->d_gridA = base + off[3];
  h->d_rec_d2 = h->rs_d2 > 0 ? (void*)(base + off[4]) : nullptr;
  rc = adjoint_positions(h, forces, torques, grad_velocities, grad_angular, grad_forces, grad_torques,
                         grad_positions, dtype, (hipStream_t)stream);
  h->d_sup_md = sup_md;
  h->d_rec_md = rec_md;
  h->d_zrange_md = zrange_md;
  h->d_gridA = gridA;
  h->d_rec_d2 = rec_d2;
  return rc;
}

int dps_stage_spread(dps_handle* h, const void* forces, const void* torques, int dtype, double* grid_out,
                     void* stream)
{
  int rc = check_handle(h, true);
  if (rc) return rc;
  if (torques && h->reg)
    return fail(DPS_ERR_ARG, "with the regular torque kernel the torque curl is spectral: no real-space torque grid");
  if (dtype != DPS_F32 && dtype != DPS_F64) return fail(DPS_ERR_ARG, "dtype must be DPS_F32 or DPS_F64");
  if (h->N > 0 && (!forces || !grid_out)) return fail(DPS_ERR_ARG, "null forces or grid");
  if (torques && !h->has_dip) return fail(DPS_ERR_ARG, "torques given but no dipole kernel (w_d = 0)");
  hipStream_t s = (hipStream_t)stream;
  if ((rc = do_spread(h, forces, torques, dtype, s))) return rc;
  if (mixed(h)) launch_grid_to_ref<float>(h->d_grid, grid_out, h->Nx, h->Ny, h->Nz, s);
  else launch_grid_to_ref<double>(h->d_grid, grid_out, h->Nx, h->Ny, h->Nz, s);
  HIPCHK(hipGetLastError());
  return DPS_OK;
}

int dps_stage_solve(dps_handle* h, const double* grid_in, double* grid_out, void* stream)
{
  int rc = check_handle(h, false);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  if (mixed(h)) launch_ref_to_grid<float>(grid_in, h->d_grid, h->Nx, h->Ny, h->Nz, s);
  else launch_ref_to_grid<double>(grid_in, h->d_grid, h->Nx, h->Ny, h->Nz, s);
  HIPCHK(hipGetLastError());
  const bool p = h->prof;
  h->prof = false;
  rc = do_solve(h, s, 1);
  h->prof = p;
  if (rc) return rc;
  if (mixed(h)) launch_grid_to_ref<float>(h->d_grid, grid_out, h->Nx, h->Ny, h->Nz, s);
  else launch_grid_to_ref<double>(h->d_grid, grid_out, h->Nx, h->Ny, h->Nz, s);
  HIPCHK(hipGetLastError());
  return DPS_OK;
}

int dps_stage_interp(dps_handle* h, const double* grid_in, void* velocities, void* angular, int dtype,
                     void* stream)
{
  int rc = check_handle(h, true);
  if (rc) return rc;
  if (angular && h->reg)
    return fail(DPS_ERR_ARG, "with the regular torque kernel angular velocities need the spectral vorticity");
  hipStream_t s = (hipStream_t)stream;
  if (mixed(h)) launch_ref_to_grid<float>(grid_in, h->d_grid, h->Nx, h->Ny, h->Nz, s);
  else launch_ref_to_grid<double>(grid_in, h->d_grid, h->Nx, h->Ny, h->Nz, s);
  HIPCHK(hipGetLastError());
  return do_interp(h, velocities, angular, dtype, s);
}

int dps_batch_create(dps_handle* const* handles, int n, int nstreams, dps_batch** out)
{
  if (!out) return fail(DPS_ERR_ARG, "null output batch");
  *out = nullptr;
  if (!handles || n < 1) return fail(DPS_ERR_ARG, "a batch needs at least one handle");
  if (nstreams < 1 || nstreams > 16) return fail(DPS_ERR_ARG, "nstreams must be in [1, 16]");
  for (int i = 0; i < n; ++i)
    if (!handles[i]) return fail(DPS_ERR_ARG, "null handle in batch");
  dps_batch* b = new dps_batch();
  b->hs.assign(handles, handles + n);
  for (int j = 0; j < nstreams; ++j) {
    hipStream_t st = nullptr;
    hipEvent_t e = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
      if (st) (void)hipStreamDestroy(st);
      b->release();
      delete b;
      return fail(DPS_ERR_HIP, "stream or event creation failed");
    }
    b->streams.push_back(st);
    b->ev_join.push_back(e);
  }
  if (hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking) != hipSuccess) {
    b->release();
    delete b;
    return fail(DPS_ERR_HIP, "event or stream creation failed");
  }
  *out = b;
  return DPS_OK;
}

int dps_batch_destroy(dps_batch* b)
{
  if (!b) return DPS_OK;
  b->release();
  delete b;
  return DPS_OK;
}

int dps_batch_mdot(dps_batch* b, const void* const* forces, const void* const* torques, void* const* velocities,
                   void* const* angular, int dtype, void* stream)
{
  if (!b) return fail(DPS_ERR_STATE, "null batch");
  if (!forces || !velocities) return fail(DPS_ERR_ARG, "null forces or velocities array");
  if (dtype != DPS_F32 && dtype != DPS_F64) return fail(DPS_ERR_ARG, "dtype must be DPS_F32 or DPS_F64");
  const int n = (int)b->hs.size();
  std::vector<const void*> key;
  key.reserve(4 * n + 1);
  for (int i = 0; i < n; ++i) {
    int rc = check_handle(b->hs[i], true);
    if (rc) return rc;
    const dps_handle* h = b->hs[i];
    if (h->N > 0 && (!forces[i] || !velocities[i])) return fail(DPS_ERR_ARG, "null forces or velocities");
    const void* t = torques ? torques[i] : nullptr;
    if (t && !h->has_dip) return fail(DPS_ERR_ARG, "torques given but no dipole kernel (w_d = 0)");
    key.push_back(forces[i]);
    key.push_back(t);
    key.push_back(velocities[i]);
    key.push_back(angular ? angular[i] : nullptr);
  }
  key.push_back((const void*)(intptr_t)(dtype + 1));
  hipStream_t s = (hipStream_t)stream;
  if (!b->exec || key != b->key) {
    b->drop_graph();
    // capturing records no work: capture on the batch's own stream, launch on
    // the caller's
    hipStream_t c = b->cap;
    HIPCHK(hipStreamBeginCapture(c, hipStreamCaptureModeThreadLocal));
    int rc = DPS_OK;
    hipError_t e = hipEventRecord(b->ev_fork, c);
    const int ns = (int)b->streams.size();
    for (int j = 0; j < ns && e == hipSuccess; ++j) e = hipStreamWaitEvent(b->streams[j], b->ev_fork, 0);
    for (int i = 0; i < n && e == hipSuccess && !rc; ++i) {
      dps_handle* h = b->hs[i];
      const bool p = h->prof;
      h->prof = false;   // no timing events inside the graph
      rc = dps_mdot(h, forces[i], torques ? torques[i] : nullptr, velocities[i], angular ? angular[i] : nullptr,
                    dtype, (void*)b->streams[i % ns]);
      h->prof = p;
      h->have_timings = false;
    }
    for (int j = 0; j < ns && e == hipSuccess; ++j) {
      e = hipEventRecord(b->ev_join[j], b->streams[j]);
      if (e == hipSuccess) e = hipStreamWaitEvent(c, b->ev_join[j], 0);
    }
    hipGraph_t g = nullptr;
    const hipError_t e2 = hipStreamEndCapture(c, &g);
    if (rc || e != hipSuccess || e2 != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      if (rc) return rc;
      return fail(DPS_ERR_HIP, std::string("batch capture failed: ") + hipGetErrorString(e != hipSuccess ? e : e2));
    }
    b->graph = g;
    const hipError_t e3 = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
    if (e3 != hipSuccess) {
      b->drop_graph();
      return fail(DPS_ERR_HIP, std::string("batch graph instantiation failed: ") + hipGetErrorString(e3));
    }
    b->key = key;
  }
  HIPCHK(hipGraphLaunch(b->exec, s));
  return DPS_OK;
}

int dps_batch_tune(dps_batch* b, const void* const* forces, const void* const* torques, int dtype, void* stream,
                   int* chunk_out)
{
  if (!b) return fail(DPS_ERR_STATE, "null batch");
  if (!forces) return fail(DPS_ERR_ARG, "null forces array");
  if (dtype != DPS_F32 && dtype != DPS_F64) return fail(DPS_ERR_ARG, "dtype must be DPS_F32 or DPS_F64");
  hipStream_t s = (hipStream_t)stream;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIPCHK(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return fail(DPS_ERR_HIP, "event creation failed");
  }
  static const int cands[] = {8, 12, 16, 24, 32, 48};
  int rc = DPS_OK;
  for (size_t i = 0; i < b->hs.size() && !rc; ++i) {
    dps_handle* h = b->hs[i];
    if ((rc = check_handle(h, true))) break;
    if (h->N > 0 && !forces[i]) { rc = fail(DPS_ERR_ARG, "null forces"); break; }
    const void* t = torques ? torques[i] : nullptr;
    if (t && !h->has_dip) { rc = fail(DPS_ERR_ARG, "torques given but no dipole kernel (w_d = 0)"); break; }
    const int keep = h->ZC;
    int best = keep;
    float best_ms = 1e30f;
    for (int zc : cands) {
      // the candidate must fit the passes do_spread will run for these inputs
      if (zc > h->Nz || zc < h->zc_min || !spread_fits(h, t != nullptr, zc)) continue;
      h->ZC = zc;
      if ((rc = build_lists(h, s))) break;
      // one untimed call, then the spreading of three calls timed together
      if ((rc = do_spread(h, forces[i], t, dtype, s))) break;
      if (hipEventRecord(e0, s) != hipSuccess) { rc = fail(DPS_ERR_HIP, "event record failed"); break; }
      for (int r = 0; r < 3 && !rc; ++r) rc = do_spread(h, forces[i], t, dtype, s);
      if (rc) break;
      float ms = 0;
      if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
          hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
        rc = fail(DPS_ERR_HIP, "timing failed");
        break;
      }
      if (ms < best_ms) { best_ms = ms; best = zc; }
    }
    h->ZC = best;
    if (!rc) rc = build_lists(h, s);
    if (chunk_out) chunk_out[i] = best;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  b->drop_graph();   // launch shapes changed
  return rc;
}

int dps_batch_set_chunks(dps_batch* b, const int* chunks, int torques, void* stream)
{
  if (!b) return fail(DPS_ERR_STATE, "null batch");
  if (!chunks) return fail(DPS_ERR_ARG, "null chunks");
  hipStream_t s = (hipStream_t)stream;
  int rc;
  // every instance is checked before any changes, so a refused call leaves
  // the batch as it was
  for (size_t i = 0; i < b->hs.size(); ++i) {
    dps_handle* h = b->hs[i];
    if ((rc = check_handle(h, true))) return rc;
    const int zc = chunks[i];
    if (zc > h->Nz || zc < h->zc_min || !spread_fits(h, torques && h->has_dip, zc))
      return fail(DPS_ERR_ARG, "spreading z chunk " + std::to_string(zc) + " does not fit instance " +
                                   std::to_string(i));
  }
  for (size_t i = 0; i < b->hs.size(); ++i) {
    dps_handle* h = b->hs[i];
    h->ZC = chunks[i];
    if ((rc = build_lists(h, s))) return rc;
  }
  b->drop_graph();   // launch shapes changed
  return DPS_OK;
}

int dps_get_info(const dps_handle* h, int* out)
{
  int rc = check_handle(h, false);
  if (rc) return rc;
  if (!out) return fail(DPS_ERR_ARG, "null output");
  out[0] = h->Nx;
  out[1] = h->Ny;
  out[2] = h->Nz;
  out[3] = h->N;
  out[4] = h->km.nzeff;
  out[5] = h->has_dip ? h->kd.nzeff : 0;
  out[6] = h->km.wzmax;
  out[7] = h->has_dip ? h->kd.wzmax : 0;
  return DPS_OK;
}

int dps_set_profiling(dps_handle* h, int enable)
{
  int rc = check_handle(h, false);
  if (rc) return rc;
  h->prof = enable != 0;
  h->have_timings = false;
  return DPS_OK;
}

int dps_get_timings(dps_handle* h, double* out)
{
  int rc = check_handle(h, false);
  if (rc) return rc;
  if (!h->have_timings) return fail(DPS_ERR_STATE, "no profiled Mdot available");
  HIPCHK(hipEventSynchronize(h->ev[7]));
  // events: 0 start, 1 spread done, 2 fft fwd, 3 z fwd, 4 bvp, 5 z inv, 6 fft inv, 7 interp
  // (fused spectral kernel: 3 and 5 are not recorded; bvp spans 2 -> 4,
  // fft inverse 4 -> 6)
  out[0] = 0.0;   // locate/sort happens in dps_set_positions
  for (int i = 0; i < 7; ++i) {
    int a = i, b = i + 1;
    // the regular torque kernel's path records every event (torque curl and
    // vorticity as the z stages) unless triply periodic
    if ((h->fused && !(h->reg && h->have_reg_events)) || h->geometry == DPS_MODE_PERIODIC) {
      if (i == 2 || i == 4) {
        out[i + 1] = 0.0;
        continue;
      }
      if (i == 3 || i == 5) a = i - 1;
    }
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, h->ev[a], h->ev[b]));
    out[i + 1] = 1000.0 * ms;
  }
  return DPS_OK;
}

const char* dps_get_precision(const dps_handle* h)
{
  if (h && h->prec == DPS_PREC_MIXED) return "single";
  return "double";
}

const char* dps_last_error(void) { return g_err.c_str(); }

}  // extern "C"
