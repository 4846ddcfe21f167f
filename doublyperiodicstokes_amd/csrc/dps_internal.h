// dps_internal.h -- kernel argument blocks and launchers shared by the host
// orchestration (dps_api.hip) and the kernels (dps_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include "dps_core.h"

#define DPS_MAX_W 32          // largest x/y support (points) a stencil may have
#define DPS_LIST_MAXC 64       // z chunks one spreading entry may reach
#define DPS_SPREAD_BATCH 16   // queued spreading entries staged together
// spreading tile: SP_TX x SP_TY x-y cells
#define SP_TX 16
#define SP_TY 8

namespace dps {

template <class T> struct C2;
template <> struct C2<float> { typedef float2 t; };
template <> struct C2<double> { typedef double2 t; };

// wavenumber of FFT index i on n points over length L (Nyquist positive)
__host__ __device__ inline double kvec(int i, int n, double L)
{
  const int nn = (i <= n / 2) ? i : i - n;
  return 2 * M_PI * (double)nn / L;
}

// Block index -> work group index such that consecutive groups run on the
// same XCD (the 8 XCDs receive blocks round robin and each has its own L2).
__device__ __forceinline__ int xcd_group(int b, int nb)
{
  const int x = b & 7, q = nb >> 3, r = nb & 7;
  return x * q + (x < r ? x : r) + (b >> 3);
}

struct StencilArgs {
  int N;
  double xmin, ymin, zmin;
  double Lx, Ly, Lz, hx, hy;
  int Nx, Ny;
  int has_dip;
  KernelDesc km, kd;
  const double *zext_m, *zw_m, *zext_d, *zw_d;
  double* pos;     // 3N wrapped/shifted positions
  int* err;        // set to 1 if a particle is outside [0, Lz]
  int4 *sup_m, *sup_d;
  void *rec_m, *rec_d;   // per-particle weight records (grid precision)
  int rs_m, rs_d;        // record strides (elements)
  int* key;              // tile index
  int TX, TY, ntx;
  const int* kt_m;       // z fold tables, to record the interior rows reached
  const int* kt_d;
  const double *fA_m, *fA_d, *fB_d;
  int4* zrange;          // per particle: (klo, khi) monopole, (klo, khi) dipole
  int periodic_z;        // triply periodic: uniform z nodes k hz, wrapped like x and y
  double hz;
  int dip_regular;       // the torque kernel is a plain ES kernel: records in the monopole layout
  int first_kind;        // 1: the second (kd) record set only (e.g. force-kernel derivative records)
};

struct SpreadArgs {
  int Nx, Ny, Nz;
  int periodic_z;   // rows of a support wrap modulo Nz
  int TX, TY, ZC, ntx, nty;
  int nch;          // z chunks of ZC rows
  // entries (2 p + kind) of every (tile, chunk) cell, cell = (ty ntx + tx) nch + chunk:
  // lentries[lstart[cell] .. lstart[cell + 1])
  const int *lstart, *lentries;
  int lcap;         // capacity of lentries
  const int4* zrange;
  KernelDesc km, kd;
  const int4 *sup_m, *sup_d;
  const void *rec_m, *rec_d;
  int rs_m, rs_d, rs_max;
  void* grid;   // [3][Nz][Ny][Nx] grid precision
  int zr_dip;   // kind-0 entries cull by the dipole z range (a pass that spreads torques
                // with the torque kernel's records given as km/sup_m/rec_m)
  // paired entries: force and torque of a particle in one entry (the two
  // kernels share their x/y supports); staged words: the spreading part of
  // the force record (pm words), of the torque record (pd words), then the
  // header at ho: both supports, the force, the torque
  int pair, pm, pd;
  int ho;       // header offset in a staged entry
  // fixed-point accumulation (FP32 grids): the pass's force and torque
  // maxima as namax partial pairs from k_absmax, and per-entry bounds
  // |contribution| <= K0 max|F| + K1 max|T| at any grid point
  const float* amax;
  int namax;
  double K0, K1;
};
size_t spread_lds_bytes(const SpreadArgs& A, size_t tsz);
// partial maxima of |F| and |T| (3n values each; T may be null) into out[2 b], out[2 b + 1]
constexpr int DPS_AMAX_BLOCKS = 64;
template <class Tin> void launch_absmax(const void* F, const void* Tq, int n3, float* out, hipStream_t s);

// Spreading entry lists of one pass, built from the located particles: every
// (particle, kind) entry is listed in each (tile, z chunk) cell its support
// reaches.  rows0: z rows of kind-0 entries, 0 = zrange.x..y (force kernel),
// 1 = zrange.z..w (torque kernel), 2 = the union of both (paired entries);
// kind-1 entries (nkind = 2) use zrange.z..w and sup1.
struct ListArgs {
  int N, Nx, Ny, Nz, TX, TY, ZC, ntx, nty, nch;
  int periodic_z;
  int nkind, rows0;
  int wx0, wy0, wx1, wy1;
  const int4 *sup0, *sup1, *zrange;
  int *counts, *fill, *start, *entries;
  int cap;   // capacity of entries
};
void launch_lists(const ListArgs& A, hipStream_t s);
// upper bound of the cells one entry of support w (points) and z span S (rows) reaches
inline int list_cells_bound(int wx, int wy, int S, int TX, int TY, int ZC)
{
  return ((wx + TX - 1) / TX + 2) * ((wy + TY - 1) / TY + 2) * ((S + ZC - 1) / ZC + 2);
}
// words staged per thread and spreading entry for records of rs_max words (0: too long)
int spread_npf(int rs_max);

struct InterpArgs {
  int N, Nx, Ny, Nz;
  const int* order;
  KernelDesc km, kd;
  const int4 *sup_m, *sup_d;
  const void *rec_m, *rec_d;
  int rs_m, rs_d;
  const int *kt_m, *kt_d;
  const double *fA_m, *fA_d, *fB_d;
  const void* grid;
  int periodic_z;   // rows of a support wrap modulo Nz
  int pair;         // velocities and angular velocities in one pass (shared x/y supports)
};

// Half the curl of a Fourier-Chebyshev field [c][k][Nm] (x, y derivatives i kx,
// i ky with the unpaired Nyquist entries zeroed; z derivative by the Chebyshev
// coefficient recurrence on [0, 2H]), times `scale`; added to `out` or stored.
template <class T>
void launch_curl_cheb(const void* in, void* out, int Nx, int Ny, int Nz, int Nm, const double* kxs,
                      const double* kys, double H, double scale, int add, hipStream_t s);
// Half the Fourier curl of an r2c spectrum [c][kz][ky][kx] (unpaired Nyquist
// entries of each direction zeroed); added to `out` or stored.
template <class T>
void launch_tp_curl(const void* in, void* out, int Nx, int Ny, int Nz, double Lx, double Ly, double Lz, int add,
                    hipStream_t s);

// y pass of the planar transforms: in-place c2c of length Ny along ky of the
// spectrum [nslices][Ny][Nkx], dir -1 forward / +1 inverse (unnormalised);
// tw holds exp(-2 pi i k / Ny), k < Ny, in grid precision.
// x pass of the planar transforms: real rows of NX = Nx points ([row][x]) to
// half-complex rows of Nx/2 + 1 values ([row][kx]), dir -1 forward (real ->
// complex), +1 inverse (complex -> real, unnormalised, imaginary parts of
// kx = 0 and kx = Nx/2 ignored); tw holds exp(-2 pi i k / Nx), k < Nx, in
// grid precision.  Out of place.
bool fft_x_supported(int Nx);
template <class T>
void launch_fft_x(const void* in, void* out, int nrows, int Nx, const void* tw, int dir, hipStream_t s);
bool fft_y_supported(int Ny);
template <class T>
void launch_fft_y(void* data, int nslices, int Nkx, int Ny, const void* tw, int dir, hipStream_t s);

// Position gradient of the adjoint: per particle p (records of the force
// kernel's derivatives, dipole layout), with G_u[j][i] = D_j-interpolation of
// component i of grid gu and likewise G_v for gv,
//   grad[p][j] = -sum_i (lam[p][i] G_u[j][i] + f[p][i] G_v[j][i]).
template <class T, class Tio>
void launch_interp_grad(const InterpArgs& A, const int4* sup, const void* rec, int rs, const KernelDesc& K,
                        const void* gu, const void* gv, const void* lam, const void* f, void* grad,
                        hipStream_t s);

// Torque-kernel records of the position adjoint (4 wx + 4 wy + 4 wzmax words
// per particle, on the torque kernel's supports sup_d) and the torque terms
// of the position gradient, added to grad:
//   grad[p][k] += 1/2 sum eps_abc (mu[p][a] J_kb u_c - t[p][c] J_kb v_a).
template <class T>
void launch_stencil_d2(const StencilArgs& A, const int4* sup_d, void* rec, int rs, hipStream_t s);
template <class T, class Tio>
void launch_interp_grad_t(const InterpArgs& A, const int4* sup, const void* rec, int rs, const KernelDesc& K,
                          const void* gu, const void* gv, const void* mu, const void* t, void* grad, hipStream_t s);

// Fourier solve of the triply periodic Stokes problem on the r2c spectrum
// [c][kz][ky][kx] (kx in [0, Nx/2]), scaled by 1/(Nx Ny Nz) for the c2r
template <class T>
void launch_tp_solve(const void* in, void* out, int Nx, int Ny, int Nz, double Lx, double Ly, double Lz,
                     double eta, hipStream_t s);

struct ZArgs {
  int N, NH, NE, NO;   // nodes, node pairs, even/odd coefficient counts
  int Nm, Nkx;         // modes per z level, kx count
  const void *TfE, *TfO, *TiE, *TiO;   // grid precision
  // wall correction (inverse only)
  int nbasis;
  const double2* corr;   // [Nm][3][4]
  const double *kxs, *kys, *zpts;
  double Hs;
  double inv_nxy;   // 1/(Nx Ny): the c2r transform is unnormalised
  int accumulate;   // matrix-core inverse only: add the values to `out`
};

struct BvpArgs {
  BvpConsts B;   // pointers are device pointers
  int Nx, Ny, Nkx, Nm;
  int Ms;                  // scratch stride: Nm + nspecial lanes
  double Lx, Ly;
  const int* special;      // modes touching a Nyquist index
  int nspecial;
  double2* u2;             // [nspecial][3][Nz] mirror solutions
  double *s_ib, *s_ab;     // [Nz+1][Ms], [Nz][Ms]
  double2 *s_P, *s_CP, *s_DP, *s_U0, *s_U1, *s_U2;   // [Nz][Ms] each
  double2* corr;           // [Nm][3][4]
  double2* corr2;          // [nspecial][3][4]
};

template <class Tin, class T> void launch_stencil(const StencilArgs& A, const void* pos, hipStream_t s);
void launch_reset_err(int* e, hipStream_t s);
template <class Tp>
void launch_advance(void* pos, const void* vel, double dt, int N, double xmin, double Lx, double ymin, double Ly,
                    hipStream_t s);
void launch_sort(int* key, int* counts, int* rank, int* start, int* order, int N, int ntiles,
                 hipStream_t s);
template <class T, class Tin> void launch_spread(const SpreadArgs& A, const void* F, const void* Tq, hipStream_t s);
template <class T, class Tout> void launch_interp(const InterpArgs& A, void* vel, void* ang, hipStream_t s);
size_t z_lds_bytes(const ZArgs& A, bool mixed);
// z transforms of the one-lane chain as MFMA GEMMs (false: shape not
// supported, the caller keeps the VALU kernels)
template <class T> bool zgemm_supported(const ZArgs& A);
template <class T> bool launch_zfwd_mma(const ZArgs& A, const void* in, void* out, hipStream_t s);
template <class T> bool launch_zinv_mma(const ZArgs& A, const void* in, void* out, hipStream_t s);
template <class T> void launch_zfwd(const ZArgs& A, const void* in, void* out, hipStream_t s);
template <class T> void launch_zinv(const ZArgs& A, const void* in, void* out, hipStream_t s);
template <class T> void launch_bvp(const BvpArgs& A, const void* in, void* out, hipStream_t s);
// Fused spectral solve: z transform -> per-mode BVP on lane groups ->
// Nyquist averaging -> inverse z transform with wall corrections.
struct SpecArgs {
  BvpConsts B;   // device pointers
  int Nx, Ny, Nkx, Nm;
  double Lx, Ly;
  int NH;                 // node pairs ceil(Nz/2)
  int lpm, e, G;          // lanes per mode, rows per lane, modes per block
  const void* TfP;        // [NH][LPM][E] forward rows in lane order (grid precision)
  const void* TiP;        // [ceil(Nz/2)][LPM][E] inverse rows in lane order (scaled by 1/(Nx Ny))
  // matrix-core operands of the FP32 transforms (16-lane shapes), in
  // v_mfma 16x16x4 lane order, zero-padded to LPM E / 8 steps: [2][MT][KSM][64]
  // forward (parity, coefficient tile, 4 node pairs) and inverse (parity,
  // node-pair tile, 4 coefficients of that parity; scaled by 1/(Nx Ny))
  const float* TfM;
  const float* TiM;
  const int* fac_idx;     // [Nm] entry of each mode in the factor tables (its |k|^2)
  const double* fac_ib;   // [nk][LPM][E] reciprocal pivots of each lane's rows
  const double* fac_ab;   // [nk][LPM][E] A^{-1}B of each lane's rows
  const double* fac_G;    // [nk][4] inverse Schur complement
  const double *kxs, *kys;   // grid wavenumbers by index
  const double* zpts;
  double Hs, inv_nxy;
};
// lane-group shape for Nz rows: returns false when Nz is too large
bool spectral_shape(int Nz, int* lpm, int* e);
// FP32 z transforms on the matrix cores: 16-lane groups of a mixed-precision
// solve (16 modes per workgroup).  8-lane groups keep the VALU transforms:
// their 24-row parities pad to 32-row tiles and the preloaded tables push
// the solve into scratch (Nz = 48: 836 against 771 us)
constexpr bool spectral_mfma(int lpm, bool mixed) { return mixed && lpm == 16; }
// LDS of one spectral workgroup: 9 row tables of LPM E doubles, 12 wall
// corrections per mode (and |k| per mode on the matrix-core path),
// 3 components x LPM E rows x (G + 1) columns
constexpr size_t spectral_lds_bytes(int lpm, int e, int G, bool mixed)
{
  return (size_t)9 * lpm * e * 8 + (size_t)G * (12 * 16 + (spectral_mfma(lpm, mixed) ? 8 : 0)) +
         (size_t)3 * lpm * e * (G + 1) * (mixed ? 8 : 16);
}
// modes per workgroup: the most (up to 256 lanes) within 64 KiB of LDS, else 1
constexpr int spectral_modes_per_block(int lpm, int e, bool mixed)
{
  int G = 256 / lpm;
  while (G > 1 && spectral_lds_bytes(lpm, e, G, mixed) > 65536) G /= 2;
  return G;
}
// waves per SIMD that the LDS of one CU (160 KiB) allows, at most 3: the
// lane-group solve needs 136 (E = 2) to 226 (E = 8) VGPRs, and a tighter
// register cap sends it to scratch
constexpr int spectral_waves_per_simd(int lpm, int e, bool mixed)
{
  const int G = spectral_modes_per_block(lpm, e, mixed);
  const int blocks = (int)(163840 / spectral_lds_bytes(lpm, e, G, mixed));
  const int w = blocks * (G * lpm / 64) / 4;
  return w < 1 ? 1 : (w > 3 ? 3 : w);
}
void launch_bvp_factor(const BvpConsts& B, const double* ksq, int nk, double* s_ib, double* s_ab, int lpm,
                       int e, double* fib, double* fab, double* fG, hipStream_t s);
template <class T> void launch_spectral(const SpecArgs& A, const void* in, void* out, hipStream_t s);
template <class T> void launch_ref_to_grid(const double* ref, void* g, int Nx, int Ny, int Nz, hipStream_t s);
template <class T> void launch_grid_to_ref(const void* g, double* ref, int Nx, int Ny, int Nz, hipStream_t s);

}  // namespace dps
