// dps_kernels.hip -- gfx950 kernels of the doubly periodic Stokes Mdot path.
//
//  k_stencil      wrap/shift positions, z-range check, per-particle support and
//                 1-D ES weights (ParticleList.cpp:394-548), tile key; four
//                 particles per wavefront, 16 lanes each
//  k_tile_*       counting sort of particles by x-y tile
//  k_spread       forces + torque curl onto the grid: one workgroup per
//                 (x-y tile, z chunk), LDS accumulation, wall images folded on
//                 the fly, one coalesced write per grid element
//                 (SpreadInterp.cpp:302-448 + BoundaryConditions.h:22-361 +
//                  SpreadInterpWrapper.cpp:66-107)
//  k_interp       grid -> particles, one wavefront per particle, derivative kernels
//                 for the angular velocity (SpreadInterp.cpp:450-600,
//                 BoundaryConditions.h:377-731, FCM.py:411-426)
//  k_zfwd/k_zinv  Chebyshev transform in z on even/odd folded node pairs
//                 (Transform.cpp:107-304); the inverse adds the analytic wall
//                 correction profiles (DPTools.cpp:590-877) in value space
//  k_bvp          per-wavenumber FP64 pressure + velocity solves, one mode per
//                 lane (DPTools.cpp:81-345)
//
// Grid layout in HBM: real grid [c][z][y][x] (x fastest), spectra
// [c][z or k][ky][kx] with kx in [0, Nx/2] (r2c half spectrum).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <type_traits>
#include "dps_core.h"
#include "dps_internal.h"

namespace dps {

__device__ __forceinline__ int wrapi(int i, int n)
{
  int r = i % n;
  return r < 0 ? r + n : r;
}

// ---------------------------------------------------------------------------
// positions -> stencils
// ---------------------------------------------------------------------------
// Periodic wrap of an offset from the cell origin into [0, L] as the
// reference does it (ParticleList.cpp:561-580): images counted by truncation,
// so x = L stays L.
__device__ __forceinline__ double wrap_periodic(double x, double L)
{
  if (x < 0) {
    const int img = 1 - ((int)(x / L));
    x += img * L;
  } else if (x > L) {
    const int img = (int)(x / L);
    x -= img * L;
  }
  return x;
}

// A step of positions kept on the device: x += dt u, then x and y wrapped
// into the unit cell and written back in place (the reference's update
// rewraps the caller's array, ParticleList.cpp:556-580).  z is left as is
// for k_stencil to range-check.
template <class Tp>
__global__ void k_advance(Tp* __restrict__ pos, const Tp* __restrict__ vel, double dt, int N, double xmin,
                          double Lx, double ymin, double Ly)
{
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  double x = (double)pos[3 * p], y = (double)pos[3 * p + 1], z = (double)pos[3 * p + 2];
  if (vel) {
    x += dt * (double)vel[3 * p];
    y += dt * (double)vel[3 * p + 1];
    z += dt * (double)vel[3 * p + 2];
  }
  pos[3 * p] = (Tp)(wrap_periodic(x - xmin, Lx) + xmin);
  pos[3 * p + 1] = (Tp)(wrap_periodic(y - ymin, Ly) + ymin);
  pos[3 * p + 2] = (Tp)z;
}

// reductions over the 16-lane group of a particle
__device__ __forceinline__ int grp_min(int v)
{
#pragma unroll
  for (int o = 8; o >= 1; o >>= 1) v = min(v, __shfl_xor(v, o, 16));
  return v;
}
__device__ __forceinline__ int grp_max(int v)
{
#pragma unroll
  for (int o = 8; o >= 1; o >>= 1) v = max(v, __shfl_xor(v, o, 16));
  return v;
}
__device__ __forceinline__ int grp_sum(int v)
{
#pragma unroll
  for (int o = 8; o >= 1; o >>= 1) v += __shfl_xor(v, o, 16);
  return v;
}

// stencil_support with the two searches of the descending extended z nodes
// done by the lane group at once: the first entry <= v of a descending array
// is the number of entries > v, counted over the lanes (no chain of
// dependent loads)
__device__ __forceinline__ Stencil grp_stencil_support(double x, double y, double z, double hx, double hy,
                                                       const KernelDesc& k, const double* zext, int sl)
{
  Stencil s;
  s.bx = support_start(x, hx, k.wx);
  s.by = support_start(y, hy, k.wy);
  const double vh = z - k.thresh, vl = z + k.thresh;
  int ch = 0, cl = 0;
  for (int e = sl; e < k.nzeff; e += 16) {
    const double ze = zext[e];
    ch += ze > vh;
    cl += ze > vl;
  }
  const int high = grp_sum(ch), low = grp_sum(cl);
  int ir = high;
  if (ir == k.nzeff) ir -= 1;
  else if (vh > zext[ir]) ir -= 1;
  s.e0 = low;
  s.nz = ir - low + 1;
  if (s.nz < 0) s.nz = 0;
  return s;
}

// Supports and weight records, four particles per wavefront (16 lanes each).
// Lanes form the x and y weights (lane i: point i) and the contributions of
// the extended z support (lane q: entry q, kept in LDS with the interior row
// it folds onto); then lane r sums the contributions of interior row r in
// entry order, in grid precision, as a serial pass adding them to a zeroed
// record would.  nq bounds the entries of a z support (the kernels' largest
// window).
template <class Tin, class T>
__global__ void __launch_bounds__(64) k_stencil(StencilArgs A, const Tin* __restrict__ pos_in, int nq)
{
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  const int sl = threadIdx.x & 15, pg = threadIdx.x >> 4;
  double* zv = reinterpret_cast<double*>(smem_raw) + pg * 4 * nq;    // [4][nq] contributions of each entry
  int* zrow = reinterpret_cast<int*>(reinterpret_cast<double*>(smem_raw) + 16 * nq) + pg * nq;   // [nq] row, -1: none
  const int p0 = blockIdx.x * 4 + pg;
  const bool valid = p0 < A.N;   // the last block's spare groups read particle N-1 and write nothing
  const int p = valid ? p0 : A.N - 1;
  const bool lead = valid && sl == 0;
  // periodic wrap in x, y (ParticleList.cpp:561-580)
  double x = wrap_periodic((double)pos_in[3 * p] - A.xmin, A.Lx);
  double y = wrap_periodic((double)pos_in[3 * p + 1] - A.ymin, A.Ly);
  double z = (double)pos_in[3 * p + 2] - A.zmin;
  if (A.periodic_z) z = wrap_periodic(z, A.Lz);   // ParticleList.cpp:583-595
  if (!(z >= 0 && z <= A.Lz)) {
    if (lead) atomicOr(A.err, 1);
    z = z < 0 ? 0.0 : A.Lz;   // keep the stencil in range; the call reports an error
  }
  if (lead) {
    A.pos[3 * p] = x;
    A.pos[3 * p + 1] = y;
    A.pos[3 * p + 2] = z;
  }
  int4 zr = A.zrange[p];
  const double hxhy = A.hx * A.hy;
  for (int kind = A.first_kind; kind < (A.has_dip ? 2 : 1); ++kind) {
    // deriv: the torque kernel's derivative records; a regular torque kernel
    // (spectral curl) keeps the monopole layout
    const bool deriv = kind && !A.dip_regular;
    const KernelDesc& K = kind ? A.kd : A.km;
    const double* zext = kind ? A.zext_d : A.zext_m;
    const double* zwe = kind ? A.zw_d : A.zw_m;
    const double a2 = K.alpha * K.alpha;
    Stencil s;
    if (A.periodic_z) {
      s.bx = support_start(x, A.hx, K.wx);
      s.by = support_start(y, A.hy, K.wy);
      s.e0 = 0;
      s.nz = 0;
    } else {
      s = grp_stencil_support(x, y, z, A.hx, A.hy, K, zext, sl);
      if (s.nz > K.wzmax) s.nz = K.wzmax;   // cannot happen (wzmax bounds the window)
    }
    T* rec = (T*)(kind ? A.rec_d : A.rec_m) + (size_t)p * (kind ? A.rs_d : A.rs_m);
    // x and y weights (1/norm^3 folded into x), derivative weights after them
    if (valid) {
      for (int i = sl; i < K.wx; i += 16) {
        const double v = clamp_off(fma((double)(s.bx + i), A.hx, -x), K.thresh);
        const double ph = es_phi(v, K.betaw, a2) * K.inv_norm3;
        rec[i] = (T)ph;
        if (deriv)
          rec[K.wx + K.wy + i] = (T)((K.pos_deriv && fabs(v) >= K.thresh) ? 0.0 : ph * es_dfac(v, K.betaw, a2));
      }
      for (int j = sl; j < K.wy; j += 16) {
        const double v = clamp_off(fma((double)(s.by + j), A.hy, -y), K.thresh);
        const double ph = es_phi(v, K.betaw, a2) * 1.0;
        rec[K.wx + j] = (T)ph;
        if (deriv)
          rec[2 * K.wx + K.wy + j] = (T)((K.pos_deriv && fabs(v) >= K.thresh) ? 0.0 : ph * es_dfac(v, K.betaw, a2));
      }
    }
    const int o = deriv ? 2 * (K.wx + K.wy) : K.wx + K.wy;
    const int W = K.wzmax;
    int4 sup;
    sup.x = s.bx;
    sup.y = s.by;
    if (A.periodic_z) {
      // uniform periodic z (ParticleList.cpp:361-368): w_z = W points from
      // support_start, rows wrapped modulo Nz where used; interpolation
      // weight hx hy hz
      const int bz = support_start(z, A.hz, W);
      if (valid)
        for (int q = sl; q < W; q += 16) {
          const double v = clamp_off(fma((double)(bz + q), A.hz, -z), K.thresh);
          const double ph = es_phi(v, K.betaw, a2);
          const double qw = hxhy * A.hz;
          if (!deriv) {
            rec[o + q] = (T)ph;
            rec[o + W + q] = (T)(ph * qw);
          } else {
            const double d = (K.pos_deriv && fabs(v) >= K.thresh) ? 0.0 : ph * es_dfac(v, K.betaw, a2);
            rec[o + q] = (T)ph;
            rec[o + W + q] = (T)d;
            rec[o + 2 * W + q] = (T)(ph * qw);
            rec[o + 3 * W + q] = (T)(d * qw);
          }
        }
      sup.z = bz;
      sup.w = W;
      if (kind) { zr.z = bz; zr.w = bz + W - 1; }
      else { zr.x = bz; zr.y = bz + W - 1; zr.z = 1 << 30; zr.w = -1; }
    } else {
      // z support, wall images folded in and merged: dense weights of the
      // interior rows klo .. klo + span - 1 (span <= W, since the direct rows
      // and the folded images form one run of rows), each image added to the
      // row it lands on with its fold factor (class A for the value kernel,
      // class B for d/dz); open-end ghosts carry factor 0 and drop out
      const int* kt = kind ? A.kt_d : A.kt_m;
      const double* fa = kind ? A.fA_d : A.fA_m;
      const double* fb = deriv ? A.fB_d : nullptr;
      int lo = 1 << 30, hi = -1;
      for (int q = sl; q < s.nz; q += 16) {
        const int e = s.e0 + q;
        const double sa = fa[e], sb = deriv ? fb[e] : 0.0;
        int row = -1;
        if (!(sa == 0.0 && sb == 0.0)) {
          const double v = clamp_off(zext[e] - z, K.thresh);
          const double ph = es_phi(v, K.betaw, a2);
          const double qw = hxhy * zwe[e];
          row = kt[e];
          lo = min(lo, row);
          hi = max(hi, row);
          if (!deriv) {
            zv[q] = ph * sa;                 // spread
            zv[nq + q] = ph * qw * sa;       // interpolation
          } else {
            const double d = (K.pos_deriv && fabs(v) >= K.thresh) ? 0.0 : ph * es_dfac(v, K.betaw, a2);
            zv[q] = ph * sa;
            zv[nq + q] = d * sb;
            zv[2 * nq + q] = ph * qw * sa;
            zv[3 * nq + q] = d * qw * sb;
          }
        }
        zrow[q] = row;
      }
      const int klo = grp_min(lo), khi = grp_max(hi);
      __syncthreads();
      if (valid)
        for (int r = sl; r < W; r += 16) {
          T c0 = (T)0, c1 = (T)0, c2 = (T)0, c3 = (T)0;
          for (int q = 0; q < s.nz; ++q) {
            if (zrow[q] != klo + r) continue;
            c0 += (T)zv[q];
            c1 += (T)zv[nq + q];
            if (deriv) {
              c2 += (T)zv[2 * nq + q];
              c3 += (T)zv[3 * nq + q];
            }
          }
          rec[o + r] = c0;
          rec[o + W + r] = c1;
          if (deriv) {
            rec[o + 2 * W + r] = c2;
            rec[o + 3 * W + r] = c3;
          }
        }
      __syncthreads();   // the entries' LDS is reused by the next kernel kind
      sup.z = (khi >= klo) ? klo : 0;
      sup.w = (khi >= klo) ? khi - klo + 1 : 0;
      if (kind) { zr.z = klo; zr.w = khi; }
      else { zr.x = klo; zr.y = khi; zr.z = 1 << 30; zr.w = -1; }
    }
    if (lead) {
      (kind ? A.sup_d : A.sup_m)[p] = sup;
      if (kind == 0) {
        // tile key from the monopole support centre
        const int cx = wrapi(s.bx + K.wx / 2, A.Nx), cy = wrapi(s.by + K.wy / 2, A.Ny);
        A.key[p] = (cy / A.TY) * A.ntx + (cx / A.TX);
      }
    }
  }
  if (lead) A.zrange[p] = zr;
}

__global__ void k_reset_err(int* e) { *e = 0; }

__global__ void k_fill_int(int* a, int n, int v)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

__global__ void k_tile_count(const int* __restrict__ key, int* counts, int* rank, int N)
{
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < N) rank[p] = atomicAdd(&counts[key[p]], 1);
}

// exclusive scan of counts[0..n) into start[0..n], single workgroup of 1024
__global__ void k_tile_scan(const int* __restrict__ counts, int* start, int n)
{
  __shared__ int part[1024];
  const int t = threadIdx.x;
  const int per = (n + 1023) / 1024;
  const int b = t * per, e = min(n, b + per);
  int s = 0;
  for (int i = b; i < e; ++i) s += counts[i];
  part[t] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    int v = (t >= off) ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = (t > 0) ? part[t - 1] : 0;
  for (int i = b; i < e; ++i) { start[i] = run; run += counts[i]; }
  if (t == 1023) start[n] = part[1023];
}

__global__ void k_tile_scatter(const int* __restrict__ key, const int* __restrict__ rank,
                               const int* __restrict__ start, int* order, int N)
{
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < N) order[start[key[p]] + rank[p]] = p;
}

// ---------------------------------------------------------------------------
// spreading entry lists: (tile, z chunk) cells of every entry
// ---------------------------------------------------------------------------
constexpr int LIST_MAXT = 8;   // distinct tiles along x or y (support <= DPS_MAX_W points)
constexpr int LIST_MAXC = DPS_LIST_MAXC;   // distinct chunks along z

// distinct tiles of size T met by the periodic run b .. b + w - 1 of n points
// (a run never covers the period twice: n >= 2 w)
__device__ __forceinline__ int run_tiles(int b, int w, int n, int T, int* out, int maxn)
{
  int k = 0;
  for (int i = 0; i < w; ++i) {
    const int t = wrapi(b + i, n) / T;
    bool seen = false;
    for (int j = 0; j < k; ++j) seen = seen || out[j] == t;
    if (!seen && k < maxn) out[k++] = t;
  }
  return k;
}

// the cells of entry (p, kind): x tiles, y tiles, z chunks
__device__ __forceinline__ void entry_cells(const ListArgs& A, int p, int kind, int* tx, int* nx, int* ty,
                                            int* ny, int* ch, int* nc)
{
  const int4 sup = (kind ? A.sup1 : A.sup0)[p];
  const int4 zr = A.zrange[p];
  int klo, khi;
  if (kind || A.rows0 == 1) {
    klo = zr.z;
    khi = zr.w;
  } else {
    klo = zr.x;
    khi = zr.y;
    if (A.rows0 == 2 && zr.z <= zr.w) {
      klo = min(klo, zr.z);
      khi = max(khi, zr.w);
    }
  }
  *nx = run_tiles(sup.x, kind ? A.wx1 : A.wx0, A.Nx, A.TX, tx, LIST_MAXT);
  *ny = run_tiles(sup.y, kind ? A.wy1 : A.wy0, A.Ny, A.TY, ty, LIST_MAXT);
  *nc = (khi >= klo) ? run_tiles(klo, khi - klo + 1, A.periodic_z ? A.Nz : (1 << 30), A.ZC, ch, LIST_MAXC) : 0;
}

template <bool SCATTER>
__global__ void k_list_cells(ListArgs A)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.N * A.nkind) return;
  const int p = i / A.nkind, kind = i - p * A.nkind;
  int tx[LIST_MAXT], ty[LIST_MAXT], ch[LIST_MAXC], nx, ny, nc;
  entry_cells(A, p, kind, tx, &nx, ty, &ny, ch, &nc);
  for (int a = 0; a < ny; ++a)
    for (int b = 0; b < nx; ++b)
      for (int c = 0; c < nc; ++c) {
        if (ty[a] >= A.nty || tx[b] >= A.ntx || ch[c] < 0 || ch[c] >= A.nch) continue;   // cannot happen
        const int cell = (ty[a] * A.ntx + tx[b]) * A.nch + ch[c];
        if (!SCATTER) {
          atomicAdd(&A.counts[cell], 1);
        } else {
          const int at = A.start[cell] + atomicAdd(&A.fill[cell], 1);
          if (at < A.cap) A.entries[at] = 2 * p + kind;
        }
      }
}

// ---------------------------------------------------------------------------
// spread
// ---------------------------------------------------------------------------
// one contiguous piece of a periodic support range intersected with a tile:
// support columns b..b+w-1 (unwrapped) against [t0, t0+n), images shifted by
// multiples of the period.  Returns count, first support index in *i0 and the
// tile-local column of that index in *l0.
__device__ __forceinline__ int clip_piece(int b, int w, int t0, int n, int period, int shift,
                                          int* i0, int* l0)
{
  const int lo = max(b + shift, t0), hi = min(b + shift + w, t0 + n);
  if (hi <= lo) return 0;
  *i0 = lo - (b + shift);
  *l0 = lo - t0;
  return hi - lo;
}

// t = col * n + q for t < 2^24: quotient and remainder by a float reciprocal
// with one correction step (no integer division in the point loops)
__device__ __forceinline__ int fdiv(int t, int n, float inv, int* rem)
{
  int c = (int)((float)t * inv);
  int r = t - c * n;
  if (r < 0) { c -= 1; r += n; }
  else if (r >= n) { c += 1; r -= n; }
  *rem = r;
  return c;
}

// Spreading.  One workgroup owns a cell: a tile of SP_TX x SP_TY columns and
// ZC interior z rows.  The cell's (particle, kernel) entries, those whose
// support reaches it, were listed when the positions were set (k_list_cells);
// they are read into an LDS queue.  The entries are staged in batches,
// 16 threads per entry loading its weight record, support and force, one
// batch ahead of the batch being spread, so a batch's global loads are in
// flight together.  A wavefront spreads one entry at a time: each lane owns
// one column of the entry's clipped support (column weights formed once) and
// walks the entry's rows inside the chunk -- with at most 32 columns the wave
// splits the rows into phases -- loading four rows' z weights ahead of their
// LDS atomics.  The z weights are dense over the rows klo .. klo + span - 1
// (folded images merged by the stencil).  Every grid element of the tile is
// written to HBM once, coalesced.
template <class T, class Tin, int NPF>
__global__ void __launch_bounds__(256) k_spread(SpreadArgs A, const Tin* __restrict__ F,
                                                const Tin* __restrict__ Tq)
{
  constexpr int B = DPS_SPREAD_BATCH;     // queued entries staged per batch, 16 threads each
  constexpr int SLOT = 16 * NPF;          // words per staged entry: record, then header at rs_max
  constexpr int TX = SP_TX, TY = SP_TY;
  // padded tile [c][z][y][x]: row stride TX+1, odd plane stride, so that the
  // lanes of one atomic instruction spread over the LDS banks
  constexpr int RS = TX + 1, PL = TY * RS + 1;
  // one dynamic LDS region, no static LDS: the accumulator starts at offset 0
  // so that it is 8-byte aligned for ds_add_f64.  The accumulator is FP64 in
  // both precisions: on gfx950 an FP32 LDS atomic add (ds_add_f32) retires
  // at ~193 cycles per 64-lane instruction per CU, an FP64 one (ds_add_f64)
  // at ~9 (measured), so FP32 weights are summed in FP64 and rounded once on
  // the way out
  //
  // With FP32 grids the tile is int32 fixed point instead: ds_add_u32 is the
  // fastest LDS atomic (~6 cycles), the tile takes half the LDS (more
  // workgroups per CU), and integer sums do not depend on the order of the
  // additions, so the spread grid is bitwise reproducible.  The scale 2^S
  // keeps every partial sum below 2^30: a point receives at most one term
  // per entry of the cell, each at most K0 max|F| + K1 max|T|.
  constexpr bool FX = sizeof(T) == 4;
  typedef typename std::conditional<FX, int, double>::type TA;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  TA* acc = reinterpret_cast<TA*>(smem_raw);
  const int vol = A.ZC * PL;
  T* bst = reinterpret_cast<T*>(acc + 3 * vol);           // [B][SLOT] staged entries
  int* queue = reinterpret_cast<int*>(bst + B * SLOT);     // p * 2 + kind
  float* scl = reinterpret_cast<float*>(queue + 2 * 256);  // fixed-point scale and its inverse
  T* __restrict__ grid = reinterpret_cast<T*>(A.grid);
  // cells in list order, consecutive cells (the chunks of a tile, then the
  // next tile) on one XCD, whose L2 then holds their shared records
  const int cell = xcd_group(blockIdx.x, gridDim.x);
  const int zc = cell % A.nch, tile = cell / A.nch;
  const int tx = tile % A.ntx, ty = tile / A.ntx;
  const int X0 = tx * TX, Y0 = ty * TY, Z0 = zc * A.ZC;
  const int nX = min(TX, A.Nx - X0), nY = min(TY, A.Ny - Y0), nZ = min(A.ZC, A.Nz - Z0);
  for (int i = threadIdx.x; i < 3 * vol; i += blockDim.x) acc[i] = (TA)0;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), nwaves = blockDim.x >> 6;
  // this cell's entries, listed when the positions were set
  const int e0 = A.lstart[cell];
  const int ncand = max(0, min(A.lstart[cell + 1], A.lcap) - e0);   // within the list's capacity
  if (FX && wave == 0) {
    // the scale: a NaN in the inputs propagates to the whole tile
    float fm = 0.0f, tm = 0.0f;
    for (int i = lane; i < A.namax; i += 64) {
      const float a = A.amax[2 * i], b = A.amax[2 * i + 1];
      fm = (a > fm || a != a) ? a : fm;
      tm = (b > tm || b != b) ? b : tm;
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
      const float a = __shfl_xor(fm, o, 64), b = __shfl_xor(tm, o, 64);
      fm = (a > fm || a != a) ? a : fm;
      tm = (b > tm || b != b) ? b : tm;
    }
    if (lane == 0) {
      const double bound = (double)ncand * (A.K0 * fm + A.K1 * tm);
      int e = 0;
      (void)frexp(bound, &e);   // bound < 2^e
      const int S = min(30 - e, 120);
      scl[0] = ldexpf(1.0f, S);
      scl[1] = isfinite(bound) ? ldexpf(1.0f, -S) : __builtin_nanf("");
    }
  }
  // the zeroed accumulator and the scale are complete before any thread adds
  // to the tile or, for a cell without entries, writes it out
  __syncthreads();
  // the scale is folded into each column's coefficients: add takes values
  // already in fixed-point units
  const T sc = FX ? (T)scl[0] : (T)1;
  auto add = [&](TA* d, T v) {
    if constexpr (FX) atomicAdd(d, __float2int_rn((float)v));
    else atomicAdd(d, (double)v);
  };
  // z weight q of a row list of n entries, zero outside [0, n): the index
  // is clamped so the LDS read stays inside the slot
  auto zw = [&](const T* z, int q, int n) {
    const bool in = (unsigned)q < (unsigned)n;
    const T v = z[in ? q : 0];
    return in ? v : (T)0;
  };
  // header: sup.x, sup.y, sup.z (klo), sup.w (span), f0, f1, f2; paired
  // entries: force support, torque support, force, torque
  const int HO = A.ho;
  // staging thread (ss, su) loads words su, su+16, ... of entry ss of a batch
  const int ss = threadIdx.x >> 4, su = threadIdx.x & 15;
  auto prefetch = [&](int qb, int nb, T (&v)[NPF]) {
    if (ss >= nb) return;
    const int ent = queue[qb + ss];
    const int p = ent >> 1, kind = ent & 1;
    if (A.pair) {
      const T* rm = (const T*)A.rec_m + (size_t)p * A.rs_m;
      const T* rd = (const T*)A.rec_d + (size_t)p * A.rs_d;
      const int* sm = reinterpret_cast<const int*>(A.sup_m + p);
      const int* sd = reinterpret_cast<const int*>(A.sup_d + p);
#pragma unroll
      for (int k = 0; k < NPF; ++k) {
        const int off = su + 16 * k;
        T xv = (T)0;
        if (off < A.pm) xv = rm[off];
        else if (off < A.pm + A.pd) xv = rd[off - A.pm];
        else if (off >= HO && off < HO + 4) xv = (T)sm[off - HO];
        else if (off >= HO + 4 && off < HO + 8) xv = (T)sd[off - HO - 4];
        else if (off >= HO + 8 && off < HO + 11) xv = (T)F[3 * (size_t)p + off - HO - 8];
        else if (off >= HO + 11 && off < HO + 14) xv = (T)Tq[3 * (size_t)p + off - HO - 11];
        v[k] = xv;
      }
      return;
    }
    const int rs = kind ? A.rs_d : A.rs_m;
    const T* rec = (const T*)(kind ? A.rec_d : A.rec_m) + (size_t)p * rs;
    const int* sp = reinterpret_cast<const int*>((kind ? A.sup_d : A.sup_m) + p);
    const Tin* fp = (kind ? Tq : F) + 3 * (size_t)p;
#pragma unroll
    for (int k = 0; k < NPF; ++k) {
      const int off = su + 16 * k;
      T xv = (T)0;
      if (off < rs) xv = rec[off];
      else if (off >= HO && off < HO + 4) xv = (T)sp[off - HO];
      else if (off >= HO + 4 && off < HO + 7) xv = (T)fp[off - HO - 4];
      v[k] = xv;
    }
  };
  constexpr int QN = 2 * 256;   // queue capacity
  for (int c0 = 0; c0 < ncand; c0 += QN) {
    const int nq_entries = min(QN, ncand - c0);
    for (int i = threadIdx.x; i < nq_entries; i += blockDim.x) queue[i] = A.lentries[e0 + c0 + i];
    __syncthreads();
    // ---- spread the queued entries in staged batches, one wavefront per entry ----
    T pf[NPF];
    prefetch(0, min(B, nq_entries), pf);
    for (int qb = 0; qb < nq_entries; qb += B) {
      const int nb = min(B, nq_entries - qb);
      if (ss < nb)
#pragma unroll
        for (int k = 0; k < NPF; ++k) bst[ss * SLOT + su + 16 * k] = pf[k];
      __syncthreads();
      if (qb + B < nq_entries) prefetch(qb + B, min(B, nq_entries - qb - B), pf);
      for (int si = wave; si < nb; si += nwaves) {
        const T* st = bst + si * SLOT;
        // the entry's header is the same for every lane: scalar registers
        int kind = __builtin_amdgcn_readfirstlane(queue[qb + si] & 1);
        const int bx = __builtin_amdgcn_readfirstlane((int)st[HO]);
        const int by = __builtin_amdgcn_readfirstlane((int)st[HO + 1]);
        int klo = __builtin_amdgcn_readfirstlane((int)st[HO + 2]);
        int span = __builtin_amdgcn_readfirstlane((int)st[HO + 3]);
        const int wx = kind ? A.kd.wx : A.km.wx, wy = kind ? A.kd.wy : A.km.wy;
        const int W = kind ? A.kd.wzmax : A.km.wzmax;
        const T* wxr = st;
        const T* wyr = st + wx;
        const T* dxr = wyr + wy;
        const T* dyr = dxr + wx;
        const T* zr = kind ? (dyr + wy) : (wyr + wy);   // z weights of rows klo + q
        T f0 = st[HO + 4], f1 = st[HO + 5], f2 = st[HO + 6];
        // paired entry: rows of the union of the force kernel's (klm, spm) and
        // the torque kernel's (kld, spd), torque t
        int om = 0, od = 0, spm = 0, spd = 0;
        T t0 = 0, t1 = 0, t2 = 0;
        const T* pd_w = st + A.pm;   // torque record: [wx][wy][dx][dy][z][dz]
        if (A.pair) {
          kind = 2;
          const int klm = klo;
          spm = span;
          const int kld = __builtin_amdgcn_readfirstlane((int)st[HO + 6]);
          spd = __builtin_amdgcn_readfirstlane((int)st[HO + 7]);
          if (spd > 0) {
            const int lo = (spm > 0) ? min(klm, kld) : kld;
            const int hi = (spm > 0) ? max(klm + spm, kld + spd) : kld + spd;
            klo = lo;
            span = hi - lo;
          }
          om = klo - klm;
          od = klo - kld;
          f0 = st[HO + 8]; f1 = st[HO + 9]; f2 = st[HO + 10];
          t0 = st[HO + 11]; t1 = st[HO + 12]; t2 = st[HO + 13];
        }
        // rows klo + q (+ a period when z is periodic) inside this chunk:
        // q in [qa, qb2)
        for (int szp = A.periodic_z ? -1 : 0; szp <= (A.periodic_z ? 1 : 0); ++szp) {
        const int kls = klo + szp * A.Nz;
        const int qa = max(0, Z0 - kls), qb2 = min(span, Z0 + nZ - kls);
        if (qb2 <= qa) continue;
        for (int sxp = -1; sxp <= 1; ++sxp) {
          int ix0, lx0;
          const int cx = clip_piece(bx, wx, X0, nX, A.Nx, sxp * A.Nx, &ix0, &lx0);
          if (!cx) continue;
          for (int syp = -1; syp <= 1; ++syp) {
            int iy0, ly0;
            const int cy = clip_piece(by, wy, Y0, nY, A.Ny, syp * A.Ny, &iy0, &ly0);
            if (!cy) continue;
            // lane -> one support column (x fastest); with at most 32 columns
            // the wave splits the rows into P phases
            const int L = cx * cy;
            const int P = (L <= 32) ? 64 / L : 1;
            for (int cb = 0; cb < L; cb += (P > 1) ? L : 64) {
              int phase = 0, col = cb + lane;
              if (P > 1) {
                phase = lane / L;
                col = lane - phase * L;
              }
              if (phase >= P || col >= L) continue;
              const int b = col / cx, a = col - b * cx;
              const int i = ix0 + a, j = iy0 + b;
              TA* dcol = acc + (kls - Z0) * PL + (ly0 + b) * RS + (lx0 + a);   // row kls + q at + q PL
              const int q1 = qb2;
              if (kind == 2) {
                // force (monopole) and torque curl at each point, one atomic
                // per component
                const T* wxd = pd_w;
                const T* wyd = wxd + wx;
                const T* dxd = wyd + wy;
                const T* dyd = dxd + wx;
                const T* zAd = dyd + wy;
                const T* zBd = zAd + A.kd.wzmax;
                // with w = wm zm and g = (ax zA, ay zA, az zB) the point
                // receives w f + 1/2 g x t: per component
                //   P_c zm + Q_c zA + R_c zB  (R_2 = 0)
                // with the column's coefficients (scale included) formed once
                const T hs = (T)0.5 * sc;
                const T wm = wxr[i] * wyr[j] * sc;
                const T ax = dxd[i] * wyd[j] * hs, ay = wxd[i] * dyd[j] * hs, az = wxd[i] * wyd[j] * hs;
                const T P0 = wm * f0, P1 = wm * f1, P2 = wm * f2;
                const T Q0 = ay * t2, Q1 = -(ax * t2), Q2 = ax * t1 - ay * t0;
                const T R0 = -(az * t1), R1 = az * t0;
                for (int q = qa + phase; q < q1; q += P) {
                  const int qm = q + om, qd = q + od;
                  const T zm = zw(zr, qm, spm), zA = zw(zAd, qd, spd), zB = zw(zBd, qd, spd);
                  TA* d = dcol + q * PL;
                  add(d, zm * P0 + zA * Q0 + zB * R0);
                  add(d + vol, zm * P1 + zA * Q1 + zB * R1);
                  add(d + 2 * vol, zm * P2 + zA * Q2);
                }
              } else if (kind == 0) {
                const T wxy = wxr[i] * wyr[j] * sc;
                const T g0 = wxy * f0, g1 = wxy * f1, g2 = wxy * f2;
                int q = qa + phase;
                for (; q + 3 * P < q1; q += 4 * P) {
                  const T w0 = zr[q], w1 = zr[q + P], w2 = zr[q + 2 * P], w3 = zr[q + 3 * P];
                  TA* d0 = dcol + q * PL;
                  add(d0, (w0 * g0)); add(d0 + vol, (w0 * g1));
                  add(d0 + 2 * vol, (w0 * g2));
                  TA* d1 = d0 + P * PL;
                  add(d1, (w1 * g0)); add(d1 + vol, (w1 * g1));
                  add(d1 + 2 * vol, (w1 * g2));
                  TA* d2 = d1 + P * PL;
                  add(d2, (w2 * g0)); add(d2 + vol, (w2 * g1));
                  add(d2 + 2 * vol, (w2 * g2));
                  TA* d3 = d2 + P * PL;
                  add(d3, (w3 * g0)); add(d3 + vol, (w3 * g1));
                  add(d3 + 2 * vol, (w3 * g2));
                }
                for (; q < q1; q += P) {
                  const T w = zr[q];
                  TA* d = dcol + q * PL;
                  add(d, (w * g0)); add(d + vol, (w * g1));
                  add(d + 2 * vol, (w * g2));
                }
              } else {
                // 1/2 g x f with g = (ax zA, ay zA, az zB): per component
                // Q_c zA + R_c zB (R_2 = 0), coefficients formed once
                const T hs = (T)0.5 * sc;
                const T ax = dxr[i] * wyr[j] * hs, ay = wxr[i] * dyr[j] * hs, az = wxr[i] * wyr[j] * hs;
                const T Q0 = ay * f2, Q1 = -(ax * f2), Q2 = ax * f1 - ay * f0;
                const T R0 = -(az * f1), R1 = az * f0;
                int q = qa + phase;
                for (; q + P < q1; q += 2 * P) {
                  const T zA0 = zr[q], zB0 = zr[W + q], zA1 = zr[q + P], zB1 = zr[W + q + P];
                  {
                    TA* d = dcol + q * PL;
                    add(d, zA0 * Q0 + zB0 * R0);
                    add(d + vol, zA0 * Q1 + zB0 * R1);
                    add(d + 2 * vol, zA0 * Q2);
                  }
                  {
                    TA* d = dcol + (q + P) * PL;
                    add(d, zA1 * Q0 + zB1 * R0);
                    add(d + vol, zA1 * Q1 + zB1 * R1);
                    add(d + 2 * vol, zA1 * Q2);
                  }
                }
                for (; q < q1; q += P) {
                  const T zA = zr[q], zB = zr[W + q];
                  TA* d = dcol + q * PL;
                  add(d, zA * Q0 + zB * R0);
                  add(d + vol, zA * Q1 + zB * R1);
                  add(d + 2 * vol, zA * Q2);
                }
              }
            }
          }
        }
        }
      }
      __syncthreads();   // the batch's slots and the queue are free again
    }
  }
  // one write per grid element of the tile: thread -> fixed x, rows (z, y)
  // in steps of blockDim / TX
  {
    const int ix = threadIdx.x % TX, rstep = blockDim.x / TX;
    const int nrow = nZ * TY;
    if (ix < nX)
      for (int c = 0; c < 3; ++c) {
        T* gc = grid + ((size_t)c * A.Nz + Z0) * A.Ny * A.Nx + (size_t)Y0 * A.Nx + X0 + ix;
        const TA* ac = acc + c * vol + ix;
        const float isc = FX ? scl[1] : 1.0f;
        for (int row = threadIdx.x / TX; row < nrow; row += rstep) {
          const int kz = row / TY, jy = row % TY;
          if (jy < nY) gc[((size_t)kz * A.Ny + jy) * A.Nx] = FX ? (T)((float)ac[kz * PL + jy * RS] * isc)
                                                                 : (T)ac[kz * PL + jy * RS];
        }
      }
  }
}

// ---------------------------------------------------------------------------
// interpolation: one wavefront per particle, one lane per support column
// ---------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Column-stationary interpolation of one particle (supports of at most 64
// columns): lane c owns support column c for the whole particle, its x/y
// weight products formed once; the wavefront walks the rows together, so
// each row's z weights are the same for every lane (scalar loads) and a row
// costs three grid loads and the multiply-adds.  Lanes past the last
// column add zeros.  Returns the lane's partial sums (wavefront-summed by
// the caller): d[0..2] the monopole sums, d[3..11] the derivative sums
// (x, y, z derivative kernels, three components each).
//   zA: z values of the monopole rows sA.z + q (sA.w rows), zB / zBd: value
//   and d/dz of the derivative rows sB.z + q; NB = 0 skips the derivative.
template <class T, int NB>
__device__ __forceinline__ void interp_columns(const InterpArgs& A, const T* __restrict__ U, int lane, int4 sA,
                                               int4 sB, int wx, int wy, const T* awx, const T* awy,
                                               const T* za, const T* bwx, const T* bwy, const T* bdx,
                                               const T* bdy, const T* zb, const T* zbd, T (&d)[12])
{
#pragma unroll
  for (int i = 0; i < 12; ++i) d[i] = 0;
  const int ncol = wx * wy;
  int lo = sA.z, hi = sA.z + sA.w;
  if (NB && sB.w > 0) {
    lo = (sA.w > 0) ? min(lo, sB.z) : sB.z;
    hi = (sA.w > 0) ? max(hi, sB.z + sB.w) : sB.z + sB.w;
  }
  const int nz = hi - lo, oa = lo - sA.z, ob = lo - sB.z;
  if (lane >= ncol) return;
  const int b = lane / wx, a = lane - b * wx;
  int x = sA.x + a, y = sA.y + b;
  x += (x < 0) ? A.Nx : ((x >= A.Nx) ? -A.Nx : 0);
  y += (y < 0) ? A.Ny : ((y >= A.Ny) ? -A.Ny : 0);
  const T wa = awx[a] * awy[b];
  T gxw = 0, gyw = 0, gzw = 0;
  if (NB) {
    gxw = bdx[a] * bwy[b];
    gyw = bwx[a] * bdy[b];
    gzw = bwx[a] * bwy[b];
  }
  const unsigned zs = (unsigned)A.Ny * (unsigned)A.Nx;
  const size_t cs = (size_t)A.Nz * zs;
  const T* U1 = U + cs;
  const T* U2 = U + 2 * cs;
  const unsigned col = (unsigned)y * (unsigned)A.Nx + (unsigned)x;
#pragma unroll 2
  for (int q = 0; q < nz; ++q) {
    int zr_ = lo + q;
    if (A.periodic_z) zr_ += (zr_ < 0) ? A.Nz : ((zr_ >= A.Nz) ? -A.Nz : 0);
    const unsigned o = (unsigned)zr_ * zs + col;
    const T v0 = U[o], v1 = U1[o], v2 = U2[o];
    const int qa = q + oa;
    const T w = ((unsigned)qa < (unsigned)sA.w) ? wa * za[qa] : (T)0;
    d[0] += w * v0; d[1] += w * v1; d[2] += w * v2;
    if (NB) {
      const int qb = q + ob;
      T zv = 0, zd = 0;
      if ((unsigned)qb < (unsigned)sB.w) {
        zv = zb[qb];
        zd = zbd[qb];
      }
      const T gx = gxw * zv, gy = gyw * zv, gz = gzw * zd;
      d[3] += gx * v0; d[4] += gx * v1; d[5] += gx * v2;
      d[6] += gy * v0; d[7] += gy * v1; d[8] += gy * v2;
      d[9] += gz * v0; d[10] += gz * v1; d[11] += gz * v2;
    }
  }
}

// Interpolation: one wavefront per particle (tile order, mapped so that each
// XCD handles a contiguous run of tiles), wavefront sums.  Supports of at
// most 64 columns take the column-stationary form above; wider ones one
// lane per (column, z) point of the folded support.  Ω from the derivative
// kernels:
// omega = -1/2 (dy vz - dz vy, dz vx - dx vz, dx vy - dy vx)  (FCM.py:424-426)
template <class T, class Tout>
__global__ void __launch_bounds__(256) k_interp(InterpArgs A, Tout* __restrict__ vel,
                                                Tout* __restrict__ ang)
{
  // the particle index is the same for the whole wavefront: scalar registers
  const int g = __builtin_amdgcn_readfirstlane(xcd_group(blockIdx.x, gridDim.x) * (blockDim.x >> 6) +
                                               (int)(threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  if (g >= A.N) return;
  const int p = A.order[g];
  const T* __restrict__ U = reinterpret_cast<const T*>(A.grid);
  const size_t cs = (size_t)A.Nz * A.Ny * A.Nx;
  const size_t zs = (size_t)A.Ny * A.Nx;
  if (A.pair && ang) {
    // both kernels over the union of their rows, each grid point read once
    const int4 sm = A.sup_m[p], sd = A.sup_d[p];
    const int wx = A.km.wx, wy = A.km.wy;
    const T* rm = (const T*)A.rec_m + (size_t)p * A.rs_m;
    const T* rd = (const T*)A.rec_d + (size_t)p * A.rs_d;
    const T* mwx = rm;
    const T* mwy = rm + wx;
    const T* mzi = mwy + wy + A.km.wzmax;   // interpolation weights, rows sm.z + q
    const T* dwx = rd;
    const T* dwy = rd + wx;
    const T* ddx = dwy + wy;
    const T* ddy = ddx + wx;
    const T* dzi = ddy + wy + 2 * A.kd.wzmax;   // value, then d/dz, rows sd.z + q
    const T* ddzi = dzi + A.kd.wzmax;
    if (wx * wy <= 64) {
      T d[12];
      interp_columns<T, 1>(A, U, lane, sm, sd, wx, wy, mwx, mwy, mzi, dwx, dwy, ddx, ddy, dzi, ddzi, d);
#pragma unroll
      for (int i = 0; i < 12; ++i) d[i] = wave_sum(d[i]);
      if (lane == 0) {
        vel[3 * p] = (Tout)d[0];
        vel[3 * p + 1] = (Tout)d[1];
        vel[3 * p + 2] = (Tout)d[2];
        ang[3 * p] = (Tout)((T)(-0.5) * (d[8] - d[10]));
        ang[3 * p + 1] = (Tout)((T)(-0.5) * (d[9] - d[5]));
        ang[3 * p + 2] = (Tout)((T)(-0.5) * (d[4] - d[6]));
      }
      return;
    }
    int lo = sm.z, hi = sm.z + sm.w;
    if (sd.w > 0) {
      lo = (sm.w > 0) ? min(lo, sd.z) : sd.z;
      hi = (sm.w > 0) ? max(hi, sd.z + sd.w) : sd.z + sd.w;
    }
    const int nz = hi - lo, om = lo - sm.z, od = lo - sd.z;
    const int ncol = wx * wy;
    const int npts = ncol * nz;
    const float incol = 1.0f / (float)ncol, iwx = 1.0f / (float)wx;
    T d[12];
#pragma unroll
    for (int i = 0; i < 12; ++i) d[i] = 0;
    for (int t = lane; t < npts; t += 64) {
      int col, a;
      const int q = fdiv(t, ncol, incol, &col);
      const int b = fdiv(col, wx, iwx, &a);
      int x = sm.x + a, y = sm.y + b;
      x += (x < 0) ? A.Nx : ((x >= A.Nx) ? -A.Nx : 0);
      y += (y < 0) ? A.Ny : ((y >= A.Ny) ? -A.Ny : 0);
      int zr_ = lo + q;
      if (A.periodic_z) zr_ += (zr_ < 0) ? A.Nz : ((zr_ >= A.Nz) ? -A.Nz : 0);
      const size_t idx = (size_t)zr_ * zs + (size_t)y * A.Nx + x;
      const T v0 = U[idx], v1 = U[idx + cs], v2 = U[idx + 2 * cs];
      const int qm = q + om, qd = q + od;
      const T zm = (qm >= 0 && qm < sm.w) ? mzi[qm] : (T)0;
      T zv = (T)0, zd = (T)0;
      if (qd >= 0 && qd < sd.w) {
        zv = dzi[qd];
        zd = ddzi[qd];
      }
      const T w = mwx[a] * mwy[b] * zm;
      const T gx = ddx[a] * dwy[b] * zv;
      const T gy = dwx[a] * ddy[b] * zv;
      const T gz = dwx[a] * dwy[b] * zd;
      d[0] += w * v0; d[1] += w * v1; d[2] += w * v2;
      d[3] += gx * v0; d[4] += gx * v1; d[5] += gx * v2;
      d[6] += gy * v0; d[7] += gy * v1; d[8] += gy * v2;
      d[9] += gz * v0; d[10] += gz * v1; d[11] += gz * v2;
    }
#pragma unroll
    for (int i = 0; i < 12; ++i) d[i] = wave_sum(d[i]);
    if (lane == 0) {
      vel[3 * p] = (Tout)d[0];
      vel[3 * p + 1] = (Tout)d[1];
      vel[3 * p + 2] = (Tout)d[2];
      ang[3 * p] = (Tout)((T)(-0.5) * (d[8] - d[10]));
      ang[3 * p + 1] = (Tout)((T)(-0.5) * (d[9] - d[5]));
      ang[3 * p + 2] = (Tout)((T)(-0.5) * (d[4] - d[6]));
    }
    return;
  }
  for (int kind = 0; kind < (ang ? 2 : 1); ++kind) {
    const int4 sup = kind ? A.sup_d[p] : A.sup_m[p];
    const KernelDesc& K = kind ? A.kd : A.km;
    const T* rec = (const T*)(kind ? A.rec_d : A.rec_m) + (size_t)p * (kind ? A.rs_d : A.rs_m);
    const int W = K.wzmax, nz = sup.w, wx = K.wx, wy = K.wy;
    const T* wxr = rec;
    const T* wyr = rec + wx;
    const T* dxr = wyr + wy;
    const T* dyr = dxr + wx;
    const T* zr = kind ? (dyr + wy) : (wyr + wy);
    const T* wzi = zr + (kind ? 2 : 1) * W;   // interpolation weights (value kernel)
    const T* dzi = zr + 3 * W;                // d/dz kernel (dipole only)
    const int ncol = wx * wy;
    if (ncol <= 64) {
      T e[12];
      if (kind == 0) {
        interp_columns<T, 0>(A, U, lane, sup, sup, wx, wy, wxr, wyr, wzi, nullptr, nullptr, nullptr, nullptr,
                             nullptr, nullptr, e);
#pragma unroll
        for (int i = 0; i < 3; ++i) e[i] = wave_sum(e[i]);
        if (lane == 0) {
          vel[3 * p] = (Tout)e[0];
          vel[3 * p + 1] = (Tout)e[1];
          vel[3 * p + 2] = (Tout)e[2];
        }
      } else {
        // derivative kernels only: no monopole rows
        const int4 s0 = make_int4(sup.x, sup.y, sup.z, 0);
        interp_columns<T, 1>(A, U, lane, s0, sup, wx, wy, wxr, wyr, wzi, wxr, wyr, dxr, dyr, wzi, dzi, e);
#pragma unroll
        for (int i = 3; i < 12; ++i) e[i] = wave_sum(e[i]);
        if (lane == 0) {
          ang[3 * p] = (Tout)((T)(-0.5) * (e[8] - e[10]));
          ang[3 * p + 1] = (Tout)((T)(-0.5) * (e[9] - e[5]));
          ang[3 * p + 2] = (Tout)((T)(-0.5) * (e[4] - e[6]));
        }
      }
      continue;
    }
    const int npts = ncol * nz;
    const float incol = 1.0f / (float)ncol, iwx = 1.0f / (float)wx;
    T d[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) d[i] = 0;
    for (int t = lane; t < npts; t += 64) {
      // x fastest (contiguous loads), then y, then z
      int col, a;
      const int q = fdiv(t, ncol, incol, &col);
      const int b = fdiv(col, wx, iwx, &a);
      int x = sup.x + a, y = sup.y + b;
      x += (x < 0) ? A.Nx : ((x >= A.Nx) ? -A.Nx : 0);
      y += (y < 0) ? A.Ny : ((y >= A.Ny) ? -A.Ny : 0);
      int zr_ = sup.z + q;
      if (A.periodic_z) zr_ += (zr_ < 0) ? A.Nz : ((zr_ >= A.Nz) ? -A.Nz : 0);
      const size_t idx = (size_t)zr_ * zs + (size_t)y * A.Nx + x;
      const T v0 = U[idx], v1 = U[idx + cs], v2 = U[idx + 2 * cs];
      if (kind == 0) {
        const T w = wxr[a] * wyr[b] * wzi[q];
        d[0] += w * v0; d[1] += w * v1; d[2] += w * v2;
      } else {
        const T gx = dxr[a] * wyr[b] * wzi[q];
        const T gy = wxr[a] * dyr[b] * wzi[q];
        const T gz = wxr[a] * wyr[b] * dzi[q];
        d[0] += gx * v0; d[1] += gx * v1; d[2] += gx * v2;
        d[3] += gy * v0; d[4] += gy * v1; d[5] += gy * v2;
        d[6] += gz * v0; d[7] += gz * v1; d[8] += gz * v2;
      }
    }
    if (kind == 0) {
#pragma unroll
      for (int i = 0; i < 3; ++i) d[i] = wave_sum(d[i]);
      if (lane == 0) {
        vel[3 * p] = (Tout)d[0];
        vel[3 * p + 1] = (Tout)d[1];
        vel[3 * p + 2] = (Tout)d[2];
      }
    } else {
#pragma unroll
      for (int i = 0; i < 9; ++i) d[i] = wave_sum(d[i]);
      if (lane == 0) {
        ang[3 * p] = (Tout)((T)(-0.5) * (d[5] - d[7]));
        ang[3 * p + 1] = (Tout)((T)(-0.5) * (d[6] - d[2]));
        ang[3 * p + 2] = (Tout)((T)(-0.5) * (d[1] - d[3]));
      }
    }
  }
}

// Position gradient of the adjoint (no counterpart in the reference): one
// wavefront per particle, the derivative records of the force kernel in the
// dipole layout ([wx][wy][dx][dy][z][dz][z interp][dz interp]); both grids
// are read at each point and the two 3x3 gradient tensors reduced over the
// wavefront.  Supports of at most 64 columns are column-stationary (a lane
// per column, its x/y products formed once, the rows walked together with
// wave-uniform z weights); wider ones take a lane per support point.
template <class T, class Tio>
__global__ void __launch_bounds__(256) k_interp_grad(InterpArgs A, const int4* __restrict__ sup_g,
                                                     const T* __restrict__ recs, int rs, KernelDesc K,
                                                     const T* __restrict__ U, const T* __restrict__ V,
                                                     const Tio* __restrict__ lam, const Tio* __restrict__ f,
                                                     Tio* __restrict__ grad)
{
  const int g = __builtin_amdgcn_readfirstlane(xcd_group(blockIdx.x, gridDim.x) * (blockDim.x >> 6) +
                                               (int)(threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  if (g >= A.N) return;
  const int p = A.order[g];
  const size_t cs = (size_t)A.Nz * A.Ny * A.Nx;
  const size_t zs = (size_t)A.Ny * A.Nx;
  const int4 sup = sup_g[p];
  const T* rec = recs + (size_t)p * rs;
  const int W = K.wzmax, nz = sup.w, wx = K.wx, wy = K.wy;
  const T* wxr = rec;
  const T* wyr = rec + wx;
  const T* dxr = wyr + wy;
  const T* dyr = dxr + wx;
  const T* zr = dyr + wy;
  const T* wzi = zr + 2 * W;
  const T* dzi = zr + 3 * W;
  const int ncol = wx * wy;
  const int npts = ncol * nz;
  const float incol = 1.0f / (float)ncol, iwx = 1.0f / (float)wx;
  T du[9], dv[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) du[i] = dv[i] = 0;
  if (ncol <= 64) {
    if (lane < ncol) {
      const int b = lane / wx, a = lane - b * wx;
      int x = sup.x + a, y = sup.y + b;
      x += (x < 0) ? A.Nx : ((x >= A.Nx) ? -A.Nx : 0);
      y += (y < 0) ? A.Ny : ((y >= A.Ny) ? -A.Ny : 0);
      const T cx = dxr[a] * wyr[b], cy = wxr[a] * dyr[b], cz = wxr[a] * wyr[b];
      const unsigned zsu = (unsigned)zs, col = (unsigned)y * (unsigned)A.Nx + (unsigned)x;
      const T* U1 = U + cs;
      const T* U2 = U + 2 * cs;
      const T* V1 = V + cs;
      const T* V2 = V + 2 * cs;
#pragma unroll 2
      for (int q = 0; q < nz; ++q) {
        int zr_ = sup.z + q;
        if (A.periodic_z) zr_ += (zr_ < 0) ? A.Nz : ((zr_ >= A.Nz) ? -A.Nz : 0);
        const unsigned o = (unsigned)zr_ * zsu + col;
        const T u0 = U[o], u1 = U1[o], u2 = U2[o];
        const T v0 = V[o], v1 = V1[o], v2 = V2[o];
        const T gx = cx * wzi[q], gy = cy * wzi[q], gz = cz * dzi[q];
        du[0] += gx * u0; du[1] += gx * u1; du[2] += gx * u2;
        du[3] += gy * u0; du[4] += gy * u1; du[5] += gy * u2;
        du[6] += gz * u0; du[7] += gz * u1; du[8] += gz * u2;
        dv[0] += gx * v0; dv[1] += gx * v1; dv[2] += gx * v2;
        dv[3] += gy * v0; dv[4] += gy * v1; dv[5] += gy * v2;
        dv[6] += gz * v0; dv[7] += gz * v1; dv[8] += gz * v2;
      }
    }
  } else
  for (int t = lane; t < npts; t += 64) {
    int col, a;
    const int q = fdiv(t, ncol, incol, &col);
    const int b = fdiv(col, wx, iwx, &a);
    int x = sup.x + a, y = sup.y + b;
    x += (x < 0) ? A.Nx : ((x >= A.Nx) ? -A.Nx : 0);
    y += (y < 0) ? A.Ny : ((y >= A.Ny) ? -A.Ny : 0);
    int zr_ = sup.z + q;
    if (A.periodic_z) zr_ += (zr_ < 0) ? A.Nz : ((zr_ >= A.Nz) ? -A.Nz : 0);
    const size_t idx = (size_t)zr_ * zs + (size_t)y * A.Nx + x;
    const T gx = dxr[a] * wyr[b] * wzi[q];
    const T gy = wxr[a] * dyr[b] * wzi[q];
    const T gz = wxr[a] * wyr[b] * dzi[q];
    const T u0 = U[idx], u1 = U[idx + cs], u2 = U[idx + 2 * cs];
    const T v0 = V[idx], v1 = V[idx + cs], v2 = V[idx + 2 * cs];
    du[0] += gx * u0; du[1] += gx * u1; du[2] += gx * u2;
    du[3] += gy * u0; du[4] += gy * u1; du[5] += gy * u2;
    du[6] += gz * u0; du[7] += gz * u1; du[8] += gz * u2;
    dv[0] += gx * v0; dv[1] += gx * v1; dv[2] += gx * v2;
    dv[3] += gy * v0; dv[4] += gy * v1; dv[5] += gy * v2;
    dv[6] += gz * v0; dv[7] += gz * v1; dv[8] += gz * v2;
  }
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    du[i] = wave_sum(du[i]);
    dv[i] = wave_sum(dv[i]);
  }
  if (lane == 0) {
    const double l0 = (double)lam[3 * p], l1 = (double)lam[3 * p + 1], l2 = (double)lam[3 * p + 2];
    const double f0 = (double)f[3 * p], f1 = (double)f[3 * p + 1], f2 = (double)f[3 * p + 2];
#pragma unroll
    for (int j = 0; j < 3; ++j)
      grad[3 * p + j] = (Tio)(-(l0 * (double)du[3 * j] + l1 * (double)du[3 * j + 1] + l2 * (double)du[3 * j + 2]) -
                              (f0 * (double)dv[3 * j] + f1 * (double)dv[3 * j + 1] + f2 * (double)dv[3 * j + 2]));
  }
}

// Torque-kernel records of the position adjoint (no counterpart in the
// reference).  Per axis, over the support of the torque kernel: the value phi,
// the forward derivative phi d (at clamped offsets too, as the forward
// weights), the position derivative (zero where the offset was clamped) and
// the second derivative phi (d^2 + d') (zero where clamped), with
// d = -beta o / (alpha^2 s), d' = -beta / (alpha^2 s^3), s = sqrt(1 - o^2 /
// alpha^2).  Layout [w x][d x][dpos x][dd x][same for y][z value, ghost
// class A][z dpos, A][z d, class B][z dd, B] (z with the interpolation
// weights hx hy w_z, on the interior rows sup.z .. sup.z + sup.w - 1 of the
// torque kernel's own records).
//
// Four particles per wavefront, 16 lanes each, as k_stencil: lanes form the x
// and y points, and the entries of the extended z support (kept in LDS with
// the interior row each folds onto); lane r then adds row r's contributions
// in entry order, in record precision, as a serial pass adding them to a
// zeroed record would.  nq bounds the entries of the z support.
template <class T>
__global__ void __launch_bounds__(64) k_stencil_d2(StencilArgs A, const int4* __restrict__ sup_d,
                                                   T* __restrict__ rec_out, int rs, int nq)
{
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  const int sl = threadIdx.x & 15, pg = threadIdx.x >> 4;
  double* zv = reinterpret_cast<double*>(smem_raw) + pg * 4 * nq;    // [4][nq] contributions of each entry
  int* zrow = reinterpret_cast<int*>(reinterpret_cast<double*>(smem_raw) + 16 * nq) + pg * nq;   // [nq] row, -1: none
  const int p0 = blockIdx.x * 4 + pg;
  const bool valid = p0 < A.N;   // the last block's spare groups read particle N-1 and write nothing
  const int p = valid ? p0 : A.N - 1;
  const KernelDesc& K = A.kd;
  const double x = A.pos[3 * p], y = A.pos[3 * p + 1], z = A.pos[3 * p + 2];
  const int4 sup = sup_d[p];
  const double a2 = K.alpha * K.alpha, bw = K.betaw;
  T* rec = rec_out + (size_t)p * rs;
  auto axis = [&](double v, double scale, double* w, double* df, double* dp, double* dd) {
    const double vc = clamp_off(v, K.thresh);
    const double ph = es_phi(vc, bw, a2) * scale;
    const double s = sqrt(1 - vc * vc / a2);
    const double d = -bw * vc / (a2 * s);
    const bool cl = fabs(vc) >= K.thresh;
    *w = ph;
    *df = ph * d;
    *dp = cl ? 0.0 : ph * d;
    *dd = cl ? 0.0 : ph * (d * d - bw / (a2 * s * s * s));
  };
  const int wx = K.wx, wy = K.wy;
  if (valid) {
    for (int i = sl; i < wx; i += 16) {
      double w, df, dp, dd;
      axis(fma((double)(sup.x + i), A.hx, -x), K.inv_norm3, &w, &df, &dp, &dd);
      rec[i] = (T)w;
      rec[wx + i] = (T)df;
      rec[2 * wx + i] = (T)dp;
      rec[3 * wx + i] = (T)dd;
    }
    for (int j = sl; j < wy; j += 16) {
      double w, df, dp, dd;
      axis(fma((double)(sup.y + j), A.hy, -y), 1.0, &w, &df, &dp, &dd);
      rec[4 * wx + j] = (T)w;
      rec[4 * wx + wy + j] = (T)df;
      rec[4 * wx + 2 * wy + j] = (T)dp;
      rec[4 * wx + 3 * wy + j] = (T)dd;
    }
  }
  const int o = 4 * (wx + wy);
  const int W = K.wzmax;
  const double hxhy = A.hx * A.hy;
  if (A.periodic_z) {
    if (valid)
      for (int q = sl; q < W; q += 16) {
        double w = 0, df = 0, dp = 0, dd = 0;
        if (q < sup.w) axis(fma((double)(sup.z + q), A.hz, -z), 1.0, &w, &df, &dp, &dd);
        const double qw = hxhy * A.hz;
        rec[o + q] = (T)(w * qw);
        rec[o + W + q] = (T)(dp * qw);
        rec[o + 2 * W + q] = (T)(df * qw);
        rec[o + 3 * W + q] = (T)(dd * qw);
      }
    return;
  }
  // doubly periodic: extended nodes folded onto the interior rows with the
  // value kernel's factors (class A) or the z-derivative's (class B)
  Stencil st = grp_stencil_support(x, y, z, A.hx, A.hy, K, A.zext_d, sl);
  if (st.nz > nq) st.nz = nq;   // cannot happen (nq bounds the window)
  for (int q = sl; q < st.nz; q += 16) {
    const int e = st.e0 + q;
    const double sa = A.fA_d[e], sb = A.fB_d[e];
    int r = -1;
    if (!(sa == 0.0 && sb == 0.0)) {
      r = A.kt_d[e] - sup.z;
      if (r < 0 || r >= sup.w) r = -1;   // cannot happen: sup spans every row with a factor
    }
    if (r >= 0) {
      double w, df, dp, dd;
      axis(A.zext_d[e] - z, 1.0, &w, &df, &dp, &dd);
      const double qw = hxhy * A.zw_d[e];
      zv[q] = w * qw * sa;
      zv[nq + q] = dp * qw * sa;
      zv[2 * nq + q] = df * qw * sb;
      zv[3 * nq + q] = dd * qw * sb;
    }
    zrow[q] = r;
  }
  __syncthreads();
  if (valid)
    for (int r = sl; r < W; r += 16) {
      T c0 = (T)0, c1 = (T)0, c2 = (T)0, c3 = (T)0;
      for (int q = 0; q < st.nz; ++q) {
        if (zrow[q] != r) continue;
        c0 += (T)zv[q];
        c1 += (T)zv[nq + q];
        c2 += (T)zv[2 * nq + q];
        c3 += (T)zv[3 * nq + q];
      }
      rec[o + r] = c0;
      rec[o + W + r] = c1;
      rec[o + 2 * W + r] = c2;
      rec[o + 3 * W + r] = c3;
    }
}

// Torque terms of the position gradient: with J_kb the interpolation with the
// position derivative along k of the forward weight phi d_b,
//   grad[p][k] += 1/2 sum eps_abc (mu[p][a] J_kb u_c - t[p][c] J_kb v_a).
template <class T, class Tio>
__global__ void __launch_bounds__(256) k_interp_grad_t(InterpArgs A, const int4* __restrict__ sup_g,
                                                       const T* __restrict__ recs, int rs, KernelDesc K,
                                                       const T* __restrict__ U, const T* __restrict__ V,
                                                       const Tio* __restrict__ mu, const Tio* __restrict__ tq,
                                                       Tio* __restrict__ grad)
{
  const int g = __builtin_amdgcn_readfirstlane(xcd_group(blockIdx.x, gridDim.x) * (blockDim.x >> 6) +
                                               (int)(threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  if (g >= A.N) return;
  const int p = A.order[g];
  const size_t cs = (size_t)A.Nz * A.Ny * A.Nx;
  const size_t zs = (size_t)A.Ny * A.Nx;
  const int4 sup = sup_g[p];
  const T* rec = recs + (size_t)p * rs;
  const int W = K.wzmax, nz = sup.w, wx = K.wx, wy = K.wy;
  const T* xw = rec;
  const T* xf = xw + wx;
  const T* xp = xf + wx;
  const T* xdd = xp + wx;
  const T* yw = xdd + wx;
  const T* yf = yw + wy;
  const T* yp = yf + wy;
  const T* ydd = yp + wy;
  const T* zA = ydd + wy;
  const T* zpA = zA + W;
  const T* zfB = zpA + W;
  const T* zddB = zfB + W;
  const int ncol = wx * wy;
  const int npts = ncol * nz;
  const float incol = 1.0f / (float)ncol, iwx = 1.0f / (float)wx;
  // J[k][b] applied to component c of u (c != b) and of v (a != b): acc[k][b][i]
  T au[27], av[27];
#pragma unroll
  for (int i = 0; i < 27; ++i) au[i] = av[i] = 0;
  if (ncol <= 64) {
    // column-stationary, as in k_interp_grad: the nine x/y products of J
    // formed once per lane, the z factor per row uniform over the wavefront
    if (lane < ncol) {
      const int b = lane / wx, a = lane - b * wx;
      int x = sup.x + a, y = sup.y + b;
      x += (x < 0) ? A.Nx : ((x >= A.Nx) ? -A.Nx : 0);
      y += (y < 0) ? A.Ny : ((y >= A.Ny) ? -A.Ny : 0);
      T C[9];
      C[0] = xdd[a] * yw[b]; C[1] = xp[a] * yf[b]; C[2] = xp[a] * yw[b];
      C[3] = xf[a] * yp[b]; C[4] = xw[a] * ydd[b]; C[5] = xw[a] * yp[b];
      C[6] = xf[a] * yw[b]; C[7] = xw[a] * yf[b]; C[8] = xw[a] * yw[b];
      const unsigned zsu = (unsigned)zs, col = (unsigned)y * (unsigned)A.Nx + (unsigned)x;
      for (int q = 0; q < nz; ++q) {
        int zr_ = sup.z + q;
        if (A.periodic_z) zr_ += (zr_ < 0) ? A.Nz : ((zr_ >= A.Nz) ? -A.Nz : 0);
        const unsigned o = (unsigned)zr_ * zsu + col;
        const T u[3] = {U[o], U[o + cs], U[o + 2 * cs]};
        const T v[3] = {V[o], V[o + cs], V[o + 2 * cs]};
        const T za = zA[q], zpa = zpA[q], zfb = zfB[q], zddb = zddB[q];
        T J[9];
        J[0] = C[0] * za; J[1] = C[1] * za; J[2] = C[2] * zfb;
        J[3] = C[3] * za; J[4] = C[4] * za; J[5] = C[5] * zfb;
        J[6] = C[6] * zpa; J[7] = C[7] * zpa; J[8] = C[8] * zddb;
#pragma unroll
        for (int kb = 0; kb < 9; ++kb)
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            au[kb * 3 + c] += J[kb] * u[c];
            av[kb * 3 + c] += J[kb] * v[c];
          }
      }
    }
  } else
  for (int t = lane; t < npts; t += 64) {
    int col, a;
    const int q = fdiv(t, ncol, incol, &col);
    const int b = fdiv(col, wx, iwx, &a);
    int x = sup.x + a, y = sup.y + b;
    x += (x < 0) ? A.Nx : ((x >= A.Nx) ? -A.Nx : 0);
    y += (y < 0) ? A.Ny : ((y >= A.Ny) ? -A.Ny : 0);
    int zr_ = sup.z + q;
    if (A.periodic_z) zr_ += (zr_ < 0) ? A.Nz : ((zr_ >= A.Nz) ? -A.Nz : 0);
    const size_t idx = (size_t)zr_ * zs + (size_t)y * A.Nx + x;
    T J[9];
    J[0] = xdd[a] * yw[b] * zA[q];    // k x, b x
    J[1] = xp[a] * yf[b] * zA[q];     // k x, b y
    J[2] = xp[a] * yw[b] * zfB[q];    // k x, b z
    J[3] = xf[a] * yp[b] * zA[q];     // k y, b x
    J[4] = xw[a] * ydd[b] * zA[q];    // k y, b y
    J[5] = xw[a] * yp[b] * zfB[q];    // k y, b z
    J[6] = xf[a] * yw[b] * zpA[q];    // k z, b x
    J[7] = xw[a] * yf[b] * zpA[q];    // k z, b y
    J[8] = xw[a] * yw[b] * zddB[q];   // k z, b z
    const T u[3] = {U[idx], U[idx + cs], U[idx + 2 * cs]};
    const T v[3] = {V[idx], V[idx + cs], V[idx + 2 * cs]};
#pragma unroll
    for (int kb = 0; kb < 9; ++kb)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        au[kb * 3 + c] += J[kb] * u[c];
        av[kb * 3 + c] += J[kb] * v[c];
      }
  }
#pragma unroll
  for (int i = 0; i < 27; ++i) {
    au[i] = wave_sum(au[i]);
    av[i] = wave_sum(av[i]);
  }
  if (lane == 0) {
    // a NULL torque or angular cotangent counts as zero
    double m[3] = {0, 0, 0}, tt[3] = {0, 0, 0};
    for (int c = 0; c < 3; ++c) {
      if (mu) m[c] = (double)mu[3 * p + c];
      if (tq) tt[c] = (double)tq[3 * p + c];
    }
    // (a, b, c, sign) of the Levi-Civita symbol
    const int E[6][4] = {{0, 1, 2, 1}, {1, 2, 0, 1}, {2, 0, 1, 1}, {0, 2, 1, -1}, {1, 0, 2, -1}, {2, 1, 0, -1}};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      double acc = 0;
#pragma unroll
      for (int e = 0; e < 6; ++e) {
        const int ea = E[e][0], eb = E[e][1], ec = E[e][2];
        acc += 0.5 * E[e][3] * (m[ea] * (double)au[(k * 3 + eb) * 3 + ec] - tt[ec] * (double)av[(k * 3 + eb) * 3 + ea]);
      }
      grad[3 * p + k] = (Tio)((double)grad[3 * p + k] + acc);
    }
  }
}

// ---------------------------------------------------------------------------
// Chebyshev transform in z (folded even/odd node pairs)
// Each workgroup: 64 complex columns (modes) of one component, all nodes.
// ---------------------------------------------------------------------------
template <class T> struct ZCols { enum { value = 64 }; };
template <> struct ZCols<double> { enum { value = 32 }; };

template <class T>
__global__ void __launch_bounds__(256) k_zfwd(ZArgs A, const typename C2<T>::t* __restrict__ in,
                                              typename C2<T>::t* __restrict__ out)
{
  typedef typename C2<T>::t V;
  constexpr int CW = ZCols<T>::value, RS = 256 / CW;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  V* E = reinterpret_cast<V*>(smem_raw);            // [NH][CW]
  V* O = E + (size_t)A.NH * CW;                      // [NH][CW]
  const int c = blockIdx.y;
  const int m0 = blockIdx.x * CW;
  const int col = threadIdx.x % CW;
  const int m = m0 + col;
  const int N = A.N;
  const V* src = in + (size_t)c * N * A.Nm;
  for (int j = threadIdx.x / CW; j < A.NH; j += RS) {
    V a, b;
    a.x = a.y = b.x = b.y = (T)0;
    if (m < A.Nm) {
      a = src[(size_t)j * A.Nm + m];
      b = src[(size_t)(N - 1 - j) * A.Nm + m];
    }
    V e, o;
    if (j == N - 1 - j) { e = a; o.x = o.y = (T)0; }
    else { e.x = a.x + b.x; e.y = a.y + b.y; o.x = a.x - b.x; o.y = a.y - b.y; }
    E[j * CW + col] = e;
    O[j * CW + col] = o;
  }
  __syncthreads();
  if (m >= A.Nm) return;
  V* dst = out + (size_t)c * N * A.Nm;
  // RS is even, so all outputs of one thread share the parity of its row;
  // each pass computes RB of them from one read of the folded column
  constexpr int RB = 4;
  const int row = threadIdx.x / CW;
  const V* X = (row & 1) ? O : E;
  const T* Tbase = (row & 1) ? (const T*)A.TfO : (const T*)A.TfE;
  for (int kb = row; kb < N; kb += RS * RB) {
    T sr[RB], si[RB];
    const T* Tr[RB];
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      const int kr = kb + r * RS;
      Tr[r] = Tbase + (size_t)((kr < N ? kr : kb) >> 1) * A.NH;
      sr[r] = 0;
      si[r] = 0;
    }
    for (int j = 0; j < A.NH; ++j) {
      const V x = X[j * CW + col];
#pragma unroll
      for (int r = 0; r < RB; ++r) {
        const T t = Tr[r][j];
        sr[r] += t * x.x;
        si[r] += t * x.y;
      }
    }
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      const int kr = kb + r * RS;
      if (kr < N) {
        V v;
        v.x = sr[r];
        v.y = si[r];
        dst[(size_t)kr * A.Nm + m] = v;
      }
    }
  }
}

// inverse: values at nodes from coefficients, plus wall-correction profiles
template <class T>
__global__ void __launch_bounds__(256) k_zinv(ZArgs A, const typename C2<T>::t* __restrict__ in,
                                              typename C2<T>::t* __restrict__ out)
{
  typedef typename C2<T>::t V;
  constexpr int CW = ZCols<T>::value, RS = 256 / CW;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  V* Ck = reinterpret_cast<V*>(smem_raw);   // [N][CW]
  const int c = blockIdx.y;
  const int m0 = blockIdx.x * CW;
  const int col = threadIdx.x % CW;
  const int m = m0 + col;
  const int N = A.N;
  const V* src = in + (size_t)c * N * A.Nm;
  for (int k = threadIdx.x / CW; k < N; k += RS) {
    V a;
    a.x = a.y = (T)0;
    if (m < A.Nm) a = src[(size_t)k * A.Nm + m];
    Ck[k * CW + col] = a;
  }
  __syncthreads();
  if (m >= A.Nm) return;
  // wall correction coefficients of this mode/component
  cplx cc[4];
  int nb = A.nbasis;
  double K = 0;
  if (nb) {
    const double2* cr = A.corr + ((size_t)m * 3 + c) * 4;
    for (int b = 0; b < 4; ++b) { cc[b].re = cr[b].x; cc[b].im = cr[b].y; }
    const int iy = m / A.Nkx, ix = m - iy * A.Nkx;
    const double kx = A.kxs[ix], ky = A.kys[iy];
    K = sqrt(ksq_of(kx, ky));
    if (m == 0) nb = 0;
  }
  V* dst = out + (size_t)c * N * A.Nm;
  constexpr int RB = 4;
  for (int jb = threadIdx.x / CW; jb < A.NH; jb += RS * RB) {
    T ar[RB], ai[RB], br[RB], bi[RB];
    const T* te[RB];
    const T* to[RB];
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      const int jr = jb + r * RS;
      const int js = jr < A.NH ? jr : jb;
      te[r] = (const T*)A.TiE + (size_t)js * A.NE;
      to[r] = (const T*)A.TiO + (size_t)js * A.NO;
      ar[r] = ai[r] = br[r] = bi[r] = 0;
    }
    for (int kk = 0; kk < A.NE; ++kk) {
      const V x = Ck[(2 * kk) * CW + col];
#pragma unroll
      for (int r = 0; r < RB; ++r) {
        ar[r] += te[r][kk] * x.x;
        ai[r] += te[r][kk] * x.y;
      }
    }
    for (int kk = 0; kk < A.NO; ++kk) {
      const V x = Ck[(2 * kk + 1) * CW + col];
#pragma unroll
      for (int r = 0; r < RB; ++r) {
        br[r] += to[r][kk] * x.x;
        bi[r] += to[r][kk] * x.y;
      }
    }
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      const int j = jb + r * RS;
      if (j >= A.NH) continue;
      const int j2 = N - 1 - j;
      for (int side = 0; side < 2; ++side) {
        const int jj = side ? j2 : j;
        if (side && j2 == j) break;
        T vr = side ? (ar[r] - br[r]) : (ar[r] + br[r]);
        T vi = side ? (ai[r] - bi[r]) : (ai[r] + bi[r]);
        if (nb) {
          const double z = A.zpts[jj];
          const double e1 = exp(-K * z);
          double sr = cc[0].re * e1, si = cc[0].im * e1;
          if (nb == 2) {
            sr += cc[1].re * (z * e1);
            si += cc[1].im * (z * e1);
          } else {
            const double e2 = exp(K * (z - A.Hs));
            sr += cc[1].re * e2 + cc[2].re * (z * e1) + cc[3].re * (z * e2);
            si += cc[1].im * e2 + cc[2].im * (z * e1) + cc[3].im * (z * e2);
          }
          vr += (T)(sr * A.inv_nxy);
          vi += (T)(si * A.inv_nxy);
        }
        V v;
        v.x = vr;
        v.y = vi;
        dst[(size_t)jj * A.Nm + m] = v;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// BVP: one mode per lane
// ---------------------------------------------------------------------------
template <class T>
struct CoefIn {
  const typename C2<T>::t* p;
  long cs, ks;
  bool conj;
  DPS_HD cplx operator()(int c, int k) const
  {
    const typename C2<T>::t v = p[c * cs + k * ks];
    return conj ? mk((double)v.x, -(double)v.y) : mk((double)v.x, (double)v.y);
  }
};
template <class T>
struct CoefOut {
  typename C2<T>::t* p;
  long cs, ks;
  DPS_HD void operator()(int c, int k, cplx v)
  {
    typename C2<T>::t r;
    r.x = (T)v.re;
    r.y = (T)v.im;
    p[c * cs + k * ks] = r;
  }
};
struct CplxOut {
  double2* p;
  long cs, ks;
  DPS_HD void operator()(int c, int k, cplx v)
  {
    double2 r;
    r.x = v.re;
    r.y = v.im;
    p[c * cs + k * ks] = r;
  }
};

__device__ void store_corr(double2* cr, const ModeResult& r)
{
  for (int c = 0; c < 3; ++c)
    for (int b = 0; b < 4; ++b) {
      double2 v;
      v.x = r.corr[c][b].re;
      v.y = r.corr[c][b].im;
      cr[c * 4 + b] = v;
    }
}

template <class T>
__global__ void __launch_bounds__(64) k_bvp(BvpArgs A, const typename C2<T>::t* __restrict__ fin,
                                            typename C2<T>::t* __restrict__ fout)
{
  // lanes [0, Nm): every mode of the half spectrum; lanes [Nm, Nm + nspecial):
  // the mirror solve of each mode that touches a Nyquist index
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= A.Nm + A.nspecial) return;
  const bool mirror = t >= A.Nm;
  const int s = t - A.Nm;
  const int m = mirror ? A.special[s] : t;
  const int iy = m / A.Nkx, ix = m - iy * A.Nkx;
  CoefIn<T> in;
  in.p = fin + m;
  in.cs = (long)A.B.Nz * A.Nm;
  in.ks = A.Nm;
  in.conj = mirror;
  ScratchView sc;
  sc.stride = A.Ms;
  sc.ib = A.s_ib + t;
  sc.ab = A.s_ab + t;
  sc.P = reinterpret_cast<cplx*>(A.s_P) + t;
  sc.CP = reinterpret_cast<cplx*>(A.s_CP) + t;
  sc.DP = reinterpret_cast<cplx*>(A.s_DP) + t;
  sc.U0 = reinterpret_cast<cplx*>(A.s_U0) + t;
  sc.U1 = reinterpret_cast<cplx*>(A.s_U1) + t;
  sc.U2 = reinterpret_cast<cplx*>(A.s_U2) + t;
  ModeResult r;
  for (int c = 0; c < 3; ++c)
    for (int b = 0; b < 4; ++b) r.corr[c][b] = mk(0, 0);
  r.nbasis = 0;
  if (!mirror) {
    CoefOut<T> out;
    out.p = fout + m;
    out.cs = in.cs;
    out.ks = in.ks;
    if (m == 0) bvp_k0(A.B, in, out, sc.U0, sc.stride);
    else {
      const double kx = kvec(ix, A.Nx, A.Lx), ky = kvec(iy, A.Ny, A.Ly);
      const int nyqx = (A.Nx % 2 == 0) && (ix == A.Nx / 2);
      const int nyqy = (A.Ny % 2 == 0) && (iy == A.Ny / 2);
      bvp_mode(A.B, kx, ky, nyqx, nyqy, in, out, sc, r);
    }
    store_corr(A.corr + (size_t)m * 12, r);
  } else {
    // the reference's c2c transform keeps only the real part of the inverse,
    // i.e. the Hermitian average of a mode and its mirror solved with the
    // mirror's own grid wavenumbers; this lane solves the mirror
    const int ixm = (A.Nx - ix) % A.Nx, iym = (A.Ny - iy) % A.Ny;
    const double kxm = kvec(ixm, A.Nx, A.Lx), kym = kvec(iym, A.Ny, A.Ly);
    const int nyqxm = (A.Nx % 2 == 0) && (ixm == A.Nx / 2);
    const int nyqym = (A.Ny % 2 == 0) && (iym == A.Ny / 2);
    CplxOut o2;
    o2.p = A.u2 + (size_t)s * 3 * A.B.Nz;
    o2.cs = A.B.Nz;
    o2.ks = 1;
    bvp_mode(A.B, kxm, kym, nyqxm, nyqym, in, o2, sc, r);
    store_corr(A.corr2 + (size_t)s * 12, r);
  }
}

// average each Nyquist-touching mode with its mirror solution
template <class T>
__global__ void k_bvp_combine(BvpArgs A, typename C2<T>::t* __restrict__ fout)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= A.nspecial) return;
  const int m = A.special[s];
  const long cs = (long)A.B.Nz * A.Nm;
  for (int c = 0; c < 3; ++c)
    for (int k = 0; k < A.B.Nz; ++k) {
      typename C2<T>::t* q = fout + c * cs + (long)k * A.Nm + m;
      const double2 v2 = A.u2[((size_t)s * 3 + c) * A.B.Nz + k];
      const double re = 0.5 * ((double)q->x + v2.x);
      const double im = 0.5 * ((double)q->y - v2.y);
      q->x = (T)re;
      q->y = (T)im;
    }
  double2* c1 = A.corr + (size_t)m * 12;
  const double2* c2 = A.corr2 + (size_t)s * 12;
  for (int i = 0; i < 12; ++i) {
    double2 v;
    v.x = 0.5 * (c1[i].x + c2[i].x);
    v.y = 0.5 * (c1[i].y - c2[i].y);
    c1[i] = v;
  }
}

// ---------------------------------------------------------------------------
// layout conversions for the stage entry points:
//   reference interior layout (Ny, Nx, Nz, 3) FP64  <->  [c][z][y][x] T
// ---------------------------------------------------------------------------
template <class T>
__global__ void k_ref_to_grid(const double* __restrict__ ref, T* __restrict__ g, int Nx, int Ny, int Nz)
{
  const size_t n = (size_t)3 * Nx * Ny * Nz;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int x = i % Nx;
  size_t r = i / Nx;
  const int y = r % Ny;
  r /= Ny;
  const int z = r % Nz;
  const int c = r / Nz;
  g[i] = (T)ref[(((size_t)y * Nx + x) * Nz + z) * 3 + c];
}

template <class T>
__global__ void k_grid_to_ref(const T* __restrict__ g, double* __restrict__ ref, int Nx, int Ny, int Nz)
{
  const size_t n = (size_t)3 * Nx * Ny * Nz;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int x = i % Nx;
  size_t r = i / Nx;
  const int y = r % Ny;
  r /= Ny;
  const int z = r % Nz;
  const int c = r / Nz;
  ref[(((size_t)y * Nx + x) * Nz + z) * 3 + c] = (double)g[i];
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <class Tin, class T>
void launch_stencil(const StencilArgs& A, const void* pos, hipStream_t s)
{
  if (A.N == 0) return;
  // four particles per 64-lane block, each with [4][nq] doubles and [nq] ints of LDS
  const int nq = std::max(1, std::max(A.km.wzmax, A.has_dip ? A.kd.wzmax : 0));
  hipLaunchKernelGGL((k_stencil<Tin, T>), dim3((A.N + 3) / 4), dim3(64),
                     (size_t)4 * nq * (4 * sizeof(double) + sizeof(int)), s, A, (const Tin*)pos, nq);
}

template <class Tp>
void launch_advance(void* pos, const void* vel, double dt, int N, double xmin, double Lx, double ymin,
                    double Ly, hipStream_t s)
{
  if (N == 0) return;
  hipLaunchKernelGGL((k_advance<Tp>), dim3((N + 127) / 128), dim3(128), 0, s, (Tp*)pos, (const Tp*)vel, dt, N,
                     xmin, Lx, ymin, Ly);
}
template void launch_advance<float>(void*, const void*, double, int, double, double, double, double, hipStream_t);
template void launch_advance<double>(void*, const void*, double, int, double, double, double, double, hipStream_t);

// Triply periodic Fourier solve (TPTools.cpp:6-46, Solvers.py:197-255), one
// thread per r2c mode: with D = i k (the unpaired Nyquist entry of each
// direction zeroed, k^2 not), p = -(D.f)/k^2 and u = (f - D p)/(eta k^2), i.e.
// u = (f - d (d.f)/k^2)/(eta k^2) for the real d = D/i; the k = 0 mode is
// dropped and the 1/(Nx Ny Nz) of the c2r transform folded in.  FP64.
template <class T>
__global__ void k_tp_solve(const typename C2<T>::t* __restrict__ in, typename C2<T>::t* __restrict__ out,
                           int Nx, int Ny, int Nz, double Lx, double Ly, double Lz, double eta)
{
  typedef typename C2<T>::t V;
  const int Nkx = Nx / 2 + 1;
  const size_t nm = (size_t)Nz * Ny * Nkx;
  const size_t m = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nm) return;
  const int ix = (int)(m % Nkx);
  const size_t r = m / Nkx;
  const int iy = (int)(r % Ny), iz = (int)(r / Ny);
  double ur[3] = {0, 0, 0}, ui[3] = {0, 0, 0};
  if (m != 0) {
    const double kx = kvec(ix, Nx, Lx), ky = kvec(iy, Ny, Ly), kz = kvec(iz, Nz, Lz);
    const double ksq = kx * kx + ky * ky + kz * kz;
    const double d[3] = {(Nx % 2 == 0 && ix == Nx / 2) ? 0.0 : kx, (Ny % 2 == 0 && iy == Ny / 2) ? 0.0 : ky,
                         (Nz % 2 == 0 && iz == Nz / 2) ? 0.0 : kz};
    double fr[3], fi[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const V v = in[c * nm + m];
      fr[c] = (double)v.x;
      fi[c] = (double)v.y;
    }
    const double sr = d[0] * fr[0] + d[1] * fr[1] + d[2] * fr[2];
    const double si = d[0] * fi[0] + d[1] * fi[1] + d[2] * fi[2];
    const double iks = 1.0 / ksq;
    const double val = iks / eta / ((double)Nx * (double)Ny * (double)Nz);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      ur[c] = val * (fr[c] - d[c] * sr * iks);
      ui[c] = val * (fi[c] - d[c] * si * iks);
    }
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    V v;
    v.x = (T)ur[c];
    v.y = (T)ui[c];
    out[c * nm + m] = v;
  }
}

template <class T>
void launch_tp_solve(const void* in, void* out, int Nx, int Ny, int Nz, double Lx, double Ly, double Lz,
                     double eta, hipStream_t s)
{
  typedef typename C2<T>::t V;
  const size_t nm = (size_t)Nz * Ny * (Nx / 2 + 1);
  hipLaunchKernelGGL((k_tp_solve<T>), dim3((unsigned)((nm + 255) / 256)), dim3(256), 0, s, (const V*)in,
                     (V*)out, Nx, Ny, Nz, Lx, Ly, Lz, eta);
}
template void launch_tp_solve<float>(const void*, void*, int, int, int, double, double, double, double,
                                     hipStream_t);
template void launch_tp_solve<double>(const void*, void*, int, int, int, double, double, double, double,
                                      hipStream_t);

// Half curl of a Fourier-Chebyshev field (the regular-kernel torque path,
// FCM.py:372-396 for the torque grid, Solvers.py:159-179 for the velocity).
// The z derivative follows the coefficient recurrence (Chebyshev.py:68-94)
// D_k = D_{k+2} + (2/H)(k+1) f_{k+1}, D_0 halved, whose even and odd levels
// form two independent chains.  Four threads per (kx, ky) mode: output x or y
// (the chain of the other component's derivative) times level parity, each
// walking its levels from the top down; the x threads also write z, which
// needs no derivative.  Loads and stores are coalesced across modes.  FP64.
template <class T>
__global__ void k_curl_cheb(const typename C2<T>::t* __restrict__ in, typename C2<T>::t* __restrict__ out,
                            int Nx, int Ny, int Nz, int Nm, const double* __restrict__ kxs,
                            const double* __restrict__ kys, double two_over_H, double scale, int add)
{
  typedef typename C2<T>::t V;
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int m = tid % Nm, role = tid / Nm;   // role: (output x | y) x parity
  if (role >= 4) return;
  const int oc = role >> 1, par = role & 1;   // oc 0: x (derivative of y), 1: y (derivative of x)
  const int Nkx = Nx / 2 + 1;
  const int iy = m / Nkx, ix = m - iy * Nkx;
  const double dx = (Nx % 2 == 0 && ix == Nx / 2) ? 0.0 : kxs[ix];
  const double dy = (Ny % 2 == 0 && iy == Ny / 2) ? 0.0 : kys[iy];
  const size_t cs = (size_t)Nz * Nm;
  const double h = 0.5 * scale;
  const V* fsrc = in + (oc == 0 ? cs : 0);   // the differentiated component
  // levels of this parity from the top: D_k = D_{k+2} + c (k+1) f_{k+1}
  double Dr = 0, Di = 0;
  int k = Nz - 1;
  if ((k & 1) != par) --k;
#pragma unroll 4
  for (; k >= 0; k -= 2) {
    const size_t o = (size_t)k * Nm + m;
    if (k + 1 < Nz) {
      const V f = fsrc[o + Nm];
      const double c = two_over_H * (double)(k + 1);
      Dr += c * (double)f.x;
      Di += c * (double)f.y;
    }
    const double zr = (k == 0) ? 0.5 * Dr : Dr, zi = (k == 0) ? 0.5 * Di : Di;
    const V az = in[o + 2 * cs];
    double w0, w1;
    if (oc == 0) {   // (i dy az - dz ay) / 2
      w0 = h * (-dy * (double)az.y - zr);
      w1 = h * (dy * (double)az.x - zi);
    } else {         // (dz ax - i dx az) / 2
      w0 = h * (zr + dx * (double)az.y);
      w1 = h * (zi - dx * (double)az.x);
    }
    V* d = out + o + (size_t)oc * cs;
    V v;
    if (add) {
      const V b = *d;
      v.x = (T)((double)b.x + w0);
      v.y = (T)((double)b.y + w1);
    } else {
      v.x = (T)w0;
      v.y = (T)w1;
    }
    *d = v;
    if (oc == 0) {   // z: (i dx ay - i dy ax) / 2
      const V ax = in[o], ay = in[o + cs];
      const double z0 = h * (-dx * (double)ay.y + dy * (double)ax.y);
      const double z1 = h * (dx * (double)ay.x - dy * (double)ax.x);
      V* dz = out + o + 2 * cs;
      V u;
      if (add) {
        const V b = *dz;
        u.x = (T)((double)b.x + z0);
        u.y = (T)((double)b.y + z1);
      } else {
        u.x = (T)z0;
        u.y = (T)z1;
      }
      *dz = u;
    }
  }
}

template <class T>
void launch_curl_cheb(const void* in, void* out, int Nx, int Ny, int Nz, int Nm, const double* kxs,
                      const double* kys, double H, double scale, int add, hipStream_t s)
{
  typedef typename C2<T>::t V;
  const int nt = 4 * Nm;
  hipLaunchKernelGGL((k_curl_cheb<T>), dim3((nt + 255) / 256), dim3(256), 0, s, (const V*)in, (V*)out, Nx, Ny, Nz,
                     Nm, kxs, kys, 2.0 / H, scale, add);
}
template void launch_curl_cheb<float>(const void*, void*, int, int, int, int, const double*, const double*, double,
                                      double, int, hipStream_t);
template void launch_curl_cheb<double>(const void*, void*, int, int, int, int, const double*, const double*,
                                       double, double, int, hipStream_t);

// Half the Fourier curl (FCM.py:376-383, Solvers.py:137-157): w = (i d) x a / 2
// with d the wavevector, the unpaired Nyquist entry of each direction zeroed.
template <class T>
__global__ void k_tp_curl(const typename C2<T>::t* __restrict__ in, typename C2<T>::t* __restrict__ out,
                          int Nx, int Ny, int Nz, double Lx, double Ly, double Lz, int add)
{
  typedef typename C2<T>::t V;
  const int Nkx = Nx / 2 + 1;
  const size_t nm = (size_t)Nz * Ny * Nkx;
  const size_t m = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nm) return;
  const int ix = (int)(m % Nkx);
  const size_t r = m / Nkx;
  const int iy = (int)(r % Ny), iz = (int)(r / Ny);
  const double d0 = (Nx % 2 == 0 && ix == Nx / 2) ? 0.0 : kvec(ix, Nx, Lx);
  const double d1 = (Ny % 2 == 0 && iy == Ny / 2) ? 0.0 : kvec(iy, Ny, Ly);
  const double d2 = (Nz % 2 == 0 && iz == Nz / 2) ? 0.0 : kvec(iz, Nz, Lz);
  double ar[3], ai[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const V v = in[c * nm + m];
    ar[c] = (double)v.x;
    ai[c] = (double)v.y;
  }
  // (d x a) componentwise, then times i / 2: (re, im) -> (-im, re) / 2
  const double cr[3] = {d1 * ar[2] - d2 * ar[1], d2 * ar[0] - d0 * ar[2], d0 * ar[1] - d1 * ar[0]};
  const double ci[3] = {d1 * ai[2] - d2 * ai[1], d2 * ai[0] - d0 * ai[2], d0 * ai[1] - d1 * ai[0]};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    double wr = -0.5 * ci[c], wi = 0.5 * cr[c];
    if (add) {
      const V b = out[c * nm + m];
      wr += (double)b.x;
      wi += (double)b.y;
    }
    V v;
    v.x = (T)wr;
    v.y = (T)wi;
    out[c * nm + m] = v;
  }
}

template <class T>
void launch_tp_curl(const void* in, void* out, int Nx, int Ny, int Nz, double Lx, double Ly, double Lz, int add,
                    hipStream_t s)
{
  typedef typename C2<T>::t V;
  const size_t nm = (size_t)Nz * Ny * (Nx / 2 + 1);
  hipLaunchKernelGGL((k_tp_curl<T>), dim3((unsigned)((nm + 255) / 256)), dim3(256), 0, s, (const V*)in, (V*)out,
                     Nx, Ny, Nz, Lx, Ly, Lz, add);
}
template void launch_tp_curl<float>(const void*, void*, int, int, int, double, double, double, int, hipStream_t);
template void launch_tp_curl<double>(const void*, void*, int, int, int, double, double, double, int, hipStream_t);

void launch_reset_err(int* e, hipStream_t s)
{
  hipLaunchKernelGGL(k_reset_err, dim3(1), dim3(1), 0, s, e);
}

void launch_sort(int* key, int* counts, int* rank, int* start, int* order, int N, int ntiles,
                 hipStream_t s)
{
  hipLaunchKernelGGL(k_fill_int, dim3((ntiles + 255) / 256), dim3(256), 0, s, counts, ntiles, 0);
  if (N > 0)
    hipLaunchKernelGGL(k_tile_count, dim3((N + 255) / 256), dim3(256), 0, s, key, counts, rank, N);
  hipLaunchKernelGGL(k_tile_scan, dim3(1), dim3(1024), 0, s, counts, start, ntiles);
  if (N > 0)
    hipLaunchKernelGGL(k_tile_scatter, dim3((N + 255) / 256), dim3(256), 0, s, key, rank, start,
                       order, N);
}

// partial maxima of |F| and |T|; NaN propagates
template <class Tin>
__global__ void __launch_bounds__(256) k_absmax(const Tin* __restrict__ F, const Tin* __restrict__ Tq, int n3,
                                                float* __restrict__ out)
{
  __shared__ float red[2][4];
  float fm = 0.0f, tm = 0.0f;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n3; i += gridDim.x * 256) {
    const float a = fabsf((float)F[i]);
    fm = (a > fm || a != a) ? a : fm;
    if (Tq) {
      const float b = fabsf((float)Tq[i]);
      tm = (b > tm || b != b) ? b : tm;
    }
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    const float a = __shfl_xor(fm, o, 64), b = __shfl_xor(tm, o, 64);
    fm = (a > fm || a != a) ? a : fm;
    tm = (b > tm || b != b) ? b : tm;
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = fm;
    red[1][threadIdx.x >> 6] = tm;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 4; ++w) {
      fm = (red[0][w] > fm || red[0][w] != red[0][w]) ? red[0][w] : fm;
      tm = (red[1][w] > tm || red[1][w] != red[1][w]) ? red[1][w] : tm;
    }
    out[2 * blockIdx.x] = fm;
    out[2 * blockIdx.x + 1] = tm;
  }
}

template <class Tin>
void launch_absmax(const void* F, const void* Tq, int n3, float* out, hipStream_t s)
{
  hipLaunchKernelGGL(k_absmax<Tin>, dim3(DPS_AMAX_BLOCKS), dim3(256), 0, s, (const Tin*)F, (const Tin*)Tq, n3, out);
}
template void launch_absmax<float>(const void*, const void*, int, float*, hipStream_t);
template void launch_absmax<double>(const void*, const void*, int, float*, hipStream_t);

void launch_lists(const ListArgs& A, hipStream_t s)
{
  const int ncells = A.ntx * A.nty * A.nch;
  const int ne = A.N * A.nkind;
  hipLaunchKernelGGL(k_fill_int, dim3((ncells + 255) / 256), dim3(256), 0, s, A.counts, ncells, 0);
  hipLaunchKernelGGL(k_fill_int, dim3((ncells + 255) / 256), dim3(256), 0, s, A.fill, ncells, 0);
  if (ne > 0) hipLaunchKernelGGL(k_list_cells<false>, dim3((ne + 255) / 256), dim3(256), 0, s, A);
  hipLaunchKernelGGL(k_tile_scan, dim3(1), dim3(1024), 0, s, A.counts, A.start, ncells);
  if (ne > 0) hipLaunchKernelGGL(k_list_cells<true>, dim3((ne + 255) / 256), dim3(256), 0, s, A);
}

int spread_npf(int rs_max)
{
  // staged words per entry: record + 7 header words, 16 threads per entry
  static const int sizes[] = {8, 10, 12, 16, 20, 24, 32};
  for (int npf : sizes)
    if (rs_max + 7 <= 16 * npf) return npf;
  return 0;
}

size_t spread_lds_bytes(const SpreadArgs& A, size_t tsz)
{
  // accumulator (int32 fixed point with FP32 grids, else FP64), staged
  // batch, queue, scale
  return (size_t)3 * A.ZC * (SP_TY * (SP_TX + 1) + 1) * (tsz == 4 ? 4 : 8) +
         (size_t)DPS_SPREAD_BATCH * 16 * spread_npf(A.rs_max) * tsz + 2 * 256 * sizeof(int) + 4 * sizeof(float);
}

template <class T, class Tin>
void launch_spread(const SpreadArgs& A, const void* F, const void* Tq, hipStream_t s)
{
  const size_t lds = spread_lds_bytes(A, sizeof(T));
  dim3 grid(A.ntx * A.nty * A.nch);
  const Tin *f = (const Tin*)F, *t = (const Tin*)Tq;
  switch (spread_npf(A.rs_max)) {
#define SPREAD_CASE(NPF)                                                                             \
  case NPF:                                                                                          \
    if (lds > 65536)                                                                                 \
      (void)hipFuncSetAttribute((const void*)k_spread<T, Tin, NPF>,                                  \
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);               \
    hipLaunchKernelGGL((k_spread<T, Tin, NPF>), grid, dim3(256), lds, s, A, f, t);                   \
    break;
    SPREAD_CASE(8) SPREAD_CASE(10) SPREAD_CASE(12) SPREAD_CASE(16) SPREAD_CASE(20) SPREAD_CASE(24)
    default:
      if (lds > 65536)
        (void)hipFuncSetAttribute((const void*)k_spread<T, Tin, 32>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds);
      hipLaunchKernelGGL((k_spread<T, Tin, 32>), grid, dim3(256), lds, s, A, f, t);
      break;
#undef SPREAD_CASE
  }
}

template <class T, class Tio>
void launch_interp_grad(const InterpArgs& A, const int4* sup, const void* rec, int rs, const KernelDesc& K,
                        const void* gu, const void* gv, const void* lam, const void* f, void* grad,
                        hipStream_t s)
{
  if (A.N == 0) return;
  const int waves_per_block = 256 / 64;
  hipLaunchKernelGGL((k_interp_grad<T, Tio>), dim3((A.N + waves_per_block - 1) / waves_per_block), dim3(256), 0,
                     s, A, sup, (const T*)rec, rs, K, (const T*)gu, (const T*)gv, (const Tio*)lam,
                     (const Tio*)f, (Tio*)grad);
}

template <class T>
void launch_stencil_d2(const StencilArgs& A, const int4* sup_d, void* rec, int rs, hipStream_t s)
{
  if (A.N == 0) return;
  // four particles per 64-lane block, each with [4][nq] doubles and [nq] ints of LDS
  const int nq = std::max(1, A.kd.wzmax);
  hipLaunchKernelGGL((k_stencil_d2<T>), dim3((A.N + 3) / 4), dim3(64),
                     (size_t)4 * nq * (4 * sizeof(double) + sizeof(int)), s, A, sup_d, (T*)rec, rs, nq);
}

template <class T, class Tio>
void launch_interp_grad_t(const InterpArgs& A, const int4* sup, const void* rec, int rs, const KernelDesc& K,
                          const void* gu, const void* gv, const void* mu, const void* t, void* grad, hipStream_t s)
{
  if (A.N == 0) return;
  const int waves_per_block = 256 / 64;
  hipLaunchKernelGGL((k_interp_grad_t<T, Tio>), dim3((A.N + waves_per_block - 1) / waves_per_block), dim3(256),
                     0, s, A, sup, (const T*)rec, rs, K, (const T*)gu, (const T*)gv, (const Tio*)mu,
                     (const Tio*)t, (Tio*)grad);
}

template <class T, class Tout>
void launch_interp(const InterpArgs& A, void* vel, void* ang, hipStream_t s)
{
  if (A.N == 0) return;
  const int waves_per_block = 256 / 64;
  hipLaunchKernelGGL((k_interp<T, Tout>), dim3((A.N + waves_per_block - 1) / waves_per_block),
                     dim3(256), 0, s, A, (Tout*)vel, (Tout*)ang);
}

template <class T>
size_t zfwd_lds(const ZArgs& A) { return (size_t)2 * A.NH * ZCols<T>::value * sizeof(typename C2<T>::t); }
template <class T>
size_t zinv_lds(const ZArgs& A) { return (size_t)A.N * ZCols<T>::value * sizeof(typename C2<T>::t); }

template <class T>
void launch_zfwd(const ZArgs& A, const void* in, void* out, hipStream_t s)
{
  if (launch_zfwd_mma<T>(A, in, out, s)) return;
  typedef typename C2<T>::t V;
  constexpr int CW = ZCols<T>::value;
  const size_t lds = zfwd_lds<T>(A);
  if (lds > 65536) (void)hipFuncSetAttribute((const void*)k_zfwd<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL((k_zfwd<T>), dim3((A.Nm + CW - 1) / CW, 3), dim3(256), lds, s, A,
                     (const V*)in, (V*)out);
}

template <class T>
void launch_zinv(const ZArgs& A, const void* in, void* out, hipStream_t s)
{
  if (launch_zinv_mma<T>(A, in, out, s)) return;
  typedef typename C2<T>::t V;
  constexpr int CW = ZCols<T>::value;
  const size_t lds = zinv_lds<T>(A);
  if (lds > 65536) (void)hipFuncSetAttribute((const void*)k_zinv<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL((k_zinv<T>), dim3((A.Nm + CW - 1) / CW, 3), dim3(256), lds, s, A,
                     (const V*)in, (V*)out);
}

size_t z_lds_bytes(const ZArgs& A, bool mixed)
{
  return mixed ? std::max(zfwd_lds<float>(A), zinv_lds<float>(A))
               : std::max(zfwd_lds<double>(A), zinv_lds<double>(A));
}

template <class T>
void launch_bvp(const BvpArgs& A, const void* in, void* out, hipStream_t s)
{
  typedef typename C2<T>::t V;
  const int n = A.Nm + A.nspecial;
  hipLaunchKernelGGL((k_bvp<T>), dim3((n + 63) / 64), dim3(64), 0, s, A, (const V*)in, (V*)out);
  if (A.nspecial > 0)
    hipLaunchKernelGGL((k_bvp_combine<T>), dim3((A.nspecial + 63) / 64), dim3(64), 0, s, A, (V*)out);
}

template <class T>
void launch_ref_to_grid(const double* ref, void* g, int Nx, int Ny, int Nz, hipStream_t s)
{
  const size_t n = (size_t)3 * Nx * Ny * Nz;
  hipLaunchKernelGGL((k_ref_to_grid<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, ref,
                     (T*)g, Nx, Ny, Nz);
}

template <class T>
void launch_grid_to_ref(const void* g, double* ref, int Nx, int Ny, int Nz, hipStream_t s)
{
  const size_t n = (size_t)3 * Nx * Ny * Nz;
  hipLaunchKernelGGL((k_grid_to_ref<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                     (const T*)g, ref, Nx, Ny, Nz);
}

#define DPS_INST(T)                                                                          \
  template void launch_stencil<float, T>(const StencilArgs&, const void*, hipStream_t);     \
  template void launch_stencil<double, T>(const StencilArgs&, const void*, hipStream_t);    \
  template void launch_spread<T, float>(const SpreadArgs&, const void*, const void*, hipStream_t); \
  template void launch_spread<T, double>(const SpreadArgs&, const void*, const void*, hipStream_t); \
  template void launch_interp<T, float>(const InterpArgs&, void*, void*, hipStream_t);      \
  template void launch_interp<T, double>(const InterpArgs&, void*, void*, hipStream_t);     \
  template void launch_interp_grad<T, float>(const InterpArgs&, const int4*, const void*, int,       \
                                             const KernelDesc&, const void*, const void*,             \
                                             const void*, const void*, void*, hipStream_t);           \
  template void launch_interp_grad<T, double>(const InterpArgs&, const int4*, const void*, int,      \
                                              const KernelDesc&, const void*, const void*,            \
                                              const void*, const void*, void*, hipStream_t);          \
  template void launch_stencil_d2<T>(const StencilArgs&, const int4*, void*, int, hipStream_t);     \
  template void launch_interp_grad_t<T, float>(const InterpArgs&, const int4*, const void*, int,     \
                                               const KernelDesc&, const void*, const void*,           \
                                               const void*, const void*, void*, hipStream_t);         \
  template void launch_interp_grad_t<T, double>(const InterpArgs&, const int4*, const void*, int,    \
                                                const KernelDesc&, const void*, const void*,          \
                                                const void*, const void*, void*, hipStream_t);        \
  template void launch_zfwd<T>(const ZArgs&, const void*, void*, hipStream_t);              \
  template void launch_zinv<T>(const ZArgs&, const void*, void*, hipStream_t);              \
  template void launch_bvp<T>(const BvpArgs&, const void*, void*, hipStream_t);             \
  template void launch_ref_to_grid<T>(const double*, void*, int, int, int, hipStream_t);    \
  template void launch_grid_to_ref<T>(const void*, double*, int, int, int, hipStream_t);

DPS_INST(float)
DPS_INST(double)

}  // namespace dps
