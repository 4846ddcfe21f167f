// dps_spectral.hip -- fused spectral solve of the doubly periodic Stokes Mdot.
//
// One launch replaces the chain Chebyshev forward transform -> per-mode BVP ->
// Nyquist averaging -> inverse transform with wall corrections
// (Transform.cpp:107-304, DPTools.cpp:81-345 and 590-877).  A workgroup owns G
// Fourier modes; each mode is solved by a group of LPM lanes of one wavefront,
// lane l owning the E consecutive Chebyshev rows [l E, l E + E) (dps_seg.h).
//
//   phase 0  the G columns of the r2c spectrum, all z levels and components,
//            are read once from HBM (rows of G consecutive modes) into LDS
//   phase 1  forward transform: lane l forms the coefficients of its rows
//            from even/odd folded node pairs (grid precision), written back
//            in place
//   phase 2  FP64 BVP on the lane group, in place; modes touching a Nyquist
//            index are solved twice (mode and mirror) and averaged, the
//            mirror reading a copy of the forcing kept in its output column
//   phase 3  inverse transform of node pairs (l + LPM r, Nz-1-l-LPM r) plus
//            the analytic wall profiles, written back in place
//   phase 4  the G columns are written back as rows of G consecutive modes
//
// The spectrum crosses HBM once in each direction; everything in between
// stays in one LDS buffer of 3 Nz G elements and registers.
#include <hip/hip_runtime.h>
#include "dps_internal.h"
#include "dps_seg.h"

// velocity components sharing each scan of the lane-group solve
#define DPS_SPEC_NRV 1

namespace dps {

namespace {

// 64-bit DPP move within a 16-lane row; lanes whose source is outside the row
// keep their own value
template <int CTRL>
__device__ __forceinline__ double dpp64(double v)
{
  int lo = __double2loint(v), hi = __double2hiint(v);
  lo = __builtin_amdgcn_update_dpp(lo, lo, CTRL, 0xF, 0xF, false);
  hi = __builtin_amdgcn_update_dpp(hi, hi, CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// Lane exchange of one mode's group.  A 16-lane group is exactly one DPP row:
// shifts are row_shr/row_shl and the butterfly partners are quad_perm
// [1,0,3,2], quad_perm [2,3,0,1], row_half_mirror and row_mirror, all VALU
// operations.  Wider groups use lane permutes.
template <int LPM>
struct DevLanes {
  int l;
  template <int D>
  __device__ double up(double v) const
  {
    if constexpr (LPM == 16) return dpp64<0x110 + D>(v);   // row_shr:D
    else return __shfl_up(v, (unsigned)D, LPM);
  }
  template <int D>
  __device__ double down(double v) const
  {
    if constexpr (LPM == 16) return dpp64<0x100 + D>(v);   // row_shl:D
    else return __shfl_down(v, (unsigned)D, LPM);
  }
  template <int D>
  __device__ double xr(double v) const
  {
    if constexpr (LPM == 16) {
      if constexpr (D == 1) return dpp64<0xB1>(v);         // quad_perm [1,0,3,2]
      else if constexpr (D == 2) return dpp64<0x4E>(v);    // quad_perm [2,3,0,1]
      else if constexpr (D == 4) return dpp64<0x141>(v);   // row_half_mirror
      else return dpp64<0x140>(v);                         // row_mirror
    } else {
      return __shfl_xor(v, D, LPM);
    }
  }
};

// coefficient rows of one mode in LDS, layout [c][row][g], row stride GS
template <class T>
struct LdsCoef {
  const typename C2<T>::t* S;
  int GS, g, Nz, r0;
  bool conj;
  __device__ cplx operator()(int c, int q) const
  {
    const int row = r0 + q;
    if (row >= Nz) return mk(0, 0);
    const typename C2<T>::t v = S[(c * Nz + row) * GS + g];
    return conj ? mk((double)v.x, -(double)v.y) : mk((double)v.x, (double)v.y);
  }
};

// store velocity coefficients; avg: average with the conjugate of v (the
// mirror solve of a Nyquist-touching mode)
template <class T>
struct LdsStore {
  typename C2<T>::t* S;
  int GS, g, Nz, r0;
  bool avg;
  __device__ void operator()(int c, int q, cplx v)
  {
    typename C2<T>::t& d = S[(c * Nz + r0 + q) * GS + g];
    if (avg) {
      d.x = (T)(0.5 * ((double)d.x + v.re));
      d.y = (T)(0.5 * ((double)d.y - v.im));
    } else {
      d.x = (T)v.re;
      d.y = (T)v.im;
    }
  }
};

// Global-memory view of the coefficients of one mode (the stash that keeps the
// forcing coefficients of a Nyquist-touching mode for its mirror solve).
template <class T>
struct GlobalCoef {
  const typename C2<T>::t* p;   // element (c, row) at p[(c * Nz + row) * Nm]
  long Nm;
  int Nz, r0;
  __device__ cplx operator()(int c, int q) const
  {
    const int row = r0 + q;
    if (row >= Nz) return mk(0, 0);
    const typename C2<T>::t v = p[(size_t)(c * Nz + row) * Nm];
    return mk((double)v.x, -(double)v.y);   // conjugated input of the mirror
  }
};

template <class T, int LPM, int E, int NRV>
__global__ void __launch_bounds__(256) k_spectral(SpecArgs A, const typename C2<T>::t* __restrict__ in,
                                                  typename C2<T>::t* __restrict__ out)
{
  typedef typename C2<T>::t V;
  constexpr int R = E / 2;   // node pairs per lane in the inverse transform
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  const int Nz = A.B.Nz, G = A.G, Nm = A.Nm, NH = A.NH;
  const int ncol = 3 * Nz * G;
  // [c][row][g] with row stride G+1: the lanes of one mode (different rows)
  // fall in different LDS banks
  const int GS = G + 1;
  V* S = reinterpret_cast<V*>(smem_raw);   // values, coefficients, values
  const int m0 = xcd_group(blockIdx.x, gridDim.x) * G;
  const int g = threadIdx.x / LPM;
  const int l = threadIdx.x % LPM;
  const int m = m0 + g;
  const int r0 = l * E;
  int ix = 0, iy = 0, nyqx = 0, nyqy = 0;
  if (m < Nm) {
    iy = m / A.Nkx;
    ix = m - iy * A.Nkx;
    nyqx = (A.Nx % 2 == 0) && (ix == A.Nx / 2);
    nyqy = (A.Ny % 2 == 0) && (iy == A.Ny / 2);
  }
  const bool special = (m > 0) && (m < Nm) && (nyqx || nyqy);

  // ---- phase 0: columns in ----
  for (int i = threadIdx.x; i < ncol; i += blockDim.x) {
    const int gi = i % G, row = i / G;
    V v;
    v.x = v.y = (T)0;
    if (m0 + gi < Nm) v = in[(size_t)row * Nm + m0 + gi];
    S[row * GS + gi] = v;
  }
  __syncthreads();

  // ---- phase 1: forward transform of this lane's rows ----
  {
    T ar[3][E], ai[3][E];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int q = 0; q < E; ++q) ar[c][q] = ai[c][q] = (T)0;
    const T* Tf = (const T*)A.TfP;
#pragma unroll 2
    for (int j = 0; j < NH; ++j) {
      T er[3], ei[3], orr[3], oi[3];
      const bool mid = (j == Nz - 1 - j);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const V a = S[(c * Nz + j) * GS + g];
        const V b = S[(c * Nz + Nz - 1 - j) * GS + g];
        if (mid) { er[c] = a.x; ei[c] = a.y; orr[c] = oi[c] = (T)0; }
        else { er[c] = a.x + b.x; ei[c] = a.y + b.y; orr[c] = a.x - b.x; oi[c] = a.y - b.y; }
      }
      T t[E];
#pragma unroll
      for (int q = 0; q < E; ++q) t[q] = Tf[((size_t)j * E + q) * LPM + l];
#pragma unroll
      for (int q = 0; q < E; ++q)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          ar[c][q] += t[q] * ((q & 1) ? orr[c] : er[c]);
          ai[c][q] += t[q] * ((q & 1) ? oi[c] : ei[c]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < E; ++q)
      if (r0 + q < Nz)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          V v;
          v.x = ar[c][q];
          v.y = ai[c][q];
          S[(c * Nz + r0 + q) * GS + g] = v;
          if (special) out[(size_t)(c * Nz + r0 + q) * Nm + m] = v;
        }
  }

  // ---- phase 2: FP64 BVP of mode m on this lane group, in place ----
  ModeResult res;
  res.nbasis = 0;
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int b = 0; b < 4; ++b) res.corr[c][b] = mk(0, 0);
  double K = 0;
  if (m < Nm) {
    const DevLanes<LPM> ln{l};
    LdsCoef<T> fin{S, GS, g, Nz, r0, false};
    LdsStore<T> fout{S, GS, g, Nz, r0, false};
    if (m == 0) {
      seg_bvp_k0<LPM, E>(A.B, ln, fin, fout);
    } else {
      const double kx = kvec(ix, A.Nx, A.Lx), ky = kvec(iy, A.Ny, A.Ly);
      K = sqrt(kx * kx + ky * ky);
      double ib[E], ab[E], Gi[4];
#pragma unroll
      for (int q = 0; q < E; ++q) {
        ib[q] = A.fac_ib[((size_t)m * E + q) * LPM + l];
        ab[q] = A.fac_ab[((size_t)m * E + q) * LPM + l];
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) Gi[i] = A.fac_G[(size_t)m * 4 + i];
      seg_bvp_mode<LPM, E, NRV>(A.B, ln, kx, ky, nyqx, nyqy, ib, ab, Gi, fin, fout, res);
      if (special) {
        // the reference's c2c transform keeps only the real part of the
        // inverse: the Hermitian average of this mode and its mirror, solved
        // with the mirror's own grid wavenumbers (same |k|, same factors) on
        // the conjugated forcing kept in the output column
        const int ixm = (A.Nx - ix) % A.Nx, iym = (A.Ny - iy) % A.Ny;
        const double kxm = kvec(ixm, A.Nx, A.Lx), kym = kvec(iym, A.Ny, A.Ly);
        const int nyqxm = (A.Nx % 2 == 0) && (ixm == A.Nx / 2);
        const int nyqym = (A.Ny % 2 == 0) && (iym == A.Ny / 2);
        // the first solve's correction waits in memory (every lane of the
        // group writes the same values), not in registers
        double2* cs = A.cstash + (size_t)m * 12;
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
          for (int b = 0; b < 4; ++b) cs[c * 4 + b] = make_double2(res.corr[c][b].re, res.corr[c][b].im);
        DPS_STAGE_FENCE();
        GlobalCoef<T> fin2{out + m, (long)Nm, Nz, r0};
        fout.avg = true;
        seg_bvp_mode<LPM, E, NRV>(A.B, ln, kxm, kym, nyqxm, nyqym, ib, ab, Gi, fin2, fout, res);
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
          for (int b = 0; b < 4; ++b) {
            const double2 a = cs[c * 4 + b];
            res.corr[c][b] = mk(0.5 * (a.x + res.corr[c][b].re), 0.5 * (a.y - res.corr[c][b].im));
          }
      }
    }
  }
  __syncthreads();

  // ---- phase 3: inverse transform + wall profiles ----
  {
    const T* Ti = (const T*)A.TiP;
    T er[3][R], ei[3][R], orr[3][R], oi[3][R];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int r = 0; r < R; ++r) er[c][r] = ei[c][r] = orr[c][r] = oi[c][r] = (T)0;
#pragma unroll 2
    for (int k = 0; k < Nz; k += 2) {
      V u[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) u[c] = S[(c * Nz + k) * GS + g];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const T t = Ti[((size_t)k * R + r) * LPM + l];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          er[c][r] += t * u[c].x;
          ei[c][r] += t * u[c].y;
        }
      }
    }
#pragma unroll 2
    for (int k = 1; k < Nz; k += 2) {
      V u[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) u[c] = S[(c * Nz + k) * GS + g];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const T t = Ti[((size_t)k * R + r) * LPM + l];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          orr[c][r] += t * u[c].x;
          oi[c][r] += t * u[c].y;
        }
      }
    }
    __syncthreads();
    const int nb = res.nbasis;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int j = l + LPM * r;
      if (j >= NH) continue;
      const int jm = Nz - 1 - j;
      // z_{Nz-1-j} = Lz - z_j, so e^{K(z - Lz)} at one node of the pair is
      // e^{-K z} at the other
      double ej = 0, em = 0;
      if (nb) {
        ej = exp(-K * A.zpts[j]);
        em = (jm == j) ? ej : exp(-K * A.zpts[jm]);
      }
      for (int side = 0; side < 2; ++side) {
        const int jj = side ? jm : j;
        if (side && jj == j) break;
        const double z = A.zpts[jj];
        const double e1 = side ? em : ej;
        const double e2 = side ? ej : em;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          V v;
          v.x = side ? (er[c][r] - orr[c][r]) : (er[c][r] + orr[c][r]);
          v.y = side ? (ei[c][r] - oi[c][r]) : (ei[c][r] + oi[c][r]);
          if (nb) {
            double sr = res.corr[c][0].re * e1, si = res.corr[c][0].im * e1;
            if (nb == 2) {
              sr += res.corr[c][1].re * (z * e1);
              si += res.corr[c][1].im * (z * e1);
            } else {
              sr += res.corr[c][1].re * e2 + res.corr[c][2].re * (z * e1) + res.corr[c][3].re * (z * e2);
              si += res.corr[c][1].im * e2 + res.corr[c][2].im * (z * e1) + res.corr[c][3].im * (z * e2);
            }
            v.x += (T)(sr * A.inv_nxy);
            v.y += (T)(si * A.inv_nxy);
          }
          S[(c * Nz + jj) * GS + g] = v;
        }
      }
    }
  }
  __syncthreads();

  // ---- phase 4: columns out ----
  for (int i = threadIdx.x; i < ncol; i += blockDim.x) {
    const int gi = i % G, row = i / G;
    if (m0 + gi < Nm) out[(size_t)row * Nm + m0 + gi] = S[row * GS + gi];
  }
}

// Static factorisation of every mode (depends on |k| only), one lane per
// mode, written in the lane order of k_spectral.
__global__ void k_bvp_factor(BvpConsts B, int Nx, int Ny, int Nkx, int Nm, double Lx, double Ly,
                             double* s_ib, double* s_ab, int lpm, int e, double* fib, double* fab,
                             double* fG)
{
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= Nm) return;
  ScratchView sc;
  sc.stride = Nm;
  sc.ib = s_ib + m;
  sc.ab = s_ab + m;
  sc.P = sc.CP = sc.DP = sc.U0 = sc.U1 = sc.U2 = nullptr;
  double Gi[4] = {0, 0, 0, 0};
  if (m != 0) {
    const int iy = m / Nkx, ix = m - iy * Nkx;
    const double kx = kvec(ix, Nx, Lx), ky = kvec(iy, Ny, Ly);
    bvp_factor(B, kx * kx + ky * ky, sc, Gi);
  }
  const int Nz = B.Nz;
  for (int row = 0; row < lpm * e; ++row) {
    const int l = row / e, q = row - l * e;
    const size_t i = ((size_t)m * e + q) * lpm + l;
    const bool live = (m != 0) && (row < Nz);
    fib[i] = live ? s_ib[(size_t)(row + 1) * Nm + m] : 0.0;
    fab[i] = live ? s_ab[(size_t)row * Nm + m] : 0.0;
  }
  for (int i = 0; i < 4; ++i) fG[(size_t)m * 4 + i] = Gi[i];
}

}  // namespace

bool spectral_shape(int Nz, int* lpm, int* e)
{
  for (int E = 2; E <= 8; E += 2)
    if (16 * E >= Nz) { *lpm = 16; *e = E; return true; }
  if (Nz <= 32 * 8) { *lpm = 32; *e = 8; return true; }
  if (Nz <= 64 * 8) { *lpm = 64; *e = 8; return true; }
  return false;
}

int spectral_modes_per_block(int Nz, int lpm, bool mixed)
{
  const size_t per_row = (size_t)3 * Nz * (mixed ? sizeof(float2) : sizeof(double2));
  int G = 256 / lpm;
  while (G > 1 && (size_t)(G + 1) * per_row > 65536) G /= 2;
  return G;
}

void launch_bvp_factor(const BvpConsts& B, int Nx, int Ny, int Nkx, int Nm, double Lx, double Ly,
                       double* s_ib, double* s_ab, int lpm, int e, double* fib, double* fab, double* fG,
                       hipStream_t s)
{
  hipLaunchKernelGGL(k_bvp_factor, dim3((Nm + 63) / 64), dim3(64), 0, s, B, Nx, Ny, Nkx, Nm, Lx, Ly,
                     s_ib, s_ab, lpm, e, fib, fab, fG);
}

template <class T>
void launch_spectral(const SpecArgs& A, const void* in, void* out, hipStream_t s)
{
  typedef typename C2<T>::t V;
  const size_t lds = (size_t)3 * A.B.Nz * (A.G + 1) * sizeof(V);
  const dim3 grid((A.Nm + A.G - 1) / A.G), block(A.G * A.lpm);
  const V* i = (const V*)in;
  V* o = (V*)out;
#define SPEC_CASE(LP, EE)                                                                    \
  if (A.lpm == LP && A.e == EE) {                                                            \
    hipLaunchKernelGGL((k_spectral<T, LP, EE, DPS_SPEC_NRV>), grid, block, lds, s, A, i, o); \
    return;                                                                                  \
  }
  SPEC_CASE(16, 2) SPEC_CASE(16, 4) SPEC_CASE(16, 6) SPEC_CASE(16, 8) SPEC_CASE(32, 8) SPEC_CASE(64, 8)
#undef SPEC_CASE
}

template void launch_spectral<float>(const SpecArgs&, const void*, void*, hipStream_t);
template void launch_spectral<double>(const SpecArgs&, const void*, void*, hipStream_t);

}  // namespace dps
