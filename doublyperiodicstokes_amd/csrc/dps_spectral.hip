// dps_spectral.hip -- fused spectral solve of the doubly periodic Stokes Mdot.
//
// One launch replaces the chain Chebyshev forward transform -> per-mode BVP ->
// Nyquist averaging -> inverse transform with wall corrections
// (Transform.cpp:107-304, DPTools.cpp:81-345 and 590-877).  A workgroup owns G
// Fourier modes; each mode is solved by a group of LPM lanes of one wavefront,
// lane l owning the E consecutive Chebyshev rows [l E, l E + E) (dps_seg.h).
//
//   phase 0  the row tables of the solve are copied to LDS (zero past Nz), and
//            the G columns of the r2c spectrum, all z levels and components,
//            are read once from HBM (rows of G modes) into LDS
//   phase 1  forward transform from even/odd folded node pairs (grid
//            precision), written back in place: FP32 16-lane shapes on the
//            matrix cores (ZMma), else lane l forms the coefficients of its
//            own rows
//   phase 2  FP64 BVP on the lane group, in place
//   phase 3  inverse transform of node pairs plus the analytic wall
//            profiles: on the matrix cores the even and odd sums go to LDS
//            and thread (jr, gl) writes node pairs jr + LPM r of mode gl
//            straight to HBM; else lane l takes the pairs (l + LPM r,
//            Nz-1-l-LPM r), written back in place, and
//   phase 4  writes the G columns back as rows of G modes
//
// The spectrum crosses HBM once in each direction; everything in between
// stays in LDS and registers.
//
// Modes on a Nyquist index take the Hermitian average with their mirror mode
// (the reference's c2c transform keeps only the real part of the inverse).
// The mirror's BVP solution is the conjugate of the mode's own, so the
// average costs one solve and a second evaluation of the wall corrections
// (dps_seg.h, seg_bvp_mode's avg).
#include <hip/hip_runtime.h>
#include "dps_internal.h"
#include "dps_mma.h"
#include "dps_seg.h"

// velocity components sharing each scan of the lane-group solve
#define DPS_SPEC_NRV 1

namespace dps {

namespace {

// row tables of the solve held in LDS, rows [0, LPM E), zero from row Nz on
enum { TB_SD0, TB_SM2, TB_SP2, TB_FM1, TB_FP1, TB_R1, TB_R2, TB_L1, TB_L2, TB_N };

__device__ __forceinline__ const double* row_table(const BvpConsts& B, int t)
{
  switch (t) {
    case TB_SD0: return B.sd0;
    case TB_SM2: return B.sm2;
    case TB_SP2: return B.sp2;
    case TB_FM1: return B.fm1;
    case TB_FP1: return B.fp1;
    case TB_R1: return B.R1;
    case TB_R2: return B.R2;
    case TB_L1: return B.L1;
    default: return B.L2;
  }
}

// 64-bit DPP move within a 16-lane row; lanes whose source is outside the row
// receive `fill` (the DPP old operand with bound_ctrl off)
template <int CTRL>
__device__ __forceinline__ double dpp64(double v, double fill)
{
  int lo = __double2loint(v), hi = __double2hiint(v);
  lo = __builtin_amdgcn_update_dpp(__double2loint(fill), lo, CTRL, 0xF, 0xF, false);
  hi = __builtin_amdgcn_update_dpp(__double2hiint(fill), hi, CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// Lane exchange of one mode's group.  A 16-lane group is exactly one DPP row:
// shifts are row_shr/row_shl and the butterfly partners are quad_perm
// [1,0,3,2], quad_perm [2,3,0,1], row_half_mirror and row_mirror, all VALU
// operations.  An 8-lane group is half a row: the same shifts, with the lanes
// whose partner lies in the other half taking the fill by a select, and the
// first three butterfly partners.  Wider groups use lane permutes.
template <int LPM>
struct DevLanes {
  int l;
  template <int D>
  __device__ double up(double v, double fill) const
  {
    if constexpr (LPM == 16) return dpp64<0x110 + D>(v, fill);   // row_shr:D
    else if constexpr (LPM == 8) {
      const double r = dpp64<0x110 + D>(v, fill);
      return (l >= D) ? r : fill;
    } else {
      const double r = __shfl_up(v, (unsigned)D, LPM);
      return (l >= D) ? r : fill;
    }
  }
  template <int D>
  __device__ double down(double v, double fill) const
  {
    if constexpr (LPM == 16) return dpp64<0x100 + D>(v, fill);   // row_shl:D
    else if constexpr (LPM == 8) {
      const double r = dpp64<0x100 + D>(v, fill);
      return (l + D < LPM) ? r : fill;
    } else {
      const double r = __shfl_down(v, (unsigned)D, LPM);
      return (l + D < LPM) ? r : fill;
    }
  }
  template <int D>
  __device__ double xr(double v) const
  {
    if constexpr (LPM == 16 || LPM == 8) {
      static_assert(D < LPM, "butterfly partner outside the group");
      if constexpr (D == 1) return dpp64<0xB1>(v, v);         // quad_perm [1,0,3,2]
      else if constexpr (D == 2) return dpp64<0x4E>(v, v);    // quad_perm [2,3,0,1]
      else if constexpr (D == 4) return dpp64<0x141>(v, v);   // row_half_mirror
      else return dpp64<0x140>(v, v);                         // row_mirror
    } else {
      return __shfl_xor(v, D, LPM);
    }
  }
};

// coefficient rows of one mode in LDS, layout [c][row][g] (rows padded to
// LPM E with zeros, row stride GS)
template <class T, int NZP, int GS>
struct LdsCoef {
  const typename C2<T>::t* S;
  int g, r0;
  __device__ cplx operator()(int c, int q) const
  {
    const typename C2<T>::t v = S[(c * NZP + r0 + q) * GS + g];
    return mk((double)v.x, (double)v.y);
  }
};

// store velocity coefficients (padding rows included)
template <class T, int NZP, int GS>
struct LdsStore {
  typename C2<T>::t* S;
  int g, r0;
  __device__ void operator()(int c, int q, cplx v)
  {
    typename C2<T>::t& d = S[(c * NZP + r0 + q) * GS + g];
    d.x = (T)v.re;
    d.y = (T)v.im;
  }
};

// FP32 z transforms on the matrix cores.  The G columns of a component are
// blocks of 16 modes; a column block is (16 modes, component c, real or
// imaginary part), NCB of them per workgroup.  Each transform is two
// products, one per parity, each a set of 16 x 16 output tiles with one
// v_mfma 16x16x4 per tile and 4 terms of the sum: wavefront w takes parity
// w & 1 and half of the column blocks, over every row tile.  The data operand
// is read from the LDS columns, the table operand from global memory in lane
// order, padded with zeros to KSM = NZP / 8 steps so that the step loop has a
// compile-time length and every table load is issued ahead of the products.
template <int NZP, int G>
struct ZMma {
  static constexpr int GS = G + 1;
  static constexpr int NCB = (G / 16) * 6;         // column blocks
  static constexpr int CBW = NCB / 2;              // column blocks per wavefront
  static constexpr int MT = (NZP / 2 + 15) / 16;   // 16-row tiles of one parity
  static constexpr int KSM = NZP / 8;              // steps of 4 terms
  typedef Mma<float>::acc acc;

  // float offset in S of this lane's column of block cb (row 0)
  static __device__ __forceinline__ int col(int cb, int lane)
  {
    const int gb = cb / 6, c = (cb % 6) >> 1, part = cb & 1;
    return 2 * (c * NZP * GS + 16 * gb + (lane & 15)) + part;
  }

  // D = Tab_P B over this wavefront's tiles; row(ks) is the S row holding
  // term 4 ks + lane / 16 of the sum (kept in range past the last term, where
  // the table is zero)
  template <class Row>
  static __device__ __forceinline__ void product(const float* __restrict__ tab, const float* S, int P, int cb0,
                                                 Row row, acc (&d)[MT][CBW])
  {
    const int lane = threadIdx.x & 63;
    const float* at = tab + (size_t)P * MT * KSM * 64 + lane;
    float a[MT][KSM];
#pragma unroll
    for (int t = 0; t < MT; ++t)
#pragma unroll
      for (int ks = 0; ks < KSM; ++ks) a[t][ks] = at[(t * KSM + ks) * 64];
    int co[CBW];
#pragma unroll
    for (int i = 0; i < CBW; ++i) co[i] = col(cb0 + i, lane);
#pragma unroll
    for (int t = 0; t < MT; ++t)
#pragma unroll
      for (int i = 0; i < CBW; ++i) d[t][i] = acc{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < KSM; ++ks) {
      const int r = row(ks);
      float b[CBW];
#pragma unroll
      for (int i = 0; i < CBW; ++i) b[i] = S[co[i] + 2 * r * GS];
#pragma unroll
      for (int t = 0; t < MT; ++t)
#pragma unroll
        for (int i = 0; i < CBW; ++i) d[t][i] = Mma<float>::step(a[t][ks], b[i], d[t][i]);
    }
  }

  // Coefficient k = 2 i + P of every column is sum_j Tf_P[i][j] B_P[j], B_0
  // the folded sums (row j) and B_1 the folded differences (row Nz-1-j; for
  // the middle node of an odd Nz the zero padding row NZP-1).  The results
  // replace the folded values in place.
  static __device__ __forceinline__ void forward(const float* __restrict__ Tf, float* S, int Nz, int NH)
  {
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int P = wv & 1, cb0 = (wv >> 1) * CBW;
    acc d[MT][CBW];
    product(Tf, S, P, cb0,
            [&](int ks) {
              const int j = min(4 * ks + (lane >> 4), NH - 1);
              return P ? ((j == Nz - 1 - j) ? NZP - 1 : Nz - 1 - j) : j;
            },
            d);
    __syncthreads();
#pragma unroll
    for (int t = 0; t < MT; ++t)
#pragma unroll
      for (int i = 0; i < CBW; ++i)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int k = 2 * (16 * t + Mma<float>::row(lane, v)) + P;
          if (k < NZP) S[col(cb0 + i, lane) + 2 * k * GS] = d[t][i][v];
        }
  }

  // Node pair (j, Nz-1-j), j < NH: SE = sum_p TiE[j][p] C_{2p} and
  // SO = sum_p TiO[j][p] C_{2p+1} (tables scaled by 1/(Nx Ny)).  Once every
  // wavefront has read the coefficients, SE goes to row j and SO to row
  // NZP/2 + j of each column.
  static __device__ __forceinline__ void inverse(const float* __restrict__ Ti, float* S, int Nz, int NH)
  {
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int P = wv & 1, cb0 = (wv >> 1) * CBW;
    const int np = P ? Nz / 2 : (Nz + 1) / 2;   // coefficients of parity P
    acc d[MT][CBW];
    product(Ti, S, P, cb0, [&](int ks) { return 2 * min(4 * ks + (lane >> 4), max(np - 1, 0)) + P; }, d);
    __syncthreads();
#pragma unroll
    for (int t = 0; t < MT; ++t)
#pragma unroll
      for (int i = 0; i < CBW; ++i)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int j = 16 * t + Mma<float>::row(lane, v);
          if (j < NH) S[col(cb0 + i, lane) + 2 * (P * (NZP / 2) + j) * GS] = d[t][i][v];
        }
  }
};

// The solve: each workgroup takes G consecutive modes (G and the waves per
// SIMD that LDS allows are fixed by the shape, so LDS offsets are
// compile-time and the register budget is capped to reach that occupancy).
template <class T, int LPM, int E>
__global__ void __launch_bounds__(spectral_modes_per_block(LPM, E, sizeof(T) == 4) * LPM,
                                  spectral_waves_per_simd(LPM, E, sizeof(T) == 4))
    k_spectral(SpecArgs A, const typename C2<T>::t* __restrict__ in, typename C2<T>::t* __restrict__ out)
{
  typedef typename C2<T>::t V;
  constexpr int R = E / 2;       // node pairs per lane in the inverse transform
  constexpr int NZP = LPM * E;   // rows held per component (zero from Nz on)
  constexpr int G = spectral_modes_per_block(LPM, E, sizeof(T) == 4);
  constexpr int NT = G * LPM;    // threads per workgroup
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  const int Nz = A.B.Nz, Nm = A.Nm, NH = A.NH;
  // [c][row][g] with row stride G+1: the lanes of one mode (different rows)
  // fall in different LDS banks
  constexpr int GS = G + 1;
  double* tab = reinterpret_cast<double*>(smem_raw);   // [TB_N][NZP]
  constexpr bool kMfma = spectral_mfma(LPM, sizeof(T) == 4);
  static_assert(!kMfma || (NT == 256 && G % 16 == 0), "matrix-core transforms take 4 wavefronts, 16-mode blocks");
  double2* Cr = reinterpret_cast<double2*>(tab + TB_N * NZP);   // [G][3][4] wall corrections
  double* Kl = reinterpret_cast<double*>(Cr + 12 * G);          // [G] |k| (matrix-core path)
  V* S = reinterpret_cast<V*>(Kl + (kMfma ? G : 0));            // values, coefficients, values
  const int g = threadIdx.x / LPM;
  const int l = threadIdx.x % LPM;
  const int r0 = l * E;
  // mode of this lane group (m) and column of this thread in phases 0 and 4
  // (mc; blockDim.x is a multiple of G, so the column is fixed per thread);
  // -1 = beyond the last mode
  const int gc = threadIdx.x % G;
  const int m0 = xcd_group(blockIdx.x, gridDim.x) * G;
  const int m = (m0 + g < Nm) ? m0 + g : -1;
  const int mc = (m0 + gc < Nm) ? m0 + gc : -1;
  // factor-table entry of this lane group's mode, loaded ahead of phase 2
  const int fu = (m > 0) ? A.fac_idx[m] : 0;

  // ---- phase 0: row tables and columns in ----
  // every load of a thread is issued before the first LDS write, so the
  // workgroup waits for HBM once rather than once per row
  {
    constexpr int NTB = (TB_N * NZP + NT - 1) / NT;
    double tv[NTB];
#pragma unroll
    for (int k = 0; k < NTB; ++k) {
      const int i = threadIdx.x + k * NT;
      const int t = i / NZP, row = i - t * NZP;
      tv[k] = (i < TB_N * NZP && row < Nz) ? row_table(A.B, t)[row] : 0.0;
    }
    // thread rows threadIdx.x / G + k LPM, k < 3 E, cover the 3 NZP rows
    constexpr int NL = 3 * E;
    V cv[NL];
#pragma unroll
    for (int k = 0; k < NL; ++k) {
      const int row = threadIdx.x / G + k * LPM;
      const int c = row / NZP, r = row - c * NZP;
      cv[k].x = cv[k].y = (T)0;
      if (mc >= 0 && r < Nz) cv[k] = in[(size_t)(c * Nz + r) * Nm + mc];
    }
#pragma unroll
    for (int k = 0; k < NTB; ++k)
      if (threadIdx.x + k * NT < TB_N * NZP) tab[threadIdx.x + k * NT] = tv[k];
#pragma unroll
    for (int k = 0; k < NL; ++k) S[(threadIdx.x / G + k * LPM) * GS + gc] = cv[k];
  }
  __syncthreads();
  // fold node pairs in place, once per column: row j < Nz-1-j takes
  // x_j + x_{Nz-1-j} and row Nz-1-j takes x_j - x_{Nz-1-j} (the middle node
  // of an odd Nz stays; its odd part is a zero padding row below)
  {
    const int np = Nz / 2;
    for (int i = threadIdx.x; i < 3 * np * G; i += NT) {
      const int gg = i % G, cj = i / G;
      const int c = cj / np, j = cj - c * np;
      V& a = S[(c * NZP + j) * GS + gg];
      V& b = S[(c * NZP + Nz - 1 - j) * GS + gg];
      const V x = a, y = b;
      a.x = x.x + y.x;
      a.y = x.y + y.y;
      b.x = x.x - y.x;
      b.y = x.y - y.y;
    }
  }
  __syncthreads();

  // ---- phase 1: forward transform ----
  // (matrix cores; else each lane forms the coefficients of its own rows)
  if constexpr (kMfma) {
    ZMma<NZP, G>::forward(A.TfM, reinterpret_cast<float*>(S), Nz, NH);
  } else {
    typedef T T2 __attribute__((ext_vector_type(2)));
    T ar[3][E], ai[3][E];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int q = 0; q < E; ++q) ar[c][q] = ai[c][q] = (T)0;
    // [NH][LPM][E]: this lane's E rows of node pair j are contiguous
    const T2* Tf = (const T2*)A.TfP + l * (E / 2);
    // NZP is even, so an odd Nz leaves row NZP-1 as zero padding for the
    // middle node's odd part
#pragma unroll 4
    for (int j = 0; j < NH; ++j) {
      T er[3], ei[3], orr[3], oi[3];
      const int jo = (j == Nz - 1 - j) ? NZP - 1 : Nz - 1 - j;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const V a = S[(c * NZP + j) * GS + g];
        const V b = S[(c * NZP + jo) * GS + g];
        er[c] = a.x;
        ei[c] = a.y;
        orr[c] = b.x;
        oi[c] = b.y;
      }
      T t[E];
#pragma unroll
      for (int q = 0; q < E; q += 2) {
        const T2 w = Tf[(size_t)j * (LPM * E / 2) + q / 2];
        t[q] = w.x;
        t[q + 1] = w.y;
      }
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
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        V v;
        v.x = ar[c][q];
        v.y = ai[c][q];
        S[(c * NZP + r0 + q) * GS + g] = v;
      }
  }
  __syncthreads();

  // ---- phase 2: FP64 BVP of mode m on this lane group, in place ----
  // (the wall corrections go to LDS, not through registers to phase 3)
  double K = 0;
  if (m >= 0) {
    ModeResult res;
    res.nbasis = 0;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int b = 0; b < 4; ++b) res.corr[c][b] = mk(0, 0);
    BvpConsts B = A.B;   // row tables from LDS
    B.sd0 = tab + TB_SD0 * NZP;
    B.sm2 = tab + TB_SM2 * NZP;
    B.sp2 = tab + TB_SP2 * NZP;
    B.fm1 = tab + TB_FM1 * NZP;
    B.fp1 = tab + TB_FP1 * NZP;
    B.R1 = tab + TB_R1 * NZP;
    B.R2 = tab + TB_R2 * NZP;
    B.L1 = tab + TB_L1 * NZP;
    B.L2 = tab + TB_L2 * NZP;
    const DevLanes<LPM> ln{l};
    LdsCoef<T, NZP, GS> fin{S, g, r0};
    LdsStore<T, NZP, GS> fout{S, g, r0};
    const int iy = m / A.Nkx, ix = m - iy * A.Nkx;
    if (m == 0) {
      seg_bvp_k0<LPM, E>(A.B, ln, fin, fout);   // unpadded global tables
    } else {
      const double kx = A.kxs[ix], ky = A.kys[iy];
      K = sqrt(ksq_of(kx, ky));
      // [nk][LPM][E] factor rows of this lane, [nk][4] inverse Schur
      // complement, entry fu of the mode's |k|^2
      double ib[E], ab[E], Gi[4];
      const double2* pib = (const double2*)(A.fac_ib + ((size_t)fu * LPM + l) * E);
      const double2* pab = (const double2*)(A.fac_ab + ((size_t)fu * LPM + l) * E);
#pragma unroll
      for (int q = 0; q < E; q += 2) {
        const double2 u = pib[q / 2], w = pab[q / 2];
        ib[q] = u.x; ib[q + 1] = u.y;
        ab[q] = w.x; ab[q + 1] = w.y;
      }
      {
        const double2* pg = (const double2*)(A.fac_G + (size_t)fu * 4);
        const double2 g0 = pg[0], g1 = pg[1];
        Gi[0] = g0.x; Gi[1] = g0.y; Gi[2] = g1.x; Gi[3] = g1.y;
      }
      const int nyqx = (A.Nx % 2 == 0) && (ix == A.Nx / 2);
      const int nyqy = (A.Ny % 2 == 0) && (iy == A.Ny / 2);
      // the mirror mode's grid wavenumbers, negated: the mirror index of a
      // Nyquist index is itself, any other index maps to minus its wavenumber
      const double kxa = nyqx ? 0.0 - kx : kx, kya = nyqy ? 0.0 - ky : ky;
      seg_bvp_mode<LPM, E, DPS_SPEC_NRV, true, sizeof(T) == 4>(B, ln, kx, ky, nyqx, nyqy, ib, ab, Gi, fin, fout, res,
                                                nyqx || nyqy, kxa, kya);
    }
    if (l == 0)
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int b = 0; b < 4; ++b) Cr[g * 12 + c * 4 + b] = make_double2(res.corr[c][b].re, res.corr[c][b].im);
  }
  // basis functions per component (bottom wall: 2, slit: 4; the k = 0 mode
  // has zero corrections)
  const int nb = (A.B.geometry == MODE_BOTTOM) ? 2 : ((A.B.geometry == MODE_SLIT) ? 4 : 0);
  if constexpr (kMfma) {
    if (l == 0) Kl[g] = K;
    __syncthreads();

    // ---- phase 3 (matrix cores): inverse sums, then node pairs out ----
    // SE and SO of every node pair land in LDS (ZMma::inverse); thread
    // (row group jr, column gl) then forms the pairs jr + LPM r of mode gl
    // for all three components, adds the wall profiles and writes the
    // values to HBM as rows of G modes (no separate column pass)
    ZMma<NZP, G>::inverse(A.TiM, reinterpret_cast<float*>(S), Nz, NH);
    const int gl = threadIdx.x % G, jr = threadIdx.x / G;
    const int mo = (m0 + gl < Nm) ? m0 + gl : -1;
    double zj[R], zm[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int j = min(jr + LPM * r, NH - 1);
      zj[r] = A.zpts[j];
      zm[r] = A.zpts[Nz - 1 - j];
    }
    __syncthreads();
    const double Kg = Kl[gl];
    const double2* cr = Cr + gl * 12;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int j = jr + LPM * r;
      if (j >= NH || mo < 0) continue;
      const int jm = Nz - 1 - j;
      // z_{Nz-1-j} = Lz - z_j, so e^{K(z - Lz)} at one node of the pair is
      // e^{-K z} at the other; FP32 exponentials for the FP32 grid
      double ej = 0, em = 0;
      if (nb) {
        ej = (double)expf((float)(-Kg * zj[r]));
        em = (jm == j) ? ej : (double)expf((float)(-Kg * zm[r]));
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const V se = S[(c * NZP + j) * GS + gl], so = S[(c * NZP + NZP / 2 + j) * GS + gl];
        V vs, vm;
        vs.x = se.x + so.x;
        vs.y = se.y + so.y;
        vm.x = se.x - so.x;
        vm.y = se.y - so.y;
        if (nb) {
          const double2 c0 = cr[c * 4], c1 = cr[c * 4 + 1];
          double sr = c0.x * ej, si = c0.y * ej, tr = c0.x * em, ti = c0.y * em;
          if (nb == 2) {
            sr += c1.x * (zj[r] * ej);
            si += c1.y * (zj[r] * ej);
            tr += c1.x * (zm[r] * em);
            ti += c1.y * (zm[r] * em);
          } else {
            const double2 c2 = cr[c * 4 + 2], c3 = cr[c * 4 + 3];
            sr += c1.x * em + c2.x * (zj[r] * ej) + c3.x * (zj[r] * em);
            si += c1.y * em + c2.y * (zj[r] * ej) + c3.y * (zj[r] * em);
            tr += c1.x * ej + c2.x * (zm[r] * em) + c3.x * (zm[r] * ej);
            ti += c1.y * ej + c2.y * (zm[r] * em) + c3.y * (zm[r] * ej);
          }
          vs.x += (T)(sr * A.inv_nxy);
          vs.y += (T)(si * A.inv_nxy);
          vm.x += (T)(tr * A.inv_nxy);
          vm.y += (T)(ti * A.inv_nxy);
        }
        out[(size_t)(c * Nz + j) * Nm + mo] = vs;
        if (jm != j) out[(size_t)(c * Nz + jm) * Nm + mo] = vm;
      }
    }
  } else {
    __syncthreads();

    // ---- phase 3: inverse transform + wall profiles ----
    {
      typedef T T2 __attribute__((ext_vector_type(2)));
      // [ceil(Nz/2)][LPM][E]: for the coefficient pair (2p, 2p+1), the R rows
      // of coefficient 2p then the R rows of 2p+1, for this lane's node pairs
      const T2* Ti = (const T2*)A.TiP + l * (E / 2);
      // coordinates of this lane's node pairs, loaded ahead of the sums
      double zj[R], zm[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int j = (l + LPM * r < NH) ? l + LPM * r : NH - 1;
        zj[r] = A.zpts[j];
        zm[r] = A.zpts[Nz - 1 - j];
      }
      T er[3][R], ei[3][R], orr[3][R], oi[3][R];
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < R; ++r) er[c][r] = ei[c][r] = orr[c][r] = oi[c][r] = (T)0;
#pragma unroll 4
      for (int p = 0; p < Nz; p += 2) {
        V u[3], w[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          u[c] = S[(c * NZP + p) * GS + g];
          w[c] = S[(c * NZP + p + 1) * GS + g];   // a zero padding row when p + 1 = Nz
        }
        T t[E];
#pragma unroll
        for (int q = 0; q < E; q += 2) {
          const T2 v = Ti[(size_t)(p / 2) * (LPM * E / 2) + q / 2];
          t[q] = v.x;
          t[q + 1] = v.y;
        }
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            er[c][r] += t[r] * u[c].x;
            ei[c][r] += t[r] * u[c].y;
            orr[c][r] += t[R + r] * w[c].x;
            oi[c][r] += t[R + r] * w[c].y;
          }
      }
      __syncthreads();
      const double2* cr = Cr + g * 12;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int j = l + LPM * r;
        if (j >= NH) continue;
        const int jm = Nz - 1 - j;
        // z_{Nz-1-j} = Lz - z_j, so e^{K(z - Lz)} at one node of the pair is
        // e^{-K z} at the other.  The profile lands in grid precision, so an
        // FP32 grid takes FP32 exponentials (the coefficients stay FP64).
        double ej = 0, em = 0;
        if (nb) {
          if constexpr (sizeof(T) == 4) {
            ej = (double)expf((float)(-K * zj[r]));
            em = (jm == j) ? ej : (double)expf((float)(-K * zm[r]));
          } else {
            ej = exp(-K * zj[r]);
            em = (jm == j) ? ej : exp(-K * zm[r]);
          }
        }
        for (int side = 0; side < 2; ++side) {
          const int jj = side ? jm : j;
          if (side && jj == j) break;
          const double z = side ? zm[r] : zj[r];
          const double e1 = side ? em : ej;
          const double e2 = side ? ej : em;
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            V v;
            v.x = side ? (er[c][r] - orr[c][r]) : (er[c][r] + orr[c][r]);
            v.y = side ? (ei[c][r] - oi[c][r]) : (ei[c][r] + oi[c][r]);
            if (nb) {
              const double2 c0 = cr[c * 4], c1 = cr[c * 4 + 1];
              double sr = c0.x * e1, si = c0.y * e1;
              if (nb == 2) {
                sr += c1.x * (z * e1);
                si += c1.y * (z * e1);
              } else {
                const double2 c2 = cr[c * 4 + 2], c3 = cr[c * 4 + 3];
                sr += c1.x * e2 + c2.x * (z * e1) + c3.x * (z * e2);
                si += c1.y * e2 + c2.y * (z * e1) + c3.y * (z * e2);
              }
              v.x += (T)(sr * A.inv_nxy);
              v.y += (T)(si * A.inv_nxy);
            }
            S[(c * NZP + jj) * GS + g] = v;
          }
        }
      }
    }
    __syncthreads();

    // ---- phase 4: columns out ----
    // (the LDS rows of a thread are read together, then stored; the thread
    // index is re-read opaquely so the row offsets are not shared with phase 0
    // and kept live through the solve)
    {
      int tid = threadIdx.x;
      asm volatile("" : "+v"(tid));
      const int gc4 = tid % G;
      const int mc4 = (m0 + gc4 < Nm) ? m0 + gc4 : -1;
      constexpr int NL = 3 * E;
      V cv[NL];
#pragma unroll
      for (int k = 0; k < NL; ++k) cv[k] = S[(tid / G + k * LPM) * GS + gc4];
#pragma unroll
      for (int k = 0; k < NL; ++k) {
        const int row = tid / G + k * LPM;
        const int c = row / NZP, r = row - c * NZP;
        if (mc4 >= 0 && r < Nz) out[(size_t)(c * Nz + r) * Nm + mc4] = cv[k];
      }
    }
  }
}

// Static factorisation of the BVP (it depends on |k|^2 only), one lane per
// distinct |k|^2, written in the lane order of k_spectral: fib/fab
// [nk][lpm][e], fG [nk][4].  The k = 0 entry (ksq = 0) stays zero.
__global__ void k_bvp_factor(BvpConsts B, const double* ksq, int nk, double* s_ib, double* s_ab, int lpm,
                             int e, double* fib, double* fab, double* fG)
{
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= nk) return;
  ScratchView sc;
  sc.stride = nk;
  sc.ib = s_ib + u;
  sc.ab = s_ab + u;
  sc.P = sc.CP = sc.DP = sc.U0 = sc.U1 = sc.U2 = nullptr;
  double Gi[4] = {0, 0, 0, 0};
  const double k2 = ksq[u];
  if (k2 != 0) bvp_factor(B, k2, sc, Gi);
  const int Nz = B.Nz;
  for (int row = 0; row < lpm * e; ++row) {
    const size_t i = (size_t)u * lpm * e + row;   // row = l e + q
    const bool live = (k2 != 0) && (row < Nz);
    fib[i] = live ? s_ib[(size_t)(row + 1) * nk + u] : 0.0;
    fab[i] = live ? s_ab[(size_t)row * nk + u] : 0.0;
  }
  for (int i = 0; i < 4; ++i) fG[(size_t)u * 4 + i] = Gi[i];
}

}  // namespace

bool spectral_shape(int Nz, int* lpm, int* e)
{
  // 8-lane groups where they leave fewer padding rows than 16-lane ones
  // (Nz <= 16: 8 x 2; 33 <= Nz <= 48: 8 x 6 rows instead of 16 x 4).  8 x 8
  // needs too many registers: at Nz = 64 the 16 x 4 shape measured faster
  if (Nz <= 16) { *lpm = 8; *e = 2; return true; }
  if (Nz > 32 && Nz <= 48) { *lpm = 8; *e = 6; return true; }
  for (int E = 2; E <= 8; E += 2)
    if (16 * E >= Nz) { *lpm = 16; *e = E; return true; }
  if (Nz <= 32 * 8) { *lpm = 32; *e = 8; return true; }
  if (Nz <= 64 * 8) { *lpm = 64; *e = 8; return true; }
  return false;
}

static_assert(TB_N == 9, "spectral_lds_bytes counts 9 row tables");

void launch_bvp_factor(const BvpConsts& B, const double* ksq, int nk, double* s_ib, double* s_ab, int lpm,
                       int e, double* fib, double* fab, double* fG, hipStream_t s)
{
  hipLaunchKernelGGL(k_bvp_factor, dim3((nk + 63) / 64), dim3(64), 0, s, B, ksq, nk, s_ib, s_ab, lpm, e,
                     fib, fab, fG);
}

template <class T>
void launch_spectral(const SpecArgs& A0, const void* in, void* out, hipStream_t s)
{
  typedef typename C2<T>::t V;
  const bool mixed = sizeof(T) == sizeof(float);
  const V* i = (const V*)in;
  V* o = (V*)out;
  const SpecArgs& A = A0;
  const int G = spectral_modes_per_block(A.lpm, A.e, mixed);
  const size_t lds = spectral_lds_bytes(A.lpm, A.e, G, mixed);
  const dim3 grid((A.Nm + G - 1) / G), block(G * A.lpm);
#define SPEC_CASE(LP, EE)                                                         \
  if (A.lpm == LP && A.e == EE) {                                                 \
    hipLaunchKernelGGL((k_spectral<T, LP, EE>), grid, block, lds, s, A, i, o); \
    return;                                                                       \
  }
  SPEC_CASE(8, 2) SPEC_CASE(8, 6)
  SPEC_CASE(16, 2) SPEC_CASE(16, 4) SPEC_CASE(16, 6) SPEC_CASE(16, 8) SPEC_CASE(32, 8) SPEC_CASE(64, 8)
#undef SPEC_CASE
}

template void launch_spectral<float>(const SpecArgs&, const void*, void*, hipStream_t);
template void launch_spectral<double>(const SpecArgs&, const void*, void*, hipStream_t);

}  // namespace dps
