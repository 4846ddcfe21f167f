// dps_zgemm.hip -- the Chebyshev z transform of the one-wavenumber-per-lane
// chain as batched GEMMs on the matrix cores.
//
// The spectrum [c][z][Nm] complex is, per component, a real matrix of N rows
// (z) and 2 Nm columns (re, im of each mode).  The forward transform
// (Transform.cpp:215-268: values at the Chebyshev nodes -> coefficients) is
// folded over node pairs: with E_j = x_j + x_{N-1-j} and O_j = x_j - x_{N-1-j}
// (j < NH), even coefficients are TfE E and odd ones TfO O.  The inverse
// (Transform.cpp:270-304) forms S_E = TiE C_even, S_O = TiO C_odd and writes
// x_j = S_E + S_O, x_{N-1-j} = S_E - S_O, adding the analytic wall-correction
// profiles of the mode (DPTools.cpp:590-877) at each node.
//
// A workgroup owns TN consecutive columns of one component: the folded input
// rows are staged in LDS (256-byte row segments), and its four wavefronts
// take the 16 x 16 output tiles round robin, each tile a chain of
// v_mfma 16x16x4 steps (FP32 or FP64) with the data operand from LDS and the
// table operand from LDS too when the tables fit beside the data (odd row
// stride), else from global memory (every workgroup reads the same tables).
#include "dps_internal.h"
#include "dps_mma.h"

namespace dps {

namespace {

__host__ __device__ constexpr int up4(int n) { return (n + 3) & ~3; }
__host__ __device__ constexpr int up16(int n) { return (n + 15) & ~15; }

// One 16 x 16 tile D = A B with A (rows mt*16.., K columns) row-major at Ag
// (row stride lda, zero beyond M rows / K columns; global memory or LDS) and
// B (K rows, TN columns) in LDS.
template <class T, int TN>
__device__ __forceinline__ typename Mma<T>::acc tile(const T* __restrict__ Ag, int lda, int M, int K, int mt,
                                                     const T* Bs, int KP, int nt, int lane)
{
  typename Mma<T>::acc d;
  d[0] = d[1] = d[2] = d[3] = (T)0;
  const int row = mt * 16 + (lane & 15);
  const int kq = lane >> 4;
  const bool rok = row < M;
  const T* ar = Ag + (size_t)row * lda;
  const T* bc = Bs + nt * 16 + (lane & 15);
#pragma unroll 4
  for (int k0 = 0; k0 < KP; k0 += 4) {
    const int k = k0 + kq;
    const T a = (rok && k < K) ? ar[k] : (T)0;
    const T b = bc[k * TN];
    d = Mma<T>::step(a, b, d);
  }
  return d;
}

// Copy a row-major table (M rows, K columns, row stride K) into LDS with an
// odd row stride (no bank conflicts between the 16 rows an MFMA step reads).
template <class T>
__device__ __forceinline__ void stage_table(T* dst, int ld, const T* __restrict__ src, int M, int K)
{
  for (int i = threadIdx.x; i < M * K; i += blockDim.x) {
    const int r = i / K, k = i - r * K;
    dst[r * ld + k] = src[i];
  }
}

template <class T, int TN, bool tab_lds>
__global__ void __launch_bounds__(256) k_zfwd_mma(ZArgs A, const T* __restrict__ in, T* __restrict__ out)
{
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  const int N = A.N, NH = A.NH, KP = up4(NH);
  T* BE = reinterpret_cast<T*>(smem_raw);   // [KP][TN]
  T* BO = BE + KP * TN;                     // [KP][TN]
  // tables in LDS (row stride NH | 1) when they fit, else read from global
  const int ld = tab_lds ? (NH | 1) : NH;
  const T* TE = (const T*)A.TfE;
  const T* TO = (const T*)A.TfO;
  if (tab_lds) {
    T* tE = BO + KP * TN;
    T* tO = tE + A.NE * ld;
    stage_table(tE, ld, TE, A.NE, NH);
    stage_table(tO, ld, TO, A.NO, NH);
    TE = tE;
    TO = tO;
  }
  const int c = blockIdx.y;
  const int ncols = 2 * A.Nm;
  const int n0 = blockIdx.x * TN;
  const T* src = in + (size_t)c * N * ncols;
  for (int i = threadIdx.x; i < KP * TN; i += blockDim.x) {
    const int j = i / TN, n = i - j * TN, col = n0 + n;
    T e = (T)0, o = (T)0;
    if (j < NH && col < ncols) {
      const T a = src[(size_t)j * ncols + col];
      if (j == N - 1 - j) {
        e = a;
      } else {
        const T b = src[(size_t)(N - 1 - j) * ncols + col];
        e = a + b;
        o = a - b;
      }
    }
    BE[i] = e;
    BO[i] = o;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
  constexpr int NT = TN / 16;
  const int MTE = up16(A.NE) / 16, MTO = up16(A.NO) / 16;
  const int ntile = (MTE + MTO) * NT;
  T* dst = out + (size_t)c * N * ncols;
  for (int t = wave; t < ntile; t += nwaves) {
    const int part = (t / NT) >= MTE;            // 0: even coefficients, 1: odd
    const int mt = part ? (t / NT - MTE) : t / NT;
    const int nt = t % NT;
    const int M = part ? A.NO : A.NE;
    const typename Mma<T>::acc d =
        tile<T, TN>(part ? TO : TE, ld, M, NH, mt, part ? BO : BE, KP, nt, lane);
    const int col = n0 + nt * 16 + (lane & 15);
    if (col >= ncols) continue;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int p = mt * 16 + Mma<T>::row(lane, v);
      if (p < M) dst[(size_t)(2 * p + part) * ncols + col] = d[v];
    }
  }
}

// Wall-profile exponentials e^{-K z} at the two nodes of a pair (j, N-1-j),
// in the grid precision: they scale FP64 correction coefficients into a
// value that is stored in that precision.
__device__ __forceinline__ void pair_exp(float K, float zj, float zm, float* ej, float* em)
{
  *ej = __expf(-K * zj);
  *em = __expf(-K * zm);
}
__device__ __forceinline__ void pair_exp(double K, double zj, double zm, double* ej, double* em)
{
  *ej = exp(-K * zj);
  *em = exp(-K * zm);
}

template <class T, int TN, bool tab_lds>
__global__ void __launch_bounds__(256) k_zinv_mma(ZArgs A, const T* __restrict__ in, T* __restrict__ out)
{
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  const int N = A.N, NH = A.NH, KE = up4(A.NE), KO = up4(A.NO);
  T* BE = reinterpret_cast<T*>(smem_raw);   // [KE][TN] even coefficients
  T* BO = BE + KE * TN;                     // [KO][TN] odd coefficients
  const int ldE = tab_lds ? (A.NE | 1) : A.NE, ldO = tab_lds ? (A.NO | 1) : A.NO;
  const T* TE = (const T*)A.TiE;
  const T* TO = (const T*)A.TiO;
  if (tab_lds) {
    T* tE = BO + KO * TN;
    T* tO = tE + NH * ldE;
    stage_table(tE, ldE, TE, NH, A.NE);
    stage_table(tO, ldO, TO, NH, A.NO);
    TE = tE;
    TO = tO;
  }
  const int c = blockIdx.y;
  const int ncols = 2 * A.Nm;
  const int n0 = blockIdx.x * TN;
  const T* src = in + (size_t)c * N * ncols;
  for (int i = threadIdx.x; i < (KE + KO) * TN; i += blockDim.x) {
    const int r = i / TN, n = i - r * TN, col = n0 + n;
    const bool odd = r >= KE;
    const int p = odd ? r - KE : r;
    const int k = 2 * p + (odd ? 1 : 0);
    T v = (T)0;
    if (p < (odd ? A.NO : A.NE) && col < ncols) v = src[(size_t)k * ncols + col];
    BE[i] = v;   // BO follows BE contiguously
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
  constexpr int NT = TN / 16;
  const int MT = up16(NH) / 16;
  T* dst = out + (size_t)c * N * ncols;
  for (int t = wave; t < MT * NT; t += nwaves) {
    const int mt = t / NT, nt = t % NT;
    const typename Mma<T>::acc se = tile<T, TN>(TE, ldE, NH, A.NE, mt, BE, KE, nt, lane);
    const typename Mma<T>::acc so = tile<T, TN>(TO, ldO, NH, A.NO, mt, BO, KO, nt, lane);
    const int col = n0 + nt * 16 + (lane & 15);
    if (col >= ncols) continue;
    // wall corrections of this column's mode and component
    const int m = col >> 1, im = col & 1;
    const int nb = (m == 0) ? 0 : A.nbasis;
    double cc[4] = {0, 0, 0, 0};
    T K = 0;
    if (nb) {
      const double2* cr = A.corr + ((size_t)m * 3 + c) * 4;
      for (int b = 0; b < 4; ++b) cc[b] = im ? cr[b].y : cr[b].x;
      const int iy = m / A.Nkx, ix = m - iy * A.Nkx;
      const double kx = A.kxs[ix], ky = A.kys[iy];
      K = (T)sqrt(ksq_of(kx, ky));
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int j = mt * 16 + Mma<T>::row(lane, v);
      if (j >= NH) continue;
      const int jm = N - 1 - j;
      // z_{N-1-j} = Lz - z_j: e^{K (z - Lz)} at one node of the pair is
      // e^{-K z} at the other, two exponentials per pair
      T ex[2] = {0, 0};
      if (nb) pair_exp(K, (T)A.zpts[j], (T)A.zpts[jm], &ex[0], &ex[1]);
      for (int side = 0; side < 2; ++side) {
        const int jj = side ? jm : j;
        if (side && jj == j) break;
        T val = side ? (se[v] - so[v]) : (se[v] + so[v]);
        if (nb) {
          const double z = A.zpts[jj];
          const double e1 = (double)ex[side], e2 = (double)ex[1 - side];
          double s = cc[0] * e1;
          if (nb == 2) s += cc[1] * (z * e1);
          else s += cc[1] * e2 + cc[2] * (z * e1) + cc[3] * (z * e2);
          val += (T)(s * A.inv_nxy);
        }
        if (A.accumulate) dst[(size_t)jj * ncols + col] += val;
        else dst[(size_t)jj * ncols + col] = val;
      }
    }
  }
}

// widest column tile (a multiple of 16) whose LDS stays within 64 KB; 0: none
template <class T>
int zmma_tile(int rows)
{
  for (int tn : {64, 32, 16})
    if ((size_t)rows * tn * sizeof(T) <= 65536 && (tn * sizeof(T) <= 256)) return tn;
  return 0;
}

}  // namespace

template <class T>
bool launch_zfwd_mma(const ZArgs& A, const void* in, void* out, hipStream_t s)
{
  const int rows = 2 * up4(A.NH);
  const int tn = zmma_tile<T>(rows);
  if (!tn) return false;
  const int ncols = 2 * A.Nm;
  const size_t data = (size_t)rows * tn * sizeof(T);
  const size_t tabs = (size_t)(A.NE + A.NO) * (A.NH | 1) * sizeof(T);
  const bool tab_lds = data + tabs <= 65536;
  const size_t lds = data + (tab_lds ? tabs : 0);
  const dim3 grid((ncols + tn - 1) / tn, 3);
  switch (tn) {
#define ZF(TN)                                                                                              \
  case TN:                                                                                                  \
    if (tab_lds) hipLaunchKernelGGL((k_zfwd_mma<T, TN, true>), grid, dim3(256), lds, s, A, (const T*)in, (T*)out); \
    else hipLaunchKernelGGL((k_zfwd_mma<T, TN, false>), grid, dim3(256), lds, s, A, (const T*)in, (T*)out);        \
    return true;
    ZF(64) ZF(32) ZF(16)
#undef ZF
    default: return false;
  }
}

template <class T>
bool launch_zinv_mma(const ZArgs& A, const void* in, void* out, hipStream_t s)
{
  const int rows = up4(A.NE) + up4(A.NO);
  const int tn = zmma_tile<T>(rows);
  if (!tn) return false;
  const int ncols = 2 * A.Nm;
  const size_t data = (size_t)rows * tn * sizeof(T);
  const size_t tabs = (size_t)A.NH * ((A.NE | 1) + (A.NO | 1)) * sizeof(T);
  const bool tab_lds = data + tabs <= 65536;
  const size_t lds = data + (tab_lds ? tabs : 0);
  const dim3 grid((ncols + tn - 1) / tn, 3);
  switch (tn) {
#define ZI(TN)                                                                                              \
  case TN:                                                                                                  \
    if (tab_lds) hipLaunchKernelGGL((k_zinv_mma<T, TN, true>), grid, dim3(256), lds, s, A, (const T*)in, (T*)out); \
    else hipLaunchKernelGGL((k_zinv_mma<T, TN, false>), grid, dim3(256), lds, s, A, (const T*)in, (T*)out);        \
    return true;
    ZI(64) ZI(32) ZI(16)
#undef ZI
    default: return false;
  }
}

template <class T>
bool zgemm_supported(const ZArgs& A)
{
  return zmma_tile<T>(2 * up4(A.NH)) > 0 && zmma_tile<T>(up4(A.NE) + up4(A.NO)) > 0;
}

template bool zgemm_supported<float>(const ZArgs&);
template bool zgemm_supported<double>(const ZArgs&);
template bool launch_zfwd_mma<float>(const ZArgs&, const void*, void*, hipStream_t);
template bool launch_zfwd_mma<double>(const ZArgs&, const void*, void*, hipStream_t);
template bool launch_zinv_mma<float>(const ZArgs&, const void*, void*, hipStream_t);
template bool launch_zinv_mma<double>(const ZArgs&, const void*, void*, hipStream_t);

}  // namespace dps
