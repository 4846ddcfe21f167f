// dps_fft.hip -- the y pass of the planar r2c / c2r transforms.
//
// The spectrum after the x pass is [slice][ky][kx] (kx fastest, Nkx = Nx/2+1
// per row; a slice is one (component, z level)).  The transform along ky is a
// strided c2c of length Ny for every (slice, kx) column.  Columns are flattened
// over (slice, kx) and a workgroup takes TW consecutive ones, so that every
// load and store of the tile touches TW consecutive complex values of a row
// (256 bytes) -- the tile may straddle two slices, which only splits a row
// access in two.  Consecutive tiles run on the same XCD.  The length N = R1 R2
// transform of each column is done in two register passes with one exchange
// through LDS (n = n1 + R1 n2, k = k2 + R2 k1):
//   Y[n1][k2] = W_N^{n1 k2} sum_n2 x[n1 + R1 n2] W_R2^{n2 k2}
//   X[k2 + R2 k1] = sum_n1 Y[n1][k2] W_R1^{n1 k1}
// in place, so the spectrum crosses HBM once each way.
#include "dps_fft.h"
#include "dps_internal.h"

namespace dps {

namespace {

template <class T, int N> struct FftTile {
  typedef typename C2<T>::t V;
  // 256 bytes of each row per access; the LDS tile (N x TW complex) within 64 KB
  static constexpr int tw0 = 256 / (int)sizeof(V);
  static constexpr int value = (N * tw0 * (int)sizeof(V) <= 65536)         ? tw0
                               : (N * (tw0 / 2) * (int)sizeof(V) <= 65536) ? tw0 / 2
                               : (N * (tw0 / 4) * (int)sizeof(V) <= 65536) ? tw0 / 4
                                                                           : tw0 / 8;
};

// threads per workgroup: one column group per row of the longer pass, at
// most 512 (more loads in flight per CU)
template <class T, int R1, int R2> struct FftThreads {
  static constexpr int TW = FftTile<T, R1 * R2>::value;
  static constexpr int R = R1 > R2 ? R1 : R2;
  static constexpr int value = TW * R < 512 ? TW * R : 512;
};

template <class T, int R1, int R2, int DIR>
__global__ void __launch_bounds__(512) k_fft_y(typename C2<T>::t* __restrict__ data, int ncols, int Nkx,
                                               const typename C2<T>::t* __restrict__ tw)
{
  typedef typename C2<T>::t V;
  constexpr int N = R1 * R2;
  constexpr int TW = FftTile<T, N>::value;
  constexpr int G1 = FftThreads<T, R1, R2>::value / TW;   // column groups
  __shared__ V buf[N * TW];       // [n1][k2][col]
  const int col = threadIdx.x % TW, g = threadIdx.x / TW;
  // neighbouring tiles share the cache lines at their row ends (rows of Nkx
  // complex values are not line aligned): keep them on one XCD's L2
  const int q = xcd_group(blockIdx.x, gridDim.x) * TW + col;
  const bool ok = q < ncols;
  const int s = ok ? q / Nkx : 0;
  const int kx = ok ? q - s * Nkx : 0;
  V* __restrict__ base = data + (size_t)s * N * Nkx + kx;
  for (int n1 = g; n1 < R1; n1 += G1) {
    T re[R2], im[R2];
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) {
      V v;
      v.x = v.y = (T)0;
      if (ok) v = base[(size_t)(n1 + R1 * n2) * Nkx];
      re[n2] = v.x;
      im[n2] = v.y;
    }
    fft_reg<R2, DIR>(re, im);
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) {
      V o;
      if (k2 == 0) {
        o.x = re[0];
        o.y = im[0];
      } else {
        // W_N^{n1 k2} (the table holds the forward twiddles)
        const V w = tw[n1 * k2];
        const T wi = (DIR > 0) ? -w.y : w.y;
        o.x = re[k2] * w.x - im[k2] * wi;
        o.y = re[k2] * wi + im[k2] * w.x;
      }
      buf[(n1 * R2 + k2) * TW + col] = o;
    }
  }
  __syncthreads();
  for (int k2 = g; k2 < R2; k2 += G1) {
    T re[R1], im[R1];
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) {
      const V v = buf[(n1 * R2 + k2) * TW + col];
      re[n1] = v.x;
      im[n1] = v.y;
    }
    fft_reg<R1, DIR>(re, im);
    if (ok)
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) {
        V v;
        v.x = re[k1];
        v.y = im[k1];
        base[(size_t)(k2 + R2 * k1) * Nkx] = v;
      }
  }
}

template <class T, int R1, int R2>
void fft_y_launch(void* data, int nslices, int Nkx, const void* tw, int dir, hipStream_t s)
{
  typedef typename C2<T>::t V;
  constexpr int TW = FftTile<T, R1 * R2>::value;
  constexpr int NT = FftThreads<T, R1, R2>::value;
  const int ncols = nslices * Nkx;
  const dim3 grid((ncols + TW - 1) / TW);
  if (dir < 0)
    hipLaunchKernelGGL((k_fft_y<T, R1, R2, -1>), grid, dim3(NT), 0, s, (V*)data, ncols, Nkx, (const V*)tw);
  else
    hipLaunchKernelGGL((k_fft_y<T, R1, R2, 1>), grid, dim3(NT), 0, s, (V*)data, ncols, Nkx, (const V*)tw);
}

}  // namespace

bool fft_y_supported(int Ny)
{
  return Ny == 16 || Ny == 32 || Ny == 64 || Ny == 128 || Ny == 256 || Ny == 512 || Ny == 1024;
}

template <class T>
void launch_fft_y(void* data, int nslices, int Nkx, int Ny, const void* tw, int dir, hipStream_t s)
{
  switch (Ny) {
    case 16: fft_y_launch<T, 4, 4>(data, nslices, Nkx, tw, dir, s); break;
    case 32: fft_y_launch<T, 8, 4>(data, nslices, Nkx, tw, dir, s); break;
    case 64: fft_y_launch<T, 8, 8>(data, nslices, Nkx, tw, dir, s); break;
    case 128: fft_y_launch<T, 16, 8>(data, nslices, Nkx, tw, dir, s); break;
    case 256: fft_y_launch<T, 16, 16>(data, nslices, Nkx, tw, dir, s); break;
    case 512: fft_y_launch<T, 32, 16>(data, nslices, Nkx, tw, dir, s); break;
    case 1024: fft_y_launch<T, 32, 32>(data, nslices, Nkx, tw, dir, s); break;
    default: break;   // callers check fft_y_supported
  }
}
template void launch_fft_y<float>(void*, int, int, int, const void*, int, hipStream_t);
template void launch_fft_y<double>(void*, int, int, int, const void*, int, hipStream_t);

}  // namespace dps
