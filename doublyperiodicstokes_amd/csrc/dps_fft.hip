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

// ---------------------------------------------------------------------------
// x pass: real <-> half-complex transforms of length NX = 2M along the
// contiguous rows of the grid ([slice][y][x], NX reals) and of the spectrum
// ([slice][ky][kx], M + 1 complex values).  A row of reals is read as M
// complex values z[n] = x[2n] + i x[2n+1]; the length-M complex transform
// Z = DFT(z) runs in two register passes with an LDS exchange (M = R1 R2, as
// the y pass), and the half spectrum follows from
//   E[k] = (Z[k] + conj Z[M-k]) / 2,  O[k] = (Z[k] - conj Z[M-k]) / (2i),
//   X[k] = E[k] + W^k O[k],  W = exp(-2 pi i / NX),  k = 0..M
// (indices of Z modulo M).  The inverse runs the same steps backwards:
//   Z[k] = (X[k] + conj X[M-k]) + i W^{-k} (X[k] - conj X[M-k]),  k < M,
// with the imaginary parts of X[0] and X[M] ignored (the real part of the
// full inverse, as a c2r transform takes it), z = IDFT(Z) unnormalised,
// x[2n] = Re z[n], x[2n+1] = Im z[n].  RW rows per workgroup, TR threads per
// row; each row crosses HBM once.
// ---------------------------------------------------------------------------
constexpr int fftx_rows(int M, int TR, int vsz)
{
  int rw = 256 / TR;
  if (rw < 1) rw = 1;
  while (rw > 1 && rw * (M + 1) * vsz > 65536) rw /= 2;
  return rw;
}

template <class T, int R1, int R2> struct FftX {
  static constexpr int M = R1 * R2;
  static constexpr int TR = R1 > R2 ? R1 : R2;
  static constexpr int RW = fftx_rows(M, TR, (int)sizeof(typename C2<T>::t));
  static constexpr int NT = RW * TR;
};

// One LDS row buffer per row, reused by both exchanges (a barrier between the
// reads of one layout and the writes of the next), so that more workgroups
// fit per CU.
template <class T, int R1, int R2>
__global__ void __launch_bounds__(256) k_fft_x_fwd(const T* __restrict__ in, typename C2<T>::t* __restrict__ out,
                                                   int nrows, const typename C2<T>::t* __restrict__ tw)
{
  typedef typename C2<T>::t V;
  typedef FftX<T, R1, R2> S;
  constexpr int M = S::M, TR = S::TR, RW = S::RW;
  __shared__ V buf[RW * M];   // [row][n1][k2] after the first pass, then [row][k] = Z[k]
  const int lr = threadIdx.x / TR, t = threadIdx.x % TR;
  const int row = xcd_group(blockIdx.x, gridDim.x) * RW + lr;
  const bool ok = row < nrows;
  V* b = buf + lr * M;
  if (t < R1) {
    const V* z = reinterpret_cast<const V*>(in + (size_t)(ok ? row : 0) * (2 * M));
    const int n1 = t;
    T re[R2], im[R2];
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) {
      V v;
      v.x = v.y = (T)0;
      if (ok) v = z[n1 + R1 * n2];
      re[n2] = v.x;
      im[n2] = v.y;
    }
    fft_reg<R2, -1>(re, im);
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) {
      V o;
      if (k2 == 0) {
        o.x = re[0];
        o.y = im[0];
      } else {
        const V w = tw[2 * n1 * k2];   // W_M^{n1 k2} = W^{2 n1 k2}
        o.x = re[k2] * w.x - im[k2] * w.y;
        o.y = re[k2] * w.y + im[k2] * w.x;
      }
      b[n1 * R2 + k2] = o;
    }
  }
  __syncthreads();
  T re[R1], im[R1];
  if (t < R2) {
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) {
      const V v = b[n1 * R2 + t];
      re[n1] = v.x;
      im[n1] = v.y;
    }
    fft_reg<R1, -1>(re, im);
  }
  __syncthreads();
  if (t < R2)
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) {
      V v;
      v.x = re[k1];
      v.y = im[k1];
      b[t + R2 * k1] = v;
    }
  __syncthreads();
  if (!ok) return;
  V* x = out + (size_t)row * (M + 1);
  for (int k = t; k <= M; k += TR) {
    const V p = b[k == M ? 0 : k], q = b[k == 0 ? 0 : M - k];   // Z[k mod M], Z[(M - k) mod M]
    const T er = (T)0.5 * (p.x + q.x), ei = (T)0.5 * (p.y - q.y);
    const T orr = (T)0.5 * (p.y + q.y), oi = (T)0.5 * (q.x - p.x);
    const V w = tw[k];
    V v;
    v.x = er + (w.x * orr - w.y * oi);
    v.y = ei + (w.x * oi + w.y * orr);
    x[k] = v;
  }
}

// Inverse: the first pass forms its inputs Z[n1 + R1 n2] from the pairs
// (X[k], X[M - k]) read from HBM in its own order (runs of R1 values, the
// mirrored reads hit the cache), then the two passes run as in the forward
// kernel, the final values written from registers.  (Tried and not kept: Z
// staged in LDS before the first pass, with the row read once or per pair,
// and the radices swapped so that the second pass writes longer runs; all
// slower.)
template <class T, int R1, int R2>
__global__ void __launch_bounds__(256) k_fft_x_inv(const typename C2<T>::t* __restrict__ in, T* __restrict__ out,
                                                   int nrows, const typename C2<T>::t* __restrict__ tw)
{
  typedef typename C2<T>::t V;
  typedef FftX<T, R1, R2> S;
  constexpr int M = S::M, TR = S::TR, RW = S::RW;
  __shared__ V buf[RW * M];   // [row][n1][k2] after the first pass
  const int lr = threadIdx.x / TR, t = threadIdx.x % TR;
  const int row = xcd_group(blockIdx.x, gridDim.x) * RW + lr;
  const bool ok = row < nrows;
  V* b = buf + lr * M;
  if (t < R1) {
    const V* X = in + (size_t)(ok ? row : 0) * (M + 1);
    const int n1 = t;
    T re[R2], im[R2];
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) {
      // Z[k] = (p + conj q) + i W^{-k} (p - conj q), p = X[k], q = X[M - k]
      const int k = n1 + R1 * n2;
      V p, q;
      p.x = p.y = q.x = q.y = (T)0;
      if (ok) {
        p = X[k];
        q = X[M - k];
      }
      if (k == 0) p.y = q.y = (T)0;   // X[0] and X[M] count by their real parts
      const T er = p.x + q.x, ei = p.y - q.y;
      const T dr = p.x - q.x, di = p.y + q.y;
      const V w = tw[k];   // W^{-k} = conj(w)
      const T orr = dr * w.x + di * w.y, oi = di * w.x - dr * w.y;
      re[n2] = er - oi;
      im[n2] = ei + orr;
    }
    fft_reg<R2, 1>(re, im);
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) {
      V o;
      if (k2 == 0) {
        o.x = re[0];
        o.y = im[0];
      } else {
        const V w = tw[2 * n1 * k2];   // conj of W_M^{n1 k2}
        o.x = re[k2] * w.x + im[k2] * w.y;
        o.y = im[k2] * w.x - re[k2] * w.y;
      }
      b[n1 * R2 + k2] = o;
    }
  }
  __syncthreads();
  if (t < R2 && ok) {
    const int k2 = t;
    T ro[R1], io[R1];
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) {
      const V v = b[n1 * R2 + k2];
      ro[n1] = v.x;
      io[n1] = v.y;
    }
    fft_reg<R1, 1>(ro, io);
    V* z = reinterpret_cast<V*>(out + (size_t)row * (2 * M));
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) {
      V v;
      v.x = ro[k1];
      v.y = io[k1];
      z[k2 + R2 * k1] = v;
    }
  }
}

// R1 >= R2: the forward's first pass reads runs of R1 values
template <class T, int R1, int R2>
void fft_x_launch(const void* in, void* out, int nrows, const void* tw, int dir, hipStream_t s)
{
  typedef typename C2<T>::t V;
  typedef FftX<T, R1, R2> S;
  static_assert(R1 >= R2, "first radix the larger");
  const dim3 grid((nrows + S::RW - 1) / S::RW);
  if (dir < 0)
    hipLaunchKernelGGL((k_fft_x_fwd<T, R1, R2>), grid, dim3(S::NT), 0, s, (const T*)in, (V*)out, nrows,
                       (const V*)tw);
  else
    hipLaunchKernelGGL((k_fft_x_inv<T, R1, R2>), grid, dim3(S::NT), 0, s, (const V*)in, (T*)out, nrows,
                       (const V*)tw);
}

}  // namespace

bool fft_x_supported(int Nx)
{
  switch (Nx) {
    case 32: case 64: case 96: case 128: case 192: case 256: case 384: case 512: case 768: case 1024:
    case 1536: case 2048:
      return true;
    default:
      return false;
  }
}

template <class T>
void launch_fft_x(const void* in, void* out, int nrows, int Nx, const void* tw, int dir, hipStream_t s)
{
  switch (Nx) {
    case 32: fft_x_launch<T, 4, 4>(in, out, nrows, tw, dir, s); break;
    case 64: fft_x_launch<T, 8, 4>(in, out, nrows, tw, dir, s); break;
    case 96: fft_x_launch<T, 8, 6>(in, out, nrows, tw, dir, s); break;
    case 128: fft_x_launch<T, 8, 8>(in, out, nrows, tw, dir, s); break;
    case 192: fft_x_launch<T, 12, 8>(in, out, nrows, tw, dir, s); break;
    case 256: fft_x_launch<T, 16, 8>(in, out, nrows, tw, dir, s); break;
    case 384: fft_x_launch<T, 16, 12>(in, out, nrows, tw, dir, s); break;
    case 512: fft_x_launch<T, 16, 16>(in, out, nrows, tw, dir, s); break;
    case 768: fft_x_launch<T, 24, 16>(in, out, nrows, tw, dir, s); break;
    case 1024: fft_x_launch<T, 32, 16>(in, out, nrows, tw, dir, s); break;
    case 1536: fft_x_launch<T, 32, 24>(in, out, nrows, tw, dir, s); break;
    case 2048: fft_x_launch<T, 32, 32>(in, out, nrows, tw, dir, s); break;
    default: break;   // callers check fft_x_supported
  }
}
template void launch_fft_x<float>(const void*, void*, int, int, const void*, int, hipStream_t);
template void launch_fft_x<double>(const void*, void*, int, int, const void*, int, hipStream_t);

bool fft_y_supported(int Ny)
{
  switch (Ny) {
    case 16: case 32: case 48: case 64: case 96: case 128: case 192: case 256: case 384: case 512: case 768:
    case 1024:
      return true;
    default:
      return false;
  }
}

template <class T>
void launch_fft_y(void* data, int nslices, int Nkx, int Ny, const void* tw, int dir, hipStream_t s)
{
  switch (Ny) {
    case 16: fft_y_launch<T, 4, 4>(data, nslices, Nkx, tw, dir, s); break;
    case 32: fft_y_launch<T, 8, 4>(data, nslices, Nkx, tw, dir, s); break;
    case 48: fft_y_launch<T, 8, 6>(data, nslices, Nkx, tw, dir, s); break;
    case 64: fft_y_launch<T, 8, 8>(data, nslices, Nkx, tw, dir, s); break;
    case 96: fft_y_launch<T, 12, 8>(data, nslices, Nkx, tw, dir, s); break;
    case 192: fft_y_launch<T, 16, 12>(data, nslices, Nkx, tw, dir, s); break;
    case 384: fft_y_launch<T, 24, 16>(data, nslices, Nkx, tw, dir, s); break;
    case 768: fft_y_launch<T, 32, 24>(data, nslices, Nkx, tw, dir, s); break;
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
