// dps_fft.h -- register-resident complex DFTs of 4..32 points, shared by the
// y-direction FFT kernel and the host-side checks.
#pragma once
#include "dps_core.h"

namespace dps {

// cos(pi j / 16), j = 0..16
constexpr double kCos16[17] = {
    1.0,
    0.98078528040323044912618223613424,
    0.92387953251128675612818318939679,
    0.83146961230254523707878837761791,
    0.70710678118654752440084436210485,
    0.55557023301960222474283081394853,
    0.38268343236508977172845998403040,
    0.19509032201612826784828486847702,
    0.0,
    -0.19509032201612826784828486847702,
    -0.38268343236508977172845998403040,
    -0.55557023301960222474283081394853,
    -0.70710678118654752440084436210485,
    -0.83146961230254523707878837761791,
    -0.92387953251128675612818318939679,
    -0.98078528040323044912618223613424,
    -1.0};

// bit reversal of i over log2(R) bits
constexpr int fft_brev(int i, int R)
{
  int r = 0;
  for (int b = 1; b < R; b <<= 1) {
    r = (r << 1) | (i & 1);
    i >>= 1;
  }
  return r;
}

// cos(2 pi j / 24), j = 0..23
constexpr double kCos24[24] = {
    1.0,
    0.96592582628906828674974319972890,
    0.86602540378443864676372317075294,
    0.70710678118654752440084436210485,
    0.5,
    0.25881904510252076234889883762405,
    0.0,
    -0.25881904510252076234889883762405,
    -0.5,
    -0.70710678118654752440084436210485,
    -0.86602540378443864676372317075294,
    -0.96592582628906828674974319972890,
    -1.0,
    -0.96592582628906828674974319972890,
    -0.86602540378443864676372317075294,
    -0.70710678118654752440084436210485,
    -0.5,
    -0.25881904510252076234889883762405,
    0.0,
    0.25881904510252076234889883762405,
    0.5,
    0.70710678118654752440084436210485,
    0.86602540378443864676372317075294,
    0.96592582628906828674974319972890};

template <int R, int DIR, class T> DPS_HD void fft_reg(T (&re)[R], T (&im)[R]);

// DFT of R = 3 P points (P = 1, 2, 4 or 8) held in registers: the three
// P-point DFTs of the decimated sequences x[3 n2 + n1], then
//   X_k = sum_{n1 < 3} exp(DIR 2 pi i n1 k / R) A_{n1}[k mod P]
// with the twiddles cos/sin(2 pi m / 24), m = (24 / R) n1 k mod 24,
// compile-time constants once the loops are unrolled.
template <int R, int DIR, class T>
DPS_HD void dft_reg3(T (&re)[R], T (&im)[R])
{
  static_assert(R == 3 || R == 6 || R == 12 || R == 24, "R must be 3 times 1, 2, 4 or 8");
  constexpr int P = R / 3;
  T ar[3][P], ai[3][P];
#pragma unroll
  for (int n1 = 0; n1 < 3; ++n1)
#pragma unroll
    for (int n2 = 0; n2 < P; ++n2) {
      ar[n1][n2] = re[3 * n2 + n1];
      ai[n1][n2] = im[3 * n2 + n1];
    }
  if constexpr (P > 1) {
#pragma unroll
    for (int n1 = 0; n1 < 3; ++n1) fft_reg<P, DIR>(ar[n1], ai[n1]);
  }
#pragma unroll
  for (int k = 0; k < R; ++k) {
    T xr = ar[0][k % P], xi = ai[0][k % P];
#pragma unroll
    for (int n1 = 1; n1 < 3; ++n1) {
      const int m = ((24 / R) * n1 * k) % 24;
      const T c = (T)kCos24[m];
      const T sn = (T)(DIR * kCos24[(m + 18) % 24]);   // sin(2 pi m / 24) = cos(2 pi (m - 6) / 24)
      const T yr = ar[n1][k % P], yi = ai[n1][k % P];
      xr += yr * c - yi * sn;
      xi += yr * sn + yi * c;
    }
    re[k] = xr;
    im[k] = xi;
  }
}

// In-place DFT of R points held in registers: X_k = sum_n x_n exp(DIR 2 pi
// i n k / R), DIR = -1 forward, +1 inverse (unnormalised).  R = 2, 4, 8, 16
// or 32: radix-2 decimation in time; every twiddle is a compile-time constant
// once the loops are unrolled, and the trivial ones are skipped.  R = 3, 6,
// 12 or 24: dft_reg3 (three R/3-point transforms and a radix-3 combination).
template <int R, int DIR, class T>
DPS_HD void fft_reg(T (&re)[R], T (&im)[R])
{
  if constexpr (R % 3 == 0) {
    dft_reg3<R, DIR>(re, im);
    return;
  } else {
  static_assert(R >= 2 && R <= 32 && (R & (R - 1)) == 0, "R must be a power of two <= 32");
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const int j = fft_brev(i, R);
    if (i < j) {
      const T a = re[i], b = im[i];
      re[i] = re[j];
      im[i] = im[j];
      re[j] = a;
      im[j] = b;
    }
  }
#pragma unroll
  for (int len = 2; len <= R; len <<= 1) {
    const int half = len >> 1;
#pragma unroll
    for (int i = 0; i < R; i += len) {
#pragma unroll
      for (int k = 0; k < half; ++k) {
        // twiddle exp(DIR 2 pi i m / 32), m = 32 k / len in [0, 16)
        const int m = k * (32 / len);
        T br = re[i + k + half], bi = im[i + k + half];
        if (m != 0) {
          const T c = (T)kCos16[m];
          const T s = (T)(DIR * kCos16[m < 8 ? 8 - m : m - 8]);
          const T tr = br * c - bi * s;
          const T ti = br * s + bi * c;
          br = tr;
          bi = ti;
        }
        re[i + k + half] = re[i + k] - br;
        im[i + k + half] = im[i + k] - bi;
        re[i + k] += br;
        im[i + k] += bi;
      }
    }
  }
  }
}

}  // namespace dps
