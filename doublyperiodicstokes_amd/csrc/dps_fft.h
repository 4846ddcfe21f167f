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

// In-place DFT of R points (R = 2, 4, 8, 16 or 32) held in registers:
// X_k = sum_n x_n exp(DIR 2 pi i n k / R), DIR = -1 forward, +1 inverse
// (unnormalised).  Radix-2 decimation in time; every twiddle is a compile-time
// constant once the loops are unrolled, and the trivial ones are skipped.
template <int R, int DIR, class T>
DPS_HD void fft_reg(T (&re)[R], T (&im)[R])
{
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

}  // namespace dps
