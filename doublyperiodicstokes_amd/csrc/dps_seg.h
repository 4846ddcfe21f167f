// dps_seg.h -- the per-wavenumber FP64 boundary-value solve of dps_core.h
// distributed over a group of LPM lanes.
//
// Lane l of a group owns the E consecutive Chebyshev rows [l*E, l*E + E)
// (E even, so a lane's rows alternate even/odd exactly like the global rows).
// Every pass of the sequential solver maps onto one of three group patterns:
//   * first-order recurrences along the even or odd rows (KBPENTA forward and
//     backward sweeps, the chebCoeffDiff suffix sums): a local pass with zero
//     carry, an affine scan of the lanes' composite maps, and a local pass
//     that replays the recurrence from the lane's true carry;
//   * sums over rows (BC dot products, wall values): local sums and an xor
//     butterfly, so every lane ends with the bit-identical total;
//   * five-point stencils (SIMat / FIMat products): two halo rows from each
//     neighbouring lane.
// Lane exchange goes through a policy object L with members
//   int l;  double up(v, d);  double down(v, d);  double xr(v, d);
// (values from lane l-d / l+d / l^d of the group, own value when outside).
// The gfx950 kernel binds it to __shfl_up/__shfl_down/__shfl_xor over LPM
// lanes; the host-side checks bind it to threads joined by a barrier, so the
// identical code runs on both.
//
// Rows >= Nz are padding: their inputs are zero, their factors are zero, and
// the rows Nz and Nz+1 carry the two Schur unknowns where a stencil reads
// them, as the sequential solver's XP(Nz) = Y0, XP(Nz+1) = Y1.
#pragma once
#include "dps_core.h"

namespace dps {

template <int LPM, int N, class L>
DPS_HD void seg_sum(const L& ln, double (&v)[N])
{
  #pragma unroll
  for (int d = 1; d < LPM; d <<= 1)
    #pragma unroll
    for (int i = 0; i < N; ++i) v[i] += ln.xr(v[i], d);
}

// Inclusive scan over the group of the affine maps y -> a + b y held by each
// lane (UP: lane 0 applies first; !UP: lane LPM-1 applies first).  NC chains
// with their own multipliers, NR right-hand sides per chain.
template <int LPM, int NC, int NR, bool UP, class L>
DPS_HD void seg_scan(const L& ln, cplx (&a)[NC][NR], double (&b)[NC])
{
  #pragma unroll
  for (int d = 1; d < LPM; d <<= 1) {
    double bs[NC];
    cplx as[NC][NR];
    #pragma unroll
    for (int c = 0; c < NC; ++c) {
      bs[c] = UP ? ln.up(b[c], d) : ln.down(b[c], d);
      #pragma unroll
      for (int r = 0; r < NR; ++r)
        as[c][r] = UP ? mk(ln.up(a[c][r].re, d), ln.up(a[c][r].im, d))
                      : mk(ln.down(a[c][r].re, d), ln.down(a[c][r].im, d));
    }
    const bool take = UP ? (ln.l >= d) : (ln.l + d < LPM);
    if (take)
      #pragma unroll
      for (int c = 0; c < NC; ++c) {
        #pragma unroll
        for (int r = 0; r < NR; ++r) a[c][r] = a[c][r] + b[c] * as[c][r];
        b[c] = b[c] * bs[c];
      }
  }
}

// Value entering this lane: the inclusive value of the preceding lane (zero
// for the first lane in scan order).
template <int LPM, int NC, int NR, bool UP, class L>
DPS_HD void seg_carry(const L& ln, const cplx (&a)[NC][NR], cplx (&cin)[NC][NR])
{
  const bool first = UP ? (ln.l == 0) : (ln.l == LPM - 1);
  #pragma unroll
  for (int c = 0; c < NC; ++c)
    #pragma unroll
    for (int r = 0; r < NR; ++r) {
      const cplx v = UP ? mk(ln.up(a[c][r].re, 1), ln.up(a[c][r].im, 1))
                        : mk(ln.down(a[c][r].re, 1), ln.down(a[c][r].im, 1));
      cin[c][r] = first ? mk(0, 0) : v;
    }
}

// KBPENTA forward sweep g_r = P_r - lz_r g_{r-2}, in place (LinearSolvers.h:106-120).
template <int LPM, int E, int NR, class L>
DPS_HD void seg_forward(const L& ln, cplx (&P)[NR][E], const double (&lz)[E])
{
  cplx a[2][NR];
  double b[2] = {1.0, 1.0};
  #pragma unroll
  for (int p = 0; p < 2; ++p)
    #pragma unroll
    for (int r = 0; r < NR; ++r) a[p][r] = mk(0, 0);
  #pragma unroll
  for (int q = 0; q < E; ++q) {
    const int p = q & 1;
    #pragma unroll
    for (int r = 0; r < NR; ++r) a[p][r] = P[r][q] - lz[q] * a[p][r];
    b[p] = b[p] * (-lz[q]);
  }
  seg_scan<LPM, 2, NR, true>(ln, a, b);
  cplx y[2][NR];
  seg_carry<LPM, 2, NR, true>(ln, a, y);
  #pragma unroll
  for (int q = 0; q < E; ++q) {
    const int p = q & 1;
    #pragma unroll
    for (int r = 0; r < NR; ++r) {
      y[p][r] = P[r][q] - lz[q] * y[p][r];
      P[r][q] = y[p][r];
    }
  }
}

// KBPENTA backward sweep x_r = ib_r g_r - bw_r x_{r+2}, in place
// (LinearSolvers.h:122-141).
template <int LPM, int E, int NR, class L>
DPS_HD void seg_backward(const L& ln, cplx (&P)[NR][E], const double (&ib)[E], const double (&bw)[E])
{
  cplx a[2][NR];
  double b[2] = {1.0, 1.0};
  #pragma unroll
  for (int p = 0; p < 2; ++p)
    #pragma unroll
    for (int r = 0; r < NR; ++r) a[p][r] = mk(0, 0);
  #pragma unroll
  for (int q = E - 1; q >= 0; --q) {
    const int p = q & 1;
    #pragma unroll
    for (int r = 0; r < NR; ++r) a[p][r] = ib[q] * P[r][q] - bw[q] * a[p][r];
    b[p] = b[p] * (-bw[q]);
  }
  seg_scan<LPM, 2, NR, false>(ln, a, b);
  cplx y[2][NR];
  seg_carry<LPM, 2, NR, false>(ln, a, y);
  #pragma unroll
  for (int q = E - 1; q >= 0; --q) {
    const int p = q & 1;
    #pragma unroll
    for (int r = 0; r < NR; ++r) {
      y[p][r] = ib[q] * P[r][q] - bw[q] * y[p][r];
      P[r][q] = y[p][r];
    }
  }
}

// Halo rows for a five-point stencil over X: lo = rows l*E-2, l*E-1 (zero on
// lane 0), hi = rows l*E+E, l*E+E+1 (on the last lane: the rows beyond the
// group, which are Y0/Y1 if they are Nz/Nz+1 and zero otherwise).
template <int LPM, int E, class L>
DPS_HD void seg_halo(const L& ln, const cplx (&X)[E], int Nz, cplx Y0, cplx Y1, cplx (&lo)[2], cplx (&hi)[2])
{
  #pragma unroll
  for (int i = 0; i < 2; ++i) {
    const cplx u = mk(ln.up(X[E - 2 + i].re, 1), ln.up(X[E - 2 + i].im, 1));
    const cplx v = mk(ln.down(X[i].re, 1), ln.down(X[i].im, 1));
    lo[i] = (ln.l == 0) ? mk(0, 0) : u;
    if (ln.l == LPM - 1) {
      const int row = LPM * E + i;
      hi[i] = (row == Nz) ? Y0 : ((row == Nz + 1) ? Y1 : mk(0, 0));
    } else {
      hi[i] = v;
    }
  }
}

// S-band product of DPTools.cpp:266-284 at one row (same association order as
// the sequential solver).
DPS_HD cplx s_band(const BvpConsts& B, int row, cplx xm2, cplx x0, cplx x2, cplx Y0, cplx Y1)
{
  cplx v = B.sd0[row] * x0;
  if (row >= 2) v = B.sm2[row] * xm2 + v;
  v = v + B.sp2[row] * x2;
  if (row == 0) v = v + B.S0N * Y0;
  if (row == 1) v = v + B.S1N1 * Y1;
  return v;
}

// Rows of the Schur solution x - A^{-1}B Y, extended by Y0, Y1 at rows Nz, Nz+1.
template <int E>
DPS_HD cplx seg_xrow(int row, int Nz, cplx x, double ab, cplx Y0, cplx Y1)
{
  if (row < Nz) return x - ((row & 1) ? ab * Y1 : ab * Y0);
  if (row == Nz) return Y0;
  if (row == Nz + 1) return Y1;
  return mk(0, 0);
}

// One mode k != 0 (counterpart of bvp_mode).  ib/ab: this lane's rows of the
// static factorisation (ib = 1/pivot of the row, i.e. IB(row+1)); Gi: inverse
// Schur complement.  fin(c, q) returns the forcing coefficient of component c
// at row l*E+q (zero for rows >= Nz); fout(c, q, v) receives the velocity
// coefficients of the rows < Nz; res gets the wall correction (identical on
// every lane of the group).
template <int LPM, int E, class L, class FIn, class FOut>
DPS_HD void seg_bvp_mode(const BvpConsts& B, const L& ln, double kx, double ky, int nyq_x, int nyq_y,
                         const double (&ib)[E], const double (&ab)[E], const double* Gi, const FIn& fin,
                         FOut& fout, ModeResult& res)
{
  const int Nz = B.Nz;
  const double H = B.H;
  const double Ksq = kx * kx + ky * ky;
  const double K = sqrt(Ksq);
  const double H2K = (H * H) * K;
  const cplx Dx = nyq_x ? mk(0, 0) : mk(0, kx);
  const cplx Dy = nyq_y ? mk(0, 0) : mk(0, ky);
  const int r0 = ln.l * E;
#define C0(j) (B.R2[(j)] * H2K + (B.R2[(j)] + (H * B.R1[(j)] - B.R2[(j)])))
#define C1(j) (B.L2[(j)] * H2K + (-B.L2[(j)] + (H * B.L1[(j)] + B.L2[(j)])))

  // sweep multipliers of this lane's rows
  double lz[E], bw[E];
  {
    const double ibm0 = ln.up(ib[E - 2], 1), ibm1 = ln.up(ib[E - 1], 1);
    #pragma unroll
    for (int q = 0; q < E; ++q) {
      const int row = r0 + q;
      const double ibprev = (q >= 2) ? ib[q - 2] : (q == 0 ? ibm0 : ibm1);
      lz[q] = (row >= 2 && row < Nz) ? (0.0 - B.sm2[row] * Ksq) * ibprev : 0.0;
      bw[q] = (row < Nz - 2) ? (0.0 - B.sp2[row] * Ksq) * ib[q] : 0.0;
    }
  }

  // ---- pressure rhs: Dx fx + Dy fy + chebCoeffDiff(fz) (DPTools.cpp:134-195) ----
  cplx X[1][E];
  {
    cplx suf[E];
    cplx s[2][1] = {{mk(0, 0)}, {mk(0, 0)}};
    #pragma unroll
    for (int q = E - 1; q >= 0; --q) {
      const int row = r0 + q;
      const cplx w = (row < Nz) ? ((double)(row * 2) / H) * fin(2, q) : mk(0, 0);
      s[q & 1][0] = s[q & 1][0] + w;
      suf[q] = s[q & 1][0];
    }
    double one[2] = {1.0, 1.0};
    seg_scan<LPM, 2, 1, false>(ln, s, one);
    cplx above[2][1];
    seg_carry<LPM, 2, 1, false>(ln, s, above);
    #pragma unroll
    for (int q = 0; q < E; ++q) {
      const int row = r0 + q;
      cplx df = (q + 1 < E) ? suf[q + 1] + above[(q + 1) & 1][0] : above[0][0];
      if (row == 0) df = df / 2;
      X[0][q] = (row < Nz) ? Dx * fin(0, q) + Dy * fin(1, q) + df : mk(0, 0);
    }
  }
  seg_forward<LPM, E, 1>(ln, X, lz);
  seg_backward<LPM, E, 1>(ln, X, ib, bw);
  cplx Y0, Y1;
  {
    double y[4] = {0, 0, 0, 0};
    #pragma unroll
    for (int q = 0; q < E; ++q) {
      const int row = r0 + q;
      if (row < Nz) {
        const double c0 = C0(row), c1 = C1(row);
        y[0] += X[0][q].re * c0; y[1] += X[0][q].im * c0;
        y[2] += X[0][q].re * c1; y[3] += X[0][q].im * c1;
      }
    }
    seg_sum<LPM>(ln, y);
    const cplx y0 = mk(y[0], y[1]), y1 = mk(y[2], y[3]);
    Y0 = Gi[0] * y0 + Gi[1] * y1;
    Y1 = Gi[2] * y0 + Gi[3] * y1;
  }
  // x = xp - A^{-1}B Y ; Cp = SIMat x ; Dp = FIMat x  (DPTools.cpp:266-284)
  #pragma unroll
  for (int q = 0; q < E; ++q) X[0][q] = seg_xrow<E>(r0 + q, Nz, X[0][q], ab[q], Y0, Y1);
  cplx Cp[E], Dp[E];
  cplx psum, palt;
  {
    cplx lo[2], hi[2];
    seg_halo<LPM, E>(ln, X[0], Nz, Y0, Y1, lo, hi);
    double red[4] = {0, 0, 0, 0};
    #pragma unroll
    for (int q = 0; q < E; ++q) {
      const int row = r0 + q;
      if (row < Nz) {
        const cplx xm2 = (q >= 2) ? X[0][q - 2] : lo[q];
        const cplx xm1 = (q >= 1) ? X[0][q - 1] : lo[1];
        const cplx x1 = (q + 1 < E) ? X[0][q + 1] : hi[q + 1 - E];
        const cplx x2 = (q + 2 < E) ? X[0][q + 2] : hi[q + 2 - E];
        Cp[q] = s_band(B, row, xm2, X[0][q], x2, Y0, Y1);
        cplx dp = B.fp1[row] * x1;
        if (row >= 1) dp = B.fm1[row] * xm1 + dp;
        if (row == 0) dp = dp + B.F0N1 * Y1;
        Dp[q] = dp;
        red[0] += Cp[q].re; red[1] += Cp[q].im;
        const double sg = (row & 1) ? -1.0 : 1.0;
        red[2] += sg * Cp[q].re; red[3] += sg * Cp[q].im;
      } else {
        Cp[q] = mk(0, 0);
        Dp[q] = mk(0, 0);
      }
    }
    seg_sum<LPM>(ln, red);
    psum = mk(red[0], red[1]);
    palt = mk(red[2], red[3]);
  }
  // BC rows for the velocity solves (DPTools.cpp:218-234)
  const double fac1 = 2 * B.eta;
  const double facK = fac1 * K;
  cplx bc[6];
  bc[0] = ((-1.0) * Dx) * psum / facK;
  bc[1] = Dx * palt / facK;
  bc[2] = ((-1.0) * Dy) * psum / facK;
  bc[3] = Dy * palt / facK;
  bc[4] = psum / fac1;
  bc[5] = palt / fac1;

  // ---- velocity: three Schur solves (DPTools.cpp:197-216, 287-345) ----
  cplx U[3][E];
  #pragma unroll
  for (int q = 0; q < E; ++q) {
    const int row = r0 + q;
    if (row < Nz) {
      U[0][q] = (Dx * Cp[q] - fin(0, q)) / B.eta;
      U[1][q] = (Dy * Cp[q] - fin(1, q)) / B.eta;
      U[2][q] = (Dp[q] - fin(2, q)) / B.eta;
    } else {
      U[0][q] = U[1][q] = U[2][q] = mk(0, 0);
    }
  }
  seg_forward<LPM, E, 3>(ln, U, lz);
  seg_backward<LPM, E, 3>(ln, U, ib, bw);
  cplx Ya[3], Yb[3];
  {
    double y[12];
    #pragma unroll
    for (int i = 0; i < 12; ++i) y[i] = 0;
    #pragma unroll
    for (int q = 0; q < E; ++q) {
      const int row = r0 + q;
      if (row < Nz) {
        const double c0 = C0(row), c1 = C1(row);
        #pragma unroll
        for (int c = 0; c < 3; ++c) {
          y[4 * c + 0] += U[c][q].re * c0; y[4 * c + 1] += U[c][q].im * c0;
          y[4 * c + 2] += U[c][q].re * c1; y[4 * c + 3] += U[c][q].im * c1;
        }
      }
    }
    seg_sum<LPM>(ln, y);
    #pragma unroll
    for (int c = 0; c < 3; ++c) {
      const cplx a = mk(y[4 * c], y[4 * c + 1]) - bc[2 * c];
      const cplx b2 = mk(y[4 * c + 2], y[4 * c + 3]) - bc[2 * c + 1];
      Ya[c] = Gi[0] * a + Gi[1] * b2;
      Yb[c] = Gi[2] * a + Gi[3] * b2;
    }
  }
  cplx wall_bot[3], wall_top[3];
  {
    double red[12];
    #pragma unroll
    for (int i = 0; i < 12; ++i) red[i] = 0;
    #pragma unroll
    for (int c = 0; c < 3; ++c) {
      #pragma unroll
      for (int q = 0; q < E; ++q) U[c][q] = seg_xrow<E>(r0 + q, Nz, U[c][q], ab[q], Ya[c], Yb[c]);
      cplx lo[2], hi[2];
      seg_halo<LPM, E>(ln, U[c], Nz, Ya[c], Yb[c], lo, hi);
      cplx cu[E];
      #pragma unroll
      for (int q = 0; q < E; ++q) {
        const int row = r0 + q;
        if (row < Nz) {
          const cplx xm2 = (q >= 2) ? U[c][q - 2] : lo[q];
          const cplx x2 = (q + 2 < E) ? U[c][q + 2] : hi[q + 2 - E];
          cu[q] = s_band(B, row, xm2, U[c][q], x2, Ya[c], Yb[c]);
          red[4 * c + 0] += cu[q].re; red[4 * c + 1] += cu[q].im;
          const double sg = (row & 1) ? -1.0 : 1.0;
          red[4 * c + 2] += sg * cu[q].re; red[4 * c + 3] += sg * cu[q].im;
        }
      }
      #pragma unroll
      for (int q = 0; q < E; ++q)
        if (r0 + q < Nz) fout(c, q, cu[q]);
    }
    seg_sum<LPM>(ln, red);
    #pragma unroll
    for (int c = 0; c < 3; ++c) {
      wall_top[c] = mk(red[4 * c], red[4 * c + 1]);
      wall_bot[c] = mk(red[4 * c + 2], red[4 * c + 3]);
    }
  }
#undef C0
#undef C1
  wall_correction(B, kx, ky, wall_bot, wall_top, res);
}

// k = 0 mode (counterpart of bvp_k0).
template <int LPM, int E, class L, class FIn, class FOut>
DPS_HD void seg_bvp_k0(const BvpConsts& B, const L& ln, const FIn& fin, FOut& fout)
{
  const int Nz = B.Nz;
  const int r0 = ln.l * E;
  if (B.geometry != MODE_BOTTOM && B.geometry != MODE_SLIT) {
    #pragma unroll
    for (int c = 0; c < 3; ++c)
      #pragma unroll
      for (int q = 0; q < E; ++q)
        if (r0 + q < Nz) fout(c, q, mk(0, 0));
    return;
  }
  const double* RA = (B.geometry == MODE_BOTTOM) ? B.R1 : B.R2;
  const double* RB = B.L2;
  const double D00 = -RA[Nz], D01 = -RA[Nz + 1], D10 = -RB[Nz], D11 = -RB[Nz + 1];
  const double idet = 1.0 / (D00 * D11 - D01 * D10);
  const double g00 = idet * D11, g01 = idet * (-D01), g10 = idet * (-D10), g11 = idet * D00;
  #pragma unroll
  for (int c = 0; c < 2; ++c) {
    cplx X[E];
    double red[4] = {0, 0, 0, 0};
    #pragma unroll
    for (int q = 0; q < E; ++q) {
      const int row = r0 + q;
      if (row < Nz) {
        cplx f = fin(c, q);
        if (B.geometry == MODE_SLIT) f = mk(f.re, f.re);   // Solvers.py:564-566
        X[q] = (-1.0) * f / B.eta;
        red[0] += RA[row] * X[q].re; red[1] += RA[row] * X[q].im;
        red[2] += RB[row] * X[q].re; red[3] += RB[row] * X[q].im;
      } else {
        X[q] = mk(0, 0);
      }
    }
    seg_sum<LPM>(ln, red);
    const cplx ca = mk(red[0], red[1]), cb = mk(red[2], red[3]);
    const cplx a0 = g00 * ca + g01 * cb;
    const cplx a1 = g10 * ca + g11 * cb;
    #pragma unroll
    for (int q = 0; q < E; ++q) {
      const int row = r0 + q;
      if (row == Nz) X[q] = a0;
      else if (row == Nz + 1) X[q] = a1;
    }
    cplx lo[2], hi[2];
    seg_halo<LPM, E>(ln, X, Nz, a0, a1, lo, hi);
    cplx v[E];
    #pragma unroll
    for (int q = 0; q < E; ++q) {
      const int row = r0 + q;
      if (row < Nz) {
        const cplx xm2 = (q >= 2) ? X[q - 2] : lo[q];
        const cplx x2 = (q + 2 < E) ? X[q + 2] : hi[q + 2 - E];
        v[q] = s_band(B, row, xm2, X[q], x2, a0, a1);
      }
    }
    #pragma unroll
    for (int q = 0; q < E; ++q)
      if (r0 + q < Nz) fout(c, q, v[q]);
  }
  #pragma unroll
  for (int q = 0; q < E; ++q)
    if (r0 + q < Nz) fout(2, q, mk(0, 0));
}

}  // namespace dps
