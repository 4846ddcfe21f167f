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
//   int l;  template <int D> double up(v, fill), down(v, fill), xr(v);
// up/down: the value of lane l-D / l+D of the group, or `fill` where that
// lane is outside the group; xr: the value of lane l's partner in butterfly
// stage D (an involution that pairs the two halves of each aligned 2D-lane
// block, e.g. l ^ D).  Fills are chosen so that every lane runs the same
// instructions (identity maps in scans, zero halos): no lane-dependent
// branches.  The gfx950 kernel binds the policy to DPP row operations (16-lane
// groups, the fill entering as the DPP `old` operand) or lane permutes; the
// host-side checks bind it to threads joined by a barrier, so the identical
// code runs on both.
//
// Rows >= Nz are padding: their inputs are zero, their factors are zero, and
// the rows Nz and Nz+1 carry the two Schur unknowns where a stencil reads
// them, as the sequential solver's XP(Nz) = Y0, XP(Nz+1) = Y1.
//
// Row tables.  seg_bvp_mode reads the row tables of B (sd0, sm2, sp2, fm1,
// fp1, R1, R2, L1, L2) at every row in [0, LPM*E) without bounds tests, so
// they must be readable there and zero at rows >= Nz (build_cheb_tables
// already has sm2 = 0 at rows 0, 1 and fm1 = 0 at row 0).  fin(c, q) returns
// zero for rows >= Nz; fout(c, q, v) is called for every row and must drop or
// absorb the rows >= Nz.  seg_bvp_k0 takes the unpadded tables.
#pragma once
#include "dps_core.h"

namespace dps {

// Compiler-only memory fence between solver stages: row tables are re-read
// from cache after it instead of being held in registers for the whole solve.
#define DPS_STAGE_FENCE() __asm__ __volatile__("" ::: "memory")

// Sum over the group: butterfly stages D = 1, 2, 4, ... pairing each lane with
// a partner in the other half of its aligned 2D-lane block, so every lane ends
// with the same total.
template <int LPM, int N, int D = 1, class L>
DPS_HD void seg_sum(const L& ln, double (&v)[N])
{
  if constexpr (D < LPM) {
    #pragma unroll
    for (int i = 0; i < N; ++i) v[i] += ln.template xr<D>(v[i]);
    seg_sum<LPM, N, 2 * D>(ln, v);
  }
}

template <bool UP, int D, class L>
DPS_HD double shift(const L& ln, double v, double fill)
{
  if constexpr (UP) return ln.template up<D>(v, fill);
  else return ln.template down<D>(v, fill);
}

// Inclusive scan over the group of the affine maps y -> a + b y held by each
// lane (UP: lane 0 applies first; !UP: lane LPM-1 applies first).  NC chains
// with their own multipliers, NR right-hand sides per chain.  A lane with no
// partner at stage D receives the identity map (a = 0, b = 1), which leaves
// its own map unchanged.
template <int LPM, int NC, int NR, bool UP, int D = 1, class L>
DPS_HD void seg_scan(const L& ln, cplx (&a)[NC][NR], double (&b)[NC])
{
  if constexpr (D < LPM) {
    double bs[NC];
    cplx as[NC][NR];
    #pragma unroll
    for (int c = 0; c < NC; ++c) {
      bs[c] = shift<UP, D>(ln, b[c], 1.0);
      #pragma unroll
      for (int r = 0; r < NR; ++r)
        as[c][r] = mk(shift<UP, D>(ln, a[c][r].re, 0.0), shift<UP, D>(ln, a[c][r].im, 0.0));
    }
    #pragma unroll
    for (int c = 0; c < NC; ++c) {
      #pragma unroll
      for (int r = 0; r < NR; ++r) a[c][r] = a[c][r] + b[c] * as[c][r];
      b[c] = b[c] * bs[c];
    }
    seg_scan<LPM, NC, NR, UP, 2 * D>(ln, a, b);
  }
}

// Value entering this lane: the inclusive value of the preceding lane (zero
// for the first lane in scan order).
template <int LPM, int NC, int NR, bool UP, class L>
DPS_HD void seg_carry(const L& ln, const cplx (&a)[NC][NR], cplx (&cin)[NC][NR])
{
  #pragma unroll
  for (int c = 0; c < NC; ++c)
    #pragma unroll
    for (int r = 0; r < NR; ++r)
      cin[c][r] = mk(shift<UP, 1>(ln, a[c][r].re, 0.0), shift<UP, 1>(ln, a[c][r].im, 0.0));
}

// Sweep multipliers of one lane's rows, formed on demand from the reciprocal
// pivots (so they need no registers of their own):
//   lz(q) = DM2(row-2) / pivot(row-2)   (forward elimination)
//   bw(q) = DP2(row) / pivot(row)       (back substitution)
// sm2 is zero at rows 0, 1 and >= Nz, so lz needs no row test; bw is zero
// from row Nz-2 on, where sp2 still feeds the S band.
template <int E>
struct SegMul {
  const BvpConsts* B;
  double Ksq;
  int Nz, r0;
  const double (*ib)[E];
  double ibm0, ibm1;   // reciprocal pivots of rows r0-2, r0-1 (lane below; 0 on lane 0)
  DPS_HD double lz(int q) const
  {
    const int row = r0 + q;
    const double ibprev = (q >= 2) ? (*ib)[q - 2] : (q == 0 ? ibm0 : ibm1);
    return (0.0 - B->sm2[row] * Ksq) * ibprev;
  }
  DPS_HD double bw(int q) const
  {
    const int row = r0 + q;
    const double s = B->sp2[row];
    return (0.0 - (row < Nz - 2 ? s : 0.0) * Ksq) * (*ib)[q];
  }
};

// KBPENTA forward sweep g_r = P_r - lz_r g_{r-2}, in place (LinearSolvers.h:106-120).
template <int LPM, int E, int NR, class L>
DPS_HD void seg_forward(const L& ln, cplx (&P)[NR][E], const SegMul<E>& mu)
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
    const double lz = mu.lz(q);
    #pragma unroll
    for (int r = 0; r < NR; ++r) a[p][r] = P[r][q] - lz * a[p][r];
    b[p] = b[p] * (-lz);
  }
  seg_scan<LPM, 2, NR, true>(ln, a, b);
  cplx y[2][NR];
  seg_carry<LPM, 2, NR, true>(ln, a, y);
  #pragma unroll
  for (int q = 0; q < E; ++q) {
    const int p = q & 1;
    const double lz = mu.lz(q);
    #pragma unroll
    for (int r = 0; r < NR; ++r) {
      y[p][r] = P[r][q] - lz * y[p][r];
      P[r][q] = y[p][r];
    }
  }
}

// KBPENTA backward sweep x_r = ib_r g_r - bw_r x_{r+2}, in place
// (LinearSolvers.h:122-141).
template <int LPM, int E, int NR, class L>
DPS_HD void seg_backward(const L& ln, cplx (&P)[NR][E], const SegMul<E>& mu)
{
  const double (&ib)[E] = *mu.ib;
  cplx a[2][NR];
  double b[2] = {1.0, 1.0};
  #pragma unroll
  for (int p = 0; p < 2; ++p)
    #pragma unroll
    for (int r = 0; r < NR; ++r) a[p][r] = mk(0, 0);
  #pragma unroll
  for (int q = E - 1; q >= 0; --q) {
    const int p = q & 1;
    const double bw = mu.bw(q);
    #pragma unroll
    for (int r = 0; r < NR; ++r) a[p][r] = ib[q] * P[r][q] - bw * a[p][r];
    b[p] = b[p] * (-bw);
  }
  seg_scan<LPM, 2, NR, false>(ln, a, b);
  cplx y[2][NR];
  seg_carry<LPM, 2, NR, false>(ln, a, y);
  #pragma unroll
  for (int q = E - 1; q >= 0; --q) {
    const int p = q & 1;
    const double bw = mu.bw(q);
    #pragma unroll
    for (int r = 0; r < NR; ++r) {
      y[p][r] = ib[q] * P[r][q] - bw * y[p][r];
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
    const int row = LPM * E + i;   // beyond the last lane
    const cplx past = (row == Nz) ? Y0 : ((row == Nz + 1) ? Y1 : mk(0, 0));
    lo[i] = mk(ln.template up<1>(X[E - 2 + i].re, 0.0), ln.template up<1>(X[E - 2 + i].im, 0.0));
    hi[i] = mk(ln.template down<1>(X[i].re, past.re), ln.template down<1>(X[i].im, past.im));
  }
}

// Velocity component solved at position i: one at a time the z component goes
// first, so the pressure derivative Dp is dead before the x and y solves.
template <int NRV>
DPS_HD int vel_comp(int i) { return NRV == 1 ? (i + 2) % 3 : i; }

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

// S band of a lane's row q (row = l*E + q): the row-0 and row-1 couplings to
// the Schur unknowns belong to lane 0, q = 0 and 1.  Same association order
// as s_band; the table terms vanish on padding rows.
DPS_HD cplx seg_sband(const BvpConsts& B, int row, int q, bool lane0, cplx xm2, cplx x0, cplx x2, cplx Y0,
                      cplx Y1)
{
  cplx v = B.sd0[row] * x0;
  v = B.sm2[row] * xm2 + v;   // sm2 = 0 at rows 0, 1
  v = v + B.sp2[row] * x2;
  if (q == 0) v = v + (lane0 ? B.S0N : 0.0) * Y0;
  if (q == 1) v = v + (lane0 ? B.S1N1 : 0.0) * Y1;
  return v;
}

// Rows of the Schur solution x - A^{-1}B Y, extended by Y0, Y1 at rows Nz, Nz+1.
// (x and ab are zero at rows >= Nz, and E is even, so row parity is q parity.)
DPS_HD cplx seg_xrow(int row, int Nz, cplx x, double ab, cplx Y, cplx Y0, cplx Y1)
{
  const cplx v = x - ab * Y;
  return csel(row == Nz, Y0, csel(row == Nz + 1, Y1, v));
}

// One mode k != 0 (counterpart of bvp_mode); NRV (1 or 3) velocity components
// share each scan.  ib/ab: this lane's rows of the
// static factorisation (ib = 1/pivot of the row, i.e. IB(row+1)); Gi: inverse
// Schur complement.  fin(c, q) returns the forcing coefficient of component c
// at row l*E+q (zero for rows >= Nz); fout(c, q, v) receives the velocity
// coefficients; res gets the wall correction (identical on every lane of the
// group).
// avg (modes on a Nyquist index): the corrections are the mean of those at
// (kx, ky) and at (kxa, kya) = minus the mirror mode's grid wavenumbers, from
// the same wall values -- the Hermitian average of the mode and its mirror,
// whose BVP part is the conjugate of the mode's own.
// STAGE: the x and y velocity right-hand sides (D_c p - f_c) / eta are formed
// as soon as the pressure is known and written through fout over the x and y
// forcing, then read back through fin by their solves, so the pressure
// coefficients hold no registers during the velocity solves.  fin and fout
// must then address the same storage, and fout must store what it is given.
// RREG: the wall correction is formed by wall_correction_regs (res held in
// registers) rather than wall_correction; the results are identical.
template <int LPM, int E, int NRV = 3, bool STAGE = false, bool RREG = false, class L, class FIn, class FOut>
DPS_HD void seg_bvp_mode(const BvpConsts& B, const L& ln, double kx, double ky, int nyq_x, int nyq_y,
                         const double (&ib)[E], const double (&ab)[E], const double* Gi, const FIn& fin,
                         FOut& fout, ModeResult& res, bool avg = false, double kxa = 0, double kya = 0)
{
  const int Nz = B.Nz;
  const double H = B.H;
  const double Ksq = ksq_of(kx, ky);
  const double K = sqrt(Ksq);
  const double H2K = (H * H) * K;
  const cplx Dx = nyq_x ? mk(0, 0) : mk(0, kx);
  const cplx Dy = nyq_y ? mk(0, 0) : mk(0, ky);
  const int r0 = ln.l * E;
  const double iH = 1.0 / H;
#define C0(j) (B.R2[(j)] * H2K + (B.R2[(j)] + (H * B.R1[(j)] - B.R2[(j)])))
#define C1(j) (B.L2[(j)] * H2K + (-B.L2[(j)] + (H * B.L1[(j)] + B.L2[(j)])))

  // sweep multipliers of this lane's rows
  SegMul<E> mu;
  mu.B = &B;
  mu.Ksq = Ksq;
  mu.Nz = Nz;
  mu.r0 = r0;
  mu.ib = &ib;
  mu.ibm0 = ln.template up<1>(ib[E - 2], 0.0);
  mu.ibm1 = ln.template up<1>(ib[E - 1], 0.0);
  const bool lane0 = (ln.l == 0);

  // ---- pressure rhs: Dx fx + Dy fy + chebCoeffDiff(fz) (DPTools.cpp:134-195) ----
  cplx X[1][E];
  {
    cplx suf[E];
    cplx s[2][1] = {{mk(0, 0)}, {mk(0, 0)}};
    #pragma unroll
    for (int q = E - 1; q >= 0; --q) {
      const int row = r0 + q;
      const cplx w = ((double)(row * 2) * iH) * fin(2, q);
      s[q & 1][0] = s[q & 1][0] + w;
      suf[q] = s[q & 1][0];
    }
    double one[2] = {1.0, 1.0};
    seg_scan<LPM, 2, 1, false>(ln, s, one);
    cplx above[2][1];
    seg_carry<LPM, 2, 1, false>(ln, s, above);
    #pragma unroll
    for (int q = 0; q < E; ++q) {
      cplx df = (q + 1 < E) ? suf[q + 1] + above[(q + 1) & 1][0] : above[0][0];
      if (q == 0) df = df * (lane0 ? 0.5 : 1.0);   // row 0
      X[0][q] = Dx * fin(0, q) + Dy * fin(1, q) + df;
    }
  }
  DPS_STAGE_FENCE();
  seg_forward<LPM, E, 1>(ln, X, mu);
  DPS_STAGE_FENCE();
  seg_backward<LPM, E, 1>(ln, X, mu);
  DPS_STAGE_FENCE();
  cplx Y0, Y1;
  {
    double y[4] = {0, 0, 0, 0};
    #pragma unroll
    for (int q = 0; q < E; ++q) {
      const int row = r0 + q;
      const double c0 = C0(row), c1 = C1(row);
      y[0] += X[0][q].re * c0; y[1] += X[0][q].im * c0;
      y[2] += X[0][q].re * c1; y[3] += X[0][q].im * c1;
    }
    seg_sum<LPM>(ln, y);
    const cplx y0 = mk(y[0], y[1]), y1 = mk(y[2], y[3]);
    Y0 = Gi[0] * y0 + Gi[1] * y1;
    Y1 = Gi[2] * y0 + Gi[3] * y1;
  }
  // x = xp - A^{-1}B Y ; Cp = SIMat x ; Dp = FIMat x  (DPTools.cpp:266-284)
  #pragma unroll
  for (int q = 0; q < E; ++q) X[0][q] = seg_xrow(r0 + q, Nz, X[0][q], ab[q], (q & 1) ? Y1 : Y0, Y0, Y1);
  cplx Cp[E], Dp[E];
  cplx psum, palt;
  {
    cplx lo[2], hi[2];
    seg_halo<LPM, E>(ln, X[0], Nz, Y0, Y1, lo, hi);
    double red[4] = {0, 0, 0, 0};
    #pragma unroll
    for (int q = 0; q < E; ++q) {
      const int row = r0 + q;
      const cplx xm2 = (q >= 2) ? X[0][q - 2] : lo[q];
      const cplx xm1 = (q >= 1) ? X[0][q - 1] : lo[1];
      const cplx x1 = (q + 1 < E) ? X[0][q + 1] : hi[q + 1 - E];
      const cplx x2 = (q + 2 < E) ? X[0][q + 2] : hi[q + 2 - E];
      Cp[q] = seg_sband(B, row, q, lane0, xm2, X[0][q], x2, Y0, Y1);
      cplx dp = B.fp1[row] * x1;
      dp = B.fm1[row] * xm1 + dp;   // fm1 = 0 at row 0
      if (q == 0) dp = dp + (lane0 ? B.F0N1 : 0.0) * Y1;
      Dp[q] = dp;
      red[0] += Cp[q].re; red[1] += Cp[q].im;
      const double sg = (q & 1) ? -1.0 : 1.0;
      red[2] += sg * Cp[q].re; red[3] += sg * Cp[q].im;
    }
    seg_sum<LPM>(ln, red);
    psum = mk(red[0], red[1]);
    palt = mk(red[2], red[3]);
  }
  const double ieta = 1.0 / B.eta;
  if constexpr (STAGE) {
    DPS_STAGE_FENCE();
    #pragma unroll
    for (int q = 0; q < E; ++q) {
      fout(0, q, (Dx * Cp[q] - fin(0, q)) * ieta);
      fout(1, q, (Dy * Cp[q] - fin(1, q)) * ieta);
    }
  }
  // BC rows for the velocity solves (DPTools.cpp:218-234), formed where used:
  //   bc[2c] = -D_c psum / (2 eta K), bc[2c+1] = D_c palt / (2 eta K)  (c = 0, 1)
  //   bc[4] = psum / (2 eta),         bc[5] = palt / (2 eta)
  const double ifac1 = 1.0 / (2 * B.eta);
  const double ifacK = 1.0 / ((2 * B.eta) * K);

  // ---- velocity: three Schur solves (DPTools.cpp:197-216, 287-345), in
  //      batches of NRV components sharing each scan ----
  cplx wall_bot[3] = {mk(0, 0), mk(0, 0), mk(0, 0)}, wall_top[3] = {mk(0, 0), mk(0, 0), mk(0, 0)};
  #pragma unroll
  for (int cb = 0; cb < 3; cb += NRV) {
    DPS_STAGE_FENCE();
    cplx U[NRV][E];
    #pragma unroll
    for (int q = 0; q < E; ++q) {
      #pragma unroll
      for (int b = 0; b < NRV; ++b) {
        const int c = vel_comp<NRV>(cb + b);
        if (STAGE && c != 2) {
          U[b][q] = fin(c, q);
        } else {
          const cplx g = (c == 0) ? Dx * Cp[q] : ((c == 1) ? Dy * Cp[q] : Dp[q]);
          U[b][q] = (g - fin(c, q)) * ieta;
        }
      }
    }
    DPS_STAGE_FENCE();
    seg_forward<LPM, E, NRV>(ln, U, mu);
    DPS_STAGE_FENCE();
    seg_backward<LPM, E, NRV>(ln, U, mu);
    DPS_STAGE_FENCE();
    cplx Ya[NRV], Yb[NRV];
    {
      double y[4 * NRV];
      #pragma unroll
      for (int i = 0; i < 4 * NRV; ++i) y[i] = 0;
      #pragma unroll
      for (int q = 0; q < E; ++q) {
        const int row = r0 + q;
        const double c0 = C0(row), c1 = C1(row);
        #pragma unroll
        for (int b = 0; b < NRV; ++b) {
          y[4 * b + 0] += U[b][q].re * c0; y[4 * b + 1] += U[b][q].im * c0;
          y[4 * b + 2] += U[b][q].re * c1; y[4 * b + 3] += U[b][q].im * c1;
        }
      }
      seg_sum<LPM>(ln, y);
      #pragma unroll
      for (int b = 0; b < NRV; ++b) {
        const int c = vel_comp<NRV>(cb + b);
        cplx bc0, bc1;
        if (c == 2) {
          bc0 = psum * ifac1;
          bc1 = palt * ifac1;
        } else {
          const cplx Dc = (c == 0) ? Dx : Dy;
          bc0 = ((-1.0) * Dc) * psum * ifacK;
          bc1 = Dc * palt * ifacK;
        }
        const cplx a = mk(y[4 * b], y[4 * b + 1]) - bc0;
        const cplx b2 = mk(y[4 * b + 2], y[4 * b + 3]) - bc1;
        Ya[b] = Gi[0] * a + Gi[1] * b2;
        Yb[b] = Gi[2] * a + Gi[3] * b2;
      }
    }
    double red[4 * NRV];
    #pragma unroll
    for (int i = 0; i < 4 * NRV; ++i) red[i] = 0;
    #pragma unroll
    for (int b = 0; b < NRV; ++b) {
      const int c = vel_comp<NRV>(cb + b);
      #pragma unroll
      for (int q = 0; q < E; ++q) U[b][q] = seg_xrow(r0 + q, Nz, U[b][q], ab[q], (q & 1) ? Yb[b] : Ya[b], Ya[b], Yb[b]);
      cplx lo[2], hi[2];
      seg_halo<LPM, E>(ln, U[b], Nz, Ya[b], Yb[b], lo, hi);
      cplx cu[E];
      #pragma unroll
      for (int q = 0; q < E; ++q) {
        const int row = r0 + q;
        const cplx xm2 = (q >= 2) ? U[b][q - 2] : lo[q];
        const cplx x2 = (q + 2 < E) ? U[b][q + 2] : hi[q + 2 - E];
        cu[q] = seg_sband(B, row, q, lane0, xm2, U[b][q], x2, Ya[b], Yb[b]);
        red[4 * b + 0] += cu[q].re; red[4 * b + 1] += cu[q].im;
        const double sg = (q & 1) ? -1.0 : 1.0;
        red[4 * b + 2] += sg * cu[q].re; red[4 * b + 3] += sg * cu[q].im;
      }
      #pragma unroll
      for (int q = 0; q < E; ++q) fout(c, q, cu[q]);
    }
    seg_sum<LPM>(ln, red);
    // stored through selects on constant indices (a runtime index would put
    // the two arrays in private memory)
    #pragma unroll
    for (int b = 0; b < NRV; ++b) {
      const int c = vel_comp<NRV>(cb + b);
      const cplx t = mk(red[4 * b], red[4 * b + 1]), bo = mk(red[4 * b + 2], red[4 * b + 3]);
      #pragma unroll
      for (int i = 0; i < 3; ++i) {
        wall_top[i] = csel(c == i, t, wall_top[i]);
        wall_bot[i] = csel(c == i, bo, wall_bot[i]);
      }
    }
  }
#undef C0
#undef C1
  if constexpr (RREG) wall_correction_regs(B, kx, ky, wall_bot, wall_top, res);
  else wall_correction(B, kx, ky, wall_bot, wall_top, res);
  if (avg) {
    ModeResult r2;
    if constexpr (RREG) wall_correction_regs(B, kxa, kya, wall_bot, wall_top, r2);
    else wall_correction(B, kxa, kya, wall_bot, wall_top, r2);
    #pragma unroll
    for (int c = 0; c < 3; ++c)
      #pragma unroll
      for (int b = 0; b < 4; ++b)
        res.corr[c][b] = mk(0.5 * (res.corr[c][b].re + r2.corr[c][b].re), 0.5 * (res.corr[c][b].im + r2.corr[c][b].im));
  }
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
  const double ieta = 1.0 / B.eta;
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
        X[q] = f * (-ieta);
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
