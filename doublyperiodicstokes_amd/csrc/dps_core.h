// dps_core.h -- arithmetic shared by the gfx950 kernels and the host-side
// checks: ES kernels, per-particle stencils, and the per-wavenumber FP64
// boundary-value solve of the doubly periodic Stokes problem.
//
// Everything here is plain C++ marked __host__ __device__ under hipcc, so the
// exact code that runs in the kernels can also be executed on the CPU.
#pragma once
#include <math.h>
#include <stdint.h>

// Everything here is small and called from hot loops: always inlined, so that
// arrays passed by reference stay in registers.
#if defined(__HIPCC__)
#define DPS_HD __host__ __device__ inline __attribute__((always_inline))
#else
#define DPS_HD inline __attribute__((always_inline))
#endif

namespace dps {

// |k|^2 of a planar wavenumber, one rounding (fma), identical bit for bit
// under a swap of kx and ky and under sign changes: modes that share |k| on a
// square box share one factorisation of their BVP.
DPS_HD double ksq_of(double kx, double ky)
{
  const double a = fabs(kx), b = fabs(ky);
  const double hi = fmax(a, b), lo = fmin(a, b);
  return fma(hi, hi, lo * lo);
}

// ---------------------------------------------------------------------------
// minimal FP64 complex
// ---------------------------------------------------------------------------
struct cplx {
  double re, im;
};
DPS_HD cplx mk(double r, double i) { cplx c; c.re = r; c.im = i; return c; }
DPS_HD cplx operator+(cplx a, cplx b) { return mk(a.re + b.re, a.im + b.im); }
DPS_HD cplx operator-(cplx a, cplx b) { return mk(a.re - b.re, a.im - b.im); }
DPS_HD cplx operator-(cplx a) { return mk(-a.re, -a.im); }
DPS_HD cplx operator*(cplx a, cplx b) { return mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
DPS_HD cplx operator*(double s, cplx a) { return mk(s * a.re, s * a.im); }
DPS_HD cplx operator*(cplx a, double s) { return mk(a.re * s, a.im * s); }
DPS_HD cplx operator/(cplx a, double s) { return mk(a.re / s, a.im / s); }
DPS_HD cplx conjg(cplx a) { return mk(a.re, -a.im); }
DPS_HD cplx times_i(cplx a) { return mk(-a.im, a.re); }   // i * a
// Selects on values.  (A conditional expression on two lvalues is an lvalue:
// the compiler then selects between their addresses, which keeps the
// operands in memory instead of registers.)
DPS_HD double dsel(bool s, double a, double b) { return s ? a : b; }
DPS_HD cplx csel(bool s, cplx a, cplx b) { return mk(s ? a.re : b.re, s ? a.im : b.im); }

// geometry codes (same values as the C ABI)
enum { MODE_NOWALL = 1, MODE_BOTTOM = 2, MODE_SLIT = 3 };

// ---------------------------------------------------------------------------
// ES kernel (reference include/ESKernels.h:8-54)
//   phi(x)  = exp(bw (sqrt(1 - x^2/a^2) - 1))
//   phi'(x) = -bw x phi(x) / (a^2 sqrt(1 - x^2/a^2))
// The 3-D kernel is the product of three 1-D factors.
// ---------------------------------------------------------------------------
DPS_HD double es_phi(double x, double bw, double a2)
{
  return exp(bw * (sqrt(1 - x * x / a2) - 1));
}
DPS_HD double es_dfac(double x, double bw, double a2)
{
  return -bw * x / (a2 * sqrt(1 - x * x / a2));
}

// Kernel description shared by host and device.
struct KernelDesc {
  double betaw;      // beta * w
  double alpha;      // support half width
  double thresh;     // cutoff (alpha for the monopole, < alpha for the dipole)
  double inv_norm3;  // 1 / norm^3
  int wx, wy;        // support points in x and y: round(2 thresh / hx), round(2 thresh / hy)
  int pos_deriv;     // derivative weights as functions of the particle position: zero where
                     // the offset was clamped (the weight is then the constant phi(thresh))
  int nzeff;         // points of the z-extended grid
  int ext_up;        // mirrored ghost nodes above the top
  int wzmax;         // bound on the z support size
  double dmax;       // bound on |phi'| over |x| <= thresh (dipole kernels; 0 otherwise)
};

// first index e with zext[e] <= v for descending zext (std::lower_bound with
// std::greater, as in the reference's ParticleList.cpp:456-459)
DPS_HD int lower_bound_desc(const double* zext, int n, double v)
{
  int lo = 0, hi = n;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (zext[mid] > v) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// Per-particle stencil: support origin and 1-D weights.
// Follows ParticleList::locateOnGridNonUnifZ (ParticleList.cpp:453-524):
// offsets beyond +-thresh are clamped to +thresh (not zeroed), the z support is
// every extended-grid node within [z - thresh, z + thresh].
//   x, y, z : position in [0,Lx] x [0,Ly] x [0,Lz]
//   bx, by  : unwrapped grid index of the first support point
//   e0, nz  : first extended-grid z index and number of z points
// Weights written when the pointer is non-null (stride 1):
//   wx  = phi(x offsets) / norm^3        dx = wx * dfac (derivative)
//   wy  = phi(y offsets)                 dy
//   wz  = phi(z offsets)                 dz
struct Stencil {
  int bx, by, e0, nz;
};

// First support point of a uniform direction: xclose = (int)(x/h), +1 for odd
// w past the half cell, start = xclose - w/2 + even (ParticleList.cpp:337-347)
DPS_HD int support_start(double x, double h, int w)
{
  const int even = -1 * (w % 2) + 1;
  int c = (int)(x / h);
  c += ((w % 2) && (x / h - c > 1.0 / 2.0) ? 1 : 0);
  return c - w / 2 + even;
}

DPS_HD Stencil stencil_support(double x, double y, double z, double hx, double hy,
                               const KernelDesc& k, const double* zext)
{
  Stencil s;
  s.bx = support_start(x, hx, k.wx);
  s.by = support_start(y, hy, k.wy);
  const int high = lower_bound_desc(zext, k.nzeff, z - k.thresh);
  const int low = lower_bound_desc(zext, k.nzeff, z + k.thresh);
  int ir = high;
  if (ir == k.nzeff) ir -= 1;
  else if (z - k.thresh > zext[ir]) ir -= 1;
  s.e0 = low;
  s.nz = ir - low + 1;
  if (s.nz < 0) s.nz = 0;
  return s;
}

DPS_HD double clamp_off(double v, double thresh)
{
  return (v > thresh || v < -thresh) ? thresh : v;
}

// 1-D weights along x (with 1/norm^3 folded in) or y (scale = 1); w points.
DPS_HD void stencil_weights_xy(double pos, double h, int b, int w, const KernelDesc& k,
                               double scale, double* wout, double* dout)
{
  const double a2 = k.alpha * k.alpha;
  for (int i = 0; i < w; ++i) {
    const double v = clamp_off(fma((double)(b + i), h, -pos), k.thresh);
    const double p = es_phi(v, k.betaw, a2) * scale;
    if (wout) wout[i] = p;
    if (dout) dout[i] = (k.pos_deriv && fabs(v) >= k.thresh) ? 0.0 : p * es_dfac(v, k.betaw, a2);
  }
}

// 1-D weights along z on the extended grid; wint (optional) also multiplies
// by the interpolation quadrature weight hx*hy*zw_ext[e] (ParticleList.cpp:521).
DPS_HD void stencil_weights_z(double z, const Stencil& s, const KernelDesc& k, const double* zext,
                              const double* zwext, double hxhy, double* wout, double* dout,
                              double* wint, double* dint)
{
  const double a2 = k.alpha * k.alpha;
  for (int q = 0; q < s.nz; ++q) {
    const int e = s.e0 + q;
    const double v = clamp_off(zext[e] - z, k.thresh);
    const double p = es_phi(v, k.betaw, a2);
    const double d = (k.pos_deriv && fabs(v) >= k.thresh) ? 0.0 : p * es_dfac(v, k.betaw, a2);
    const double qw = hxhy * zwext[e];
    if (wout) wout[q] = p;
    if (dout) dout[q] = d;
    if (wint) wint[q] = p * qw;
    if (dint) dint[q] = d * qw;
  }
}

// ---------------------------------------------------------------------------
// Per-wavenumber boundary-value solve (FP64).
//
// Restates, for one Fourier mode k != 0, the reference chain
//   chebCoeffDiff -> assemble_p_rhs -> schurSolve_p -> assemble_u_rhs ->
//   assemble_u_bc_rhs -> schurSolve_u      (DPTools.cpp:81-345)
// with the per-k operators of Solvers.py:258-347 / LinearSolvers.cpp:11-60
// rebuilt from |k| on the fly instead of stored per mode, followed by the
// wall correction of DPTools.cpp:590-877 expressed as coefficients of the
// analytic profiles e^{-Kz}, e^{K(z-Hs)}, z e^{-Kz}, z e^{K(z-Hs)}.  The
// profiles themselves are added in value space by the inverse z transform,
// which is the same linear map as adding their Chebyshev transforms here.
// ---------------------------------------------------------------------------
struct BvpConsts {
  int Nz;
  int geometry;
  double H;      // Lz / 2
  double Hs;     // slit height (Lz)
  double eta;
  // second integral matrix band (H^2 scaled): S[j][j], S[j][j-2], S[j][j+2]
  const double* sd0;
  const double* sm2;
  const double* sp2;
  double S0N, S1N1;   // S[0][Nz], S[1][Nz+1]
  // first integral matrix band (H scaled): F[j][j-1], F[j][j+1]
  const double* fm1;
  const double* fp1;
  double F0N1;        // F[0][Nz+1]
  // no-wall BC rows, length Nz+2 (Solvers.py:603-643)
  const double *R1, *R2, *L1, *L2;
};

// Mode-local arrays are accessed through this view: element i of array `a`
// lives at a[i * stride] (stride = number of modes, so that neighbouring modes
// are contiguous and one wavefront's accesses coalesce).  On the host the same
// view runs with stride 1 over private buffers.  Every recurrence carries its
// last two values in registers, so loads never wait on the store before them.
struct ScratchView {
  double* ib;     // Nz+1  reciprocal KBPENTA pivots 1/beta[i]
  double* ab;     // Nz    A^{-1}B (column 0 on even rows, column 1 on odd rows)
  cplx* P;        // Nz    pressure rhs / sweep / solution
  cplx* CP;       // Nz    pressure coefficients
  cplx* DP;       // Nz    z-derivative of the pressure coefficients
  cplx* U0;       // Nz    velocity sweeps, one per component
  cplx* U1;
  cplx* U2;
  long stride;
};

// Input accessor: coefficient (c, k) of the forcing for this mode.
template <class In>
DPS_HD cplx fin(const In& in, int c, int k) { return in(c, k); }

// Output: Chebyshev coefficients of velocity (no wall correction) and the
// correction coefficients corr[c][b] (b over the profile basis).
struct ModeResult {
  cplx corr[3][4];
  int nbasis;
};

// Each pass over z is processed in chunks of CH steps: all loads of a chunk
// are issued first, then the dependent arithmetic runs on registers, then the
// stores -- so a lane keeps CH memory requests in flight instead of waiting
// one latency per recurrence step.
#define DPS_CH 8

// Band entries of I - Ksq S and the Robin BC rows; the functions below define
// Ksq, H, H2K, st, B and sc before using them.
#define D0(i) (1.0 - B.sd0[(i)] * Ksq)
#define DP2(i) (0.0 - B.sp2[(i)] * Ksq)
#define DM2(i) (0.0 - B.sm2[(i) + 2] * Ksq)
#define IB(i) sc.ib[(long)(i) * st]
#define AB(i) sc.ab[(long)(i) * st]
#define PP(i) sc.P[(long)(i) * st]
// Robin BC rows H^2 K BCR2 + H BCR1 and H^2 K BCL2 + H BCL1 (Solvers.py:309-313)
#define C0(j) (B.R2[(j)] * H2K + (B.R2[(j)] + (H * B.R1[(j)] - B.R2[(j)])))
#define C1(j) (B.L2[(j)] * H2K + (-B.L2[(j)] + (H * B.L1[(j)] + B.L2[(j)])))

// Factorisation that depends on |k| only: reciprocal KBPENTA pivots IB(1..Nz)
// (LinearSolvers.h:52-74), A^{-1}B (LinearSolvers.cpp:42) on even/odd rows of
// sc.ab, and the inverse of the 2x2 Schur complement G = C A^{-1}B - D
// (LinearSolvers.cpp:44-56) returned as Gi = {Gi00, Gi01, Gi10, Gi11}.
DPS_HD void bvp_factor(const BvpConsts& B, double Ksq, const ScratchView& sc, double* Gi)
{
  const int Nz = B.Nz;
  const long st = sc.stride;
  const double H = B.H;
  const double K = sqrt(Ksq);
  const double H2K = (H * H) * K;
  constexpr int CH = DPS_CH;
  // KBPENTA pivots (LinearSolvers.h:52-74), stored as reciprocals
  {
    double bm2 = D0(0), bm1 = D0(1);
    IB(1) = 1.0 / bm2;
    IB(2) = 1.0 / bm1;
    for (int i = 3; i <= Nz; ++i) {
      const double b = D0(i - 1) - DM2(i - 3) * DP2(i - 3) / bm2;
      IB(i) = 1.0 / b;
      bm2 = bm1;
      bm1 = b;
    }
  }
  // A^{-1}B (LinearSolvers.cpp:42): B = -Ksq SIMat[:, Nz:Nz+2] has one non-zero
  // per column (rows 0 and 1), so the two solutions live on even and odd rows.
  // G = C A^{-1}B - D and its explicit inverse (LinearSolvers.cpp:44-56).
  double G00 = 0, G01 = 0, G10 = 0, G11 = 0;
  {
    double xm2 = B.S0N * (-Ksq), xm1 = B.S1N1 * (-Ksq);
    AB(0) = xm2;
    AB(1) = xm1;
    for (int i0 = 2; i0 < Nz; i0 += CH) {
      double ibv[CH];
#pragma unroll
      for (int u = 0; u < CH; ++u)
        if (i0 + u < Nz) ibv[u] = IB(i0 + u - 1);
#pragma unroll
      for (int u = 0; u < CH; ++u) {
        const int i = i0 + u;
        if (i < Nz) {
          const double v = 0.0 - (DM2(i - 2) * ibv[u]) * xm2;
          AB(i) = v;
          xm2 = xm1;
          xm1 = v;
        }
      }
    }
    double xp2 = xm1 * IB(Nz), xp1 = xm2 * IB(Nz - 1);   // x[Nz-1], x[Nz-2]
    AB(Nz - 1) = xp2;
    AB(Nz - 2) = xp1;
    if ((Nz - 1) & 1) { G01 += C0(Nz - 1) * xp2; G11 += C1(Nz - 1) * xp2; G00 += C0(Nz - 2) * xp1; G10 += C1(Nz - 2) * xp1; }
    else { G00 += C0(Nz - 1) * xp2; G10 += C1(Nz - 1) * xp2; G01 += C0(Nz - 2) * xp1; G11 += C1(Nz - 2) * xp1; }
    for (int i0 = Nz - 3; i0 >= 0; i0 -= CH) {
      double abv[CH], ibv[CH];
#pragma unroll
      for (int u = 0; u < CH; ++u)
        if (i0 - u >= 0) { abv[u] = AB(i0 - u); ibv[u] = IB(i0 - u + 1); }
#pragma unroll
      for (int u = 0; u < CH; ++u) {
        const int i = i0 - u;
        if (i >= 0) {
          const double v = (abv[u] - DP2(i) * xp2) * ibv[u];
          abv[u] = v;
          if (i & 1) { G01 += C0(i) * v; G11 += C1(i) * v; }
          else { G00 += C0(i) * v; G10 += C1(i) * v; }
          xp2 = xp1;
          xp1 = v;
        }
      }
#pragma unroll
      for (int u = 0; u < CH; ++u)
        if (i0 - u >= 0) AB(i0 - u) = abv[u];
    }
  }
  G00 -= C0(Nz); G10 -= C1(Nz); G01 -= C0(Nz + 1); G11 -= C1(Nz + 1);
  const double det = 1.0 / (G00 * G11 - G01 * G10);
  Gi[0] = G11 * det;
  Gi[1] = -G01 * det;
  Gi[2] = -G10 * det;
  Gi[3] = G00 * det;
}

// Wall correction (DPTools.cpp:590-877) as coefficients of the analytic
// profiles, from the no-wall velocity at the bottom (wall_bot) and top
// (wall_top) boundaries of this mode.
DPS_HD void wall_correction(const BvpConsts& B, double kx, double ky, const cplx* wall_bot,
                            const cplx* wall_top, ModeResult& res);

template <class In, class Out>
DPS_HD void bvp_mode(const BvpConsts& B, double kx, double ky, int nyq_x, int nyq_y,
                     const In& fin_, Out& fout, const ScratchView& sc, ModeResult& res)
{
  const int Nz = B.Nz;
  const long st = sc.stride;
  const double H = B.H;
  const double Ksq = ksq_of(kx, ky);
  const double K = sqrt(Ksq);
  const double H2K = (H * H) * K;
  const cplx Dx = nyq_x ? mk(0, 0) : mk(0, kx);
  const cplx Dy = nyq_y ? mk(0, 0) : mk(0, ky);
  constexpr int CH = DPS_CH;
  double Gi[4];
  bvp_factor(B, Ksq, sc, Gi);
  const double Gi00 = Gi[0], Gi01 = Gi[1], Gi10 = Gi[2], Gi11 = Gi[3];

  // ---- pressure: chebCoeffDiff + assemble_p_rhs (DPTools.cpp:134-195) ----
  {
    cplx dfp1 = mk(0, 0), dfp2 = mk(0, 0);   // df[j+1], df[j+2]
    cplx fz_next = mk(0, 0);                 // f_z[j+1]
    for (int j0 = Nz - 1; j0 >= 0; j0 -= CH) {
      cplx fx[CH], fy[CH], fz[CH];
#pragma unroll
      for (int u = 0; u < CH; ++u)
        if (j0 - u >= 0) { fx[u] = fin(fin_, 0, j0 - u); fy[u] = fin(fin_, 1, j0 - u); fz[u] = fin(fin_, 2, j0 - u); }
#pragma unroll
      for (int u = 0; u < CH; ++u) {
        const int j = j0 - u;
        if (j >= 0) {
          cplx df;
          if (j == Nz - 1) df = mk(0, 0);
          else if (j == Nz - 2) df = ((double)((Nz - 1) * 2) / H) * fz_next;
          else df = dfp2 + ((double)((j + 1) * 2) / H) * fz_next;
          dfp2 = dfp1;
          dfp1 = df;
          fz_next = fz[u];
          if (j == 0) df = df / 2;
          fx[u] = Dx * fx[u] + Dy * fy[u] + df;
        }
      }
#pragma unroll
      for (int u = 0; u < CH; ++u)
        if (j0 - u >= 0) PP(j0 - u) = fx[u];
    }
  }
  // KBPENTA forward / backward sweeps (LinearSolvers.h:106-141)
  {
    cplx gm2 = PP(0), gm1 = PP(1);
    for (int i0 = 2; i0 < Nz; i0 += CH) {
      cplx v[CH];
      double ibv[CH];
#pragma unroll
      for (int u = 0; u < CH; ++u)
        if (i0 + u < Nz) { v[u] = PP(i0 + u); ibv[u] = IB(i0 + u - 1); }
#pragma unroll
      for (int u = 0; u < CH; ++u) {
        const int i = i0 + u;
        if (i < Nz) {
          const cplx g = v[u] - (DM2(i - 2) * ibv[u]) * gm2;
          v[u] = g;
          gm2 = gm1;
          gm1 = g;
        }
      }
#pragma unroll
      for (int u = 0; u < CH; ++u)
        if (i0 + u < Nz) PP(i0 + u) = v[u];
    }
  }
  cplx y0 = mk(0, 0), y1 = mk(0, 0);
  {
    cplx xp2 = PP(Nz - 1) * IB(Nz), xp1 = PP(Nz - 2) * IB(Nz - 1);
    PP(Nz - 1) = xp2;
    PP(Nz - 2) = xp1;
    y0 = xp2 * C0(Nz - 1) + xp1 * C0(Nz - 2);
    y1 = xp2 * C1(Nz - 1) + xp1 * C1(Nz - 2);
    for (int i0 = Nz - 3; i0 >= 0; i0 -= CH) {
      cplx v[CH];
      double ibv[CH];
#pragma unroll
      for (int u = 0; u < CH; ++u)
        if (i0 - u >= 0) { v[u] = PP(i0 - u); ibv[u] = IB(i0 - u + 1); }
#pragma unroll
      for (int u = 0; u < CH; ++u) {
        const int i = i0 - u;
        if (i >= 0) {
          const cplx x = (v[u] - DP2(i) * xp2) * ibv[u];
          v[u] = x;
          y0 = y0 + x * C0(i);
          y1 = y1 + x * C1(i);
          xp2 = xp1;
          xp1 = x;
        }
      }
#pragma unroll
      for (int u = 0; u < CH; ++u)
        if (i0 - u >= 0) PP(i0 - u) = v[u];
    }
  }
  const cplx Y0 = Gi00 * y0 + Gi01 * y1;
  const cplx Y1 = Gi10 * y0 + Gi11 * y1;
  // x = xp - A^{-1}B Y ; Cp = SIMat x ; Dp = FIMat x  (DPTools.cpp:266-284),
  // streamed with a five-point window x[m-2..m+2]
  cplx psum = mk(0, 0), palt = mk(0, 0);
  {
#define XP(j) ((j) < Nz ? (PP(j) - ((((j) & 1) ? AB(j) * Y1 : AB(j) * Y0))) : ((j) == Nz ? Y0 : Y1))
    cplx xm2 = mk(0, 0), xm1 = mk(0, 0), x0 = XP(0), x1 = XP(1), x2 = XP(2);
    for (int m0 = 0; m0 < Nz; m0 += CH) {
      cplx ahead[CH];
#pragma unroll
      for (int u = 0; u < CH; ++u) {
        const int j = m0 + u + 3;
        if (m0 + u < Nz && j <= Nz + 1) ahead[u] = XP(j);
      }
      cplx cpv[CH], dpv[CH];
#pragma unroll
      for (int u = 0; u < CH; ++u) {
        const int m = m0 + u;
        if (m < Nz) {
          cplx cp = B.sd0[m] * x0;
          if (m >= 2) cp = B.sm2[m] * xm2 + cp;
          cp = cp + B.sp2[m] * x2;
          if (m == 0) cp = cp + B.S0N * Y0;
          if (m == 1) cp = cp + B.S1N1 * Y1;
          cplx dp = B.fp1[m] * x1;
          if (m >= 1) dp = B.fm1[m] * xm1 + dp;
          if (m == 0) dp = dp + B.F0N1 * Y1;
          cpv[u] = cp;
          dpv[u] = dp;
          psum = psum + cp;
          palt = (m & 1) ? palt - cp : palt + cp;
          xm2 = xm1; xm1 = x0; x0 = x1; x1 = x2;
          if (m + 3 <= Nz + 1) x2 = ahead[u];
        }
      }
#pragma unroll
      for (int u = 0; u < CH; ++u)
        if (m0 + u < Nz) { sc.CP[(long)(m0 + u) * st] = cpv[u]; sc.DP[(long)(m0 + u) * st] = dpv[u]; }
    }
#undef XP
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

  // ---- velocity: three Schur solves (DPTools.cpp:197-216, 287-345),
  //      the three components interleaved ----
  cplx* Uc[3] = {sc.U0, sc.U1, sc.U2};
  {
    constexpr int CV = DPS_CH / 2;
    cplx gm2[3], gm1[3];
    for (int j0 = 0; j0 < Nz; j0 += CV) {
      cplx cpv[CV], dpv[CV], f0[CV], f1[CV], f2[CV];
      double ibv[CV];
#pragma unroll
      for (int u = 0; u < CV; ++u) {
        const int j = j0 + u;
        if (j < Nz) {
          cpv[u] = sc.CP[(long)j * st];
          dpv[u] = sc.DP[(long)j * st];
          f0[u] = fin(fin_, 0, j);
          f1[u] = fin(fin_, 1, j);
          f2[u] = fin(fin_, 2, j);
          ibv[u] = (j >= 2) ? IB(j - 1) : 0.0;
        }
      }
#pragma unroll
      for (int u = 0; u < CV; ++u) {
        const int j = j0 + u;
        if (j < Nz) {
          cplx r[3];
          r[0] = (Dx * cpv[u] - f0[u]) / B.eta;
          r[1] = (Dy * cpv[u] - f1[u]) / B.eta;
          r[2] = (dpv[u] - f2[u]) / B.eta;
          const double l = (j >= 2) ? DM2(j - 2) * ibv[u] : 0.0;
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            cplx g = r[c];
            if (j >= 2) g = g - l * gm2[c];
            gm2[c] = (j >= 1) ? gm1[c] : mk(0, 0);
            gm1[c] = g;
          }
          f0[u] = gm1[0];
          f1[u] = gm1[1];
          f2[u] = gm1[2];
        }
      }
#pragma unroll
      for (int u = 0; u < CV; ++u) {
        const int j = j0 + u;
        if (j < Nz) {
          Uc[0][(long)j * st] = f0[u];
          Uc[1][(long)j * st] = f1[u];
          Uc[2][(long)j * st] = f2[u];
        }
      }
    }
  }
  cplx Ya[3], Yb[3];
  {
    constexpr int CV = DPS_CH / 2;
    cplx xp2[3], xp1[3], ya[3], yb[3];
    const double i1 = IB(Nz), i2 = IB(Nz - 1);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      xp2[c] = Uc[c][(long)(Nz - 1) * st] * i1;
      xp1[c] = Uc[c][(long)(Nz - 2) * st] * i2;
      Uc[c][(long)(Nz - 1) * st] = xp2[c];
      Uc[c][(long)(Nz - 2) * st] = xp1[c];
      ya[c] = xp2[c] * C0(Nz - 1) + xp1[c] * C0(Nz - 2);
      yb[c] = xp2[c] * C1(Nz - 1) + xp1[c] * C1(Nz - 2);
    }
    for (int i0 = Nz - 3; i0 >= 0; i0 -= CV) {
      cplx v[3][CV];
      double ibv[CV];
#pragma unroll
      for (int u = 0; u < CV; ++u) {
        const int i = i0 - u;
        if (i >= 0) {
          ibv[u] = IB(i + 1);
#pragma unroll
          for (int c = 0; c < 3; ++c) v[c][u] = Uc[c][(long)i * st];
        }
      }
#pragma unroll
      for (int u = 0; u < CV; ++u) {
        const int i = i0 - u;
        if (i >= 0) {
          const double d = DP2(i), c0 = C0(i), c1 = C1(i);
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            const cplx x = (v[c][u] - d * xp2[c]) * ibv[u];
            v[c][u] = x;
            ya[c] = ya[c] + x * c0;
            yb[c] = yb[c] + x * c1;
            xp2[c] = xp1[c];
            xp1[c] = x;
          }
        }
      }
#pragma unroll
      for (int u = 0; u < CV; ++u) {
        const int i = i0 - u;
        if (i >= 0)
#pragma unroll
          for (int c = 0; c < 3; ++c) Uc[c][(long)i * st] = v[c][u];
      }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const cplx a = ya[c] - bc[2 * c], b2 = yb[c] - bc[2 * c + 1];
      Ya[c] = Gi00 * a + Gi01 * b2;
      Yb[c] = Gi10 * a + Gi11 * b2;
    }
  }
  cplx wall_bot[3], wall_top[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const cplx* Uv = Uc[c];
    const cplx ya = Ya[c], yb = Yb[c];
#define XU(j) ((j) < Nz ? (Uv[(long)(j) * st] - ((((j) & 1) ? AB(j) * yb : AB(j) * ya))) : ((j) == Nz ? ya : yb))
    cplx sb = mk(0, 0), stp = mk(0, 0);
    cplx xm2 = mk(0, 0), x0 = XU(0), x1 = XU(1), x2 = XU(2), xm1 = mk(0, 0);
    for (int m0 = 0; m0 < Nz; m0 += CH) {
      cplx ahead[CH];
#pragma unroll
      for (int u = 0; u < CH; ++u) {
        const int j = m0 + u + 3;
        if (m0 + u < Nz && j <= Nz + 1) ahead[u] = XU(j);
      }
#pragma unroll
      for (int u = 0; u < CH; ++u) {
        const int m = m0 + u;
        if (m < Nz) {
          cplx cu = B.sd0[m] * x0;
          if (m >= 2) cu = B.sm2[m] * xm2 + cu;
          cu = cu + B.sp2[m] * x2;
          if (m == 0) cu = cu + B.S0N * ya;
          if (m == 1) cu = cu + B.S1N1 * yb;
          fout(c, m, cu);
          stp = stp + cu;
          sb = (m & 1) ? sb - cu : sb + cu;
          xm2 = xm1; xm1 = x0; x0 = x1; x1 = x2;
          if (m + 3 <= Nz + 1) x2 = ahead[u];
        }
      }
    }
#undef XU
    wall_bot[c] = sb;
    wall_top[c] = stp;
  }
  wall_correction(B, kx, ky, wall_bot, wall_top, res);
}
#undef D0
#undef DP2
#undef DM2
#undef C0
#undef C1
#undef IB
#undef AB
#undef PP

// One row of a real 4x4 system with a complex right-hand side.  The solve
// below keeps every row in named scalars (no arrays, no runtime indices), so
// on the GPU the system lives in registers.
struct Row4 {
  double a0, a1, a2, a3;
  cplx b;
};

DPS_HD void row_swap_if(bool s, Row4& x, Row4& y)
{
  const Row4 t = x, u = y;
  x.a0 = dsel(s, u.a0, t.a0); y.a0 = dsel(s, t.a0, u.a0);
  x.a1 = dsel(s, u.a1, t.a1); y.a1 = dsel(s, t.a1, u.a1);
  x.a2 = dsel(s, u.a2, t.a2); y.a2 = dsel(s, t.a2, u.a2);
  x.a3 = dsel(s, u.a3, t.a3); y.a3 = dsel(s, t.a3, u.a3);
  x.b = csel(s, u.b, t.b);    y.b = csel(s, t.b, u.b);
}

// y -= (y[col] / x[col]) x for the columns after `col` (Gaussian elimination step)
DPS_HD void row_eliminate(const Row4& x, Row4& y, int col, double inv)
{
  if (col == 0) {
    const double l = y.a0 * inv;
    y.a1 -= l * x.a1; y.a2 -= l * x.a2; y.a3 -= l * x.a3; y.b = y.b - l * x.b;
  } else if (col == 1) {
    const double l = y.a1 * inv;
    y.a2 -= l * x.a2; y.a3 -= l * x.a3; y.b = y.b - l * x.b;
  } else {
    const double l = y.a2 * inv;
    y.a3 -= l * x.a3; y.b = y.b - l * x.b;
  }
}

// Gaussian elimination with partial pivoting (largest |entry| of the column,
// first one on ties) and back substitution; (x0, x1, x2, x3) receive the
// solution (named outputs: an output array whose element a branch picks is
// kept in private memory on the GPU).
DPS_HD void solve4_pivot(Row4& R0, Row4& R1, Row4& R2, Row4& R3, cplx& x0, cplx& x1, cplx& x2,
                         cplx& x3)
{
  double inv0, inv1, inv2;
  // column 0
  {
    int piv = 0;
    double best = fabs(R0.a0);
    if (fabs(R1.a0) > best) { best = fabs(R1.a0); piv = 1; }
    if (fabs(R2.a0) > best) { best = fabs(R2.a0); piv = 2; }
    if (fabs(R3.a0) > best) { best = fabs(R3.a0); piv = 3; }
    row_swap_if(piv == 1, R0, R1);
    row_swap_if(piv == 2, R0, R2);
    row_swap_if(piv == 3, R0, R3);
    inv0 = 1.0 / R0.a0;
    row_eliminate(R0, R1, 0, inv0);
    row_eliminate(R0, R2, 0, inv0);
    row_eliminate(R0, R3, 0, inv0);
  }
  // column 1
  {
    int piv = 1;
    double best = fabs(R1.a1);
    if (fabs(R2.a1) > best) { best = fabs(R2.a1); piv = 2; }
    if (fabs(R3.a1) > best) { best = fabs(R3.a1); piv = 3; }
    row_swap_if(piv == 2, R1, R2);
    row_swap_if(piv == 3, R1, R3);
    inv1 = 1.0 / R1.a1;
    row_eliminate(R1, R2, 1, inv1);
    row_eliminate(R1, R3, 1, inv1);
  }
  // column 2
  {
    const bool sw = fabs(R3.a2) > fabs(R2.a2);
    row_swap_if(sw, R2, R3);
    inv2 = 1.0 / R2.a2;
    row_eliminate(R2, R3, 2, inv2);
  }
  // back substitution with the pivots' reciprocals (one division each)
  x3 = R3.b * (1.0 / R3.a3);
  x2 = (R2.b - R2.a3 * x3) * inv2;
  x1 = ((R1.b - R1.a2 * x2) - R1.a3 * x3) * inv1;
  x0 = (((R0.b - R0.a1 * x1) - R0.a2 * x2) - R0.a3 * x3) * inv0;
}

// Wall-correction coefficients of the bottom wall (DPTools.cpp:617-633),
// basis (e^{-Kz}, z e^{-Kz}): c[comp][basis].
DPS_HD void bottom_coeffs(double kx, double ky, const cplx* wall_bot, cplx& c00, cplx& c01, cplx& c10,
                          cplx& c11, cplx& c20, cplx& c21)
{
  const double K = sqrt(ksq_of(kx, ky));
  const double iK = 1.0 / K;
  const cplx fx = -wall_bot[0], fy = -wall_bot[1], fz = -wall_bot[2];
  const cplx alpha = kx * fx + ky * fy + times_i(K * fz);
  c00 = fx; c01 = ((-kx) * alpha) * iK;
  c10 = fy; c11 = ((-ky) * alpha) * iK;
  c20 = fz; c21 = -times_i(alpha);
}

// Wall-correction unknowns of the slit (DPTools.cpp:727-812), basis
// (e^{-kz}, e^{k(z-Hs)}, z e^{-kz}, z e^{k(z-Hs)}): X0..X7 of the 8x8 system.
DPS_HD void slit_unknowns(const BvpConsts& B, double kx, double ky, const cplx* wall_bot, const cplx* wall_top,
                          cplx& X0, cplx& X1, cplx& X2, cplx& X3, cplx& X4, cplx& X5, cplx& X6, cplx& X7)
{
  const double k = sqrt(ksq_of(kx, ky));
  const double Hs = B.Hs;
  const double fac = 1.0 / 2.0 / B.eta;
  const double enkh = exp(-k * Hs);
  const cplx fbx = -wall_bot[0], fby = -wall_bot[1], fbz = -wall_bot[2];
  const cplx ftx = -wall_top[0], fty = -wall_top[1], ftz = -wall_top[2];
  // right-hand side of the 8x8 system
  const cplx rhs0 = times_i(-kx * fbx) - times_i(ky * fby);
  const cplx rhs1 = times_i(-kx * ftx) - times_i(ky * fty);
  const cplx rhs2 = fbx, rhs3 = fby, rhs4 = fbz;
  const cplx rhs5 = ftx, rhs6 = fty, rhs7 = ftz;
  // rows 0,1,4,7 of the 8x8 system couple only (x0, x1, x6, x7) and are
  // real: solve that 4x4 block with partial pivoting, then the two 2x2
  // blocks for (x2, x3) and (x4, x5).
  Row4 R0 = {fac, enkh * fac, -k, k * enkh, rhs0};
  Row4 R1 = {(1 - k * Hs) * enkh * fac, (1 + k * Hs) * fac, -k * enkh, k, rhs1};
  Row4 R2 = {0, 0, 1, enkh, rhs4};
  Row4 R3 = {Hs * enkh * fac, Hs * fac, enkh, 1, rhs7};
  solve4_pivot(R0, R1, R2, R3, X0, X1, X6, X7);
  // rows 2,5: x2 + enkh x3 = b2 ; enkh x2 + x3 = b5 - (A50 X0 + A51 X1)
  const double id2 = 1.0 / (1 - enkh * enkh);
  const double fk = fac * (1.0 / k);
  const cplx r5 = rhs5 - (times_i(-kx * Hs * enkh * fk * X0) + times_i(kx * Hs * fk * X1));
  X2 = (rhs2 - enkh * r5) * id2;
  X3 = (r5 - enkh * rhs2) * id2;
  const cplx r6 = rhs6 - (times_i(-ky * Hs * enkh * fk * X0) + times_i(ky * Hs * fk * X1));
  X4 = (rhs3 - enkh * r6) * id2;
  X5 = (r6 - enkh * rhs3) * id2;
}

// Slit coefficients c[comp][basis] from the unknowns.
DPS_HD void slit_coeffs(const BvpConsts& B, double kx, double ky, const cplx* wall_bot, const cplx* wall_top,
                        cplx& c00, cplx& c01, cplx& c02, cplx& c03, cplx& c10, cplx& c11, cplx& c12,
                        cplx& c13, cplx& c20, cplx& c21, cplx& c22, cplx& c23)
{
  cplx X0, X1, X2, X3, X4, X5, X6, X7;
  slit_unknowns(B, kx, ky, wall_bot, wall_top, X0, X1, X2, X3, X4, X5, X6, X7);
  const double k = sqrt(ksq_of(kx, ky));
  const double fac = 1.0 / 2.0 / B.eta;
  const double fk = fac * (1.0 / k);
  c00 = X2; c01 = X3;
  c02 = times_i(-(kx * fk) * X0); c03 = times_i((kx * fk) * X1);
  c10 = X4; c11 = X5;
  c12 = times_i(-(ky * fk) * X0); c13 = times_i((ky * fk) * X1);
  c20 = X6; c21 = X7;
  c22 = fac * X0; c23 = fac * X1;
}

// Wall correction of one mode into res: zeroed, then the entries of the
// geometry's basis written on its branch.
DPS_HD void wall_correction(const BvpConsts& B, double kx, double ky, const cplx* wall_bot,
                            const cplx* wall_top, ModeResult& res)
{
  for (int c = 0; c < 3; ++c)
    for (int b = 0; b < 4; ++b) res.corr[c][b] = mk(0, 0);
  res.nbasis = 0;
  if (B.geometry == MODE_BOTTOM) {
    res.nbasis = 2;
    bottom_coeffs(kx, ky, wall_bot, res.corr[0][0], res.corr[0][1], res.corr[1][0], res.corr[1][1],
                  res.corr[2][0], res.corr[2][1]);
  } else if (B.geometry == MODE_SLIT) {
    res.nbasis = 4;
    slit_coeffs(B, kx, ky, wall_bot, wall_top, res.corr[0][0], res.corr[0][1], res.corr[0][2], res.corr[0][3],
                res.corr[1][0], res.corr[1][1], res.corr[1][2], res.corr[1][3], res.corr[2][0],
                res.corr[2][1], res.corr[2][2], res.corr[2][3]);
  }
}

// The same into named scalars, every entry of res assigned once at the end.
// In a GPU kernel res then stays in registers; the form above keeps it in
// private memory (the branches write different entries).  The FP32 spectral
// kernel takes this form (no private-memory traffic, and faster); the FP64
// one the form above, which measures faster there.
DPS_HD void wall_correction_regs(const BvpConsts& B, double kx, double ky, const cplx* wall_bot,
                                 const cplx* wall_top, ModeResult& res)
{
  const cplx z0 = mk(0, 0);
  cplx c00 = z0, c01 = z0, c02 = z0, c03 = z0;
  cplx c10 = z0, c11 = z0, c12 = z0, c13 = z0;
  cplx c20 = z0, c21 = z0, c22 = z0, c23 = z0;
  int nb = 0;
  if (B.geometry == MODE_BOTTOM) {
    nb = 2;
    bottom_coeffs(kx, ky, wall_bot, c00, c01, c10, c11, c20, c21);
  } else if (B.geometry == MODE_SLIT) {
    nb = 4;
    slit_coeffs(B, kx, ky, wall_bot, wall_top, c00, c01, c02, c03, c10, c11, c12, c13, c20, c21, c22, c23);
  }
  res.nbasis = nb;
  res.corr[0][0] = c00; res.corr[0][1] = c01; res.corr[0][2] = c02; res.corr[0][3] = c03;
  res.corr[1][0] = c10; res.corr[1][1] = c11; res.corr[1][2] = c12; res.corr[1][3] = c13;
  res.corr[2][0] = c20; res.corr[2][1] = c21; res.corr[2][2] = c22; res.corr[2][3] = c23;
}

// k = 0 mode for the walled geometries (Solvers.py:482-516, 563-599): the
// velocity BVP u'' = -f/eta with u(0)=0 and u'(top)=0 (bottom wall) or
// u(0)=u(Lz)=0 (slit); w and the no-wall k=0 mode stay zero.
template <class In, class Out>
DPS_HD void bvp_k0(const BvpConsts& B, const In& fin_, Out& fout, cplx* work, long ws)
{
  const int Nz = B.Nz;
  if (B.geometry != MODE_BOTTOM && B.geometry != MODE_SLIT) {
    for (int c = 0; c < 3; ++c)
      for (int m = 0; m < Nz; ++m) fout(c, m, mk(0, 0));
    return;
  }
  const double* RA = (B.geometry == MODE_BOTTOM) ? B.R1 : B.R2;   // first BC row
  const double* RB = B.L2;                                        // second BC row
  // D = -BCs[:, Nz:]; Ginv = 1/det [[D11, -D01], [-D10, D00]]
  const double D00 = -RA[Nz], D01 = -RA[Nz + 1], D10 = -RB[Nz], D11 = -RB[Nz + 1];
  const double idet = 1.0 / (D00 * D11 - D01 * D10);
  const double g00 = idet * D11, g01 = idet * (-D01), g10 = idet * (-D10), g11 = idet * D00;
  for (int c = 0; c < 2; ++c) {
    cplx ca = mk(0, 0), cb = mk(0, 0);
    for (int j = 0; j < Nz; ++j) {
      cplx f = fin(fin_, c, j);
      if (B.geometry == MODE_SLIT) f = mk(f.re, f.re);   // Solvers.py:564-566
      const cplx r = (-1.0) * f / B.eta;
      work[(long)j * ws] = r;
      ca = ca + RA[j] * r;
      cb = cb + RB[j] * r;
    }
    const cplx a0 = g00 * ca + g01 * cb;
    const cplx a1 = g10 * ca + g11 * cb;
    for (int m = 0; m < Nz; ++m) {
      cplx v = B.sd0[m] * work[(long)m * ws];
      if (m >= 2) v = B.sm2[m] * work[(long)(m - 2) * ws] + v;
      if (m + 2 < Nz) v = v + B.sp2[m] * work[(long)(m + 2) * ws];
      else if (m + 2 == Nz) v = v + B.sp2[m] * a0;
      else v = v + B.sp2[m] * a1;
      if (m == 0) v = v + B.S0N * a0;
      if (m == 1) v = v + B.S1N1 * a1;
      fout(c, m, v);
    }
  }
  for (int m = 0; m < Nz; ++m) fout(2, m, mk(0, 0));
}

}  // namespace dps
