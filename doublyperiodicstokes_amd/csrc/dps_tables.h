// dps_tables.h -- host-side construction of every constant table the solver
// needs: Clenshaw-Curtis nodes/weights, ES normalisation and the dipole cutoff,
// the mirrored z-extended grids with their image maps, Chebyshev integral
// bands, boundary rows and the z-transform matrices.  Host-only C++ (no HIP),
// shared by the library and its CPU checks.
#pragma once
#include <cmath>
#include <complex>
#include <vector>
#include "dps_core.h"

namespace dps {

// Clenshaw-Curtis nodes (descending) and weights on [a,b]; the reference's
// src/Quadrature.cpp:11-46 and python/Chebyshev.py:4-38 (same arithmetic).
inline void clencurt(std::vector<double>& pts, std::vector<double>& wts, double a, double b,
                     unsigned Np1)
{
  const unsigned N = Np1 - 1;
  const double H = (b - a) / 2, bpad2 = (b + a) / 2;
  pts.assign(Np1, 0.0);
  wts.assign(Np1, 0.0);
  std::vector<double> v(N > 1 ? N - 1 : 1, 1.0);
  const double w1 = (N % 2 ? H / (N * N) : H / (N * N - 1));
  const unsigned end = (N % 2 ? (N - 1) / 2 + 1 : N / 2);
  for (unsigned i = 0; i < Np1; ++i) pts[i] = H * std::cos(M_PI * i / N) + bpad2;
  wts[0] = w1;
  wts[N] = w1;
  for (unsigned k = 1; k < end; ++k)
    for (unsigned ii = 1; ii < N; ++ii) v[ii - 1] -= 2 * std::cos(2 * k * M_PI * ii / N) / (4 * k * k - 1);
  if (!(N % 2))
    for (unsigned ii = 1; ii < N; ++ii) v[ii - 1] -= std::cos(N * M_PI * ii / N) / (N * N - 1);
  for (unsigned ii = 1; ii < N; ++ii) wts[ii] = 2 * H * v[ii - 1] / N;
}

// integral of the 1-D ES kernel by 1000-point Clenshaw-Curtis
// (ParticleList.cpp:217-241)
inline double es_norm(double alpha, double betaw)
{
  std::vector<double> p, w;
  clencurt(p, w, -1 * alpha, alpha, 1000);
  double norm = 0;
  for (unsigned j = 0; j < 1000; ++j)
    norm += std::exp(betaw * (std::sqrt(1 - p[j] * p[j] / (alpha * alpha)) - 1)) * w[j];
  return norm;
}

// support cutoff of the derivative kernel (ParticleList.cpp:253-265): the
// near-edge extremum of |phi'|
inline double dipole_thresh(double alpha, double betaw)
{
  typedef std::complex<double> C;
  const C inner = std::sqrt(C(-12 * betaw * betaw + 81, 0));
  const C base = (C(8 * betaw * betaw, 0) + 0.12e2 * inner - 0.108e3) * betaw;
  const C t = std::sqrt(0.6e1 * std::pow(base, 0.1e1 / 0.3e1) / betaw +
                        0.24e2 * betaw * std::pow(base, -0.1e1 / 0.3e1) + 0.12e2);
  return (t * alpha / 0.6e1).real();
}

// z-extended grid with mirrored ghosts (ParticleList.cpp:419-451) plus, for
// every extended node e, the interior node it folds onto and the factor it
// carries for the two boundary classes:
//   class A: velocity/force components (mirror_inv at a wall),
//   class B: the z-derivative kernel (mirror at a wall).
// A ghost at a wall folds with s (-1 or +1) onto its mirror; the wall node
// carries 1+s (BoundaryConditions.h:284-341 fold, :652-729 copy).  At an open
// end (BC none) ghosts carry 0 and the end node 1.
struct ZExt {
  std::vector<double> z, w;     // node values (descending) and CC weights
  std::vector<int> kt;          // interior target
  std::vector<double> facA, facB;
  int nzeff = 0, ext_up = 0, ext_down = 0, wzmax = 0;
};

inline ZExt build_zext(const std::vector<double>& zG, const std::vector<double>& zW, double thresh,
                       bool wall_top, bool wall_bottom)
{
  ZExt X;
  const int Nz = (int)zG.size();
  int i = 1;
  while (i < Nz && zG[0] - zG[i] <= thresh) { X.ext_up += 1; i += 1; }
  i = Nz - 2;
  while (i >= 0 && zG[i] - zG[Nz - 1] <= thresh) { X.ext_down += 1; i -= 1; }
  X.nzeff = Nz + X.ext_up + X.ext_down;
  X.z.assign(X.nzeff, 0.0);
  X.w.assign(X.nzeff, 0.0);
  X.kt.assign(X.nzeff, 0);
  X.facA.assign(X.nzeff, 0.0);
  X.facB.assign(X.nzeff, 0.0);
  for (int e = X.ext_up; e < X.nzeff - X.ext_down; ++e) {
    X.z[e] = zG[e - X.ext_up];
    X.w[e] = zW[e - X.ext_up];
    X.kt[e] = e - X.ext_up;
    X.facA[e] = 1.0;
    X.facB[e] = 1.0;
  }
  int j = 0;
  for (int ii = X.ext_up; ii > 0; --ii) {
    X.z[j] = 2.0 * zG[0] - zG[ii];
    X.w[j] = zW[ii];
    X.kt[j] = ii;
    X.facA[j] = wall_top ? -1.0 : 0.0;
    X.facB[j] = wall_top ? 1.0 : 0.0;
    j += 1;
  }
  j = X.nzeff - X.ext_down;
  for (int ii = Nz - 2; ii > Nz - 2 - X.ext_down; --ii) {
    X.z[j] = -1.0 * zG[ii];
    X.w[j] = zW[ii];
    X.kt[j] = ii;
    X.facA[j] = wall_bottom ? -1.0 : 0.0;
    X.facB[j] = wall_bottom ? 1.0 : 0.0;
    j += 1;
  }
  // wall nodes: (1+s) with s = -1 (class A) or +1 (class B)
  if (wall_top) { X.facA[X.ext_up] = 0.0; X.facB[X.ext_up] = 2.0; }
  if (wall_bottom) { X.facA[X.ext_up + Nz - 1] = 0.0; X.facB[X.ext_up + Nz - 1] = 2.0; }
  // largest possible z support: nodes inside a closed window of width 2 thresh
  int best = 0;
  for (int a = 0; a < X.nzeff; ++a) {
    int c = 0;
    for (int b = a; b < X.nzeff && X.z[a] - X.z[b] <= 2 * thresh * (1 + 1e-12); ++b) ++c;
    if (c > best) best = c;
  }
  X.wzmax = best + 1;
  return X;
}

// Chebyshev integral bands and BC rows (Chebyshev.py:119-160, Solvers.py:603-643)
struct ChebTables {
  std::vector<double> sd0, sm2, sp2, fm1, fp1;   // length Nz
  double S0N = 0, S1N1 = 0, F0N1 = 0;
  std::vector<double> R1, R2, L1, L2;            // length Nz+2
};

inline ChebTables build_cheb_tables(int N, double H)
{
  ChebTables T;
  T.sd0.assign(N, 0.0); T.sm2.assign(N, 0.0); T.sp2.assign(N, 0.0);
  T.fm1.assign(N, 0.0); T.fp1.assign(N, 0.0);
  const double H2 = H * H;   // numpy H**2
  // second integral: diagonals -2, 0, +2 (unscaled), then * H^2
  for (int j = 2; j < N; ++j) {
    // colm2 = [0.25, 1/(2jj(2jj-2)) for jj=3..N-1] placed at (j, j-2)
    const double v = (j == 2) ? 0.25 : 1.0 / (double)(2 * j * (2 * j - 2));
    T.sm2[j] = H2 * v;
  }
  for (int j = 0; j < N; ++j) {
    double v;
    if (j == 0) v = 0;
    else if (j == 1) v = -0.125;
    else if (j == 2) v = -1. / 8. - 1. / 24.;
    else v = -1.0 / (double)(2 * j * (2 * j - 2)) - 1.0 / (double)(2 * j * (2 * j + 2)) * (double)(j < N - 1);
    T.sd0[j] = H2 * v;
    double p;
    if (j == 0) p = 0;
    else if (j == 1) p = 0.125;
    else if (j == 2) p = 1. / 24.;
    else p = 1.0 / (double)(2 * j * (2 * j + 2)) * (double)(j < N - 2);
    T.sp2[j] = H2 * p;
  }
  T.S0N = H2 * 1.0;
  T.S1N1 = H2 * 1.0;
  // first integral: diagonals -1, +1 (unscaled), then * H
  for (int j = 1; j < N; ++j) T.fm1[j] = H * ((j == 1) ? 1.0 : 1.0 / (double)(2 * j));
  for (int j = 0; j < N; ++j) {
    double v;
    if (j == 0) v = 0;
    else if (j == 1) v = -1.0 / 2;
    else v = -1.0 / (double)(2 * j) * (double)(j < N - 1);
    T.fp1[j] = H * v;
  }
  T.F0N1 = H * 1.0;
  // BC rows
  T.R1.assign(N + 2, 0.0); T.R2.assign(N + 2, 0.0);
  T.L1.assign(N + 2, 0.0); T.L2.assign(N + 2, 0.0);
  std::vector<double>&R1 = T.R1, &R2 = T.R2, &L1 = T.L1, &L2 = T.L2;
  R1[N + 1] = 1; R2[N] = 1; R1[0] = 1; R1[2] = -1.0 / 2;
  R2[N + 1] += 1; R2[1] = -1.0 / 8; R2[3] = 1.0 / 8;
  R1[1] += 1.0 / 4; R1[3] -= 1.0 / 4; R2[0] += 1.0 / 4; R2[2] -= (1.0 / 8 + 1.0 / 24);
  R2[4] += 1.0 / 24;
  L1[N + 1] = 1; L2[N] = -1; L1[0] = -1; L1[2] = 1.0 / 2;
  L2[N + 1] += 1; L2[1] = -1.0 / 8; L2[3] = 1.0 / 8;
  L1[1] += 1.0 / 4; L1[3] -= 1.0 / 4; L2[0] -= 1.0 / 4; L2[2] += 1.0 / 8 + 1.0 / 24;
  L2[4] -= 1.0 / 24;
  for (int jj = 3; jj < N; ++jj) {
    const double sg = (jj % 2) ? -1.0 : 1.0;           // (-1)^jj
    const double inv2j = 1.0 / (double)(2 * jj);
    R1[jj - 1] += inv2j;                                // BCR1[2:N-1] += 1/(2jj)
    L1[jj - 1] += sg / (double)(2 * jj);
    R1[jj + 1] -= inv2j * (double)(jj < N - 1);         // BCR1[4:N+1]
    L1[jj + 1] -= sg / (double)(2 * jj) * (double)(jj < N - 1);
    R2[jj - 2] += inv2j * 1.0 / (double)(2 * jj - 2);   // BCR2[1:N-2]
    L2[jj - 2] -= inv2j * 1.0 / (double)(2 * jj - 2) * sg;
    R2[jj + 2] += inv2j * 1.0 / (double)(2 * jj + 2) * (double)(jj < N - 2);   // BCR2[5:N+2]
    L2[jj + 2] -= inv2j * 1.0 / (double)(2 * jj + 2) * sg * (double)(jj < N - 2);
    R2[jj] -= (inv2j * 1.0 / (double)(2 * jj - 2) + inv2j * 1.0 / (double)(2 * jj + 2) * (double)(jj < N - 1));
    L2[jj] += (inv2j * 1.0 / (double)(2 * jj - 2) + inv2j * 1.0 / (double)(2 * jj + 2) * (double)(jj < N - 1)) * sg;
  }
  return T;
}

// cos(pi r / (N-1)) with exact values at the symmetric points
inline double cos_pi_frac(long r, int N)
{
  const long M = 2L * (N - 1);
  r %= M;
  if (r < 0) r += M;
  if (r == 0) return 1.0;
  if (r == N - 1) return -1.0;
  if (2 * r == N - 1 || 2 * r == 3 * (N - 1)) return 0.0;
  return std::cos(M_PI * (double)r / (double)(N - 1));
}

// Chebyshev transform matrices on N Gauss-Lobatto nodes (node 0 at the top):
//   forward  c_k = sum_j Tf[k][j] f_j   (DCT-I of Transform.cpp:215-268)
//   inverse  f_j = sum_k Ti[j][k] c_k   (Transform.cpp:154-213, 270-304)
// stored "folded": for even/odd k only the first ceil(N/2) node pairs matter.
//   TfE[kk][j] = Tf[2kk][j],   j < NH (NH = ceil(N/2)), with the pair weight
//   TfO[kk][j] = Tf[2kk+1][j], j < NH
//   TiE[j][kk] = Ti[j][2kk],   TiO[j][kk] = Ti[j][2kk+1],  j < NH
struct ChebTransform {
  int N, NH, NE, NO;   // nodes, node pairs, even and odd mode counts
  std::vector<double> TfE, TfO, TiE, TiO;
};

inline ChebTransform build_cheb_transform(int N)
{
  ChebTransform C;
  C.N = N;
  C.NH = (N + 1) / 2;
  C.NE = (N + 1) / 2;
  C.NO = N / 2;
  const double Nu = 2.0 * N - 2.0;
  C.TfE.assign((size_t)C.NE * C.NH, 0.0);
  C.TfO.assign((size_t)C.NO * C.NH, 0.0);
  C.TiE.assign((size_t)C.NH * C.NE, 0.0);
  C.TiO.assign((size_t)C.NH * C.NO, 0.0);
  for (int k = 0; k < N; ++k) {
    const double sk = (k == 0 || k == N - 1) ? 1.0 : 2.0;
    for (int j = 0; j < C.NH; ++j) {
      const double wj = (j == 0 || j == N - 1) ? 1.0 : 2.0;
      const double c = cos_pi_frac((long)j * k, N);
      const double v = sk * wj * c / Nu;
      if (k % 2 == 0) C.TfE[(size_t)(k / 2) * C.NH + j] = v;
      else C.TfO[(size_t)(k / 2) * C.NH + j] = v;
      if (k % 2 == 0) C.TiE[(size_t)j * C.NE + k / 2] = c;
      else C.TiO[(size_t)j * C.NO + k / 2] = c;
    }
  }
  return C;
}

}  // namespace dps
