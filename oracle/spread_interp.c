/*
 * FP64 CPU oracle for the particle <-> grid stages of the doubly periodic
 * Stokes Mdot path.  TEST INFRASTRUCTURE ONLY: nothing in the product links or
 * loads this file; only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg use it, as a checker.
 *
 * It restates the reference's CPU algorithm (all citations relative to the
 * reference tree, source/cpu/):
 *   - ES kernel and its derivative ........ include/ESKernels.h:8-54
 *   - kernel normalisation, dipole cutoff . src/ParticleList.cpp:211-276
 *   - Clenshaw-Curtis nodes/weights ....... src/Quadrature.cpp:11-46
 *   - particle location on the grid ....... src/ParticleList.cpp:394-548
 *   - spreading on the extended grid ...... src/SpreadInterp.cpp:302-448,
 *                                           include/SpreadInterp.h:99-164
 *   - ghost fold after spreading .......... include/BoundaryConditions.h:22-361
 *   - ghost copy before interpolation ..... include/BoundaryConditions.h:377-731
 *   - interpolation ....................... src/SpreadInterp.cpp:450-600,
 *                                           include/SpreadInterp.h:186-212
 *
 * Layouts follow the reference: extended grid (Nzeff, Nyeff, Nxeff, dof) with
 * dof fastest then x; interior grid (Ny, Nx, Nz, dof) with dof fastest then z.
 * Only one kernel (one radius) per particle list, as in the joint interface.
 *
 * Spreading is parallelised like the reference: particles are bucketed by
 * x-y column and columns of one w x w colour class never overlap, so each
 * colour class runs in parallel without atomics.
 */
#include <math.h>
#include <complex.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

enum { BC_MIRROR = 0, BC_MIRROR_INV = 1, BC_NONE = 2 };

static inline unsigned at3(unsigned i, unsigned j, unsigned k, unsigned Nx, unsigned Ny)
{
  return i + Nx * (j + Ny * k);
}

/* Clenshaw-Curtis nodes on [a,b] (descending) and weights; src/Quadrature.cpp:11-46 */
void orc_clencurt(double* cpts, double* cwts, double a, double b, unsigned Np1)
{
  const unsigned N = Np1 - 1;
  const double H = (b - a) / 2, bpad2 = (b + a) / 2;
  double* v = (double*)malloc(sizeof(double) * (N > 1 ? N - 1 : 1));
  const double w1 = (N % 2 ? H / (N * N) : H / (N * N - 1));
  const unsigned end = (N % 2 ? (N - 1) / 2 + 1 : N / 2);
  for (unsigned i = 0; i < Np1; ++i) cpts[i] = H * cos(M_PI * i / N) + bpad2;
  for (unsigned i = 0; i + 1 < N; ++i) v[i] = 1;
  cwts[0] = w1; cwts[N] = w1;
  for (unsigned k = 1; k < end; ++k)
    for (unsigned ii = 1; ii < N; ++ii)
      v[ii - 1] -= 2 * cos(2 * k * M_PI * ii / N) / (4 * k * k - 1);
  if (!(N % 2))
    for (unsigned ii = 1; ii < N; ++ii)
      v[ii - 1] -= cos(N * M_PI * ii / N) / (N * N - 1);
  for (unsigned ii = 1; ii < N; ++ii) cwts[ii] = 2 * H * v[ii - 1] / N;
  free(v);
}

/* 1-D ES kernel used for the normalisation; include/ESKernels.h:51-54 */
static inline double eskernel1(double x, double betaw, double alpha)
{
  return exp(betaw * (sqrt(1 - x * x / (alpha * alpha)) - 1));
}

/* Normalisation of the ES kernel by 1000-point Clenshaw-Curtis;
   src/ParticleList.cpp:217-241 */
double orc_es_norm(double alpha, double betaw)
{
  enum { NQ = 1000 };
  double cpts[NQ], cwts[NQ];
  orc_clencurt(cpts, cwts, -1 * alpha, alpha, NQ);
  double norm = 0;
  for (unsigned j = 0; j < NQ; ++j) norm += eskernel1(cpts[j], betaw, alpha) * cwts[j];
  return norm;
}

/* Support cutoff of the derivative kernel; src/ParticleList.cpp:253-265 */
double orc_dipole_thresh(double alpha, double betaw)
{
  double complex inner = csqrt((double complex)(-12 * betaw * betaw + 81));
  double complex base = ((8 * betaw * betaw) + 0.12e2 * inner - 0.108e3) * betaw;
  double complex t = csqrt(0.6e1 * cpow(base, 0.1e1 / 0.3e1) / betaw +
                           0.24e2 * betaw * cpow(base, -0.1e1 / 0.3e1) + 0.12e2);
  return creal(t * alpha / 0.6e1);
}

/*
 * Extended Chebyshev z grid with mirrored ghost nodes;
 * src/ParticleList.cpp:419-451.  zG is descending (zG[0] = Lz, zG[Nz-1] = 0).
 * Returns Nzeff; writes ext_up/ext_down.  zext/zext_w must hold Nz + 2*(Nz-1).
 */
int orc_zext(const double* zG, const double* zW, int Nz, double thresh_max,
             double* zext, double* zext_w, int* ext_up_out, int* ext_down_out)
{
  int ext_up = 0, ext_down = 0;
  int i = 1;
  while (i < Nz && zG[0] - zG[i] <= thresh_max) { ext_up += 1; i += 1; }
  i = Nz - 2;
  while (i >= 0 && zG[i] - zG[Nz - 1] <= thresh_max) { ext_down += 1; i -= 1; }
  const int Nzeff = Nz + ext_up + ext_down;
  for (int e = ext_up; e < Nzeff - ext_down; ++e) {
    zext[e] = zG[e - ext_up];
    zext_w[e] = zW[e - ext_up];
  }
  int j = 0;
  for (int ii = ext_up; ii > 0; --ii) {
    zext[j] = 2.0 * zG[0] - zG[ii];
    zext_w[j] = zW[ii];
    j += 1;
  }
  j = Nzeff - ext_down;
  for (int ii = Nz - 2; ii > Nz - 2 - ext_down; --ii) {
    zext[j] = -1.0 * zG[ii];
    zext_w[j] = zW[ii];
    j += 1;
  }
  *ext_up_out = ext_up;
  *ext_down_out = ext_down;
  return Nzeff;
}

/* first index e with zext[e] <= v (zext descending); std::lower_bound with std::greater */
static int lower_bound_desc(const double* zext, int n, double v)
{
  int lo = 0, hi = n;
  while (lo < hi) {
    int mid = (lo + hi) / 2;
    if (zext[mid] > v) lo = mid + 1; else hi = mid;
  }
  return lo;
}

/*
 * Locate particles; src/ParticleList.cpp:394-548.
 * xP: 3N positions already shifted to [0,Lx]x[0,Ly]x[0,Lz].
 * Outputs per particle: xclose, yclose, indl, indr (ext-grid z range),
 * offsets xunwrap[N*wx], yunwrap[N*wy], zunwrap[N*zstride], weights pt_wts[N*zstride].
 * Returns 0, or -1 if a particle is outside the unit cell.
 */
int orc_locate(int nP, const double* xP, double hx, double hy, double Lx, double Ly, double Lz,
               double thresh, const double* zext, const double* zext_w, int Nzeff, int zstride,
               int* wx_out, int* wy_out, int* xclose, int* yclose, int* indl, int* indr,
               double* xunwrap, double* yunwrap, double* zunwrap, double* pt_wts)
{
  const int wx = (int)round(2 * thresh / hx);
  const int wy = (int)round(2 * thresh / hy);
  *wx_out = wx; *wy_out = wy;
  for (int i = 0; i < nP; ++i) {
    const double x = xP[3 * i], y = xP[3 * i + 1], z = xP[3 * i + 2];
    if (x < 0 || x > Lx || y < 0 || y > Ly || z < 0 || z > Lz) return -1;
  }
  const int evenx = -1 * (wx % 2) + 1, eveny = -1 * (wy % 2) + 1;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nP; ++i) {
    const double x = xP[3 * i], y = xP[3 * i + 1], z = xP[3 * i + 2];
    /* z support; ParticleList.cpp:453-465 */
    int high = lower_bound_desc(zext, Nzeff, z - thresh);
    int low = lower_bound_desc(zext, Nzeff, z + thresh);
    int il = low, ir = high;
    if (ir == Nzeff) ir -= 1;
    else if (z - thresh > zext[ir]) ir -= 1;
    indl[i] = il; indr[i] = ir;
    /* x/y support with clamp to +thresh; ParticleList.cpp:493-512.  The
       reference is built with -O3 -march=native (source/cpu/CMakeLists.txt:63),
       where GCC contracts (index * h - x) into one fused multiply-add.  The
       offset is rounded once, so fma() here.  It matters at the support edge:
       phi has a square-root singularity at |r| = alpha, and a one-ulp change
       of an offset landing there moves phi by ~1e-7 relative (positions on
       grid nodes, e.g. the reference's own x = L/2 mobility runs). */
    int xc = (int)(x / hx), yc = (int)(y / hy);
    xc += ((wx % 2) && (x / hx - xc > 1.0 / 2.0) ? 1 : 0);
    yc += ((wy % 2) && (y / hy - yc > 1.0 / 2.0) ? 1 : 0);
    xclose[i] = xc; yclose[i] = yc;
    for (int j = 0; j < wx; ++j) {
      double v = fma((double)xc + j - wx / 2 + evenx, hx, -x);
      if (v > thresh || v < -thresh) v = thresh;
      xunwrap[j + i * wx] = v;
    }
    for (int j = 0; j < wy; ++j) {
      double v = fma((double)yc + j - wy / 2 + eveny, hy, -y);
      if (thresh < v || v < -thresh) v = thresh;
      yunwrap[j + i * wy] = v;
    }
    /* z offsets and weights; ParticleList.cpp:513-523 */
    int k = 0;
    for (int e = il; e <= ir; ++e) {
      double v = zext[e] - z;
      if (thresh < v || v < -thresh) v = thresh;
      zunwrap[k + i * zstride] = v;
      pt_wts[k + i * zstride] = hx * hy * zext_w[e];
      k += 1;
    }
  }
  return 0;
}

/* ES kernel (kind 0) or its x/y/z derivative (kind 1/2/3); include/ESKernels.h:8-46 */
static inline double kern3(int kind, const double x[3], double beta, double alpha)
{
  const double alpha2 = alpha * alpha;
  const double e = exp(beta * (sqrt(1 - x[0] * x[0] / alpha2) + sqrt(1 - x[1] * x[1] / alpha2) +
                               sqrt(1 - x[2] * x[2] / alpha2) - 3));
  if (kind == 0) return e;
  const int c = kind - 1;
  return -beta * x[c] * e / (alpha2 * sqrt(1 - x[c] * x[c] / (alpha2)));
}

/* Weights as functions of the particle position (no counterpart in the
   reference; the adjoint with respect to positions).  kind 1..3: first
   derivative d/do_k of the kernel; kind 10 + 3k + b: second derivative
   d2/(do_k do_b), the derivative along k of the forward weight
   phi d_b.  A coordinate whose offset was clamped to thresh is constant: a
   derivative along k is zero when o_k was clamped, while a clamped o_b keeps
   its value d_b(thresh) in the weight.  With
   s_i = sqrt(1 - o_i^2/alpha^2) and d_i = -beta o_i / (alpha^2 s_i):
   d/do_k = phi d_k, d2/(do_k do_b) = phi (d_k d_b + [k == b] d'_k),
   d'_k = -beta / (alpha^2 s_k^3). */
static inline double kern3_pos(int kind, const double x[3], double beta, double alpha, double thresh)
{
  int k, b;
  if (kind >= 10) { k = (kind - 10) / 3; b = (kind - 10) % 3; }
  else { k = kind - 1; b = -1; }
  if (fabs(x[k]) >= thresh) return 0.0;
  const double alpha2 = alpha * alpha;
  double sq[3], d[3];
  for (int i = 0; i < 3; ++i) {
    sq[i] = sqrt(1 - x[i] * x[i] / alpha2);
    d[i] = -beta * x[i] / (alpha2 * sq[i]);
  }
  const double e = exp(beta * (sq[0] + sq[1] + sq[2] - 3));
  if (b < 0) return e * d[k];
  double r = d[k] * d[b];
  if (k == b) r += -beta / (alpha2 * sq[k] * sq[k] * sq[k]);
  return e * r;
}

/* fold of the extended grid into the interior grid; include/BoundaryConditions.h:22-361
   (x and y are always periodic on this path) */
static void fold(double* Fe, double* Fw, unsigned Nxw, unsigned Nyw, unsigned Nzw,
                 unsigned Nx, unsigned Ny, unsigned Nz, unsigned dof, const int* BCs)
{
  const unsigned lend = (Nx - Nxw) / 2, rbeg = Nx - lend;
  const unsigned bend = (Ny - Nyw) / 2, tbeg = Ny - bend;
  const unsigned dend = (Nz - Nzw) / 2, ubeg = Nz - dend;
#pragma omp parallel for collapse(2)
  for (unsigned k = 0; k < Nz; ++k)
    for (unsigned j = 0; j < Ny; ++j) {
      for (unsigned i = 0; i < lend; ++i)
        for (unsigned d = 0; d < dof; ++d)
          Fe[d + dof * at3(i + Nxw, j, k, Nx, Ny)] += Fe[d + dof * at3(i, j, k, Nx, Ny)];
      for (unsigned i = rbeg; i < Nx; ++i)
        for (unsigned d = 0; d < dof; ++d)
          Fe[d + dof * at3(i - Nxw, j, k, Nx, Ny)] += Fe[d + dof * at3(i, j, k, Nx, Ny)];
    }
#pragma omp parallel for
  for (unsigned k = 0; k < Nz; ++k) {
    for (unsigned j = 0; j < bend; ++j)
      for (unsigned i = 0; i < Nx; ++i)
        for (unsigned d = 0; d < dof; ++d)
          Fe[d + dof * at3(i, j + Nyw, k, Nx, Ny)] += Fe[d + dof * at3(i, j, k, Nx, Ny)];
    for (unsigned j = tbeg; j < Ny; ++j)
      for (unsigned i = 0; i < Nx; ++i)
        for (unsigned d = 0; d < dof; ++d)
          Fe[d + dof * at3(i, j - Nyw, k, Nx, Ny)] += Fe[d + dof * at3(i, j, k, Nx, Ny)];
  }
  /* z ends: left = top (index 0), right = bottom; BoundaryConditions.h:284-341 */
  for (unsigned d = 0; d < dof; ++d) {
    const int bl = BCs[4 * dof + d];
    if (bl != BC_NONE) {
      const double s = (bl == BC_MIRROR) ? 1.0 : -1.0;
      for (unsigned k = 0; k <= dend; ++k)
        for (unsigned j = 0; j < Ny; ++j)
          for (unsigned i = 0; i < Nx; ++i)
            Fe[d + dof * at3(i, j, 2 * dend - k, Nx, Ny)] += s * Fe[d + dof * at3(i, j, k, Nx, Ny)];
    }
    const int br = BCs[5 * dof + d];
    if (br != BC_NONE) {
      const double s = (br == BC_MIRROR) ? 1.0 : -1.0;
      for (unsigned k = ubeg - 1; k < Nz; ++k)
        for (unsigned j = 0; j < Ny; ++j)
          for (unsigned i = 0; i < Nx; ++i)
            Fe[d + dof * at3(i, j, 2 * ubeg - k - 2, Nx, Ny)] += s * Fe[d + dof * at3(i, j, k, Nx, Ny)];
    }
  }
  /* permute to interior layout (Ny, Nx, Nz, dof); BoundaryConditions.h:342-360 */
#pragma omp parallel for collapse(2)
  for (unsigned j = bend; j < tbeg; ++j)
    for (unsigned i = lend; i < rbeg; ++i)
      for (unsigned k = dend; k < ubeg; ++k)
        for (unsigned d = 0; d < dof; ++d)
          Fw[d + dof * at3(k - dend, i - lend, j - bend, Nzw, Nxw)] = Fe[d + dof * at3(i, j, k, Nx, Ny)];
}

/* copy interior -> extended with ghosts; include/BoundaryConditions.h:377-731.
   Ghost regions at a 'none' end keep what Fe held before (the caller decides). */
static void copy_ghost(double* Fe, const double* Fw, unsigned Nxw, unsigned Nyw, unsigned Nzw,
                       unsigned Nx, unsigned Ny, unsigned Nz, unsigned dof, const int* BCs)
{
  const unsigned lend = (Nx - Nxw) / 2, rbeg = Nx - lend;
  const unsigned bend = (Ny - Nyw) / 2, tbeg = Ny - bend;
  const unsigned dend = (Nz - Nzw) / 2, ubeg = Nz - dend;
#pragma omp parallel for collapse(2)
  for (unsigned k = 0; k < Nzw; ++k)
    for (unsigned j = 0; j < Nyw; ++j)
      for (unsigned i = 0; i < Nxw; ++i)
        for (unsigned d = 0; d < dof; ++d)
          Fe[d + dof * at3(i + lend, j + bend, k + dend, Nx, Ny)] = Fw[d + dof * at3(k, i, j, Nzw, Nxw)];
#pragma omp parallel for collapse(2)
  for (unsigned k = dend; k < ubeg; ++k)
    for (unsigned j = bend; j < tbeg; ++j) {
      for (unsigned i = 0; i < lend; ++i)
        for (unsigned d = 0; d < dof; ++d)
          Fe[d + dof * at3(i, j, k, Nx, Ny)] = Fe[d + dof * at3(i + Nxw, j, k, Nx, Ny)];
      for (unsigned i = rbeg; i < Nx; ++i)
        for (unsigned d = 0; d < dof; ++d)
          Fe[d + dof * at3(i, j, k, Nx, Ny)] = Fe[d + dof * at3(i - Nxw, j, k, Nx, Ny)];
    }
#pragma omp parallel for
  for (unsigned k = dend; k < ubeg; ++k) {
    for (unsigned j = tbeg; j < Ny; ++j)
      for (unsigned i = 0; i < Nx; ++i)
        for (unsigned d = 0; d < dof; ++d)
          Fe[d + dof * at3(i, j, k, Nx, Ny)] = Fe[d + dof * at3(i, j - Nyw, k, Nx, Ny)];
    for (unsigned j = 0; j < bend; ++j)
      for (unsigned i = 0; i < Nx; ++i)
        for (unsigned d = 0; d < dof; ++d)
          Fe[d + dof * at3(i, j, k, Nx, Ny)] = Fe[d + dof * at3(i, j + Nyw, k, Nx, Ny)];
  }
  for (unsigned d = 0; d < dof; ++d) {
    const int bl = BCs[4 * dof + d];
    if (bl != BC_NONE) {
      const double s = (bl == BC_MIRROR) ? 1.0 : -1.0;
      for (unsigned k = 0; k < dend; ++k)
        for (unsigned j = 0; j < Ny; ++j)
          for (unsigned i = 0; i < Nx; ++i)
            Fe[d + dof * at3(i, j, k, Nx, Ny)] = s * Fe[d + dof * at3(i, j, 2 * dend - k, Nx, Ny)];
      for (unsigned j = 0; j < Ny; ++j)
        for (unsigned i = 0; i < Nx; ++i)
          Fe[d + dof * at3(i, j, dend, Nx, Ny)] += s * Fe[d + dof * at3(i, j, dend, Nx, Ny)];
    }
    const int br = BCs[5 * dof + d];
    if (br != BC_NONE) {
      const double s = (br == BC_MIRROR) ? 1.0 : -1.0;
      for (unsigned k = ubeg - 1; k < Nz; ++k)
        for (unsigned j = 0; j < Ny; ++j)
          for (unsigned i = 0; i < Nx; ++i)
            Fe[d + dof * at3(i, j, k, Nx, Ny)] = s * Fe[d + dof * at3(i, j, 2 * ubeg - k - 2, Nx, Ny)];
      for (unsigned j = 0; j < Ny; ++j)
        for (unsigned i = 0; i < Nx; ++i)
          Fe[d + dof * at3(i, j, ubeg - 1, Nx, Ny)] += s * Fe[d + dof * at3(i, j, ubeg - 1, Nx, Ny)];
    }
  }
}

typedef struct {
  int nP, wx, wy, zstride;
  const int *xclose, *yclose, *indl, *indr;
  const double *xunwrap, *yunwrap, *zunwrap, *pt_wts;
  int Nx, Ny, Nz, Nxeff, Nyeff, Nzeff;
} orc_locs;

/* bucket particles by extended-grid column; src/ParticleList.cpp:527-545 */
static void bucket(const orc_locs* L, int* first, int* next)
{
  const int N2 = L->Nxeff * L->Nyeff;
  for (int c = 0; c < N2; ++c) first[c] = -1;
  int* last = (int*)malloc(sizeof(int) * N2);
  for (int i = 0; i < L->nP; ++i) {
    const int col = (L->yclose[i] + (L->Nyeff - L->Ny) / 2) + (L->xclose[i] + (L->Nxeff - L->Nx) / 2) * L->Nyeff;
    next[i] = -1;
    if (first[col] < 0) first[col] = i; else next[last[col]] = i;
    last[col] = i;
  }
  free(last);
}

static void spread_ext(int kind, const orc_locs* L, const double* fP, double betaw, double alpha,
                       double norm, double* Fe)
{
  const int wx = L->wx, wy = L->wy;
  const int evenx = -1 * (wx % 2) + 1, eveny = -1 * (wy % 2) + 1;
  const int Nxeff = L->Nxeff, Nyeff = L->Nyeff;
  const int N2 = Nxeff * Nyeff;
  int* first = (int*)malloc(sizeof(int) * N2);
  int* next = (int*)malloc(sizeof(int) * (L->nP > 0 ? L->nP : 1));
  bucket(L, first, next);
  const double norm3 = norm * (norm * norm);
  /* colour classes of columns: SpreadInterp.cpp:313-321 */
  for (int izero = 0; izero < wx; ++izero)
    for (int jzero = 0; jzero < wy; ++jzero) {
#pragma omp parallel for collapse(2) schedule(dynamic, 16)
      for (int ii = izero; ii < Nxeff; ii += wx)
        for (int jj = jzero; jj < Nyeff; jj += wy) {
          for (int p = first[jj + ii * Nyeff]; p >= 0; p = next[p]) {
            const double* f = fP + 3 * p;
            const int wz = L->indr[p] - L->indl[p] + 1;
            for (int k = 0; k < wz; ++k)
              for (int j = 0; j < wy; ++j)
                for (int i = 0; i < wx; ++i) {
                  double x[3];
                  x[0] = L->xunwrap[i + p * wx];
                  x[1] = L->yunwrap[j + p * wy];
                  x[2] = L->zunwrap[k + p * L->zstride];
                  const double delta = kern3(kind, x, betaw, alpha) / norm3;
                  const unsigned g = at3(ii + i - wx / 2 + evenx, jj + j - wy / 2 + eveny,
                                         L->indl[p] + k, Nxeff, Nyeff);
                  for (int d = 0; d < 3; ++d) Fe[d + 3 * g] += delta * f[d];
                }
          }
        }
    }
  free(first); free(next);
}

/*
 * Spread (kind 0) or derivative-spread (kind 1..3) 3-vectors fP onto the grid
 * and fold with the given BCs.  Writes the interior grid fG (Ny, Nx, Nz, 3).
 * If Fe_keep is non-null it receives the extended grid after the fold (the
 * reference keeps it, and reads its 'none'-end ghost region again in
 * interpolation).
 */
void orc_spread(int kind, int nP, const double* fP, double betaw, double alpha, double norm,
                int wx, int wy, const int* xclose, const int* yclose, const int* indl,
                const int* indr, const double* xunwrap, const double* yunwrap,
                const double* zunwrap, int zstride, int Nx, int Ny, int Nz, int Nzeff,
                const int* BCs, double* fG, double* Fe_keep)
{
  orc_locs L = {nP, wx, wy, zstride, xclose, yclose, indl, indr, xunwrap, yunwrap, zunwrap, 0,
                Nx, Ny, Nz, Nx + wx + (wx % 2), Ny + wy + (wy % 2), Nzeff};
  const size_t n3 = (size_t)L.Nxeff * L.Nyeff * L.Nzeff * 3;
  double* Fe = Fe_keep ? Fe_keep : (double*)malloc(sizeof(double) * n3);
  memset(Fe, 0, sizeof(double) * n3);
  spread_ext(kind, &L, fP, betaw, alpha, norm, Fe);
  fold(Fe, fG, Nx, Ny, Nz, L.Nxeff, L.Nyeff, L.Nzeff, 3, BCs);
  if (!Fe_keep) free(Fe);
}

/*
 * Interpolate (kind 0) or derivative-interpolate (kind 1..3) the interior
 * grid fG (Ny, Nx, Nz, 3) onto the particles, writing uP (3N).  Fe_io is the
 * extended grid: its ghost region at a 'none' end is read as-is (the reference
 * reads whatever the last spread left there); pass a zeroed buffer to read 0.
 */
static void interp_impl(int kind, int posmode, double thresh, int nP, const double* fG, double betaw, double alpha,
                        double norm, int wx, int wy, const int* xclose, const int* yclose, const int* indl,
                        const int* indr, const double* xunwrap, const double* yunwrap,
                        const double* zunwrap, const double* pt_wts, int zstride, int Nx, int Ny,
                        int Nz, int Nzeff, const int* BCs, double* Fe_io, double* uP);

void orc_interp(int kind, int nP, const double* fG, double betaw, double alpha, double norm,
                int wx, int wy, const int* xclose, const int* yclose, const int* indl,
                const int* indr, const double* xunwrap, const double* yunwrap,
                const double* zunwrap, const double* pt_wts, int zstride, int Nx, int Ny,
                int Nz, int Nzeff, const int* BCs, double* Fe_io, double* uP)
{
  interp_impl(kind, 0, 0.0, nP, fG, betaw, alpha, norm, wx, wy, xclose, yclose, indl, indr, xunwrap, yunwrap,
              zunwrap, pt_wts, zstride, Nx, Ny, Nz, Nzeff, BCs, Fe_io, uP);
}

/* Interpolation with position-derivative weights (kern3_pos kinds). */
void orc_interp_pos(int kind, double thresh, int nP, const double* fG, double betaw, double alpha, double norm,
                    int wx, int wy, const int* xclose, const int* yclose, const int* indl,
                    const int* indr, const double* xunwrap, const double* yunwrap,
                    const double* zunwrap, const double* pt_wts, int zstride, int Nx, int Ny,
                    int Nz, int Nzeff, const int* BCs, double* Fe_io, double* uP)
{
  interp_impl(kind, 1, thresh, nP, fG, betaw, alpha, norm, wx, wy, xclose, yclose, indl, indr, xunwrap, yunwrap,
              zunwrap, pt_wts, zstride, Nx, Ny, Nz, Nzeff, BCs, Fe_io, uP);
}

static void interp_impl(int kind, int posmode, double thresh, int nP, const double* fG, double betaw, double alpha,
                        double norm, int wx, int wy, const int* xclose, const int* yclose, const int* indl,
                        const int* indr, const double* xunwrap, const double* yunwrap,
                        const double* zunwrap, const double* pt_wts, int zstride, int Nx, int Ny,
                        int Nz, int Nzeff, const int* BCs, double* Fe_io, double* uP)
{
  const int Nxeff = Nx + wx + (wx % 2), Nyeff = Ny + wy + (wy % 2);
  const int evenx = -1 * (wx % 2) + 1, eveny = -1 * (wy % 2) + 1;
  copy_ghost(Fe_io, fG, Nx, Ny, Nz, Nxeff, Nyeff, Nzeff, 3, BCs);
  const double norm3 = norm * (norm * norm);
#pragma omp parallel for schedule(dynamic, 64)
  for (int p = 0; p < nP; ++p) {
    const int ii = xclose[p] + (Nxeff - Nx) / 2, jj = yclose[p] + (Nyeff - Ny) / 2;
    const int wz = indr[p] - indl[p] + 1;
    double acc[3] = {0, 0, 0};
    for (int k = 0; k < wz; ++k)
      for (int j = 0; j < wy; ++j)
        for (int i = 0; i < wx; ++i) {
          double x[3];
          x[0] = xunwrap[i + p * wx];
          x[1] = yunwrap[j + p * wy];
          x[2] = zunwrap[k + p * zstride];
          const double delta = (posmode ? kern3_pos(kind, x, betaw, alpha, thresh) : kern3(kind, x, betaw, alpha)) /
                               norm3;
          const unsigned g = at3(ii + i - wx / 2 + evenx, jj + j - wy / 2 + eveny, indl[p] + k,
                                 Nxeff, Nyeff);
          const double wgt = pt_wts[k + p * zstride];
          for (int d = 0; d < 3; ++d) acc[d] += Fe_io[d + 3 * g] * delta * wgt;
        }
    for (int d = 0; d < 3; ++d) uP[3 * p + d] = acc[d];
  }
}

/* off[i*w + j] = clamp(fma(close[i] + j - w/2 + even, h, -x[i])): the
   uniform-direction offsets of ParticleList.cpp:339-368 (locateOnGridUnifZ),
   rounded once as the reference build computes them (see orc_locate). */
void orc_fma_offsets(int n, int w, const double* x, const long long* close, double h, double thresh,
                     double* off)
{
  const int even = -1 * (w % 2) + 1;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < w; ++j) {
      double v = fma((double)close[i] + j - w / 2 + even, h, -x[i]);
      if (v > thresh || v < -thresh) v = thresh;
      off[i * w + j] = v;
    }
}

int orc_num_threads(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
