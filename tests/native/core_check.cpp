// Host-side harness over the solver's shared core (dps_core.h / dps_tables.h):
// runs the exact per-mode BVP, stencil and table code that the gfx950 kernels
// run, on the CPU, so tests can compare it with the FP64 oracle without a GPU.
// Test infrastructure only (built by tests/conftest.py with g++).
#include <barrier>
#include <cstring>
#include <thread>
#include <algorithm>
#include <vector>
#include "dps_core.h"
#include "dps_fft.h"
#include "dps_seg.h"
#include "dps_tables.h"

using namespace dps;

namespace {
struct HostIn {
  const double* p;   // [3][Nz][2]
  int Nz;
  bool conj;
  cplx operator()(int c, int k) const
  {
    const double* q = p + 2 * (c * Nz + k);
    return conj ? mk(q[0], -q[1]) : mk(q[0], q[1]);
  }
};
struct HostOut {
  double* p;
  int Nz;
  void operator()(int c, int k, cplx v)
  {
    double* q = p + 2 * (c * Nz + k);
    q[0] = v.re;
    q[1] = v.im;
  }
};

// Lane exchange for the group solver: one host thread per lane, every
// exchange is a write, a barrier, a read and a barrier.
struct HostGroup {
  std::barrier<> bar;
  double buf[64];
  explicit HostGroup(int n) : bar(n) {}
};
struct HostLanes {
  int l, n;
  HostGroup* g;
  double exch(double v, int src) const
  {
    g->buf[l] = v;
    g->bar.arrive_and_wait();
    const double r = g->buf[src];
    g->bar.arrive_and_wait();
    return r;
  }
  template <int D> double up(double v, double fill) const
  {
    const double r = exch(v, l - D >= 0 ? l - D : l);
    return l - D >= 0 ? r : fill;
  }
  template <int D> double down(double v, double fill) const
  {
    const double r = exch(v, l + D < n ? l + D : l);
    return l + D < n ? r : fill;
  }
  template <int D> double xr(double v) const { return exch(v, l ^ D); }
};
struct SegIn {
  const double* p;   // [3][Nz][2] of this mode
  int Nz, r0;
  cplx operator()(int c, int q) const
  {
    const int row = r0 + q;
    if (row >= Nz) return mk(0, 0);
    const double* v = p + 2 * (c * Nz + row);
    return mk(v[0], v[1]);
  }
};
struct SegOut {
  double* p;
  int Nz, r0;
  void operator()(int c, int q, cplx v)
  {
    if (r0 + q >= Nz) return;   // padding rows
    double* o = p + 2 * (c * Nz + r0 + q);
    o[0] = v.re;
    o[1] = v.im;
  }
};

// Copies of the row tables of B readable at every row a lane group touches,
// zero from row Nz on (what seg_bvp_mode requires).
struct PaddedTables {
  std::vector<double> t[9];
  BvpConsts B;
  PaddedTables(const BvpConsts& B0, int rows)
  {
    B = B0;
    const double* src[9] = {B0.sd0, B0.sm2, B0.sp2, B0.fm1, B0.fp1, B0.R1, B0.R2, B0.L1, B0.L2};
    const int n = std::max(rows, B0.Nz + 2);
    for (int k = 0; k < 9; ++k) {
      t[k].assign(n, 0.0);
      for (int r = 0; r < B0.Nz; ++r) t[k][r] = src[k][r];
    }
    B.sd0 = t[0].data(); B.sm2 = t[1].data(); B.sp2 = t[2].data(); B.fm1 = t[3].data();
    B.fp1 = t[4].data(); B.R1 = t[5].data(); B.R2 = t[6].data(); B.L1 = t[7].data(); B.L2 = t[8].data();
  }
};

// STAGE: fin and fout address one buffer (the forcing is copied to uout
// first), as in the kernel's in-place column.
template <int LPM, int E, int NRV, bool STAGE>
void run_seg(const BvpConsts& B, int M, const double* kx, const double* ky, const int* nyqx,
             const int* nyqy, const double* fin, double* uout, double* corr)
{
  const int Nz = B.Nz;
  const PaddedTables pad(B, LPM * E);
  std::vector<double> ibv(Nz + 1), abv(Nz);
  std::vector<cplx> dummy(Nz);
  ScratchView sc;
  sc.ib = ibv.data(); sc.ab = abv.data(); sc.stride = 1;
  sc.P = sc.CP = sc.DP = sc.U0 = sc.U1 = sc.U2 = dummy.data();
  for (int m = 0; m < M; ++m) {
    const bool k0 = (kx[m] == 0 && ky[m] == 0);
    double Gi[4] = {0, 0, 0, 0};
    if (!k0) bvp_factor(B, kx[m] * kx[m] + ky[m] * ky[m], sc, Gi);
    HostGroup grp(LPM);
    ModeResult res0;
    if (STAGE) std::copy(fin + (size_t)m * 6 * Nz, fin + (size_t)(m + 1) * 6 * Nz, uout + (size_t)m * 6 * Nz);
    std::vector<std::thread> th;
    for (int l = 0; l < LPM; ++l)
      th.emplace_back([&, l]() {
        HostLanes ln{l, LPM, &grp};
        double ib[E], ab[E];
        for (int q = 0; q < E; ++q) {
          const int row = l * E + q;
          ib[q] = (row < Nz && !k0) ? ibv[row + 1] : 0.0;
          ab[q] = (row < Nz && !k0) ? abv[row] : 0.0;
        }
        SegIn in{(STAGE ? uout : fin) + (size_t)m * 6 * Nz, Nz, l * E};
        SegOut out{uout + (size_t)m * 6 * Nz, Nz, l * E};
        ModeResult r;
        for (int c = 0; c < 3; ++c)
          for (int b = 0; b < 4; ++b) r.corr[c][b] = mk(0, 0);
        r.nbasis = 0;
        if (k0) seg_bvp_k0<LPM, E>(B, ln, in, out);
        else seg_bvp_mode<LPM, E, NRV, STAGE>(pad.B, ln, kx[m], ky[m], nyqx[m], nyqy[m], ib, ab, Gi, in, out, r);
        if (l == 0) res0 = r;
      });
    for (auto& t : th) t.join();
    for (int c = 0; c < 3; ++c)
      for (int b = 0; b < 4; ++b) {
        corr[(size_t)m * 24 + (c * 4 + b) * 2] = k0 ? 0.0 : res0.corr[c][b].re;
        corr[(size_t)m * 24 + (c * 4 + b) * 2 + 1] = k0 ? 0.0 : res0.corr[c][b].im;
      }
  }
}

BvpConsts host_consts(const ChebTables& T, int Nz, double H, double Hs, double eta, int geometry)
{
  BvpConsts B;
  B.Nz = Nz; B.geometry = geometry; B.H = H; B.Hs = Hs; B.eta = eta;
  B.sd0 = T.sd0.data(); B.sm2 = T.sm2.data(); B.sp2 = T.sp2.data();
  B.S0N = T.S0N; B.S1N1 = T.S1N1;
  B.fm1 = T.fm1.data(); B.fp1 = T.fp1.data(); B.F0N1 = T.F0N1;
  B.R1 = T.R1.data(); B.R2 = T.R2.data(); B.L1 = T.L1.data(); B.L2 = T.L2.data();
  return B;
}
}  // namespace

extern "C" {

// Chebyshev bands / BC rows: out arrays of length Nz (bands) and Nz+2 (rows)
void cc_cheb_tables(int Nz, double H, double* sd0, double* sm2, double* sp2, double* fm1, double* fp1,
                    double* scal /*S0N,S1N1,F0N1*/, double* R1, double* R2, double* L1, double* L2)
{
  ChebTables T = build_cheb_tables(Nz, H);
  std::memcpy(sd0, T.sd0.data(), Nz * 8);
  std::memcpy(sm2, T.sm2.data(), Nz * 8);
  std::memcpy(sp2, T.sp2.data(), Nz * 8);
  std::memcpy(fm1, T.fm1.data(), Nz * 8);
  std::memcpy(fp1, T.fp1.data(), Nz * 8);
  scal[0] = T.S0N; scal[1] = T.S1N1; scal[2] = T.F0N1;
  std::memcpy(R1, T.R1.data(), (Nz + 2) * 8);
  std::memcpy(R2, T.R2.data(), (Nz + 2) * 8);
  std::memcpy(L1, T.L1.data(), (Nz + 2) * 8);
  std::memcpy(L2, T.L2.data(), (Nz + 2) * 8);
}

// dense forward/inverse Chebyshev matrices rebuilt from the folded tables
void cc_cheb_transform(int N, double* Tf /*N*N*/, double* Ti /*N*N*/)
{
  ChebTransform C = build_cheb_transform(N);
  for (int k = 0; k < N; ++k)
    for (int j = 0; j < N; ++j) {
      const int jj = (j < C.NH) ? j : N - 1 - j;
      const double sgn = (j < C.NH) ? 1.0 : ((k % 2) ? -1.0 : 1.0);
      const double f = (k % 2) ? C.TfO[(size_t)(k / 2) * C.NH + jj] : C.TfE[(size_t)(k / 2) * C.NH + jj];
      const double i = (k % 2) ? C.TiO[(size_t)jj * C.NO + k / 2] : C.TiE[(size_t)jj * C.NE + k / 2];
      Tf[k * N + j] = sgn * f;
      Ti[j * N + k] = sgn * i;
    }
}

// register DFT of the y-pass kernel: R in {2, 4, 8, 16, 32}, dir -1 / +1
int cc_fft_reg(int R, int dir, double* re, double* im)
{
#define CC_FFT(RR)                                                    \
  if (R == RR) {                                                      \
    double a[RR], b[RR];                                              \
    for (int i = 0; i < RR; ++i) { a[i] = re[i]; b[i] = im[i]; }      \
    if (dir < 0) fft_reg<RR, -1>(a, b);                               \
    else fft_reg<RR, 1>(a, b);                                        \
    for (int i = 0; i < RR; ++i) { re[i] = a[i]; im[i] = b[i]; }      \
    return 0;                                                         \
  }
  CC_FFT(2) CC_FFT(4) CC_FFT(8) CC_FFT(16) CC_FFT(32)
#undef CC_FFT
  return -1;
}

double cc_es_norm(double alpha, double betaw) { return es_norm(alpha, betaw); }
double cc_dipole_thresh(double alpha, double betaw) { return dipole_thresh(alpha, betaw); }

// z-extended grid tables; arrays sized 3*Nz
int cc_zext(const double* zG, const double* zW, int Nz, double thresh, int wall_top, int wall_bot,
            double* z, double* w, int* kt, double* facA, double* facB, int* info /*up,down,wzmax*/)
{
  std::vector<double> g(zG, zG + Nz), ww(zW, zW + Nz);
  ZExt X = build_zext(g, ww, thresh, wall_top != 0, wall_bot != 0);
  for (int e = 0; e < X.nzeff; ++e) {
    z[e] = X.z[e]; w[e] = X.w[e]; kt[e] = X.kt[e]; facA[e] = X.facA[e]; facB[e] = X.facB[e];
  }
  info[0] = X.ext_up; info[1] = X.ext_down; info[2] = X.wzmax;
  return X.nzeff;
}

// per-particle supports and weights for N particles (positions already in box)
void cc_stencil(int N, const double* pos, double hx, double hy, double betaw, double alpha, double thresh,
                double inv_norm3, int wx, int wy, const double* zext, const double* zwext, int nzeff,
                int wzmax, int* sup /*4N*/, double* wxo, double* dxo, double* wyo, double* dyo,
                double* wzo, double* dzo, double* wzio, double* dzio)
{
  KernelDesc K{};
  K.betaw = betaw; K.alpha = alpha; K.thresh = thresh; K.inv_norm3 = inv_norm3;
  K.wx = wx; K.wy = wy; K.nzeff = nzeff; K.ext_up = 0; K.wzmax = wzmax;
  for (int p = 0; p < N; ++p) {
    const double x = pos[3 * p], y = pos[3 * p + 1], z = pos[3 * p + 2];
    Stencil s = stencil_support(x, y, z, hx, hy, K, zext);
    sup[4 * p] = s.bx; sup[4 * p + 1] = s.by; sup[4 * p + 2] = s.e0; sup[4 * p + 3] = s.nz;
    stencil_weights_xy(x, hx, s.bx, wx, K, inv_norm3, wxo + p * wx, dxo + p * wx);
    stencil_weights_xy(y, hy, s.by, wy, K, 1.0, wyo + p * wy, dyo + p * wy);
    stencil_weights_z(z, s, K, zext, zwext, hx * hy, wzo + p * wzmax, dzo + p * wzmax,
                      wzio + p * wzmax, dzio + p * wzmax);
  }
}

// BVP for M modes given their wavenumbers.  fin/uout: [M][3][Nz][2];
// corr: [M][3][4][2].  Mode with kx = ky = 0 runs the k = 0 branch.
void cc_bvp(int Nz, double H, double Hs, double eta, int geometry, int M, const double* kx,
            const double* ky, const int* nyqx, const int* nyqy, const double* fin, double* uout,
            double* corr)
{
  ChebTables T = build_cheb_tables(Nz, H);
  BvpConsts B = host_consts(T, Nz, H, Hs, eta, geometry);
  std::vector<double> ib(Nz + 1), ab(Nz);
  std::vector<cplx> P(Nz), CP(Nz), DP(Nz), U0(Nz), U1(Nz), U2(Nz);
  ScratchView sc;
  sc.ib = ib.data(); sc.ab = ab.data(); sc.P = P.data(); sc.CP = CP.data(); sc.DP = DP.data();
  sc.U0 = U0.data(); sc.U1 = U1.data(); sc.U2 = U2.data(); sc.stride = 1;
  for (int m = 0; m < M; ++m) {
    HostIn in{fin + (size_t)m * 6 * Nz, Nz, false};
    HostOut out{uout + (size_t)m * 6 * Nz, Nz};
    ModeResult r;
    for (int c = 0; c < 3; ++c)
      for (int b = 0; b < 4; ++b) r.corr[c][b] = mk(0, 0);
    if (kx[m] == 0 && ky[m] == 0) bvp_k0(B, in, out, U0.data(), 1);
    else bvp_mode(B, kx[m], ky[m], nyqx[m], nyqy[m], in, out, sc, r);
    for (int c = 0; c < 3; ++c)
      for (int b = 0; b < 4; ++b) {
        corr[(size_t)m * 24 + (c * 4 + b) * 2] = r.corr[c][b].re;
        corr[(size_t)m * 24 + (c * 4 + b) * 2 + 1] = r.corr[c][b].im;
      }
  }
}

// The same modes through the lane-group solver with LPM lanes of E rows
// (LPM * E >= Nz), nrv velocity components per scan.  Returns -1 for an
// unsupported shape.
int cc_bvp_seg(int Nz, double H, double Hs, double eta, int geometry, int M, const double* kx,
               const double* ky, const int* nyqx, const int* nyqy, const double* fin, double* uout,
               double* corr, int lpm, int e, int nrv, int stage)
{
  ChebTables T = build_cheb_tables(Nz, H);
  BvpConsts B = host_consts(T, Nz, H, Hs, eta, geometry);
  if (lpm * e < Nz) return -1;
#define SEG_RUN(LP, EE, NR, ST) \
  if (lpm == LP && e == EE && nrv == NR && stage == ST) { run_seg<LP, EE, NR, ST>(B, M, kx, ky, nyqx, nyqy, fin, uout, corr); return 0; }
#define SEG_CASE(LP, EE) SEG_RUN(LP, EE, 3, 0) SEG_RUN(LP, EE, 1, 0) SEG_RUN(LP, EE, 1, 1) SEG_RUN(LP, EE, 3, 1)
  SEG_CASE(4, 2) SEG_CASE(4, 4) SEG_CASE(8, 4) SEG_CASE(16, 2) SEG_CASE(16, 4) SEG_CASE(16, 6)
  SEG_CASE(16, 8) SEG_CASE(32, 8)
#undef SEG_CASE
#undef SEG_RUN
  return -1;
}
}
