"""FP64 CPU oracle for the doubly periodic (DP) Stokes ``Mdot`` path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and
the CPU-baseline leg of ``bench.py`` may import this module, and only as the
checker; the product (``doublyperiodicstokes_amd``) never imports it.

This is a restatement of the reference's CPU algorithm (paths relative to the
reference tree):

* ``source/cpu/python/FCM.py:149-441``      -- set-up, BC tables, Mdot sequence
* ``source/cpu/python/Chebyshev.py:4-160``  -- Clenshaw-Curtis, Chebyshev
  derivative / integral matrices
* ``source/cpu/python/Solvers.py:258-643``  -- DP precompute, per-geometry
  solves, k=0 handling, BC rows
* ``source/cpu/src/LinearSolvers.cpp:11-60`` and
  ``include/LinearSolvers.h:52-141``        -- KBPENTA and Schur precompute
* ``source/cpu/src/DPTools.cpp:81-877``     -- per-wavenumber BVP chain, wall
  corrections
* ``source/cpu/src/Transform.cpp:107-304``  -- Fourier-Chebyshev transform
* particle <-> grid stages: ``oracle/spread_interp.c`` (C, OpenMP)

Library substitutions (the reference's libraries are absent here): FFTW ->
``scipy.fft`` (pocketfft); ``LAPACKE_zgesv`` -> ``numpy.linalg.solve``
(LAPACK gesv, partial pivoting); ``cblas_dgemm`` -> numpy.  Every per-mode
loop of the reference is vectorised over modes; the z recurrences keep the
reference's order of operations.

Known reference behaviours reproduced on purpose: the clamp of out-of-support
kernel offsets to ``+thresh``; the ``(1+s)`` factor on wall nodes in the
ghost fold/copy; Nyquist zeroing only in ``Dx``/``Dy``; the slit k=0 forcing
built from the real part twice (``Solvers.py:564-566``, harmless because only
the real part of the k=0 mode survives the back-transform).

Parity pins live in ``tests/test_oracle_pins.py`` and
``tests/test_oracle_golden.py``.
"""
import ctypes
import os
import numpy as np
import scipy.fft as sfft

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBPATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

BC_MIRROR, BC_MIRROR_INV, BC_NONE = 0, 1, 2


def build_oracle_lib(force=False):
    """Compile ``spread_interp.c`` with gcc/OpenMP into ``oracle/_build``."""
    import subprocess
    if os.path.exists(_LIBPATH) and not force:
        src = os.path.join(_HERE, "spread_interp.c")
        if os.path.getmtime(_LIBPATH) >= os.path.getmtime(src):
            return _LIBPATH
    subprocess.check_call(["make", "-s", "-C", _HERE])
    return _LIBPATH


def lib():
    global _lib
    if _lib is None:
        build_oracle_lib()
        L = ctypes.CDLL(_LIBPATH)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.orc_clencurt.argtypes = [dp, dp, ctypes.c_double, ctypes.c_double, ctypes.c_uint]
        L.orc_es_norm.argtypes = [ctypes.c_double, ctypes.c_double]
        L.orc_es_norm.restype = ctypes.c_double
        L.orc_dipole_thresh.argtypes = [ctypes.c_double, ctypes.c_double]
        L.orc_dipole_thresh.restype = ctypes.c_double
        L.orc_zext.argtypes = [dp, dp, ctypes.c_int, ctypes.c_double, dp, dp, ip, ip]
        L.orc_zext.restype = ctypes.c_int
        L.orc_locate.argtypes = [ctypes.c_int, dp, ctypes.c_double, ctypes.c_double,
                                 ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                 ctypes.c_double, dp, dp, ctypes.c_int, ctypes.c_int,
                                 ip, ip, ip, ip, ip, ip, dp, dp, dp, dp]
        L.orc_locate.restype = ctypes.c_int
        L.orc_spread.argtypes = [ctypes.c_int, ctypes.c_int, dp, ctypes.c_double,
                                 ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                 ip, ip, ip, ip, dp, dp, dp, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_int, ctypes.c_int, ctypes.c_int, ip, dp, dp]
        L.orc_interp.argtypes = [ctypes.c_int, ctypes.c_int, dp, ctypes.c_double,
                                 ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                 ip, ip, ip, ip, dp, dp, dp, dp, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_int, ctypes.c_int, ctypes.c_int, ip, dp, dp]
        L.orc_interp_pos.argtypes = [ctypes.c_int, ctypes.c_double] + L.orc_interp.argtypes[1:]
        L.orc_num_threads.restype = ctypes.c_int
        L.orc_fma_offsets.argtypes = [ctypes.c_int, ctypes.c_int, dp,
                                      ctypes.POINTER(ctypes.c_longlong), ctypes.c_double,
                                      ctypes.c_double, dp]
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    if a.dtype == np.float64:
        return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


# ----------------------------------------------------------------------------
# Chebyshev tools (source/cpu/python/Chebyshev.py)
# ----------------------------------------------------------------------------
def clencurt(Np1, a, b):
    """Clenshaw-Curtis nodes (descending) and weights; Chebyshev.py:4-38."""
    N = Np1 - 1
    H = (b - a) / 2
    bpad2 = (b + a) / 2
    cwts = np.zeros((Np1,), dtype=np.double)
    if N % 2:
        w1 = H / (N * N)
        end = int((N - 1) / 2 + 1)
    else:
        w1 = H / (N * N - 1)
        end = int(N / 2)
    cpts = H * np.cos(np.pi * np.arange(0, Np1) / N) + bpad2
    cwts[0] = w1
    cwts[N] = w1
    v = np.ones((N - 1,), dtype=np.double)
    iis = np.arange(1, N)
    for k in range(1, end):
        v -= 2 * np.cos(2.0 * k * np.pi * iis / N) / (4.0 * k * k - 1.0)
    if N % 2 == 0:
        v -= np.cos(N * np.pi * iis / N) / (N * N - 1)
    cwts[1:N] = 2 * H * v / N
    return cpts, cwts


def cheb_coeff_diff_py(fhat, Nz, H):
    """z-derivative of Chebyshev coefficients along axis 0; Chebyshev.py:40-66."""
    Df = np.zeros(fhat.shape, dtype=np.complex128)
    Df[-2] = 2 / H * (Nz - 1) * fhat[-1]
    for j in range(2, Nz):
        Df[Nz - j - 1] = Df[Nz - j + 1] + 2 / H * (Nz - j) * fhat[Nz - j]
    Df[0] /= 2
    return Df


def curl_cheb(C, Dx, Dy, H):
    """Half the curl of a Fourier-Chebyshev field C (modes, Nz, 3): x and y
    derivatives by Dx, Dy per mode, the z derivative by the coefficient
    recurrence (FCM.py:372-396 for the torque grid, Solvers.py:159-179 for the
    velocity; Chebyshev.py:68-94)."""
    Nz = C.shape[1]

    def dz(a):
        return cheb_coeff_diff_py(a.T, Nz, H).T

    Dx = Dx[:, None]
    Dy = Dy[:, None]
    out = np.empty_like(C)
    out[..., 0] = 0.5 * (Dy * C[..., 2] - dz(C[..., 1]))
    out[..., 1] = 0.5 * (dz(C[..., 0]) - Dx * C[..., 2])
    out[..., 2] = 0.5 * (Dx * C[..., 1] - Dy * C[..., 0])
    return out


def first_integral_matrix(N, H):
    """First Chebyshev integral matrix (N x N+2), H-scaled; Chebyshev.py:119-137."""
    F = np.zeros((N, N + 2))
    jj = np.arange(2, N)
    colm1 = np.concatenate(([1], 1 / (2 * jj)))
    colp1 = np.concatenate(([0, -1 / 2], -1 / (2 * jj) * (jj < N - 1)))
    F[np.arange(1, N), np.arange(0, N - 1)] = colm1
    F[np.arange(0, N), np.arange(1, N + 1)] = colp1
    F[0, N + 1] = 1
    return H * F


def second_integral_matrix(N, H):
    """Second Chebyshev integral matrix (N x N+2), H^2-scaled; Chebyshev.py:139-160."""
    S = np.zeros((N, N + 2))
    jj = np.arange(3, N)
    colm2 = np.concatenate(([0.25], 1 / (2 * jj * (2 * jj - 2))))
    colp2 = np.concatenate(([0, 0.125, 1. / 24.], 1 / (2 * jj * (2 * jj + 2)) * (jj < N - 2)))
    col0 = np.concatenate(([0, -0.125, -1. / 8. - 1. / 24.],
                           -1 / (2 * jj * (2 * jj - 2)) - 1 / (2 * jj * (2 * jj + 2)) * (jj < N - 1)))
    S[np.arange(2, N), np.arange(0, N - 2)] = colm2
    S[np.arange(0, N), np.arange(0, N)] = col0
    S[np.arange(0, N), np.arange(2, N + 2)] = colp2
    S[0, N] = 1
    S[1, N + 1] = 1
    return H ** 2 * S


def no_wall_bc_rows(N):
    """BC rows BCR1, BCR2, BCL1, BCL2 (length N+2); Solvers.py:603-643."""
    BCR1 = np.zeros((N + 2,))
    BCR2 = np.zeros((N + 2,))
    BCL1 = np.zeros((N + 2,))
    BCL2 = np.zeros((N + 2,))
    BCR1[N + 1] = 1; BCR2[N] = 1; BCR1[0] = 1; BCR1[2] = -1 / 2
    BCR2[N + 1] += 1; BCR2[1] = -1 / 8; BCR2[3] = 1 / 8
    BCR1[1] += 1 / 4; BCR1[3] -= 1 / 4; BCR2[0] += 1 / 4; BCR2[2] -= (1 / 8 + 1 / 24)
    BCR2[4] += 1 / 24
    BCL1[N + 1] = 1; BCL2[N] = -1; BCL1[0] = -1; BCL1[2] = 1 / 2
    BCL2[N + 1] += 1; BCL2[1] = -1 / 8; BCL2[3] = 1 / 8
    BCL1[1] += 1 / 4; BCL1[3] -= 1 / 4; BCL2[0] -= 1 / 4; BCL2[2] += 1 / 8 + 1 / 24
    BCL2[4] -= 1 / 24
    jj = np.arange(3, N)
    BCR1[2:N - 1] += 1 / (2 * jj)
    BCL1[2:N - 1] += np.power(-1, jj) / (2 * jj)
    BCR1[4:N + 1] -= 1 / (2 * jj) * (jj < N - 1)
    BCL1[4:N + 1] -= np.power(-1, jj) / (2 * jj) * (jj < N - 1)
    BCR2[1:N - 2] += 1 / (2 * jj) * 1 / (2 * jj - 2)
    BCL2[1:N - 2] -= 1 / (2 * jj) * 1 / (2 * jj - 2) * np.power(-1, jj)
    BCR2[5:N + 2] += 1 / (2 * jj) * 1 / (2 * jj + 2) * (jj < N - 2)
    BCL2[5:N + 2] -= 1 / (2 * jj) * 1 / (2 * jj + 2) * np.power(-1, jj) * (jj < N - 2)
    BCR2[3:N] -= (1 / (2 * jj) * 1 / (2 * jj - 2) + 1 / (2 * jj) * 1 / (2 * jj + 2) * (jj < N - 1))
    BCL2[3:N] += (1 / (2 * jj) * 1 / (2 * jj - 2) + 1 / (2 * jj) * 1 / (2 * jj + 2) * (jj < N - 1)) * np.power(-1, jj)
    return BCR1, BCR2, BCL1, BCL2


# ----------------------------------------------------------------------------
# KBPENTA (include/LinearSolvers.h:52-141), vectorised over modes (axis 0)
# ----------------------------------------------------------------------------
def kbpenta_precompute(d0, dp2, dm2):
    """beta storage of the 0/+-2-diagonal solver; LinearSolvers.h:52-74."""
    M, nz = d0.shape
    beta = np.zeros((M, nz + 1))
    beta[:, 1] = d0[:, 0]
    beta[:, 2] = d0[:, 1]
    for i in range(3, nz + 1):
        beta[:, i] = d0[:, i - 1] - dm2[:, i - 3] * dp2[:, i - 3] / beta[:, i - 2]
    return beta


def kbpenta_solve(beta, dp2, dm2, rhs):
    """Solve with stored beta; rhs (M, nz, nrhs) real or complex.
    LinearSolvers.h:76-141 (same recurrence for the real and complex solver)."""
    nz = rhs.shape[1]
    # rows first, so that every step of the recurrences reads contiguous memory
    x = np.ascontiguousarray(np.moveaxis(rhs, 1, 0))
    bT = np.ascontiguousarray(beta.T)[:, :, None]
    pT = np.ascontiguousarray(dp2.T)[:, :, None]
    mT = np.ascontiguousarray(dm2.T)[:, :, None]
    for i in range(2, nz):
        x[i] = x[i] - (mT[i - 2] * x[i - 2]) / bT[i - 1]
    x[nz - 1] /= bT[nz]
    x[nz - 2] /= bT[nz - 1]
    for i in range(nz - 3, -1, -1):
        x[i] = (x[i] - pT[i] * x[i + 2]) / bT[i + 1]
    return np.ascontiguousarray(np.moveaxis(x, 0, 1))


# ----------------------------------------------------------------------------
# DP solver precompute (Solvers.py:258-347, LinearSolvers.cpp:11-60)
# ----------------------------------------------------------------------------
class DPOperators:
    def __init__(self, Nx, Ny, Nz, Lx, Ly, H):
        self.Nx, self.Ny, self.Nz, self.H = Nx, Ny, Nz, H
        kvec_x = 2 * np.pi * np.concatenate((np.arange(0, np.floor(Nx / 2) + 1),
                                             np.arange(-1 * np.ceil(Nx / 2) + 1, 0)), axis=None) / Lx
        kvec_y = 2 * np.pi * np.concatenate((np.arange(0, np.floor(Ny / 2) + 1),
                                             np.arange(-1 * np.ceil(Ny / 2) + 1, 0)), axis=None) / Ly
        Kx, Ky = [K.flatten() for K in np.meshgrid(kvec_x, kvec_y, indexing='xy')]
        Ksq = Kx ** 2 + Ky ** 2
        K = np.sqrt(Ksq)
        Dx = 1j * Kx
        Dy = 1j * Ky
        if Nx % 2 == 0:
            Dx.reshape((Ny, Nx))[:, int(Nx / 2)] = 0
        if Ny % 2 == 0:
            Dy.reshape((Ny, Nx))[int(Ny / 2), :] = 0
        self.Kx, self.Ky, self.Ksq, self.K, self.Dx, self.Dy = Kx, Ky, Ksq, K, Dx, Dy
        self.FIMat = first_integral_matrix(Nz, H)
        self.SIMat = second_integral_matrix(Nz, H)
        S0 = self.SIMat[:, :Nz]
        dS, dpS, dmS = np.diag(S0), np.diag(S0, k=2), np.diag(S0, k=-2)
        # A = diagsEye - Ksq * diagsSIMat (Solvers.py:315)
        self.d0 = 1.0 - dS[None, :] * Ksq[:, None]
        self.dp2 = 0.0 - dpS[None, :] * Ksq[:, None]
        self.dm2 = 0.0 - dmS[None, :] * Ksq[:, None]
        BCR1, BCR2, BCL1, BCL2 = no_wall_bc_rows(Nz)
        self.BCR1, self.BCR2, self.BCL1, self.BCL2 = BCR1, BCR2, BCL1, BCL2
        BCs_k0 = np.stack((BCR2, -BCL2), axis=0)
        BCs_k = np.einsum('ij, k->kij', np.stack((BCR2, BCL2), axis=0), H ** 2 * K)
        BCs_k += BCs_k0 + np.stack((H * BCR1 - BCR2, H * BCL1 + BCL2), axis=0)
        self.C = BCs_k[:, :, 0:Nz]          # (M, 2, Nz)
        D = BCs_k[:, :, Nz:]                # (M, 2, 2) [row, col]
        B = self.SIMat[None, :, Nz:] * (-Ksq[:, None, None])   # (M, Nz, 2)
        self.beta = kbpenta_precompute(self.d0, self.dp2, self.dm2)
        self.AinvB = kbpenta_solve(self.beta, self.dp2, self.dm2, B)   # (M, Nz, 2)
        G = np.einsum('mrj,mjs->mrs', self.C, self.AinvB) - D
        det = 1.0 / (G[:, 0, 0] * G[:, 1, 1] - G[:, 0, 1] * G[:, 1, 0])
        Ginv = np.empty_like(G)
        Ginv[:, 0, 0] = G[:, 1, 1] * det
        Ginv[:, 1, 0] = -G[:, 1, 0] * det
        Ginv[:, 0, 1] = -G[:, 0, 1] * det
        Ginv[:, 1, 1] = G[:, 0, 0] * det
        self.Ginv = Ginv


def _banded_S(S, x, Nz):
    """SIMat @ x for x (M, Nz+2, ...) using its band (0, +-2) and extra columns."""
    out = np.zeros(x.shape[:1] + (Nz,) + x.shape[2:], dtype=x.dtype)
    idx = np.arange(Nz)
    lo = np.arange(2, Nz)
    ex = (None, slice(None)) + (None,) * (x.ndim - 2)
    out += S[idx, idx][ex] * x[:, :Nz]
    out[:, 2:Nz] += S[lo, lo - 2][ex] * x[:, 0:Nz - 2]
    out[:, 0:Nz] += S[idx, idx + 2][ex] * x[:, 2:Nz + 2]
    # null-space columns (constant and linear terms): SIMat[0, N], SIMat[1, N+1]
    out[:, 0] += S[0, Nz] * x[:, Nz]
    out[:, 1] += S[1, Nz + 1] * x[:, Nz + 1]
    return out


def _banded_F(F, x, Nz):
    """FIMat @ x (x: (M, Nz+2)) using its band (+-1) and the extra column."""
    out = np.zeros((x.shape[0], Nz), dtype=x.dtype)
    lo = np.arange(1, Nz)
    hi = np.arange(0, Nz)
    out[:, 1:Nz] += F[lo, lo - 1][None, :] * x[:, 0:Nz - 1]
    out[:, 0:Nz] += F[hi, hi + 1][None, :] * x[:, 1:Nz + 1]
    out[:, 0] += F[0, Nz + 1] * x[:, Nz + 1]
    return out


def dp_solve_no_wall(ops, Cf, eta):
    """Per-mode pressure + velocity BVP chain for modes 1..NyNx-1.
    DPTools.cpp:81-132 with parts :134-345.  Cf: (NyNx, Nz, 3) complex.
    Returns U_hat (NyNx, Nz, 3), P_hat (NyNx, Nz); mode 0 is zero."""
    Nz, H = ops.Nz, ops.H
    M = Cf.shape[0]
    sl = slice(1, M)
    cf = Cf[sl]
    beta, dp2, dm2 = ops.beta[sl], ops.dp2[sl], ops.dm2[sl]
    C, Ginv, AinvB = ops.C[sl], ops.Ginv[sl], ops.AinvB[sl]
    Dx, Dy, K = ops.Dx[sl], ops.Dy[sl], ops.K[sl]
    # chebCoeffDiff (DPTools.cpp:134-174), all three components
    cfT = np.ascontiguousarray(np.moveaxis(cf, 1, 0))     # rows first
    dfT = np.zeros_like(cfT)
    fac = (Nz - 1) * 2 / H
    dfT[Nz - 2] = fac * cfT[Nz - 1]
    for k in range(2, Nz):
        fac1 = (Nz - k) * 2 / H
        dfT[Nz - k - 1] = dfT[Nz - k + 1] + fac1 * cfT[Nz - k]
    dfT[0] /= 2
    df = np.moveaxis(dfT, 0, 1)
    # assemble_p_rhs (DPTools.cpp:177-195)
    p_rhs = Dx[:, None] * cf[:, :, 0] + Dy[:, None] * cf[:, :, 1] + df[:, :, 2]
    # schurSolve_p (DPTools.cpp:236-285)
    xp = kbpenta_solve(beta, dp2, dm2, p_rhs[:, :, None])[:, :, 0]
    y0 = np.sum(xp * C[:, 0, :], axis=1)
    y1 = np.sum(xp * C[:, 1, :], axis=1)
    Y0 = Ginv[:, 0, 0] * y0 + Ginv[:, 0, 1] * y1
    Y1 = Ginv[:, 1, 0] * y0 + Ginv[:, 1, 1] * y1
    x = np.zeros((M - 1, Nz + 2), dtype=np.complex128)
    x[:, :Nz] = xp - (AinvB[:, :, 0] * Y0[:, None] + AinvB[:, :, 1] * Y1[:, None])
    x[:, Nz] = Y0
    x[:, Nz + 1] = Y1
    cp = _banded_S(ops.SIMat, x, Nz)
    dp = _banded_F(ops.FIMat, x, Nz)
    # assemble_u_rhs (DPTools.cpp:197-216)
    u_rhs = np.empty((M - 1, Nz, 3), dtype=np.complex128)
    u_rhs[:, :, 0] = (Dx[:, None] * cp - cf[:, :, 0]) / eta
    u_rhs[:, :, 1] = (Dy[:, None] * cp - cf[:, :, 1]) / eta
    u_rhs[:, :, 2] = (dp - cf[:, :, 2]) / eta
    # assemble_u_bc_rhs (DPTools.cpp:218-234)
    s = np.sum(cp, axis=1)
    alt = np.sum(cp * np.power(-1.0, np.arange(Nz))[None, :], axis=1)
    fac1 = 2 * eta
    facK = fac1 * K
    bc = np.stack((-Dx * s / facK, Dx * alt / facK, -Dy * s / facK, Dy * alt / facK,
                   s / fac1, alt / fac1), axis=1)
    # schurSolve_u (DPTools.cpp:287-345)
    xpu = kbpenta_solve(beta, dp2, dm2, u_rhs)
    cu = np.empty((M - 1, Nz, 3), dtype=np.complex128)
    for c in range(3):
        ya = np.sum(xpu[:, :, c] * C[:, 0, :], axis=1) - bc[:, 2 * c]
        yb = np.sum(xpu[:, :, c] * C[:, 1, :], axis=1) - bc[:, 2 * c + 1]
        Ya = Ginv[:, 0, 0] * ya + Ginv[:, 0, 1] * yb
        Yb = Ginv[:, 1, 0] * ya + Ginv[:, 1, 1] * yb
        xc = np.zeros((M - 1, Nz + 2), dtype=np.complex128)
        xc[:, :Nz] = xpu[:, :, c] - (AinvB[:, :, 0] * Ya[:, None] + AinvB[:, :, 1] * Yb[:, None])
        xc[:, Nz] = Ya
        xc[:, Nz + 1] = Yb
        cu[:, :, c] = _banded_S(ops.SIMat, xc, Nz)
    U = np.zeros((M, Nz, 3), dtype=np.complex128)
    P = np.zeros((M, Nz), dtype=np.complex128)
    U[sl] = cu
    P[sl] = cp
    return U, P


def eval_theta(U, theta):
    """Chebyshev series at angle theta for every mode; DPTools.cpp:348-378."""
    Nz = U.shape[1]
    w = np.cos(np.arange(Nz) * theta)
    return np.einsum('mjl,j->ml', U, w)


def _cheb_of_profiles(prof):
    """Chebyshev coefficients of samples prof (M, Nz, ...) at the nodes by the
    mirrored FFT of length 2Nz-2 (DPTools.cpp:636-697)."""
    Nz = prof.shape[1]
    n = 2 * Nz - 2
    ext = np.concatenate((prof, prof[:, Nz - 2:0:-1]), axis=1)
    X = sfft.fft(ext, axis=1, workers=-1)
    out = np.empty(prof.shape, dtype=np.complex128)
    out[:, 0] = X[:, 0] / n
    out[:, 1:Nz - 1] = (X[:, 1:Nz - 1] + X[:, n - 1:Nz - 1:-1]) / n
    out[:, Nz - 1] = X[:, Nz - 1] / n
    return out


def bottom_wall_correction(ops, U, P, z, eta):
    """Analytic no-slip correction at z=0; DPTools.cpp:590-700."""
    phi = eval_theta(U, np.pi)
    sl = slice(1, U.shape[0])
    kx, ky, k = ops.Kx[sl], ops.Ky[sl], ops.K[sl]
    fx, fy, fz = -phi[sl, 0], -phi[sl, 1], -phi[sl, 2]
    alpha = kx * fx + ky * fy + 1j * k * fz
    enkz = np.exp(-k[:, None] * z[None, :])
    zenkz = z[None, :] * enkz
    Cp = -1j * 2 * eta * alpha[:, None] * enkz
    Cu = -kx[:, None] * alpha[:, None] * zenkz / k[:, None] + fx[:, None] * enkz
    Cv = -ky[:, None] * alpha[:, None] * zenkz / k[:, None] + fy[:, None] * enkz
    Cw = -1j * alpha[:, None] * zenkz + fz[:, None] * enkz
    corr = _cheb_of_profiles(np.stack((Cu, Cv, Cw, Cp), axis=2))
    U[sl] += corr[:, :, 0:3]
    P[sl] += corr[:, :, 3]


def slit_channel_correction(ops, U, P, z, Hs, eta):
    """No-slip correction at z=0 and z=Hs; DPTools.cpp:704-877 (Hs = Lz)."""
    phib = eval_theta(U, np.pi)
    phit = eval_theta(U, 0.0)
    sl = slice(1, U.shape[0])
    kx, ky = ops.Kx[sl], ops.Ky[sl]
    k = np.sqrt(kx * kx + ky * ky)
    M = kx.shape[0]
    fac = 1.0 / 2.0 / eta
    enkh = np.exp(-k * Hs)
    A = np.zeros((M, 8, 8), dtype=np.complex128)
    A[:, 0, 0] = fac; A[:, 0, 1] = enkh * fac; A[:, 0, 6] = -k; A[:, 0, 7] = k * enkh
    A[:, 1, 0] = (1 - k * Hs) * enkh * fac; A[:, 1, 1] = (1 + k * Hs) * fac
    A[:, 1, 6] = -k * enkh; A[:, 1, 7] = k
    A[:, 2, 2] = 1; A[:, 2, 3] = enkh
    A[:, 3, 4] = 1; A[:, 3, 5] = enkh
    A[:, 4, 6] = 1; A[:, 4, 7] = enkh
    A[:, 5, 0] = -1j * kx * Hs * enkh * fac / k; A[:, 5, 1] = 1j * kx * Hs * fac / k
    A[:, 5, 2] = enkh; A[:, 5, 3] = 1
    A[:, 6, 0] = -1j * ky * Hs * enkh * fac / k; A[:, 6, 1] = 1j * ky * Hs * fac / k
    A[:, 6, 4] = enkh; A[:, 6, 5] = 1
    A[:, 7, 0] = Hs * enkh * fac; A[:, 7, 1] = Hs * fac
    A[:, 7, 6] = enkh; A[:, 7, 7] = 1
    fb = -phib[sl]
    ft = -phit[sl]
    rhs = np.stack((-1j * kx * fb[:, 0] - 1j * ky * fb[:, 1],
                    -1j * kx * ft[:, 0] - 1j * ky * ft[:, 1],
                    fb[:, 0], fb[:, 1], fb[:, 2], ft[:, 0], ft[:, 1], ft[:, 2]), axis=1)
    xs = np.linalg.solve(A, rhs[:, :, None])[:, :, 0]
    enkz = np.exp(-k[:, None] * z[None, :])
    ekzmh = enkh[:, None] / enkz
    zekzmh = z[None, :] * ekzmh
    zenkz = z[None, :] * enkz
    kx_, ky_, k_ = kx[:, None], ky[:, None], k[:, None]
    x = [xs[:, i][:, None] for i in range(8)]
    Cp = x[0] * enkz + x[1] * ekzmh
    Cu = (-1j * x[0] * kx_ * zenkz * fac / k_ + 1j * x[1] * kx_ * zekzmh * fac / k_
          + x[2] * enkz + x[3] * ekzmh)
    Cv = (-1j * x[0] * ky_ * zenkz * fac / k_ + 1j * x[1] * ky_ * zekzmh * fac / k_
          + x[4] * enkz + x[5] * ekzmh)
    Cw = x[0] * zenkz * fac + x[1] * zekzmh * fac + x[6] * enkz + x[7] * ekzmh
    corr = _cheb_of_profiles(np.stack((Cu, Cv, Cw, Cp), axis=2))
    U[sl] += corr[:, :, 0:3]
    P[sl] += corr[:, :, 3]


def _ginv2(D):
    return 1.0 / (D[0, 0] * D[1, 1] - D[0, 1] * D[1, 0]) * np.array([[D[1, 1], -D[0, 1]],
                                                                     [-D[1, 0], D[0, 0]]])


def k0_wall_velocity(ops, Cf0, U, eta, geometry):
    """k=0 velocity BVP for bottom wall / slit; Solvers.py:482-516, 563-599.
    Cf0: (Nz, 3) complex k=0 forcing as the reference builds it."""
    Nz, H = ops.Nz, ops.H
    if geometry == 'bottom':
        BCs = np.stack((ops.BCR1, ops.BCL2), axis=0)
    else:
        BCs = np.stack((ops.BCR2, ops.BCL2), axis=0)
    Ck = BCs[:, 0:Nz]
    Dk = -BCs[:, Nz:]
    Gk = _ginv2(Dk)
    S = ops.SIMat.astype(np.complex128)
    out = []
    for c in range(2):
        rhs = -Cf0[:, c] / eta
        a = np.concatenate((rhs, Gk @ (Ck @ rhs)))
        out.append(S @ a)
    U[0, :, 0] += out[0]
    U[0, :, 1] += out[1]


# ----------------------------------------------------------------------------
# Fourier-Chebyshev transform (Transform.cpp:107-304), c2c like the reference
# ----------------------------------------------------------------------------
def fc_forward(F):
    """F: real (Ny, Nx, Nz, 3) -> complex Fourier-Chebyshev coefficients."""
    Ny, Nx, Nz, dof = F.shape
    Nu = 2 * Nz - 2
    ext = np.concatenate((F, F[:, :, Nz - 2:0:-1, :]), axis=2).astype(np.complex128)
    X = sfft.fftn(ext, axes=(0, 1, 2), workers=-1)
    out = np.empty((Ny, Nx, Nz, dof), dtype=np.complex128)
    out[:, :, 0] = X[:, :, 0] / Nu
    out[:, :, 1:Nz - 1] = (X[:, :, 1:Nz - 1] + X[:, :, Nu - 1:Nz - 1:-1]) / Nu
    out[:, :, Nz - 1] = X[:, :, Nz - 1] / Nu
    return out


def fc_backward(C):
    """C: complex coefficients (Ny, Nx, Nz, 3) -> real values (Ny, Nx, Nz, 3)."""
    Ny, Nx, Nz, dof = C.shape
    Nu = 2 * Nz - 2
    ext = np.empty((Ny, Nx, Nu, dof), dtype=np.complex128)
    ext[:, :, 0] = Nu * C[:, :, 0]
    ext[:, :, 1:Nz - 1] = Nu * C[:, :, 1:Nz - 1] / 2
    ext[:, :, Nz - 1] = Nu * C[:, :, Nz - 1]
    ext[:, :, Nz:] = Nu * C[:, :, Nz - 2:0:-1] / 2
    X = sfft.ifftn(ext, axes=(0, 1, 2), workers=-1)
    return np.ascontiguousarray(X[:, :, :Nz].real)


# ----------------------------------------------------------------------------
# The Mdot sequence (FCM.py:149-441)
# ----------------------------------------------------------------------------
class KernelSpec:
    """One ES kernel: nominal width w, beta*w, support alpha (physical)."""

    def __init__(self, w, betaw, alpha, dipole=False):
        self.w, self.betaw, self.alpha, self.dipole = int(w), float(betaw), float(alpha), dipole
        L = lib()
        self.norm = L.orc_es_norm(self.alpha, self.betaw)
        self.thresh = L.orc_dipole_thresh(self.alpha, self.betaw) if dipole else self.alpha


class OracleDPStokes:
    """FP64 CPU Mdot for geometries 'nowall' / 'bottom' / 'slit'.

    Parameters mirror the joint interface's explicit-grid form: box Lx, Ly,
    z-domain [0, Lz] (positions are shifted by zmin), grid nx, ny, nz, the
    monopole kernel (w, beta*w, alpha) and optionally the dipole kernel used
    for torques.  Velocities follow FCM.Mdot (FCM.py:334-441).
    """

    def __init__(self, geometry, Lx, Ly, zmin, zmax, nx, ny, nz, eta,
                 w, betaw, alpha=None, w_d=None, betaw_d=None, alpha_d=None, has_torque=False,
                 reg_kernel=False):
        if geometry not in ('nowall', 'bottom', 'slit'):
            raise ValueError('geometry must be nowall, bottom or slit')
        self.geometry = geometry
        self.Lx, self.Ly, self.zmin = float(Lx), float(Ly), float(zmin)
        self.Lz = float(zmax) - float(zmin)
        self.Nx, self.Ny, self.Nz = int(nx), int(ny), int(nz)
        self.hx = self.Lx / self.Nx
        self.hy = self.Ly / self.Ny
        self.eta = float(eta)
        self.H = self.Lz / 2
        if alpha is None or alpha < 0:
            alpha = w * self.hx / 2
        self.mono = KernelSpec(w, betaw, alpha)
        self.has_torque = has_torque
        if has_torque:
            if alpha_d is None or alpha_d < 0:
                alpha_d = w_d * self.hx / 2
            # useRegKernel (FCM.py:313-319): the torque kernel is the plain ES
            # kernel of the dipole width, support alpha_d
            self.dip = KernelSpec(w_d, betaw_d, alpha_d, dipole=not reg_kernel)
        self.reg_kernel = bool(reg_kernel) and has_torque
        # BC tables (FCM.py:215-248): [end*3 + comp], ends x-,x+,y-,y+,z-(top),z+(bottom)
        self.BCs = BC_NONE * np.ones(18, dtype=np.int32)
        if geometry == 'bottom':
            self.BCs[15:18] = BC_MIRROR_INV
        elif geometry == 'slit':
            self.BCs[12:18] = BC_MIRROR_INV
        self.BCst_z = BC_NONE * np.ones(18, dtype=np.int32)
        if geometry != 'nowall':
            self.BCst_z[15:18] = BC_MIRROR
            if geometry == 'slit':
                self.BCst_z[12:15] = BC_MIRROR
        self.zpts, self.zwts = clencurt(self.Nz, 0, self.Lz)
        self.ops = DPOperators(self.Nx, self.Ny, self.Nz, self.Lx, self.Ly, self.H)
        self._loc = None

    # -- particles ---------------------------------------------------------
    def _zext(self, thresh):
        L = lib()
        n = self.Nz + 2 * (self.Nz - 1)
        zext = np.zeros(n)
        zw = np.zeros(n)
        up = ctypes.c_int()
        dn = ctypes.c_int()
        Nzeff = L.orc_zext(_p(self.zpts), _p(self.zwts), self.Nz, thresh, _p(zext), _p(zw),
                           ctypes.byref(up), ctypes.byref(dn))
        return zext[:Nzeff].copy(), zw[:Nzeff].copy(), Nzeff, up.value, dn.value

    def _locate(self, xs, ks):
        L = lib()
        zext, zw, Nzeff, up, dn = self._zext(ks.thresh)
        if up != dn:
            # the fold and ghost copy split the extension into equal halves
            # (BoundaryConditions.h:22-361); a node at exactly the support
            # distance from one end but not, after rounding, from the other
            # breaks that
            raise ValueError('z extension is asymmetric (%d above, %d below): the support '
                             'edge falls on a Chebyshev node' % (up, dn))
        N = xs.shape[0] // 3
        wx = int(round(2 * ks.thresh / self.hx))
        wy = int(round(2 * ks.thresh / self.hy))
        loc = dict(Nzeff=Nzeff, ext_up=up, ext_down=dn, zext=zext, zext_w=zw, zstride=Nzeff)
        ia = lambda: np.zeros(N, dtype=np.int32)
        loc.update(xclose=ia(), yclose=ia(), indl=ia(), indr=ia(),
                   xunwrap=np.zeros(N * wx), yunwrap=np.zeros(N * wy),
                   zunwrap=np.zeros(N * Nzeff), pt_wts=np.zeros(N * Nzeff))
        cwx = ctypes.c_int()
        cwy = ctypes.c_int()
        rc = L.orc_locate(N, _p(xs), self.hx, self.hy, self.Lx, self.Ly, self.Lz, ks.thresh,
                          _p(zext), _p(zw), Nzeff, Nzeff, ctypes.byref(cwx), ctypes.byref(cwy),
                          _p(loc['xclose']), _p(loc['yclose']), _p(loc['indl']), _p(loc['indr']),
                          _p(loc['xunwrap']), _p(loc['yunwrap']), _p(loc['zunwrap']),
                          _p(loc['pt_wts']))
        if rc != 0:
            raise ValueError('Initial positions must be within unit cell')
        loc['wx'], loc['wy'] = cwx.value, cwy.value
        return loc

    def set_positions(self, positions):
        """Shift z by zmin (x, y are taken as given in [0, L]) and locate.
        FCM.SetPositions first-call path (FCM.py:284-324); every call relocates
        from scratch (see the module notes on ParticleList::update)."""
        xs = np.array(positions, dtype=np.float64).copy()
        xs[2::3] -= self.zmin
        self.N = xs.shape[0] // 3
        self.xs = xs
        self._loc = self._locate(xs, self.mono)
        self._tloc = self._locate(xs, self.dip) if self.has_torque else None

    # -- grid stages ---------------------------------------------------------
    def _spread(self, kind, loc, ks, f, BCs, keep=None):
        L = lib()
        fG = np.zeros((self.Ny, self.Nx, self.Nz, 3))
        L.orc_spread(kind, self.N, _p(np.ascontiguousarray(f, dtype=np.float64)), ks.betaw,
                     ks.alpha, ks.norm, loc['wx'], loc['wy'], _p(loc['xclose']),
                     _p(loc['yclose']), _p(loc['indl']), _p(loc['indr']), _p(loc['xunwrap']),
                     _p(loc['yunwrap']), _p(loc['zunwrap']), loc['zstride'], self.Nx, self.Ny,
                     self.Nz, loc['Nzeff'], _p(BCs), _p(fG), _p(keep))
        return fG

    def _ext_buffer(self, loc):
        wx, wy = loc['wx'], loc['wy']
        return np.zeros((self.Nx + wx + wx % 2) * (self.Ny + wy + wy % 2) * loc['Nzeff'] * 3)

    def _interp(self, kind, loc, ks, fG, BCs, ext):
        L = lib()
        uP = np.zeros(3 * self.N)
        L.orc_interp(kind, self.N, _p(np.ascontiguousarray(fG)), ks.betaw, ks.alpha, ks.norm,
                     loc['wx'], loc['wy'], _p(loc['xclose']), _p(loc['yclose']),
                     _p(loc['indl']), _p(loc['indr']), _p(loc['xunwrap']), _p(loc['yunwrap']),
                     _p(loc['zunwrap']), _p(loc['pt_wts']), loc['zstride'], self.Nx, self.Ny,
                     self.Nz, loc['Nzeff'], _p(BCs), _p(ext), _p(uP))
        return uP

    def spread_forces(self, forces, torques=None):
        """Spread + fold of forces, plus the torque curl (FCM.py:354-371,
        SpreadInterpWrapper.cpp:66-107).  Returns the interior grid.  With the
        regular torque kernel the torques are spread as a vector field into
        self.tgrid and their curl is taken spectrally by solve()."""
        self._ext_f = self._ext_buffer(self._loc)
        F = self._spread(0, self._loc, self.mono, forces, self.BCs, self._ext_f)
        self.tgrid = None
        if self.has_torque and torques is not None and np.size(torques) > 0:
            self._ext_t = self._ext_buffer(self._tloc)
            if self.reg_kernel:
                self.tgrid = self._spread(0, self._tloc, self.dip, torques, self.BCs, self._ext_t)
                return F
            dxT = self._spread(1, self._tloc, self.dip, torques, self.BCs, self._ext_t)
            F[..., 1] -= 0.5 * dxT[..., 2]
            F[..., 2] += 0.5 * dxT[..., 1]
            dyT = self._spread(2, self._tloc, self.dip, torques, self.BCs, self._ext_t)
            F[..., 0] += 0.5 * dyT[..., 2]
            F[..., 2] -= 0.5 * dyT[..., 0]
            dzT = self._spread(3, self._tloc, self.dip, torques, self.BCst_z, self._ext_t)
            F[..., 0] -= 0.5 * dzT[..., 1]
            F[..., 1] += 0.5 * dzT[..., 0]
        return F

    def solve(self, F, tgrid=None):
        """Forward transform, per-mode solve, corrections, k=0, back transform.
        FCM.py:398-407 with Solvers.py:99-136.  With a torque grid (regular
        torque kernel, FCM.py:372-396 and 427-433) half its spectral curl is
        added to the forcing, and half the curl of the velocity coefficients
        is transformed back into self.utG."""
        Ny, Nx, Nz = self.Ny, self.Nx, self.Nz
        Fh = fc_forward(F)
        Cf = Fh.reshape(Ny * Nx, Nz, 3)
        if tgrid is not None:
            Th = fc_forward(tgrid).reshape(Ny * Nx, Nz, 3)
            Cf = Cf + curl_cheb(Th, self.ops.Dx, self.ops.Dy, self.H)
        U, P = dp_solve_no_wall(self.ops, Cf, self.eta)
        if self.geometry == 'bottom':
            bottom_wall_correction(self.ops, U, P, self.zpts, self.eta)
            k0_wall_velocity(self.ops, Cf[0], U, self.eta, 'bottom')
        elif self.geometry == 'slit':
            slit_channel_correction(self.ops, U, P, self.zpts, self.Lz, self.eta)
            cf0 = Cf[0].real + 1j * Cf[0].real   # Solvers.py:564-566
            k0_wall_velocity(self.ops, cf0, U, self.eta, 'slit')
        self.U_hat = U
        self.utG = None
        if tgrid is not None:
            W = curl_cheb(U, self.ops.Dx, self.ops.Dy, self.H)
            self.utG = fc_backward(W.reshape(Ny, Nx, Nz, 3))
        return fc_backward(U.reshape(Ny, Nx, Nz, 3))

    def interpolate(self, uG, torques_present, utG=None):
        v = self._interp(0, self._loc, self.mono, uG, self.BCs, self._ext_f)
        if not torques_present:
            return v, None
        ext = self._ext_t
        if self.reg_kernel:
            # angular velocity: the regular kernel applied to half the curl
            # grid (FCM.py:437-439)
            return v, self._interp(0, self._tloc, self.dip, utG, self.BCs, ext)
        dxv = self._interp(1, self._tloc, self.dip, uG, self.BCs, ext)
        dyv = self._interp(2, self._tloc, self.dip, uG, self.BCs, ext)
        dzv = self._interp(3, self._tloc, self.dip, uG, self.BCst_z, ext)
        om = np.empty(3 * self.N)
        om[0::3] = -1 / 2 * (dyv[2::3] - dzv[1::3])
        om[1::3] = -1 / 2 * (dzv[0::3] - dxv[2::3])
        om[2::3] = -1 / 2 * (dxv[1::3] - dyv[0::3])
        return v, om

    def mdot_phases(self, forces, torques=None):
        """mdot (derivative torque kernels) with the phase timers of the
        reference's timing harness, source/cpu/examples/mobility_bench.py:312-442:
        'SI force' (force spread + velocity interpolation), 'SI torque' (the
        three derivative spreads with the curl + the derivative interpolations
        and Omega), 'FCT+iFCT' (forward and backward Fourier-Chebyshev
        transforms), 'Solve' (per-mode BVP, wall correction, k = 0).  Returns
        (U, W, seconds per phase); the arithmetic is that of mdot."""
        import time
        if self.reg_kernel:
            raise ValueError('phase timing covers the derivative torque kernels only')
        t = {'SI force': 0.0, 'SI torque': 0.0, 'FCT+iFCT': 0.0, 'Solve': 0.0}
        c = time.perf_counter
        torque = self.has_torque and torques is not None and np.size(torques) > 0
        t0 = c()
        self._ext_f = self._ext_buffer(self._loc)
        F = self._spread(0, self._loc, self.mono, forces, self.BCs, self._ext_f)
        t['SI force'] += c() - t0
        if torque:
            t0 = c()
            self._ext_t = self._ext_buffer(self._tloc)
            dxT = self._spread(1, self._tloc, self.dip, torques, self.BCs, self._ext_t)
            F[..., 1] -= 0.5 * dxT[..., 2]
            F[..., 2] += 0.5 * dxT[..., 1]
            dyT = self._spread(2, self._tloc, self.dip, torques, self.BCs, self._ext_t)
            F[..., 0] += 0.5 * dyT[..., 2]
            F[..., 2] -= 0.5 * dyT[..., 0]
            dzT = self._spread(3, self._tloc, self.dip, torques, self.BCst_z, self._ext_t)
            F[..., 0] -= 0.5 * dzT[..., 1]
            F[..., 1] += 0.5 * dzT[..., 0]
            t['SI torque'] += c() - t0
        Ny, Nx, Nz = self.Ny, self.Nx, self.Nz
        t0 = c()
        Cf = fc_forward(F).reshape(Ny * Nx, Nz, 3)
        t['FCT+iFCT'] += c() - t0
        t0 = c()
        U, P = dp_solve_no_wall(self.ops, Cf, self.eta)
        if self.geometry == 'bottom':
            bottom_wall_correction(self.ops, U, P, self.zpts, self.eta)
            k0_wall_velocity(self.ops, Cf[0], U, self.eta, 'bottom')
        elif self.geometry == 'slit':
            slit_channel_correction(self.ops, U, P, self.zpts, self.Lz, self.eta)
            k0_wall_velocity(self.ops, Cf[0].real + 1j * Cf[0].real, U, self.eta, 'slit')
        t['Solve'] += c() - t0
        t0 = c()
        uG = fc_backward(U.reshape(Ny, Nx, Nz, 3))
        t['FCT+iFCT'] += c() - t0
        t0 = c()
        v = self._interp(0, self._loc, self.mono, uG, self.BCs, self._ext_f)
        t['SI force'] += c() - t0
        om = None
        if torque:
            t0 = c()
            dxv = self._interp(1, self._tloc, self.dip, uG, self.BCs, self._ext_t)
            dyv = self._interp(2, self._tloc, self.dip, uG, self.BCs, self._ext_t)
            dzv = self._interp(3, self._tloc, self.dip, uG, self.BCst_z, self._ext_t)
            om = np.empty(3 * self.N)
            om[0::3] = -1 / 2 * (dyv[2::3] - dzv[1::3])
            om[1::3] = -1 / 2 * (dzv[0::3] - dxv[2::3])
            om[2::3] = -1 / 2 * (dxv[1::3] - dyv[0::3])
            t['SI torque'] += c() - t0
        return v, om, t

    def _interp_pos(self, kind, loc, ks, fG, BCs):
        L = lib()
        uP = np.zeros(3 * self.N)
        L.orc_interp_pos(kind, ks.thresh, self.N, _p(np.ascontiguousarray(fG)), ks.betaw, ks.alpha, ks.norm,
                         loc['wx'], loc['wy'], _p(loc['xclose']), _p(loc['yclose']),
                         _p(loc['indl']), _p(loc['indr']), _p(loc['xunwrap']), _p(loc['yunwrap']),
                         _p(loc['zunwrap']), _p(loc['pt_wts']), loc['zstride'], self.Nx, self.Ny,
                         self.Nz, loc['Nzeff'], _p(BCs), _p(self._ext_buffer(loc)), _p(uP))
        return uP

    def interp_force_kernel(self, kind, uG):
        """Interpolate uG with the force kernel (kind 0) or its x/y/z
        derivative (kind 1/2/3) as a function of the particle position (zero
        where an offset was clamped), folds and ghosts as for the force
        kernel."""
        if kind == 0:
            return self._interp(0, self._loc, self.mono, uG, self.BCs, self._ext_buffer(self._loc))
        return self._interp_pos(kind, self._loc, self.mono, uG, self.BCs)

    def interp_torque_kernel_d2(self, k, b, uG):
        """Interpolate uG with d2/(do_k do_b) of the torque kernel as a
        function of the particle position (k, b in 0..2), ghosts of the class
        the first derivative along b uses (BCst_z for z, BCs otherwise)."""
        return self._interp_pos(10 + 3 * k + b, self._tloc, self.dip, uG, self.BCst_z if b == 2 else self.BCs)

    def mdot(self, forces, torques=None):
        """Linear and angular velocities (FCM.py:334-441)."""
        forces = np.asarray(forces, dtype=np.float64)
        if forces.size != 3 * self.N:
            raise ValueError('Dimension mismatch in particle force array')
        tp = self.has_torque and torques is not None and np.size(torques) > 0
        if tp and np.size(torques) != 3 * self.N:
            raise ValueError('Dimension mismatch in particle force array')
        F = self.spread_forces(forces, torques if tp else None)
        uG = self.solve(F, self.tgrid)
        self.uG = uG
        return self.interpolate(uG, tp, self.utG)


_EPS = ((0, 1, 2, 1.0), (1, 2, 0, 1.0), (2, 0, 1, 1.0), (0, 2, 1, -1.0), (1, 0, 2, -1.0), (2, 1, 0, -1.0))


def adjoint(o, forces, lam_u, lam_w=None, torques=None):
    """Adjoint of Mdot for either oracle (no counterpart in the reference).
    With L = <lam_u, U> + <lam_w, W> and (U, W) = Mdot(F, T) at the current
    positions:
      dL/dF, dL/dT = Mdot(lam_u, lam_w)   (the solve taken as self-adjoint)
      dL/dx_p[k] = - sum_i lam_u[p,i] I_k u_i - sum_i F[p,i] I_k v_i
                   + 1/2 sum eps_abc lam_w[p,a] J_kb u_c
                   - 1/2 sum eps_abc T[p,c] J_kb v_a
    where u is the velocity grid of (F, T), v that of (lam_u, lam_w), I_k
    interpolation with the k-th derivative of the force kernel and J_kb with
    the (k, b) second derivative of the torque kernel, all as functions of
    the particle position: a weight phi(x_g - x_p) changes with x_p as minus
    the kernel derivative, ghost folds and copies keep the factors of the
    forward weights (ghost nodes do not move) and a coordinate whose offset
    was clamped contributes no derivative.  The torque terms follow from
    W = -1/2 eps_abc D_b u_c (derivative-kernel interpolation, FCM.py:411-426)
    and the torque forcing 1/2 eps_abc D_b T_c (SpreadInterpWrapper.cpp:66-107).
    Position gradients are None with the regular torque kernel."""
    forces = np.asarray(forces, dtype=np.float64)
    lam_u = np.asarray(lam_u, dtype=np.float64)
    tq = torques is not None and np.size(torques) > 0
    lw = lam_w is not None and np.size(lam_w) > 0
    gF, gT = o.mdot(lam_u, lam_w if lw else None)
    if (tq or lw) and getattr(o, 'reg_kernel', False):
        return gF, gT, None
    T = np.asarray(torques, dtype=np.float64) if tq else None
    mu = np.asarray(lam_w, dtype=np.float64) if lw else None
    uG = o.solve(o.spread_forces(forces, T))
    vG = o.solve(o.spread_forces(lam_u, mu))
    N = o.N
    lam = lam_u.reshape(N, 3)
    F3 = forces.reshape(N, 3)
    gx = np.empty((N, 3))
    for j in (1, 2, 3):
        du = o.interp_force_kernel(j, uG).reshape(N, 3)
        dv = o.interp_force_kernel(j, vG).reshape(N, 3)
        gx[:, j - 1] = -(lam * du).sum(1) - (F3 * dv).sum(1)
    if tq or lw:
        T3 = T.reshape(N, 3) if tq else np.zeros((N, 3))
        M3 = mu.reshape(N, 3) if lw else np.zeros((N, 3))
        for k in range(3):
            for a, b, c, sgn in _EPS:
                Ju = o.interp_torque_kernel_d2(k, b, uG).reshape(N, 3)
                Jv = o.interp_torque_kernel_d2(k, b, vG).reshape(N, 3)
                gx[:, k] += 0.5 * sgn * (M3[:, a] * Ju[:, c] - T3[:, c] * Jv[:, a])
    return gF, gT, gx.ravel()


def num_threads():
    return lib().orc_num_threads()
