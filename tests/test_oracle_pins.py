"""Pins of the FP64 oracle that do not use the oracle's own formulas.

Each function of oracle/ is checked against what the mathematics fixes:
closed forms, library routines (numpy.polynomial.chebyshev, scipy.fft,
scipy.integrate, numpy.linalg), physical invariants of the Stokes problem
(ODE residuals, decay conditions, no-slip, incompressibility), brute force on
tiny inputs, and exact adjointness of spreading and interpolation.  The
reference's own stored mobilities pin the composition
(tests/test_oracle_golden.py).
"""
import numpy as np
import pytest
from numpy.polynomial import chebyshev as C

from oracle import dp_oracle as o

np.seterr(all="ignore")


# ---------------------------------------------------------------------------
# Clenshaw-Curtis (Chebyshev.py:4-38, ParticleList.cpp / Quadrature.cpp)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("N", [5, 8, 17, 33, 61])
def test_clencurt_integrates_chebyshev_polynomials_exactly(N):
    a, b = 0.3, 4.1
    H, c = (b - a) / 2, (b + a) / 2
    z, w = o.clencurt(N, a, b)
    assert z[0] == pytest.approx(b) and z[-1] == pytest.approx(a)
    assert np.all(np.diff(z) < 0)
    for k in range(N):
        t = (z - c) / H
        exact = 0.0 if k == 1 else H * (1 + (-1) ** k) / (1 - k * k)
        assert np.dot(w, np.cos(k * np.arccos(np.clip(t, -1, 1)))) == pytest.approx(exact, abs=1e-13 * H)


def test_clencurt_c_matches_python():
    L = o.lib()
    for N in (9, 24, 1000):
        zc, wc = np.zeros(N), np.zeros(N)
        L.orc_clencurt(o._p(zc), o._p(wc), -1.3, 2.2, N)
        zp, wp = o.clencurt(N, -1.3, 2.2)
        np.testing.assert_allclose(zc, zp, rtol=0, atol=1e-14)
        np.testing.assert_allclose(wc, wp, rtol=0, atol=1e-14)


# ---------------------------------------------------------------------------
# ES kernel normalisation and dipole cutoff (ParticleList.cpp:217-265)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("alpha,betaw", [(1.5, 10.71), (2.0, 8.0), (0.9, 13.0), (3.0, 9.2)])
def test_es_norm_matches_adaptive_quadrature(alpha, betaw):
    from scipy.integrate import quad
    exact, _ = quad(lambda x: np.exp(betaw * (np.sqrt(1 - x * x / alpha ** 2) - 1)), -alpha, alpha,
                    epsabs=1e-14, epsrel=1e-13, limit=200)
    assert o.lib().orc_es_norm(alpha, betaw) == pytest.approx(exact, rel=1e-7)


@pytest.mark.parametrize("alpha,betaw", [(1.5, 10.71), (2.0, 8.0), (0.9, 13.0), (3.0, 9.2)])
def test_dipole_cutoff_is_the_edge_extremum(alpha, betaw):
    """The cutoff is the critical point of r e^{bw s}/s, s = sqrt(1 - r^2/a^2),
    near the edge (where |phi'| stops decreasing)."""
    from scipy.optimize import brentq

    def dlog(r):
        s = np.sqrt(1 - r * r / alpha ** 2)
        return 1 / r - betaw * r / (alpha ** 2 * s) + r / (alpha ** 2 * s * s)

    root = brentq(dlog, 0.5 * alpha, alpha * (1 - 1e-12), xtol=1e-15, rtol=1e-15)
    assert o.lib().orc_dipole_thresh(alpha, betaw) == pytest.approx(root, rel=1e-12)


# ---------------------------------------------------------------------------
# Chebyshev integral matrices and BC rows (Chebyshev.py:96-160, Solvers.py:603-643)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("N", [8, 13, 24])
def test_integral_matrices_invert_chebyshev_differentiation(N):
    H = 1.7
    rng = np.random.RandomState(N)
    S = o.second_integral_matrix(N, H)
    F = o.first_integral_matrix(N, H)
    c = np.zeros(N + 2)
    c[:N - 2] = rng.normal(size=N - 2)
    # d^2/dz^2 (S c) = c and d/dz (F c) = c, z = H t + const
    np.testing.assert_allclose(C.chebder(S @ c, 2, scl=1 / H), c[:N - 2], rtol=0, atol=1e-12)
    np.testing.assert_allclose(C.chebder(F @ c, 1, scl=1 / H)[:N - 2], c[:N - 2], rtol=0, atol=1e-12)
    # the two extra columns carry the null space: constant and linear terms
    e = np.zeros(N + 2)
    e[N] = 1
    np.testing.assert_allclose(S @ e, H * H * np.eye(N)[0], atol=1e-15)
    e[:] = 0
    e[N + 1] = 1
    np.testing.assert_allclose(S @ e, H * H * np.eye(N)[1], atol=1e-15)
    np.testing.assert_allclose(F @ e, H * np.eye(N)[0], atol=1e-15)


@pytest.mark.parametrize("N", [9, 16, 33])
def test_bc_rows_evaluate_boundary_values(N):
    H = 1.3
    rng = np.random.RandomState(7)
    x = rng.normal(size=N + 2)
    x[N - 2:N] = 0
    S, F = o.second_integral_matrix(N, H), o.first_integral_matrix(N, H)
    R1, R2, L1, L2 = o.no_wall_bc_rows(N)
    u, du = S @ x, F @ x
    assert R2 @ x == pytest.approx(C.chebval(1, u) / H ** 2, rel=1e-12)
    assert L2 @ x == pytest.approx(-C.chebval(-1, u) / H ** 2, rel=1e-12)
    assert R1 @ x == pytest.approx(C.chebval(1, du) / H, rel=1e-12)
    assert L1 @ x == pytest.approx(C.chebval(-1, du) / H, rel=1e-12)


# ---------------------------------------------------------------------------
# KBPENTA (LinearSolvers.h:52-141)
# ---------------------------------------------------------------------------
def test_kbpenta_matches_dense_solve():
    ops = o.DPOperators(8, 6, 19, 5.0, 4.0, 1.4)
    rng = np.random.RandomState(1)
    M, N = ops.d0.shape
    rhs = rng.normal(size=(M, N, 2)) + 1j * rng.normal(size=(M, N, 2))
    x = o.kbpenta_solve(ops.beta, ops.dp2, ops.dm2, rhs)
    for m in (1, 7, 23, M - 1):
        A = np.diag(ops.d0[m]) + np.diag(ops.dp2[m, :N - 2], 2) + np.diag(ops.dm2[m, :N - 2], -2)
        np.testing.assert_allclose(x[m], np.linalg.solve(A, rhs[m]), rtol=1e-12, atol=1e-12)


# ---------------------------------------------------------------------------
# Per-mode BVP (DPTools.cpp:81-345): Stokes equations + free-space decay
# ---------------------------------------------------------------------------
def _setup(Nx=6, Ny=5, Nz=33, Lx=5.0, Ly=4.0, Lz=3.0, seed=3):
    """Random smooth forcing (coefficients decaying fast enough that the
    Chebyshev truncation stays below 1e-13) that vanishes at both ends of the
    slab, as the spread forcing does: a forcing that jumps to zero outside
    would add a surface source to div f that no interior solve sees."""
    H = Lz / 2
    ops = o.DPOperators(Nx, Ny, Nz, Lx, Ly, H)
    rng = np.random.RandomState(seed)
    M = Nx * Ny
    decay = np.exp(-1.2 * np.arange(Nz - 2))[None, :, None]
    G = (rng.normal(size=(M, Nz - 2, 3)) + 1j * rng.normal(size=(M, Nz - 2, 3))) * decay
    Cf = np.zeros((M, Nz, 3), dtype=np.complex128)
    for m in range(M):
        for c in range(3):
            Cf[m, :, c] = C.chebmul(G[m, :, c], [0.5, 0, -0.5])   # (1 - t^2) g(t)
    return ops, Cf, H


def _divergence(ops, U, m, H):
    return ops.Dx[m] * U[m, :-1, 0] + ops.Dy[m] * U[m, :-1, 1] + _d(U[m, :, 2], H)


def _d(c, H, m=1):
    return C.chebder(c, m, scl=1 / H)


def test_no_wall_bvp_solves_stokes_with_decay_conditions():
    """eta (u'' - K^2 u) = grad p - f, p'' - K^2 p = div f inside; outside the
    slab f = 0 and everything decays, which fixes p' = -K p at the top,
    p' = K p at the bottom, and for the velocity
    u' + K u = -D p/(2 eta K) (x, y), w' + K w = p/(2 eta) at the top and
    u' - K u = D p/(2 eta K), w' - K w = p/(2 eta) at the bottom."""
    ops, Cf, H = _setup()
    eta = 0.7
    U, P = o.dp_solve_no_wall(ops, Cf, eta)
    worst = 0.0
    for m in range(1, Cf.shape[0]):
        K, Dx, Dy = ops.K[m], ops.Dx[m], ops.Dy[m]
        f, u, p = Cf[m], U[m], P[m]
        scale = np.abs(u).max() + np.abs(p).max()
        nz = len(p)
        res_p = _d(p, H, 2) - K * K * p[:nz - 2] - (Dx * f[:nz - 2, 0] + Dy * f[:nz - 2, 1] + _d(f[:, 2], H)[:nz - 2])
        grads = (Dx * p, Dy * p, np.concatenate((_d(p, H), [0])))
        res_u = [eta * (_d(u[:, c], H, 2) - K * K * u[:nz - 2, c]) - (grads[c][:nz - 2] - f[:nz - 2, c])
                 for c in range(3)]
        top = lambda c: C.chebval(1, c)
        bot = lambda c: C.chebval(-1, c)
        bcs = [top(_d(p, H)) + K * top(p), bot(_d(p, H)) - K * bot(p)]
        for c, D in ((0, Dx), (1, Dy)):
            bcs.append(top(_d(u[:, c], H)) + K * top(u[:, c]) + D * top(p) / (2 * eta * K))
            bcs.append(bot(_d(u[:, c], H)) - K * bot(u[:, c]) - D * bot(p) / (2 * eta * K))
        bcs.append(top(_d(u[:, 2], H)) + K * top(u[:, 2]) - top(p) / (2 * eta))
        bcs.append(bot(_d(u[:, 2], H)) - K * bot(u[:, 2]) - bot(p) / (2 * eta))
        worst = max(worst, np.abs(res_p).max() / scale, max(np.abs(r).max() for r in res_u) / scale,
                    max(abs(b) for b in bcs) / scale)
        if Dx == 1j * ops.Kx[m] and Dy == 1j * ops.Ky[m]:   # Nyquist derivatives are zeroed
            worst = max(worst, np.abs(_divergence(ops, U, m, H)).max() / scale / (1 + K))
    assert worst < 1e-9, worst


def _values_at(c, t):
    return C.chebval(t, c)


@pytest.mark.parametrize("geometry", ["bottom", "slit"])
def test_wall_corrections_enforce_no_slip_and_incompressibility(geometry):
    ops, Cf, H = _setup(seed=5)
    eta, Lz = 0.7, 2 * H
    z = o.clencurt(ops.Nz, 0, Lz)[0]
    U, P = o.dp_solve_no_wall(ops, Cf, eta)
    if geometry == "bottom":
        o.bottom_wall_correction(ops, U, P, z, eta)
    else:
        o.slit_channel_correction(ops, U, P, z, Lz, eta)
    for m in range(1, Cf.shape[0]):
        u, p = U[m], P[m]
        scale = np.abs(u).max()
        # no slip at z = 0 (t = -1), and at z = Lz (t = 1) for the slit
        for c in range(3):
            assert abs(_values_at(u[:, c], -1)) <= 1e-10 * scale
            if geometry == "slit":
                assert abs(_values_at(u[:, c], 1)) <= 1e-10 * scale
        # Nyquist modes: the no-wall solve zeroes Dx/Dy there while the
        # correction uses the grid wavenumbers (as the reference does), so the
        # field equations below hold on the other modes
        if not (ops.Dx[m] == ops.Kx[m] * 1j and ops.Dy[m] == ops.Ky[m] * 1j):
            continue
        # incompressible: Dx u + Dy v + w' = 0
        div = _divergence(ops, U, m, H)
        assert np.abs(div).max() <= 1e-9 * scale * (1 + ops.K[m])
        # the corrected field still solves the momentum equation
        K = ops.K[m]
        grads = (ops.Dx[m] * p, ops.Dy[m] * p, np.concatenate((_d(p, H), [0])))
        nz = len(p)
        for c in range(3):
            r = eta * (_d(u[:, c], H, 2) - K * K * u[:nz - 2, c]) - (grads[c][:nz - 2] - Cf[m, :nz - 2, c])
            assert np.abs(r).max() <= 1e-8 * (np.abs(u).max() + np.abs(p).max()) * (1 + K * K)


@pytest.mark.parametrize("geometry", ["bottom", "slit"])
def test_k0_velocity_solves_two_point_bvp(geometry):
    """k = 0: eta u'' = -f with u(0) = 0 and u'(Lz) = 0 (bottom) or u(Lz) = 0 (slit)."""
    ops, Cf, H = _setup(seed=9)
    eta = 0.7
    U = np.zeros((Cf.shape[0], ops.Nz, 3), dtype=np.complex128)
    f0 = Cf[0].real.astype(np.complex128)
    o.k0_wall_velocity(ops, f0, U, eta, geometry)
    for c in range(2):
        u = U[0, :, c]
        scale = np.abs(u).max()
        np.testing.assert_allclose(eta * _d(u, H, 2), -f0[:ops.Nz - 2, c], atol=1e-10 * scale)
        assert abs(_values_at(u, -1)) <= 1e-12 * scale
        if geometry == "bottom":
            assert abs(_values_at(_d(u, H), 1)) <= 1e-10 * scale
        else:
            assert abs(_values_at(u, 1)) <= 1e-12 * scale
    assert np.all(U[0, :, 2] == 0)


# ---------------------------------------------------------------------------
# Fourier-Chebyshev transform (Transform.cpp:107-304)
# ---------------------------------------------------------------------------
def test_fc_transform_is_fft2_times_chebyshev_interpolation():
    Ny, Nx, Nz = 6, 8, 11
    rng = np.random.RandomState(2)
    t = np.cos(np.pi * np.arange(Nz) / (Nz - 1))            # node 0 is the top
    F = rng.normal(size=(Ny, Nx, Nz, 3))
    got = o.fc_forward(F)
    for c in range(3):
        xy = np.fft.fft2(F[..., c], axes=(0, 1))             # unnormalised in x, y
        for iy in (0, 3, 5):
            for ix in (0, 1, 4):
                coef_re = C.chebfit(t, xy[iy, ix].real, Nz - 1)
                coef_im = C.chebfit(t, xy[iy, ix].imag, Nz - 1)
                np.testing.assert_allclose(got[iy, ix, :, c], coef_re + 1j * coef_im, atol=1e-11)
    np.testing.assert_allclose(o.fc_backward(got), F, atol=1e-13)


# ---------------------------------------------------------------------------
# Spreading and interpolation (SpreadInterp.cpp, BoundaryConditions.h)
# ---------------------------------------------------------------------------
def _small(geometry, torque=False, N=3, seed=4):
    Lx, Ly, Lz, n, nz = 8.0, 8.0, 6.0, 16, 17
    s = o.OracleDPStokes(geometry, Lx, Ly, 0.0, Lz, n, n, nz, 0.8, 6, 10.71, has_torque=torque,
                         w_d=6 if torque else None, betaw_d=10.71 if torque else None)
    rng = np.random.RandomState(seed)
    zlo = 0.0 if geometry != "nowall" else 1.6
    zhi = Lz if geometry == "slit" else Lz - 1.6
    pos = np.stack([rng.uniform(0, Lx, N), rng.uniform(0, Ly, N), rng.uniform(zlo, zhi, N)], 1).ravel()
    s.set_positions(pos)
    return s, pos


def test_spread_matches_brute_force_image_sum():
    """Bottom wall: every interior node gets phi(node - particle) from each
    particle whose support reaches it, minus the same kernel evaluated at the
    particle's mirror image below the wall (no-slip images, odd in z), with
    x-y periodic wrapping; open top.  Written as a sum over nodes and images
    instead of the reference's ghost-layer fold."""
    s, pos = _small("bottom", N=2, seed=11)
    pos = pos.reshape(-1, 3)
    pos[0, 2] = 0.4     # one particle close to the wall
    s.set_positions(pos.ravel())
    rng = np.random.RandomState(0)
    F = rng.normal(size=6)
    G = s.spread_forces(F)
    k = s.mono
    hx, hy = s.hx, s.hy
    z = s.zpts
    a2 = k.alpha ** 2
    phi = lambda r: np.exp(k.betaw * (np.sqrt(1 - r * r / a2) - 1))
    ref = np.zeros_like(G)
    w = int(round(2 * k.thresh / hx))
    for p in range(2):
        x, y, zp = pos[p]
        xc, yc = int(x // hx), int(y // hy)
        even = 1 - w % 2
        for i in range(w):
            for j in range(w):
                gx, gy = xc - w // 2 + even + i, yc - w // 2 + even + j
                dx, dy = gx * hx - x, gy * hy - y
                dx = dx if abs(dx) <= k.thresh else k.thresh       # the reference clamps offsets
                dy = dy if abs(dy) <= k.thresh else k.thresh
                wxy = phi(dx) * phi(dy) / k.norm ** 3
                for kz in range(s.Nz):
                    v = 0.0
                    if abs(z[kz] - zp) <= k.thresh:
                        v += phi(z[kz] - zp)
                    if abs(-z[kz] - zp) <= k.thresh:
                        v -= phi(-z[kz] - zp)
                    ref[gy % s.Ny, gx % s.Nx, kz, :] += wxy * v * F[3 * p:3 * p + 3]
    np.testing.assert_allclose(G, ref, rtol=0, atol=1e-13 * np.abs(ref).max())


@pytest.mark.parametrize("geometry", ["nowall", "bottom", "slit"])
def test_interpolation_is_the_adjoint_of_spreading(geometry):
    """<J u, F> = <u, S F>_W with W = hx hy w_z (the quadrature of the
    interpolation), for forces and, through the curls, for torques."""
    s, pos = _small(geometry, torque=True)
    N = s.N
    rng = np.random.RandomState(1)
    F = rng.normal(size=3 * N)
    T = rng.normal(size=3 * N)
    u = rng.normal(size=(s.Ny, s.Nx, s.Nz, 3))
    W = s.hx * s.hy * s.zwts[None, None, :, None]
    SF = s.spread_forces(F)
    v, _ = s.interpolate(u, False)
    assert np.dot(v, F) == pytest.approx(np.sum(W * u * SF), rel=1e-12)
    ST = s.spread_forces(np.zeros(3 * N), T)
    _, om = s.interpolate(u, True)
    assert np.dot(om, T) == pytest.approx(np.sum(W * u * ST), rel=1e-11)


# ---------------------------------------------------------------------------
# Whole Mdot: symmetric positive definite mobility (figures/pair_mobility_matrix_asym.py)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("geometry", ["bottom", "slit"])
def test_mobility_is_symmetric_positive_definite(geometry):
    s, pos = _small(geometry, torque=True, N=2, seed=8)
    n = 6 * s.N
    M = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1
        U, Om = s.mdot(e[:3 * s.N], e[3 * s.N:])
        M[:, j] = np.concatenate((U, Om))
    asym = np.linalg.norm(M - M.T) / np.linalg.norm(M)
    assert asym < 1e-3, asym
    assert np.linalg.eigvalsh(0.5 * (M + M.T)).min() > 0


def test_asymmetric_z_extension_is_refused():
    """Lz = 6, Nz = 13, support 1.5: Chebyshev nodes lie exactly one support
    width from each wall, rounding extends the grid on one side only, and the
    reference's equal-halves fold would run past the grid; the oracle refuses
    (as the GPU build does) instead."""
    ref = o.OracleDPStokes('nowall', 8.0, 8.0, 0.0, 6.0, 16, 16, 13, 1.0, 6, 1.785 * 6)
    with pytest.raises(ValueError, match="asymmetric"):
        ref.set_positions(np.array([4.0, 4.0, 3.0]))
