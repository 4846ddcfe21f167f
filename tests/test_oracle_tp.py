"""Pins for the triply periodic oracle (oracle/tp_oracle.py), none of which
retypes its formulas: the Fourier solve satisfies the Stokes equations and
incompressibility; spreading and interpolation are exact adjoints; a shift by
whole grid cells leaves velocities unchanged (the periodic wrap); the mobility
is symmetric positive definite; and the self-mobility of one particle follows
Hasimoto's periodic correction mu(L) = mu0 (1 - 2.837297 R/L + O(R/L)^3) with
an effective radius R that matches the free-space mu0 = 1/(6 pi eta R)."""
import numpy as np
import pytest

from oracle.tp_oracle import OracleTPStokes

ETA = 1 / (6 * np.pi)


def oracle(L=8.0, n=16, torque=True, nz=None, Lz=None):
    nz = nz or n
    Lz = Lz or L
    return OracleTPStokes(L, L, 0.0, Lz, n, n, nz, ETA, 6, 1.785 * 6, has_torque=torque,
                          w_d=6 if torque else None, betaw_d=1.785 * 6 if torque else None)


def test_solve_satisfies_stokes_and_incompressibility():
    """Away from the Nyquist planes (where the reference zeroes the unpaired
    derivative but not k^2, Solvers.py:213-219) the solution is divergence
    free and f - eta k^2 u is a pure gradient (parallel to k)."""
    o = oracle()
    rng = np.random.RandomState(0)
    F = rng.normal(size=(o.Ny, o.Nx, o.Nz, 3))
    F -= F.mean(axis=(0, 1, 2))
    U = o.solve(F)
    Fh = np.fft.fftn(F, axes=(0, 1, 2))
    Uh = np.fft.fftn(U, axes=(0, 1, 2))
    Ksq, Dx, Dy, Dz = o._operators()
    keep = np.ones(Ksq.shape, dtype=bool)
    keep[o.Ny // 2, :, :] = keep[:, o.Nx // 2, :] = keep[:, :, o.Nz // 2] = False
    keep[0, 0, 0] = False
    scale = np.abs(Fh).max()
    div = Dx * Uh[..., 0] + Dy * Uh[..., 1] + Dz * Uh[..., 2]
    assert np.abs(div[keep]).max() <= 1e-12 * scale
    R = Fh - ETA * Ksq[..., None] * Uh
    k = np.stack([Dx.imag, Dy.imag, Dz.imag], -1)
    cross = np.cross(R, k)
    assert np.abs(cross[keep]).max() <= 1e-11 * scale * np.abs(k).max()
    # and the k = 0 mode is dropped
    assert np.abs(Uh[0, 0, 0]).max() <= 1e-12 * scale


def test_spread_interp_adjoint():
    o = oracle()
    rng = np.random.RandomState(1)
    N = 7
    o.set_positions(rng.uniform(0, 8.0, 3 * N))
    f = rng.normal(size=(N, 3))
    uG = rng.normal(size=(o.Ny, o.Nx, o.Nz, 3))
    for kind, loc, ks in ((0, o._loc, o.mono), (1, o._tloc, o.dip), (2, o._tloc, o.dip), (3, o._tloc, o.dip)):
        lhs = np.sum(o._interp(kind, loc, ks, uG) * f)
        rhs = np.sum(uG * o._spread(kind, loc, ks, f)) * o.hx * o.hy * o.hz
        assert abs(lhs - rhs) <= 1e-12 * abs(rhs)


def test_shift_by_whole_cells_is_invariant():
    o = oracle(L=8.0, n=16)
    rng = np.random.RandomState(2)
    N = 6
    pos = rng.uniform(0, 8.0, 3 * N)
    F, T = rng.normal(size=3 * N), rng.normal(size=3 * N)
    o.set_positions(pos)
    v0, w0 = o.mdot(F, T)
    sh = pos.copy()
    sh[0::3] += 3 * o.hx
    sh[1::3] -= 5 * o.hy
    sh[2::3] += 7 * o.hz
    sh[2::3] = np.mod(sh[2::3], 8.0)
    o.set_positions(sh)
    v1, w1 = o.mdot(F, T)
    assert np.abs(v1 - v0).max() <= 1e-11 * np.abs(v0).max()
    assert np.abs(w1 - w0).max() <= 1e-11 * np.abs(w0).max()


def test_mobility_symmetric_positive_definite():
    o = oracle(L=6.0, n=12)
    rng = np.random.RandomState(3)
    N = 4
    o.set_positions(rng.uniform(0, 6.0, 3 * N))
    M = np.zeros((6 * N, 6 * N))
    for j in range(6 * N):
        e = np.zeros(6 * N)
        e[j] = 1
        v, w = o.mdot(e[:3 * N], e[3 * N:])
        M[:, j] = np.concatenate([v, w])
    assert np.abs(M - M.T).max() <= 1e-10 * np.abs(M).max()
    assert np.linalg.eigvalsh(0.5 * (M + M.T)).min() > 0


def test_self_mobility_follows_hasimoto():
    h = 0.5
    mus = {}
    for n in (48, 64):
        o = oracle(L=n * h, n=n, torque=False)
        o.set_positions([0.37 * n * h, 0.41 * n * h, 0.53 * n * h])
        v, _ = o.mdot([0.0, 0.0, 1.0])
        mus[n * h] = v[2]
    (L1, m1), (L2, m2) = sorted(mus.items())
    c = 2.837297
    # m = mu0 (1 - c R / L) at two box sizes -> mu0 and R
    mu0 = (m2 * L2 - m1 * L1) / (L2 - L1)
    R = (1 - m1 / mu0) * L1 / c
    # the ES kernel with w = 6, beta w = 10.71 has a hydrodynamic radius of
    # about 1.5 grid cells
    assert 1.3 * h < R < 1.8 * h
    assert abs(6 * np.pi * ETA * mu0 * R - 1) < 1e-2
