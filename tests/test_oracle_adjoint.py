"""Pins for the oracle's adjoint of Mdot (oracle.dp_oracle.adjoint), none of
which retypes its formula:

* each derivative term alone -- d/dx_p of <lam, J(x) G> and of <G, S(x) F>_W
  for a fixed random grid G -- equals its central finite difference to 1e-7
  in every geometry (kernel derivative sign, fold factors of the derivative
  weights, quadrature weights);
* triply periodic: the solve is exactly self-adjoint, so the whole position
  gradient equals the central finite difference of the oracle's Mdot to
  FD accuracy, and the force gradient equals M^T lam to rounding;
* doubly periodic: the solve is self-adjoint only up to the discretisation
  (Chebyshev collocation against Clenshaw-Curtis weights), so the gradient
  equals the finite difference to that asymmetry (observed 1e-6 .. 9e-4 on
  the grids below; bound 2e-3);
* with torques, the second-derivative term alone matches the finite
  difference of the torque kernel's derivative interpolation (1e-6), and
  the whole position gradient matches FD as above (TP observed 2e-9, DP
  8e-6 .. 1.1e-4).
"""
import numpy as np
import pytest

from oracle.dp_oracle import OracleDPStokes, adjoint
from oracle.tp_oracle import OracleTPStokes

ETA = 1 / (6 * np.pi)


def _setup(geom, L, Lz, n, nz, N, seed):
    if geom == "periodic":
        o = OracleTPStokes(L, L, 0.0, Lz, n, n, nz, ETA, 6, 1.785 * 6)
        zlo, zhi = 0.0, Lz
    else:
        o = OracleDPStokes(geom, L, L, 0.0, Lz, n, n, nz, ETA, 6, 1.785 * 6)
        zlo = 1.6 if geom == "nowall" else 0.5
        zhi = Lz - (0.5 if geom == "slit" else 1.6)
    rng = np.random.RandomState(seed)
    pos = np.stack([rng.uniform(0, L, N), rng.uniform(0, L, N), rng.uniform(zlo, zhi, N)], 1).ravel()
    return o, pos, rng.normal(size=3 * N), rng.normal(size=3 * N), rng


def _fd_positions(o, pos, F, lam, h=1e-5):
    fd = np.zeros(pos.size)
    for k in range(pos.size):
        p = pos.copy()
        p[k] += h
        o.set_positions(p)
        up, _ = o.mdot(F)
        p[k] -= 2 * h
        o.set_positions(p)
        um, _ = o.mdot(F)
        fd[k] = lam @ (up - um) / (2 * h)
    o.set_positions(pos)
    return fd


@pytest.mark.parametrize("geom", ["nowall", "bottom", "slit", "periodic"])
def test_derivative_terms_match_finite_differences(geom):
    o, pos, F, lam, rng = _setup(geom, 8.0, 6.3, 16, 17 if geom != "periodic" else 16, 5, 1)
    N = 5
    o.set_positions(pos)
    G = rng.normal(size=(o.Ny, o.Nx, o.Nz, 3))
    g1 = np.stack([-(lam.reshape(N, 3) * o.interp_force_kernel(j, G).reshape(N, 3)).sum(1)
                   for j in (1, 2, 3)], 1).ravel()
    g2 = np.stack([-(F.reshape(N, 3) * o.interp_force_kernel(j, G).reshape(N, 3)).sum(1)
                   for j in (1, 2, 3)], 1).ravel()
    if geom == "periodic":
        wz = np.full(o.Nz, o.hz)
    else:
        wz = o.zwts
    h = 1e-6
    fd1 = np.zeros(3 * N)
    fd2 = np.zeros(3 * N)
    for k in range(3 * N):
        vals = []
        for sgn in (1, -1):
            p = pos.copy()
            p[k] += sgn * h
            o.set_positions(p)
            vals.append((lam @ o.interp_force_kernel(0, G), o.spread_forces(F, None)))
        fd1[k] = (vals[0][0] - vals[1][0]) / (2 * h)
        fd2[k] = np.sum((vals[0][1] - vals[1][1]) / (2 * h) * G * wz[None, None, :, None]) * o.hx * o.hy
    assert np.linalg.norm(g1 - fd1) <= 1e-7 * np.linalg.norm(fd1)
    assert np.linalg.norm(g2 - fd2) <= 1e-7 * np.linalg.norm(fd2)


def test_periodic_adjoint_is_exact():
    o, pos, F, lam, _ = _setup("periodic", 8.0, 8.0, 16, 16, 5, 2)
    o.set_positions(pos)
    gF, gT, gx = adjoint(o, F, lam)
    fd = _fd_positions(o, pos, F, lam)
    assert np.linalg.norm(gx - fd) <= 1e-6 * np.linalg.norm(fd)
    M = np.stack([o.mdot(e)[0] for e in np.eye(15)], 1)
    assert np.linalg.norm(gF - M.T @ lam) <= 1e-12 * np.linalg.norm(gF)
    assert gT is None


@pytest.mark.parametrize("geom,nz", [("bottom", 49), ("slit", 33), ("nowall", 49)])
def test_doubly_periodic_adjoint_within_solve_asymmetry(geom, nz):
    o, pos, F, lam, _ = _setup(geom, 16.0, 10.0, 32, nz, 4, 1)
    o.set_positions(pos)
    gF, gT, gx = adjoint(o, F, lam)
    fd = _fd_positions(o, pos, F, lam)
    assert np.linalg.norm(gx - fd) <= 2e-3 * np.linalg.norm(fd)


@pytest.mark.parametrize("geom", ["nowall", "bottom", "slit", "periodic"])
def test_second_derivative_term_matches_finite_differences(geom):
    """d/dx_p,k of the torque kernel's b-derivative interpolation of a fixed
    grid equals minus the (k, b) second-derivative interpolation, with the
    ghost class of b (BCst_z for z, BCs otherwise)."""
    kw = dict(has_torque=True, w_d=6, betaw_d=1.785 * 6)
    if geom == "periodic":
        o = OracleTPStokes(8.0, 8.0, 0.0, 6.3, 16, 16, 16, ETA, 6, 1.785 * 6, **kw)
        zlo, zhi = 0.0, 6.3
    else:
        o = OracleDPStokes(geom, 8.0, 8.0, 0.0, 6.3, 16, 16, 17, ETA, 6, 1.785 * 6, **kw)
        zlo, zhi = (1.6, 4.7) if geom == "nowall" else (0.3, 4.7 if geom == "bottom" else 6.0)
    N = 4
    rng = np.random.RandomState(9)
    pos = np.stack([rng.uniform(0, 8, N), rng.uniform(0, 8, N), rng.uniform(zlo, zhi, N)], 1).ravel()
    G = rng.normal(size=(o.Ny, o.Nx, o.Nz, 3))

    def first(b):
        if geom == "periodic":
            return o._interp(b + 1, o._tloc, o.dip, G).ravel()
        return o._interp(b + 1, o._tloc, o.dip, G, o.BCst_z if b == 2 else o.BCs, o._ext_buffer(o._tloc))

    h = 1e-6
    for b in range(3):
        for k in range(3):
            o.set_positions(pos)
            got = -o.interp_torque_kernel_d2(k, b, G)
            fd = np.zeros(3 * N)
            for p in range(N):
                pp = pos.copy()
                pp[3 * p + k] += h
                o.set_positions(pp)
                up = first(b)[3 * p:3 * p + 3]
                pp[3 * p + k] -= 2 * h
                o.set_positions(pp)
                um = first(b)[3 * p:3 * p + 3]
                fd[3 * p:3 * p + 3] = (up - um) / (2 * h)
            assert np.linalg.norm(got - fd) <= 1e-6 * np.linalg.norm(fd), (k, b)


@pytest.mark.parametrize("geom,nz,tol", [("periodic", 16, 1e-6), ("bottom", 49, 2e-3), ("slit", 33, 2e-3),
                                         ("nowall", 49, 2e-3)])
def test_adjoint_with_torques_matches_finite_differences(geom, nz, tol):
    kw = dict(has_torque=True, w_d=6, betaw_d=1.785 * 6)
    if geom == "periodic":
        o = OracleTPStokes(8.0, 8.0, 0.0, 8.0, 16, 16, 16, ETA, 6, 1.785 * 6, **kw)
        L, zlo, zhi, N = 8.0, 0.0, 8.0, 5
    else:
        o = OracleDPStokes(geom, 16.0, 16.0, 0.0, 10.0, 32, 32, nz, ETA, 6, 1.785 * 6, **kw)
        L, N = 16.0, 4
        zlo, zhi = (1.6, 8.4) if geom == "nowall" else (0.5, 8.4 if geom == "bottom" else 9.5)
    rng = np.random.RandomState(2)
    pos = np.stack([rng.uniform(0, L, N), rng.uniform(0, L, N), rng.uniform(zlo, zhi, N)], 1).ravel()
    F, T, lu, lw = (rng.normal(size=3 * N) for _ in range(4))
    o.set_positions(pos)
    gF, gT, gx = adjoint(o, F, lu, lw, T)
    h = 1e-5
    fd = np.zeros(3 * N)
    for k in range(3 * N):
        vals = []
        for sgn in (1, -1):
            p = pos.copy()
            p[k] += sgn * h
            o.set_positions(p)
            u, w = o.mdot(F, T)
            vals.append(lu @ u + lw @ w)
        fd[k] = (vals[0] - vals[1]) / (2 * h)
    assert np.linalg.norm(gx - fd) <= tol * np.linalg.norm(fd)


def test_regular_torque_kernel_gives_no_position_gradient():
    o = OracleDPStokes("slit", 8.0, 8.0, 0.0, 6.3, 16, 16, 17, ETA, 6, 1.785 * 6, has_torque=True, w_d=6,
                       betaw_d=1.785 * 6, reg_kernel=True)
    rng = np.random.RandomState(3)
    o.set_positions(np.stack([rng.uniform(0, 8, 3), rng.uniform(0, 8, 3), rng.uniform(0.5, 5.8, 3)], 1).ravel())
    F, T, lu, lw = (rng.normal(size=9) for _ in range(4))
    gF, gT, gx = adjoint(o, F, lu, lw, T)
    v, w = o.mdot(lu, lw)
    assert gx is None and np.array_equal(gF, v) and np.array_equal(gT, w)
