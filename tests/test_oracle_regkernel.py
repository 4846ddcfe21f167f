"""Pins for the oracle's regular-kernel torque path (useRegKernel,
FCM.py:372-396 and 427-439; Solvers.py:137-179), none of which retypes its
formulas:

* the Chebyshev coefficient derivative equals numpy's ``chebder`` scaled by
  1/H;
* half the spectral curl of a smooth field (trigonometric in x, y; a
  polynomial of degree < Nz in z, or trigonometric in z when periodic),
  carried through the oracle's forward transform, curl and back transform,
  equals half its analytic curl at the grid points -- this fixes the node
  orientation, the signs and the component order;
* the triply periodic mobility with the regular torque kernel is symmetric
  (spreading and interpolation are adjoint, the Fourier curl is self-adjoint)
  and its rotational block positive definite;
* far from the source, the rotation a force induces and the translation a
  torque induces agree with the derivative-kernel path (both are half the
  curl of the same flow; only the kernel shapes, which the far field barely
  sees, differ).
"""
import numpy as np
import pytest
from numpy.polynomial import chebyshev as npcheb

from oracle.dp_oracle import (DPOperators, OracleDPStokes, cheb_coeff_diff_py, clencurt, curl_cheb,
                              fc_backward, fc_forward)
from oracle.tp_oracle import OracleTPStokes

ETA = 1 / (6 * np.pi)


def test_cheb_coeff_diff_is_chebder():
    rng = np.random.RandomState(3)
    for Nz, H in ((5, 1.0), (12, 2.5), (33, 0.7)):
        c = rng.normal(size=Nz) + 1j * rng.normal(size=Nz)
        got = cheb_coeff_diff_py(c[:, None], Nz, H)[:, 0]
        want = np.zeros(Nz, dtype=complex)
        want[:Nz - 1] = npcheb.chebder(c) / H
        assert np.allclose(got, want, rtol=0, atol=1e-12 * np.abs(want).max())


def _field(X, Y, Z, Lx, Ly):
    """A smooth vector field and its analytic curl at the given points."""
    a, b = 2 * np.pi / Lx, 2 * np.pi / Ly
    A = np.stack([np.cos(a * X) * Z ** 2,
                  np.sin(b * Y) * Z ** 3 + np.cos(a * X + b * Y) * Z,
                  np.cos(a * X + 2 * b * Y) * Z + np.sin(2 * a * X) * Z ** 2], -1)
    # curl = (dAz/dy - dAy/dz, dAx/dz - dAz/dx, dAy/dx - dAx/dy)
    dAz_dy = -2 * b * np.sin(a * X + 2 * b * Y) * Z
    dAz_dx = -a * np.sin(a * X + 2 * b * Y) * Z + 2 * a * np.cos(2 * a * X) * Z ** 2
    dAy_dz = 3 * np.sin(b * Y) * Z ** 2 + np.cos(a * X + b * Y)
    dAy_dx = -a * np.sin(a * X + b * Y) * Z
    dAx_dz = 2 * np.cos(a * X) * Z
    dAx_dy = 0 * X
    curl = np.stack([dAz_dy - dAy_dz, dAx_dz - dAz_dx, dAy_dx - dAx_dy], -1)
    return A, curl


@pytest.mark.parametrize("Nx,Ny,Nz", [(16, 12, 9), (15, 16, 14)])
def test_curl_cheb_analytic(Nx, Ny, Nz):
    Lx, Ly, Lz = 7.0, 5.0, 3.0
    H = Lz / 2
    z, _ = clencurt(Nz, 0, Lz)
    Y, X, Z = np.meshgrid(np.arange(Ny) * Ly / Ny, np.arange(Nx) * Lx / Nx, z, indexing="ij")
    A, curl = _field(X, Y, Z, Lx, Ly)
    ops = DPOperators(Nx, Ny, Nz, Lx, Ly, H)
    C = fc_forward(A).reshape(Ny * Nx, Nz, 3)
    got = fc_backward(curl_cheb(C, ops.Dx, ops.Dy, H).reshape(Ny, Nx, Nz, 3))
    assert np.abs(got - 0.5 * curl).max() <= 1e-11 * np.abs(curl).max()


@pytest.mark.parametrize("n", [(16, 12, 10), (15, 14, 9)])
def test_curl_fourier_analytic(n):
    Nx, Ny, Nz = n
    L = (7.0, 5.0, 3.0)
    o = OracleTPStokes(L[0], L[1], 0.0, L[2], Nx, Ny, Nz, ETA, 4, 1.785 * 4)
    Y, X, Z = np.meshgrid(*(np.arange(m) * l / m for m, l in ((Ny, L[1]), (Nx, L[0]), (Nz, L[2]))),
                          indexing="ij")
    a, b, c = (2 * np.pi / l for l in L)
    A = np.stack([np.cos(a * X + c * Z), np.sin(b * Y) * np.cos(2 * c * Z),
                  np.cos(a * X + 2 * b * Y) + np.sin(c * Z)], -1)
    curl = np.stack([-2 * b * np.sin(a * X + 2 * b * Y) + 2 * c * np.sin(b * Y) * np.sin(2 * c * Z),
                     -c * np.sin(a * X + c * Z) + a * np.sin(a * X + 2 * b * Y),
                     0 * X], -1)
    Ksq, Dx, Dy, Dz = o._operators()
    got = np.real(np.fft.ifftn(o.curl_hat(np.fft.fftn(A, axes=(0, 1, 2)), Dx, Dy, Dz), axes=(0, 1, 2)))
    assert np.abs(got - 0.5 * curl).max() <= 1e-11 * np.abs(curl).max()


def _mobility(o, N):
    M = np.zeros((6 * N, 6 * N))
    for j in range(6 * N):
        e = np.zeros(6 * N)
        e[j] = 1
        v, w = o.mdot(e[:3 * N], e[3 * N:])
        M[:3 * N, j] = v
        M[3 * N:, j] = w
    return M


def test_tp_regkernel_mobility_symmetric_positive():
    L, n, N = 10.0, 20, 3
    o = OracleTPStokes(L, L, 0.0, L, n, n, n, ETA, 6, 1.785 * 6, has_torque=True, w_d=6,
                       betaw_d=1.785 * 6, reg_kernel=True)
    rng = np.random.RandomState(5)
    o.set_positions(rng.uniform(0, L, (N, 3)).ravel())
    M = _mobility(o, N)
    assert np.abs(M - M.T).max() <= 1e-12 * np.abs(M).max()
    assert np.abs(M[3 * N:, :3 * N]).max() > 1e-3 * np.abs(M).max()   # coupling is present
    assert np.linalg.eigvalsh(M[3 * N:, 3 * N:]).min() > 0


def _pair(geom, reg, L, n, nz, zmax):
    kw = dict(has_torque=True, w_d=6, betaw_d=1.785 * 6, reg_kernel=reg)
    if geom == "periodic":
        return OracleTPStokes(L, L, 0.0, zmax, n, n, nz, ETA, 6, 1.785 * 6, **kw)
    return OracleDPStokes(geom, L, L, 0.0, zmax, n, n, nz, ETA, 6, 1.785 * 6, **kw)


@pytest.mark.parametrize("geom,n,nz,zmax", [("periodic", 64, 64, 32.0), ("slit", 64, 48, 16.0),
                                            ("nowall", 64, 48, 16.0)])
def test_far_field_matches_derivative_kernel(geom, n, nz, zmax):
    L = 32.0
    pos = np.array([8.0, 9.0, zmax / 2 - 1.0, 18.5, 14.0, zmax / 2 + 1.5])
    F = np.array([0.3, -1.0, 0.6, 0.0, 0.0, 0.0])
    T = np.array([-0.4, 0.8, 1.0, 0.0, 0.0, 0.0])
    out = {}
    for reg in (False, True):
        o = _pair(geom, reg, L, n, nz, zmax)
        o.set_positions(pos)
        _, w_from_f = o.mdot(F, np.zeros(6))
        v_from_t, _ = o.mdot(np.zeros(6), T)
        out[reg] = (w_from_f[3:], v_from_t[3:])
    for k in range(2):
        a, b = out[False][k], out[True][k]
        assert np.linalg.norm(b - a) <= 0.03 * np.linalg.norm(a), (k, a, b)
