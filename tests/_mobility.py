"""Shared set-up for the golden mobility tests (not a test module).

The fixtures tests/golden/mobility_{bw,sc}.npz hold the reference's stored
self and pair mobilities (see tests/golden/make_golden.py).  This helper
configures grid and kernels exactly as FCMJoint.Initialize does for those
runs and builds normalised mobility columns the way
figures/mobility_matrix.py:5-40 does (unit force/torque per column).
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DOM = {"bw": ("DPBW", "bottom"), "sc": ("DPSC", "slit")}


def load(tag):
    return dict(np.load(os.path.join(HERE, "golden", "mobility_%s.npz" % tag), allow_pickle=False))


def config(tag, nP):
    """Grid/kernel choice of FCMJoint.Initialize for the stored runs."""
    from doublyperiodicstokes_amd.gridconfig import (configure_grid_and_kernels_xy,
                                                     configure_grid_and_kernels_z)
    g = load(tag)
    L, zmin, zmax, rad = float(g["L"]), float(g["zmin"]), float(g["zmax"]), float(g["radius"])
    Lx, Ly, hx, hy, nx, ny, w, wd, cb, cbd, b, bd = configure_grid_and_kernels_xy(
        L, L, rad * np.ones(nP), np.zeros(nP, dtype=int), 0, True, False)
    Lz, hz, nz, a = configure_grid_and_kernels_z(zmin, zmax, hx, w, wd, DOM[tag][0], 1.5, False)
    return dict(geometry=DOM[tag][1], domType=DOM[tag][0], Lx=Lx, Ly=Ly, nx=nx, ny=ny, nz=nz,
                zlo=a, zhi=a + Lz, eta=float(g["eta"]), w=int(w[0]), betaw=float(b[0] * w[0]),
                alpha=rad / (2 * cb[0]), w_d=int(wd[0]), betaw_d=float(bd[0] * wd[0]),
                alpha_d=rad / (2 * cbd[0]), L=L, zmin=zmin, zmax=zmax, radius=rad)


def oracle(tag, nP):
    from oracle import dp_oracle as o
    c = config(tag, nP)
    return o.OracleDPStokes(c["geometry"], c["Lx"], c["Ly"], c["zlo"], c["zhi"], c["nx"], c["ny"],
                            c["nz"], c["eta"], c["w"], c["betaw"], alpha=c["alpha"], w_d=c["w_d"],
                            betaw_d=c["betaw_d"], alpha_d=c["alpha_d"], has_torque=True)


def norm_matrix(nP, eta, rad=1.0):
    n = 3 * nP
    N = np.zeros((2 * n, 2 * n))
    N[:n, :n] = 1 / (6 * np.pi * eta * rad)
    N[:n, n:] = N[n:, :n] = 1 / (6 * np.pi * eta * rad ** 2)
    N[n:, n:] = 1 / (8 * np.pi * eta * rad ** 3)
    return N


def self_position(tag, iH):
    g = load(tag)
    L = float(g["L"])
    return np.array([L / 2, L / 2, g["self_heights"][iH]])


def pair_position(tag, iS, iH):
    g = load(tag)
    L, s, h = float(g["L"]), g["pair_seps"][iS], g["pair_heights"][iH]
    return np.array([L / 2 + s / 2, L / 2, h, L / 2 - s / 2, L / 2, h])


def columns(mdot, nP, cols, eta):
    """Normalised mobility columns: mdot(F, T) -> (U, W) for unit inputs."""
    n = 3 * nP
    Nm = norm_matrix(nP, eta)
    out = {}
    for j in cols:
        e = np.zeros(2 * n)
        e[j] = 1.0
        U, W = mdot(e[:n], e[n:])
        out[j] = np.concatenate((U, W)) / Nm[:, j]
    return out
