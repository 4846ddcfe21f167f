"""Shared set-up for the golden mobility tests (not a test module).

The fixtures tests/golden/mobility_*.npz hold the reference's stored self and
pair mobilities and the FCMJoint.Initialize arguments of the run that wrote
each (see tests/golden/make_golden.py).  This helper configures grid and
kernels exactly as FCMJoint.Initialize does for those runs and builds
normalised mobility columns the way figures/mobility_matrix.py:5-40 does (unit
force/torque per column).

A fixture is named by its file stem ('mobility_bw', 'mobility_sc_ref', ...);
`kind` is 'self' (one particle) or 'pair' (two particles).
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DOM = {"bw": ("DPBW", "bottom"), "sc": ("DPSC", "slit")}
NP = {"self": 1, "pair": 2}


def load(name):
    return dict(np.load(os.path.join(HERE, "golden", name + ".npz"), allow_pickle=False))


def _dom(name):
    return DOM["bw" if "_bw" in name else "sc"]


def init_args(name, kind):
    g = load(name)
    return dict(kernType=int(g[kind + "_kernType"]), has_torque=bool(g[kind + "_has_torque"]),
                ref=bool(g[kind + "_ref"]))


def config(name, kind):
    """Grid/kernel choice of FCMJoint.Initialize for the stored run."""
    from doublyperiodicstokes_amd.gridconfig import (configure_grid_and_kernels_xy,
                                                     configure_grid_and_kernels_z)
    g = load(name)
    a = init_args(name, kind)
    nP = NP[kind]
    L, zmin, zmax, rad = float(g["L"]), float(g["zmin"]), float(g["zmax"]), float(g["radius"])
    Lx, Ly, hx, hy, nx, ny, w, wd, cb, cbd, b, bd = configure_grid_and_kernels_xy(
        L, L, rad * np.ones(nP), a["kernType"] * np.ones(nP, dtype=int), 0, a["has_torque"],
        a["ref"])
    dom, geom = _dom(name)
    Lz, hz, nz, z0 = configure_grid_and_kernels_z(zmin, zmax, hx, w, wd, dom, 1.5, a["ref"])
    c = dict(geometry=geom, domType=dom, Lx=Lx, Ly=Ly, nx=nx, ny=ny, nz=nz, zlo=z0, zhi=z0 + Lz,
             eta=float(g["eta"]), w=int(w[0]), betaw=float(b[0] * w[0]), alpha=rad / (2 * cb[0]),
             L=L, zmin=zmin, zmax=zmax, radius=rad, nP=nP, **a)
    if a["has_torque"]:
        c.update(w_d=int(wd[0]), betaw_d=float(bd[0] * wd[0]), alpha_d=rad / (2 * cbd[0]))
    return c


def oracle(name, kind):
    from oracle import dp_oracle as o
    c = config(name, kind)
    kw = dict(alpha=c["alpha"])
    if c["has_torque"]:
        kw.update(w_d=c["w_d"], betaw_d=c["betaw_d"], alpha_d=c["alpha_d"], has_torque=True)
    return o.OracleDPStokes(c["geometry"], c["Lx"], c["Ly"], c["zlo"], c["zhi"], c["nx"], c["ny"],
                            c["nz"], c["eta"], c["w"], c["betaw"], **kw)


def norm_matrix(nP, eta, has_torque=True, rad=1.0):
    n = 3 * nP
    N = np.zeros((2 * n, 2 * n))
    N[:n, :n] = 1 / (6 * np.pi * eta * rad)
    N[:n, n:] = N[n:, :n] = 1 / (6 * np.pi * eta * rad ** 2)
    N[n:, n:] = 1 / (8 * np.pi * eta * rad ** 3)
    return N if has_torque else N[:n, :n]


def self_position(name, iH):
    g = load(name)
    L = float(g["L"])
    return np.array([L / 2, L / 2, g["self_heights"][iH]])


def pair_position(name, iS, iH):
    g = load(name)
    L, s, h = float(g["L"]), g["pair_seps"][iS], g["pair_heights"][iH]
    return np.array([L / 2 + s / 2, L / 2, h, L / 2 - s / 2, L / 2, h])


def _apply(mdot, nP, v, has_torque):
    n = 3 * nP
    if has_torque:
        U, W = mdot(v[:n], v[n:])
        return np.concatenate((U, W))
    U, _ = mdot(v)
    return np.asarray(U)


def columns(mdot, nP, cols, eta, has_torque=True):
    """Normalised mobility columns: mdot(F[, T]) -> (U, W) for unit inputs."""
    Nm = norm_matrix(nP, eta, has_torque)
    out = {}
    for j in cols:
        e = np.zeros(Nm.shape[0])
        e[j] = 1.0
        out[j] = _apply(mdot, nP, e, has_torque) / Nm[:, j]
    return out


def combination_error(mdot, nP, eta, has_torque, Mref, v):
    """One Mdot of the combination v of unit inputs against the stored matrix:
    max-norm of (M v - Mref v) relative to max|Mref v| (un-normalised)."""
    Nm = norm_matrix(nP, eta, has_torque)
    want = (Mref * Nm) @ v
    got = _apply(mdot, nP, v, has_torque)
    return np.abs(got - want).max() / np.abs(want).max()
