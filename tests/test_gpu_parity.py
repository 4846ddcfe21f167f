"""GPU parity: the HIP path against the FP64 CPU oracle on identical seeded inputs.

Tolerances (relative L2 over all particles, stated per quantity):
  FP64 path  <= 1e-11   (north-star bound for the all-FP64 path)
  mixed path <= 1e-5    (FP32 grids/FFT/z-transform, FP64 BVP)
Stage tests use the FP64 path with the same bound.
"""
import sys

import numpy as np
import pytest

try:
    import torch  # noqa: F401  (first import can be slow on a fresh box)
except Exception:
    torch = None

from oracle import dp_oracle as o

pytestmark = pytest.mark.gpu

TOL64 = 1e-11
TOLMIX = 1e-5

# name, geometry, Lx, Ly, Lz, nx, ny, nz, N, torque
CASES = [
    ("nowall_even", "nowall", 16.0, 16.0, 10.0, 32, 32, 33, 48, True),
    ("bottom_even", "bottom", 16.0, 16.0, 10.0, 32, 32, 33, 48, True),
    ("slit_even", "slit", 16.0, 16.0, 10.0, 32, 32, 33, 48, True),
    ("bottom_odd", "bottom", 15.5, 14.0, 9.0, 31, 28, 24, 40, True),
    ("slit_odd", "slit", 13.5, 15.5, 8.0, 27, 31, 25, 40, True),
    ("nowall_noT", "nowall", 12.0, 16.0, 9.0, 24, 32, 20, 30, False),
    ("slit_nz96", "slit", 16.0, 16.0, 10.0, 32, 32, 96, 48, True),      # 16 lanes x 6 rows
    ("bottom_nz48", "bottom", 16.0, 16.0, 8.0, 32, 32, 48, 40, True),   # 8 lanes x 6 rows
    ("slit_nz14", "slit", 16.0, 16.0, 6.0, 32, 32, 14, 40, True),       # 8 lanes x 2 rows
    ("bottom_nz150", "bottom", 12.0, 12.0, 9.0, 24, 24, 150, 30, True),  # 32 lanes x 8 rows
    ("bottom_nz120", "bottom", 12.0, 12.0, 9.0, 24, 24, 120, 30, True),  # 16 lanes x 8 rows
    # 64 lanes x 8 rows; a tall box (the reference's Nz rule gives ~300 for
    # H = 48, h = 0.5), so that wall supports stay ~24 nodes.  A bottom wall:
    # a slit this tall is outside the reference's range, whose correction
    # divides exp(-k Lz) by exp(-k z) (DPTools.cpp:742, 801), 0/0 once
    # k Lz > ~709
    ("bottom_nz300", "bottom", 12.0, 12.0, 96.0, 24, 24, 300, 30, True),
    # odd Nz on the matrix-core transforms: 16 lanes x 6 and 16 lanes x 4 rows
    ("slit_nz85", "slit", 16.0, 16.0, 10.0, 32, 32, 85, 40, True),
    ("bottom_nz57", "bottom", 16.0, 16.0, 10.0, 32, 32, 57, 40, True),
]


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def make_case(case, seed=0):
    name, geom, Lx, Ly, Lz, nx, ny, nz, N, torque = case
    rng = np.random.RandomState(seed)
    h = Lx / nx
    a = 3 * h   # w = 6 support half width
    zlo = a if geom == "nowall" else 0.0
    zhi = Lz - a if geom in ("nowall", "bottom") else Lz
    x = rng.uniform(0, Lx, N)
    y = rng.uniform(0, Ly, N)
    z = rng.uniform(zlo, zhi, N)
    pos = np.stack([x, y, z], 1).ravel()
    F = rng.normal(size=3 * N)
    T = rng.normal(size=3 * N) if torque else None
    return pos, F, T


def oracle_for(case):
    name, geom, Lx, Ly, Lz, nx, ny, nz, N, torque = case
    return o.OracleDPStokes(geom, Lx, Ly, 0.0, Lz, nx, ny, nz, 1 / (6 * np.pi), 6, 1.785 * 6,
                            has_torque=torque, w_d=6 if torque else None,
                            betaw_d=1.785 * 6 if torque else None)


def gpu_for(case, precision, solver="auto"):
    from doublyperiodicstokes_amd import DPStokes, StokesParameters
    name, geom, Lx, Ly, Lz, nx, ny, nz, N, torque = case
    par = StokesParameters(viscosity=1 / (6 * np.pi), Lx=Lx, Ly=Ly, zmin=0.0, zmax=Lz, w=6,
                           w_d=6 if torque else 0, beta=1.785 * 6, beta_d=1.785 * 6,
                           nx=nx, ny=ny, nz=nz, mode=geom, precision=precision, solver=solver)
    s = DPStokes()
    s.initialize(par, N)
    return s


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_mdot_fp64_matches_oracle(case):
    say("case %s fp64 mdot" % case[0])
    pos, F, T = make_case(case)
    ref = oracle_for(case)
    ref.set_positions(pos)
    v_ref, w_ref = ref.mdot(F, T)
    s = gpu_for(case, "double")
    s.setPositions(pos)
    v, w = s.Mdot(F, T)
    ev = rel_l2(v, v_ref)
    say("  rel-L2 U = %.3e" % ev)
    assert ev <= TOL64
    if case[-1]:
        ew = rel_l2(w, w_ref)
        say("  rel-L2 W = %.3e" % ew)
        assert ew <= TOL64
    s.clear()


@pytest.mark.parametrize("precision", ["double", "mixed"])
@pytest.mark.parametrize("case", CASES[2:], ids=[c[0] for c in CASES[2:]])
def test_lane_solver_matches_oracle(case, precision):
    """The one-wavenumber-per-lane spectral path: the z transforms as MFMA
    GEMMs (odd and even Nz, Nz = 20 .. 150), the BVP one mode per lane."""
    say("case %s %s mdot, lane solver" % (case[0], precision))
    pos, F, T = make_case(case, seed=3)
    ref = oracle_for(case)
    ref.set_positions(pos)
    v_ref, w_ref = ref.mdot(F, T)
    s = gpu_for(case, precision, solver="lane")
    s.setPositions(pos)
    v, w = s.Mdot(F, T)
    ev = rel_l2(v, v_ref)
    ew = rel_l2(w, w_ref) if T is not None else 0.0
    say("  rel-L2 U = %.3e  W = %.3e" % (ev, ew))
    tol = TOL64 if precision == "double" else TOLMIX
    assert ev <= tol and ew <= tol
    s.clear()


# the first five cases, plus the 16-lane shapes whose FP32 transforms run on
# the matrix cores with even and odd Nz
MIXED_CASES = CASES[:5] + [c for c in CASES if c[0] in ("slit_nz96", "bottom_nz120", "slit_nz85", "bottom_nz57")]


@pytest.mark.parametrize("case", MIXED_CASES, ids=[c[0] for c in MIXED_CASES])
def test_mdot_mixed_matches_oracle(case):
    say("case %s mixed mdot" % case[0])
    pos, F, T = make_case(case, seed=1)
    ref = oracle_for(case)
    ref.set_positions(pos)
    v_ref, w_ref = ref.mdot(F, T)
    s = gpu_for(case, "mixed")
    s.setPositions(pos)
    v, w = s.Mdot(F, T)
    ev = rel_l2(v, v_ref)
    ew = rel_l2(w, w_ref)
    say("  rel-L2 U = %.3e  W = %.3e" % (ev, ew))
    assert ev <= TOLMIX and ew <= TOLMIX
    s.clear()


@pytest.mark.parametrize("case", CASES[:4], ids=[c[0] for c in CASES[:4]])
def test_stages_fp64(case):
    say("case %s fp64 stages" % case[0])
    pos, F, T = make_case(case, seed=2)
    ref = oracle_for(case)
    ref.set_positions(pos)
    Fg = ref.spread_forces(F, T)
    s = gpu_for(case, "double")
    s.setPositions(pos)
    G = s.stage_spread(F, T).cpu().numpy()
    es = rel_l2(G, Fg)
    say("  spread rel-L2 = %.3e" % es)
    assert es <= 1e-12
    import torch
    uG_ref = ref.solve(Fg)
    uG = s.stage_solve(torch.from_numpy(Fg)).cpu().numpy()
    eu = rel_l2(uG, uG_ref)
    say("  solve rel-L2 = %.3e" % eu)
    assert eu <= TOL64
    v_ref, w_ref = ref.interpolate(uG_ref, T is not None)
    v, w = s.stage_interp(torch.from_numpy(uG_ref), with_angular=T is not None)
    ei = rel_l2(v.cpu().numpy(), v_ref)
    say("  interp rel-L2 = %.3e" % ei)
    assert ei <= 1e-12
    if T is not None:
        assert rel_l2(w.cpu().numpy(), w_ref) <= 1e-12
    s.clear()
