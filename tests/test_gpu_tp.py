"""Triply periodic mode ('periodic' / FCMJoint 'TP') on the GPU against the FP64
oracle (oracle/tp_oracle.py).  Bounds as for the doubly periodic path:
rel-L2 <= 1e-11 in FP64, <= 1e-5 in mixed precision."""
import numpy as np
import pytest

from oracle.tp_oracle import OracleTPStokes
from test_gpu_parity import rel_l2, say

pytestmark = pytest.mark.gpu

ETA = 1 / (6 * np.pi)
# name, Lx, Ly, Lz, nx, ny, nz, N, torque
CASES = [
    ("tp_cube", 8.0, 8.0, 8.0, 16, 16, 16, 24, True),
    ("tp_odd", 7.5, 8.5, 9.0, 15, 17, 18, 20, True),
    ("tp_tall", 8.0, 8.0, 16.0, 16, 16, 32, 30, True),    # two z chunks
    ("tp_noT", 8.0, 8.0, 6.0, 16, 16, 12, 20, False),
]


def inputs(case, seed):
    name, Lx, Ly, Lz, nx, ny, nz, N, torque = case
    rng = np.random.RandomState(seed)
    pos = np.stack([rng.uniform(0, Lx, N), rng.uniform(0, Ly, N), rng.uniform(0, Lz, N)], 1).ravel()
    pos[2] = 0.01          # supports wrapping through z = 0 and z = Lz
    pos[5] = Lz - 0.02
    F = rng.normal(size=3 * N)
    T = rng.normal(size=3 * N) if torque else None
    return pos, F, T


def oracle(case):
    name, Lx, Ly, Lz, nx, ny, nz, N, torque = case
    return OracleTPStokes(Lx, Ly, 0.0, Lz, nx, ny, nz, ETA, 6, 1.785 * 6, has_torque=torque,
                          w_d=6 if torque else None, betaw_d=1.785 * 6 if torque else None)


def gpu(case, precision):
    from doublyperiodicstokes_amd import DPStokes, StokesParameters
    name, Lx, Ly, Lz, nx, ny, nz, N, torque = case
    par = StokesParameters(viscosity=ETA, Lx=Lx, Ly=Ly, zmin=0.0, zmax=Lz, w=6, w_d=6 if torque else 0,
                           beta=1.785 * 6, beta_d=1.785 * 6, nx=nx, ny=ny, nz=nz, mode="periodic",
                           precision=precision)
    s = DPStokes()
    s.initialize(par, N)
    return s


@pytest.mark.parametrize("precision", ["double", "mixed"])
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_tp_mdot_matches_oracle(case, precision):
    say("case %s %s triply periodic mdot" % (case[0], precision))
    pos, F, T = inputs(case, 1)
    ref = oracle(case)
    ref.set_positions(pos)
    v_ref, w_ref = ref.mdot(F, T)
    s = gpu(case, precision)
    s.setPositions(pos)
    v, w = s.Mdot(F, T)
    tol = 1e-11 if precision == "double" else 1e-5
    ev = rel_l2(v, v_ref)
    ew = rel_l2(w, w_ref) if T is not None else 0.0
    say("  rel-L2 U = %.3e  W = %.3e" % (ev, ew))
    assert ev <= tol and ew <= tol
    s.clear()


def test_tp_through_joint_interface():
    from doublyperiodicstokes_amd import FCMJoint
    say("FCMJoint('hip') TP")
    nP, R = 12, 1.0
    j = FCMJoint("hip", precision="double")
    j.Initialize(nP, R, ETA, "TP", True, 24.0, 24.0, 0.0, 20.0)
    p = j.par
    rng = np.random.RandomState(4)
    pos = np.stack([rng.uniform(0, 24, nP), rng.uniform(0, 24, nP), rng.uniform(0, 20, nP)], 1).ravel()
    F, T = rng.normal(size=3 * nP), rng.normal(size=3 * nP)
    j.SetPositions(pos)
    MF, MT = j.Mdot(F, T)
    ref = OracleTPStokes(p.Lx, p.Ly, p.zmin, p.zmax, p.nx, p.ny, p.nz, ETA, p.w, p.beta, alpha=p.alpha,
                         w_d=p.w_d, betaw_d=p.beta_d, alpha_d=p.alpha_d, has_torque=True)
    ref.set_positions(pos)
    v, w = ref.mdot(F, T)
    say("  grid %dx%dx%d  rel-L2 U = %.3e  W = %.3e" % (p.nx, p.ny, p.nz, rel_l2(MF, v), rel_l2(MT, w)))
    assert rel_l2(MF, v) <= 1e-11 and rel_l2(MT, w) <= 1e-11
    j.Clean()
