"""The regular-kernel torque path (useRegKernel: torques spread with the plain
ES kernel of the dipole width, curls taken spectrally; FCM.py:372-396 and
427-439) on the GPU against the FP64 oracle's same path.  Bounds as for the
derivative-kernel path: rel-L2 <= 1e-11 in FP64, <= 1e-5 in mixed precision."""
import numpy as np
import pytest

from oracle import dp_oracle as o
from oracle.tp_oracle import OracleTPStokes
from test_gpu_parity import rel_l2, say

pytestmark = pytest.mark.gpu

ETA = 1 / (6 * np.pi)
# name, geometry, Lx, Ly, Lz, nx, ny, nz, N, w_d
CASES = [
    ("reg_nowall", "nowall", 16.0, 16.0, 10.0, 32, 32, 33, 40, 6),
    ("reg_bottom", "bottom", 15.5, 14.0, 9.0, 31, 28, 24, 40, 6),
    ("reg_slit", "slit", 16.0, 16.0, 10.0, 32, 32, 33, 48, 6),
    ("reg_slit_wd5", "slit", 13.5, 15.5, 8.0, 27, 31, 25, 40, 5),
    ("reg_tp", "periodic", 8.0, 8.0, 8.0, 16, 16, 16, 24, 6),
    ("reg_tp_odd", "periodic", 7.5, 8.5, 9.0, 15, 17, 18, 20, 5),
]


def inputs(case, seed):
    name, geom, Lx, Ly, Lz, nx, ny, nz, N, wd = case
    rng = np.random.RandomState(seed)
    a = 3 * Lx / nx
    zlo = a if geom == "nowall" else 0.0
    zhi = Lz - a if geom in ("nowall", "bottom") else Lz
    pos = np.stack([rng.uniform(0, Lx, N), rng.uniform(0, Ly, N), rng.uniform(zlo, zhi, N)], 1).ravel()
    return pos, rng.normal(size=3 * N), rng.normal(size=3 * N)


def oracle(case):
    name, geom, Lx, Ly, Lz, nx, ny, nz, N, wd = case
    kw = dict(has_torque=True, w_d=wd, betaw_d=1.785 * wd, reg_kernel=True)
    if geom == "periodic":
        return OracleTPStokes(Lx, Ly, 0.0, Lz, nx, ny, nz, ETA, 6, 1.785 * 6, **kw)
    return o.OracleDPStokes(geom, Lx, Ly, 0.0, Lz, nx, ny, nz, ETA, 6, 1.785 * 6, **kw)


def gpu(case, precision, reg=True, solver="auto"):
    from doublyperiodicstokes_amd import DPStokes, StokesParameters
    name, geom, Lx, Ly, Lz, nx, ny, nz, N, wd = case
    par = StokesParameters(viscosity=ETA, Lx=Lx, Ly=Ly, zmin=0.0, zmax=Lz, w=6, w_d=wd, beta=1.785 * 6,
                           beta_d=1.785 * wd, nx=nx, ny=ny, nz=nz, mode=geom, precision=precision,
                           useRegKernel=reg, solver=solver)
    s = DPStokes()
    s.initialize(par, N)
    return s


@pytest.mark.parametrize("precision", ["double", "mixed"])
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_regkernel_mdot_matches_oracle(case, precision):
    say("case %s %s regular torque kernel mdot" % (case[0], precision))
    pos, F, T = inputs(case, 2)
    ref = oracle(case)
    ref.set_positions(pos)
    v_ref, w_ref = ref.mdot(F, T)
    s = gpu(case, precision)
    s.setPositions(pos)
    v, w = s.Mdot(F, T)
    tol = 1e-11 if precision == "double" else 1e-5
    ev, ew = rel_l2(v, v_ref), rel_l2(w, w_ref)
    say("  rel-L2 U = %.3e  W = %.3e" % (ev, ew))
    assert ev <= tol and ew <= tol
    # without torques the path is the plain force Mdot of the same handle
    v0, _ = s.Mdot(F, None)
    ref.set_positions(pos)
    v0_ref, _ = ref.mdot(F, None)
    assert rel_l2(v0, v0_ref) <= tol
    s.clear()


@pytest.mark.parametrize("precision", ["double", "mixed"])
@pytest.mark.parametrize("case", CASES[1:3], ids=[c[0] for c in CASES[1:3]])
def test_regkernel_lane_solver_matches_oracle(case, precision):
    """The same with the one-lane spectral chain (curls between the z
    transforms and the BVP kernels instead of around the fused solve)."""
    say("case %s %s regular torque kernel mdot, lane solver" % (case[0], precision))
    pos, F, T = inputs(case, 2)
    ref = oracle(case)
    ref.set_positions(pos)
    v_ref, w_ref = ref.mdot(F, T)
    s = gpu(case, precision, solver="lane")
    s.setPositions(pos)
    v, w = s.Mdot(F, T)
    tol = 1e-11 if precision == "double" else 1e-5
    ev, ew = rel_l2(v, v_ref), rel_l2(w, w_ref)
    say("  rel-L2 U = %.3e  W = %.3e" % (ev, ew))
    assert ev <= tol and ew <= tol
    s.clear()


def test_regkernel_stage_entry_points_refuse_torque_grids():
    case = CASES[2]
    say("case %s stage entry points with the regular torque kernel" % case[0])
    pos, F, T = inputs(case, 3)
    s = gpu(case, "double")
    s.setPositions(pos)
    with pytest.raises(ValueError, match="regular torque kernel"):
        s.stage_spread(F, T)
    g = s.stage_spread(F)
    with pytest.raises(ValueError, match="regular torque kernel"):
        s.stage_interp(g, with_angular=True)
    s.clear()


def test_regkernel_through_joint_interface():
    from doublyperiodicstokes_amd import FCMJoint
    say("FCMJoint('hip') DPSC useRegKernel")
    nP, R = 16, 1.0
    j = FCMJoint("hip", precision="double")
    j.Initialize(nP, R, ETA, "DPSC", True, 24.0, 24.0, 0.0, 12.0, useRegKernel=True)
    p = j.par
    rng = np.random.RandomState(6)
    pos = np.stack([rng.uniform(0, 24, nP), rng.uniform(0, 24, nP), rng.uniform(1.5, 10.5, nP)], 1).ravel()
    F, T = rng.normal(size=3 * nP), rng.normal(size=3 * nP)
    j.SetPositions(pos)
    MF, MT = j.Mdot(F, T)
    ref = o.OracleDPStokes("slit", p.Lx, p.Ly, p.zmin, p.zmax, p.nx, p.ny, p.nz, ETA, p.w, p.beta,
                           alpha=p.alpha, w_d=p.w_d, betaw_d=p.beta_d, alpha_d=p.alpha_d, has_torque=True,
                           reg_kernel=True)
    ref.set_positions(pos)
    v, w = ref.mdot(F, T)
    say("  grid %dx%dx%d  rel-L2 U = %.3e  W = %.3e" % (p.nx, p.ny, p.nz, rel_l2(MF, v), rel_l2(MT, w)))
    assert rel_l2(MF, v) <= 1e-11 and rel_l2(MT, w) <= 1e-11
    j.Clean()
