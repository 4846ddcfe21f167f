"""The adjoint of Mdot (dps_mdot_adjoint and the torch.autograd wrapper) on the
GPU:

* against the oracle's adjoint (oracle.dp_oracle.adjoint) on the same inputs:
  rel-L2 <= 1e-11 in FP64, <= 1e-5 in mixed precision, for position, force
  and torque gradients in every geometry;
* torch.autograd gradients against central finite differences of the GPU's
  own FP64 Mdot: to FD accuracy when triply periodic (the solve is exactly
  self-adjoint), within 2e-3 for a resolved bottom-wall grid (the solve's
  discretisation asymmetry, as pinned for the oracle).
"""
import numpy as np
import pytest

from oracle.dp_oracle import OracleDPStokes, adjoint
from oracle.tp_oracle import OracleTPStokes
from test_gpu_parity import rel_l2, say

pytestmark = pytest.mark.gpu

ETA = 1 / (6 * np.pi)
# name, geometry, L, Lz, n, nz, N
CASES = [
    ("adj_nowall", "nowall", 16.0, 10.0, 32, 33, 24),
    ("adj_bottom", "bottom", 16.0, 10.0, 32, 33, 24),
    ("adj_slit", "slit", 15.5, 9.0, 31, 25, 24),
    ("adj_tp", "periodic", 8.0, 8.0, 16, 16, 20),
]


def _inputs(case, seed, torque=False):
    name, geom, L, Lz, n, nz, N = case
    rng = np.random.RandomState(seed)
    zlo = 1.6 if geom == "nowall" else (0.0 if geom == "periodic" else 0.3)
    zhi = Lz - 1.6 if geom in ("nowall", "bottom") else Lz - (0.0 if geom == "periodic" else 0.3)
    pos = np.stack([rng.uniform(0, L, N), rng.uniform(0, L, N), rng.uniform(zlo, zhi, N)], 1).ravel()
    F, lam = rng.normal(size=3 * N), rng.normal(size=3 * N)
    T = rng.normal(size=3 * N) if torque else None
    lw = rng.normal(size=3 * N) if torque else None
    return pos, F, lam, T, lw


def _oracle(case, torque=False):
    name, geom, L, Lz, n, nz, N = case
    kw = dict(has_torque=torque, w_d=6 if torque else None, betaw_d=1.785 * 6 if torque else None)
    if geom == "periodic":
        return OracleTPStokes(L, L, 0.0, Lz, n, n, nz, ETA, 6, 1.785 * 6, **kw)
    return OracleDPStokes(geom, L, L, 0.0, Lz, n, n, nz, ETA, 6, 1.785 * 6, **kw)


def _gpu(case, precision, torque=False):
    from doublyperiodicstokes_amd import DPStokes, StokesParameters
    name, geom, L, Lz, n, nz, N = case
    par = StokesParameters(viscosity=ETA, Lx=L, Ly=L, zmin=0.0, zmax=Lz, w=6, w_d=6 if torque else 0,
                           beta=1.785 * 6, beta_d=1.785 * 6, nx=n, ny=n, nz=nz, mode=geom, precision=precision)
    s = DPStokes()
    s.initialize(par, N)
    return s


@pytest.mark.parametrize("precision", ["double", "mixed"])
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_position_adjoint_matches_oracle(case, precision):
    import torch
    say("case %s %s position adjoint" % (case[0], precision))
    pos, F, lam, _, _ = _inputs(case, 4)
    ref = _oracle(case)
    ref.set_positions(pos)
    gF_ref, _, gx_ref = adjoint(ref, F, lam)
    s = _gpu(case, precision)
    s.setPositions(pos)
    dt = torch.float64 if precision == "double" else torch.float32
    dev = torch.device("cuda", 0)
    gF, gT, gX = s.mdot_adjoint_device(torch.tensor(F, dtype=dt, device=dev), torch.tensor(lam, dtype=dt, device=dev))
    ex, ef = rel_l2(gX.double().cpu().numpy(), gx_ref), rel_l2(gF.double().cpu().numpy(), gF_ref)
    say("  rel-L2 dL/dx = %.3e  dL/dF = %.3e" % (ex, ef))
    tol = 1e-11 if precision == "double" else 1e-5
    assert gT is None and ex <= tol and ef <= tol
    s.clear()


@pytest.mark.parametrize("case", CASES[1:3], ids=[c[0] for c in CASES[1:3]])
def test_torque_adjoint_matches_oracle(case):
    import torch
    say("case %s torque adjoint" % case[0])
    pos, F, lam, T, lw = _inputs(case, 5, torque=True)
    ref = _oracle(case, torque=True)
    ref.set_positions(pos)
    gF_ref, gT_ref, gx_ref = adjoint(ref, F, lam, lw, T)
    s = _gpu(case, "double", torque=True)
    s.setPositions(pos)
    dev = torch.device("cuda", 0)
    t = lambda a: torch.tensor(a, dtype=torch.float64, device=dev)
    gF, gT, gX = s.mdot_adjoint_device(t(F), t(lam), torques=t(T), grad_angular=t(lw), positions=False)
    ef, et = rel_l2(gF.cpu().numpy(), gF_ref), rel_l2(gT.cpu().numpy(), gT_ref)
    say("  rel-L2 dL/dF = %.3e  dL/dT = %.3e" % (ef, et))
    assert gx_ref is None and gX is None and ef <= 1e-11 and et <= 1e-11
    with pytest.raises(ValueError, match="forces only"):
        s.mdot_adjoint_device(t(F), t(lam), torques=t(T), grad_angular=t(lw), positions=True)
    s.clear()


@pytest.mark.parametrize("case,tol", [(("fd_tp", "periodic", 8.0, 8.0, 16, 16, 5), 1e-6),
                                      (("fd_bottom", "bottom", 16.0, 10.0, 32, 49, 4), 2e-3)],
                         ids=["fd_tp", "fd_bottom"])
def test_autograd_matches_finite_differences(case, tol):
    import torch
    from doublyperiodicstokes_amd.autograd import mdot
    say("case %s autograd vs central FD" % case[0])
    pos, F, lam, _, _ = _inputs(case, 6)
    s = _gpu(case, "double")
    dev = torch.device("cuda", 0)
    x = torch.tensor(pos, dtype=torch.float64, device=dev, requires_grad=True)
    f = torch.tensor(F, dtype=torch.float64, device=dev, requires_grad=True)
    c = torch.tensor(lam, dtype=torch.float64, device=dev)
    U, W = mdot(s, x, f)
    (U * c).sum().backward()
    h = 1e-5
    fd = np.zeros(pos.size)
    for k in range(pos.size):
        vals = []
        for sgn in (1, -1):
            p = pos.copy()
            p[k] += sgn * h
            s.setPositions(p)
            vals.append(float((s.mdot_device(torch.tensor(F, dtype=torch.float64, device=dev))[0] * c).sum()))
        fd[k] = (vals[0] - vals[1]) / (2 * h)
    ex = rel_l2(x.grad.cpu().numpy(), fd)
    # the force gradient: M^T c against columns of M
    s.setPositions(pos)
    M = np.stack([s.mdot_device(torch.tensor(e, dtype=torch.float64, device=dev))[0].cpu().numpy()
                  for e in np.eye(pos.size)], 1)
    ef = rel_l2(f.grad.cpu().numpy(), M.T @ lam)
    say("  rel dL/dx vs FD = %.3e  dL/dF vs M^T c = %.3e" % (ex, ef))
    assert W is None and ex <= tol and ef <= max(tol, 1e-12)
    s.clear()
