"""The adjoint of Mdot (dps_mdot_adjoint and the torch.autograd wrapper) on the
GPU:

* against the oracle's adjoint (oracle.dp_oracle.adjoint) on the same inputs:
  rel-L2 <= 1e-11 in FP64, <= 1e-5 in mixed precision, for position, force
  and torque gradients in every geometry, with and without torques;
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


@pytest.mark.parametrize("precision", ["double", "mixed"])
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_torque_adjoint_matches_oracle(case, precision):
    import torch
    say("case %s %s torque adjoint" % (case[0], precision))
    pos, F, lam, T, lw = _inputs(case, 5, torque=True)
    ref = _oracle(case, torque=True)
    ref.set_positions(pos)
    gF_ref, gT_ref, gx_ref = adjoint(ref, F, lam, lw, T)
    s = _gpu(case, precision, torque=True)
    s.setPositions(pos)
    dev = torch.device("cuda", 0)
    dt = torch.float64 if precision == "double" else torch.float32
    t = lambda a: torch.tensor(a, dtype=dt, device=dev)
    gF, gT, gX = s.mdot_adjoint_device(t(F), t(lam), torques=t(T), grad_angular=t(lw), positions=True)
    c = lambda g: g.double().cpu().numpy()
    ex, ef, et = rel_l2(c(gX), gx_ref), rel_l2(c(gF), gF_ref), rel_l2(c(gT), gT_ref)
    say("  rel-L2 dL/dx = %.3e  dL/dF = %.3e  dL/dT = %.3e" % (ex, ef, et))
    tol = 1e-11 if precision == "double" else 1e-5
    assert ex <= tol and ef <= tol and et <= tol
    gF2, gT2, gX2 = s.mdot_adjoint_device(t(F), t(lam), torques=t(T), grad_angular=t(lw), positions=False)
    assert gX2 is None and rel_l2(c(gF2), gF_ref) <= tol and rel_l2(c(gT2), gT_ref) <= tol
    s.clear()


def test_regular_torque_kernel_refuses_position_gradients_with_torques():
    import torch
    from doublyperiodicstokes_amd import DPStokes, StokesParameters
    case = CASES[1]
    say("case %s regular torque kernel adjoint refusal" % case[0])
    pos, F, lam, T, lw = _inputs(case, 5, torque=True)
    name, geom, L, Lz, n, nz, N = case
    par = StokesParameters(viscosity=ETA, Lx=L, Ly=L, zmin=0.0, zmax=Lz, w=6, w_d=6, beta=1.785 * 6,
                           beta_d=1.785 * 6, nx=n, ny=n, nz=nz, mode=geom, precision="double", useRegKernel=True)
    s = DPStokes()
    s.initialize(par, N)
    s.setPositions(pos)
    t = lambda a: torch.tensor(a, dtype=torch.float64, device=torch.device("cuda", 0))
    with pytest.raises(ValueError, match="derivative torque kernels"):
        s.mdot_adjoint_device(t(F), t(lam), torques=t(T), grad_angular=t(lw), positions=True)
    s.clear()


@pytest.mark.parametrize("torque", [False, True], ids=["F", "FT"])
@pytest.mark.parametrize("case,tol", [(("fd_tp", "periodic", 8.0, 8.0, 16, 16, 5), 1e-6),
                                      (("fd_bottom", "bottom", 16.0, 10.0, 32, 49, 4), 2e-3)],
                         ids=["fd_tp", "fd_bottom"])
def test_autograd_matches_finite_differences(case, tol, torque):
    import torch
    from doublyperiodicstokes_amd.autograd import mdot
    say("case %s torque=%s autograd vs central FD" % (case[0], torque))
    pos, F, lam, T, lw = _inputs(case, 6, torque=torque)
    s = _gpu(case, "double", torque=torque)
    dev = torch.device("cuda", 0)
    x = torch.tensor(pos, dtype=torch.float64, device=dev, requires_grad=True)
    f = torch.tensor(F, dtype=torch.float64, device=dev, requires_grad=True)
    tq = torch.tensor(T, dtype=torch.float64, device=dev, requires_grad=True) if torque else None
    c = torch.tensor(lam, dtype=torch.float64, device=dev)
    cw = torch.tensor(lw, dtype=torch.float64, device=dev) if torque else None
    U, W = mdot(s, x, f, tq)
    loss = (U * c).sum() + ((W * cw).sum() if torque else 0)
    loss.backward()

    def value(p):
        s.setPositions(p)
        u, w = s.mdot_device(torch.tensor(F, dtype=torch.float64, device=dev),
                             torch.tensor(T, dtype=torch.float64, device=dev) if torque else None)
        return float((u * c).sum()) + (float((w * cw).sum()) if torque else 0.0)

    h = 1e-5
    fd = np.zeros(pos.size)
    for k in range(pos.size):
        p = pos.copy()
        p[k] += h
        vp = value(p)
        p[k] -= 2 * h
        fd[k] = (vp - value(p)) / (2 * h)
    ex = rel_l2(x.grad.cpu().numpy(), fd)
    if torque:
        s.clear()
        say("  rel dL/dx vs FD = %.3e" % ex)
        assert ex <= tol and tq.grad is not None
        return
    # the force gradient: M^T c against columns of M
    s.setPositions(pos)
    M = np.stack([s.mdot_device(torch.tensor(e, dtype=torch.float64, device=dev))[0].cpu().numpy()
                  for e in np.eye(pos.size)], 1)
    ef = rel_l2(f.grad.cpu().numpy(), M.T @ lam)
    say("  rel dL/dx vs FD = %.3e  dL/dF vs M^T c = %.3e" % (ex, ef))
    assert W is None and ex <= tol and ef <= max(tol, 1e-12)
    s.clear()
