"""The benchmark workloads at their full sizes on the GPU.

Inputs are the seeded ones bench.py times (numpy RandomState(0), draw order
x, y, z, F, T; bench.make_inputs / bench.batch_workload):
  bottom2k       bottom wall, N = 2048, 128 x 128 x 64, forces only
  slit16k        slit, N = 16384, 256 x 256 x 96, forces and torques
  bottom8k       bottom wall, N = 8192, 192 x 192 x 80, forces and torques
                 (differentiation: L = sum U . g, g iid N(0, 1), RandomState(1))
  batch64        64 bottom-wall instances, N log-uniform in [256, 32768]
  monolayer256k  bottom wall, N = 262144, 1024 x 1024 x 48, z = 1.1 + |N(0, 0.3)|

Bars (rel-L2 over all particles, per quantity): FP64 <= 1e-11 and mixed
<= 1e-5 against the FP64 oracle where the oracle runs in seconds (bottom2k,
slit16k, bottom8k, three batch instances).  For the monolayer, whose oracle
would take minutes, size-independent properties: mixed vs FP64 <= 1e-5,
finite outputs, no-slip of the velocity grid at the wall, linearity.
"""
import numpy as np
import pytest

import bench
from oracle import dp_oracle as o
from test_gpu_parity import rel_l2, say

pytestmark = pytest.mark.gpu

ETA = 1 / (6 * np.pi)


def oracle_for(wl):
    geom, N, Lx, Ly, zmax, zlo, zhi, nx, ny, nz, torque = wl
    np.seterr(all="ignore")
    return o.OracleDPStokes(geom, Lx, Ly, 0.0, zmax, nx, ny, nz, ETA, 6, 1.785 * 6, has_torque=torque,
                            w_d=6 if torque else None, betaw_d=1.785 * 6 if torque else None)


def dev_tensors(torch, dt, *arrs):
    d = torch.device("cuda", 0)
    return [torch.from_numpy(np.ascontiguousarray(a)).to(d, dt) if a is not None else None for a in arrs]


@pytest.mark.parametrize("name", ["bottom2k", "slit16k"])
def test_forward_matches_oracle_at_full_size(name):
    import torch
    wl = bench.WORKLOADS[name]
    say("workload %s: oracle Mdot" % name)
    pos, F, T = bench.make_inputs(wl)
    ref = oracle_for(wl)
    ref.set_positions(pos)
    U_ref, W_ref = ref.mdot(F, T)
    for prec, dt, tol in (("double", torch.float64, 1e-11), ("mixed", torch.float32, 1e-5)):
        say("workload %s: %s GPU Mdot" % (name, prec))
        s = bench.gpu_solver(wl, prec)
        s.setPositions(pos)
        Fd, Td = dev_tensors(torch, dt, F, T)
        U, W = s.mdot_device(Fd, Td)
        eu = rel_l2(U.double().cpu().numpy(), U_ref)
        ew = rel_l2(W.double().cpu().numpy(), W_ref) if T is not None else 0.0
        say("  %s rel-L2 U %.3e W %.3e" % (prec, eu, ew))
        assert eu <= tol and ew <= tol
        s.clear()


def _support_changes(x, h, thresh, znodes, delta):
    """True where moving a coordinate by +-delta changes the particle's
    support (a grid node enters or leaves [x - thresh, x + thresh], or the
    first support index int(x/h) changes): the discrete map jumps there by
    ~e^{-beta w}, so central differences do not measure the derivative."""
    out = np.zeros(x.shape, dtype=bool)
    for k in range(3):
        lo, hi = x[:, k] - delta, x[:, k] + delta
        if k < 2:
            out[:, k] = (np.floor(lo / h) != np.floor(hi / h)) | \
                        (np.floor((lo - thresh) / h) != np.floor((hi - thresh) / h)) | \
                        (np.floor((lo + thresh) / h) != np.floor((hi + thresh) / h))
        else:
            for e in (-thresh, thresh):
                a, b = lo + e, hi + e
                out[:, k] |= np.any((znodes[None, :] >= a[:, None]) & (znodes[None, :] <= b[:, None]), axis=1)
    return out


def test_differentiation_workload():
    """bottom8k: the hand-written adjoint (positions, forces, torques) against
    the oracle's adjoint at N = 8192 (FP64 <= 1e-11, mixed <= 1e-5), and
    against central differences (h = 1e-4) of the GPU's FP64 Mdot on 64
    particles drawn by RandomState(1).choice.  The FP64 Mdot matches the
    oracle to 1e-15, so these differences are the oracle's to 1e-11.
    Gradients come from the solve taken as self-adjoint (the north star's
    "reusing the self-adjoint solve"), so they match differences of L up to
    the discrete operator's asymmetry: dL/dF differs from the exact
    transpose M^T g by the asymmetry of M_tt (bound 1e-4; the reference's own
    stored bottom-wall asymmetry is <= 3.4e-5), dL/dT by that of the
    force-rotation block M_rt against M_tr^T, which is larger relative to
    that small block (bound 2e-3, measured 5.7e-4), and dL/dx carries both
    (bound 2e-3, measured 8.9e-4).  Where the solve is exactly symmetric
    (triply periodic) the same adjoint matches differences to 1e-10
    (test_gpu_adjoint.py).  Position coordinates whose +-h window changes
    the support are skipped (reported)."""
    import torch
    wl = bench.WORKLOADS["bottom8k"]
    geom, N, Lx, Ly, zmax, zlo, zhi, nx, ny, nz, torque = wl
    pos, F, T = bench.make_inputs(wl)
    g = np.random.RandomState(1).normal(size=3 * N)
    zeros = np.zeros(3 * N)
    say("workload bottom8k: oracle adjoint")
    ref = oracle_for(wl)
    ref.set_positions(pos)
    gF_ref, gT_ref, gX_ref = o.adjoint(ref, F, g, zeros, T)
    for prec, dt, tol in (("double", torch.float64, 1e-11), ("mixed", torch.float32, 1e-5)):
        say("workload bottom8k: %s GPU adjoint" % prec)
        s = bench.gpu_solver(wl, prec)
        s.setPositions(pos)
        Fd, Td, gd, zd = dev_tensors(torch, dt, F, T, g, zeros)
        gF, gT, gX = s.mdot_adjoint_device(Fd, gd, torques=Td, grad_angular=zd)
        c = lambda t: t.double().cpu().numpy()
        ex, ef, et = rel_l2(c(gX), gX_ref), rel_l2(c(gF), gF_ref), rel_l2(c(gT), gT_ref)
        say("  %s rel-L2 dL/dx %.3e dL/dF %.3e dL/dT %.3e" % (prec, ex, ef, et))
        assert ex <= tol and ef <= tol and et <= tol
        if prec == "mixed":
            s.clear()
            continue
        # central differences on 64 particles
        say("workload bottom8k: central differences")
        idx = np.random.RandomState(1).choice(N, 64, replace=False)
        hfd = 1e-4
        gd_t = gd

        def loss(p, f, t):
            s.setPositions(p)
            U, _ = s.mdot_device(*dev_tensors(torch, dt, f, t))
            return float((U * gd_t).sum())

        xs = pos.reshape(N, 3)
        h = Lx / nx
        thresh_d = o.lib().orc_dipole_thresh(3 * h, 1.785 * 6)
        skip = _support_changes(xs[idx], h, 3 * h, ref.zpts, hfd) | \
            _support_changes(xs[idx], h, thresh_d, ref.zpts, hfd)
        fd_x, an_x = [], []
        for a, p_ in enumerate(idx):
            for k in range(3):
                if skip[a, k]:
                    continue
                pp, pm = pos.copy(), pos.copy()
                pp[3 * p_ + k] += hfd
                pm[3 * p_ + k] -= hfd
                fd_x.append((loss(pp, F, T) - loss(pm, F, T)) / (2 * hfd))
                an_x.append(c(gX)[3 * p_ + k])
        s.setPositions(pos)
        fd_f, fd_t = [], []
        for p_ in idx:
            for k in range(3):
                for arr, out in ((F, fd_f), (T, fd_t)):
                    ap, am = arr.copy(), arr.copy()
                    ap[3 * p_ + k] += hfd
                    am[3 * p_ + k] -= hfd
                    if arr is F:
                        out.append((loss(pos, ap, T) - loss(pos, am, T)) / (2 * hfd))
                    else:
                        out.append((loss(pos, F, ap) - loss(pos, F, am)) / (2 * hfd))
        sel = (3 * idx[:, None] + np.arange(3)[None, :]).ravel()
        efd_x = rel_l2(an_x, fd_x)
        efd_f = rel_l2(c(gF)[sel], fd_f)
        efd_t = rel_l2(c(gT)[sel], fd_t)
        say("  FD (h=1e-4, 64 particles, %d of 192 position coordinates skipped): rel-L2 dL/dx %.3e "
            "dL/dF %.3e dL/dT %.3e" % (int(skip.sum()), efd_x, efd_f, efd_t))
        assert len(fd_x) >= 150
        assert efd_f <= 1e-4 and efd_x <= 2e-3 and efd_t <= 2e-3
        s.clear()


def test_batched_workload():
    """All 64 instances: graph replay against an eager loop over the same
    solvers (mixed; spreading accumulates with LDS float atomics, so replay
    and eager agree to rounding, <= 1e-6), and the three smallest instances
    against the oracle (<= 1e-5)."""
    import torch
    from doublyperiodicstokes_amd import DPStokes, DPStokesBatch, StokesParameters
    insts = bench.batch_workload()
    dev = torch.device("cuda", 0)
    solvers, F, T = [], [], []
    say("workload batch64: creating 64 instances (%d particles)" % sum(d["N"] for d in insts))
    for d in insts:
        par = StokesParameters(viscosity=ETA, Lx=d["L"], Ly=d["L"], zmin=d["zmin"], zmax=d["zmax"],
                               w=6, w_d=6, beta=1.785 * 6, beta_d=1.785 * 6, nx=d["n"], ny=d["n"],
                               nz=d["nz"], mode="bottom", precision="mixed")
        s = DPStokes()
        s.initialize(par, d["N"])
        s.setPositions(d["pos"])
        solvers.append(s)
        F.append(torch.from_numpy(d["F"].astype(np.float32)).to(dev))
        T.append(torch.from_numpy(d["T"].astype(np.float32)).to(dev))
    b = DPStokesBatch(solvers, nstreams=4)
    say("workload batch64: tune, replay, eager")
    b.tune(F, T)
    b.mdot_device(F, T)
    U, W = b.mdot_device(F, T)
    U = [u.double().cpu().numpy() for u in U]
    W = [w.double().cpu().numpy() for w in W]
    worst = 0.0
    for i, (s, f, t) in enumerate(zip(solvers, F, T)):
        u, w = s.mdot_device(f, t)
        worst = max(worst, rel_l2(U[i], u.double().cpu().numpy()), rel_l2(W[i], w.double().cpu().numpy()))
    say("  replay vs eager worst rel-L2 %.3e" % worst)
    assert worst <= 1e-6
    for i in np.argsort([d["N"] for d in insts])[:3]:
        d = insts[i]
        say("  instance %d: N=%d grid %d^2 x %d vs oracle" % (i, d["N"], d["n"], d["nz"]))
        ref = o.OracleDPStokes("bottom", d["L"], d["L"], d["zmin"], d["zmax"], d["n"], d["n"], d["nz"], ETA,
                               6, 1.785 * 6, has_torque=True, w_d=6, betaw_d=1.785 * 6)
        ref.set_positions(d["pos"])
        u_ref, w_ref = ref.mdot(d["F"].astype(np.float32).astype(np.float64),
                                d["T"].astype(np.float32).astype(np.float64))
        eu, ew = rel_l2(U[i], u_ref), rel_l2(W[i], w_ref)
        say("    rel-L2 U %.3e W %.3e" % (eu, ew))
        assert eu <= 1e-5 and ew <= 1e-5
    b.clear()
    for s in solvers:
        s.clear()


def test_monolayer_forward_and_adjoint():
    """monolayer256k: forward and position adjoint, mixed against FP64
    (<= 1e-5), finite; the FP64 velocity grid vanishes on the wall node
    (no-slip, <= 1e-12 of its largest value); Mdot is linear in the forces."""
    import torch
    wl = bench.WORKLOADS["monolayer256k"]
    geom, N, Lx, Ly, zmax, zlo, zhi, nx, ny, nz, torque = wl
    pos, F, T = bench.make_inputs(wl)
    g = np.random.RandomState(1).normal(size=3 * N)
    gw = np.random.RandomState(2).normal(size=3 * N)
    res = {}
    for prec, dt in (("double", torch.float64), ("mixed", torch.float32)):
        say("workload monolayer256k: %s forward + adjoint" % prec)
        s = bench.gpu_solver(wl, prec)
        s.setPositions(pos)
        Fd, Td, gd, gwd = dev_tensors(torch, dt, F, T, g, gw)
        U, W = s.mdot_device(Fd, Td)
        gF, gT, gX = s.mdot_adjoint_device(Fd, gd, torques=Td, grad_angular=gwd)
        out = [t.double().cpu().numpy() for t in (U, W, gF, gT, gX)]
        assert all(np.isfinite(a).all() for a in out)
        res[prec] = out
        if prec == "double":
            U2, W2 = s.mdot_device(2.5 * Fd, 2.5 * Td)
            assert rel_l2(U2.cpu().numpy(), 2.5 * out[0]) <= 1e-13
            G = s.stage_solve(s.stage_spread(Fd, Td))
            wall = G[:, :, nz - 1, :].abs().max().item()
            top = G.abs().max().item()
            say("  no-slip: max |u| on the wall node %.3e of %.3e" % (wall, top))
            assert wall <= 1e-12 * top
            del G
        s.clear()
    names = ("U", "W", "dL/dF", "dL/dT", "dL/dx")
    errs = [rel_l2(m, d) for m, d in zip(res["mixed"], res["double"])]
    say("  mixed vs FP64 rel-L2: " + " ".join("%s %.3e" % (n, e) for n, e in zip(names, errs)))
    assert max(errs) <= 1e-5
