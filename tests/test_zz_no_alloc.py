"""dps_mdot_adjoint_noalloc against dps_mdot_adjoint, through the C ABI.

Two calls in a row on one GPU share its scratch array: a bottom-wall solver,
then a smaller slit solver (fewer particles, a smaller grid), each in FP64
(<= 1e-12 rel-L2: the same kernels, differing only in the order of the
spreading's float atomics) and mixed precision (<= 1e-5), for position, force
and torque gradients.  The no-allocation call leaves the device's free memory
unchanged after a first forward Mdot has loaded the code objects and FFT
plans (the allocating entry point's first position-gradient call does not
leave it unchanged), and a solver whose adjoint exceeds the scratch is refused with
DPS_ERR_ARG.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ETA = 1 / (6 * np.pi)


def _solver(geom, L, Lz, n, nz, N, precision):
    from doublyperiodicstokes_amd import DPStokes, StokesParameters
    par = StokesParameters(viscosity=ETA, Lx=L, Ly=L, zmin=0.0, zmax=Lz, w=6, w_d=6, beta=1.785 * 6,
                           beta_d=1.785 * 6, nx=n, ny=n, nz=nz, mode=geom, precision=precision)
    s = DPStokes()
    s.initialize(par, N)
    return s


def _inputs(torch, N, L, zlo, zhi, dt, seed):
    rng = np.random.RandomState(seed)
    pos = np.stack([rng.uniform(0, L, N), rng.uniform(0, L, N), rng.uniform(zlo, zhi, N)], 1).ravel()
    d = torch.device("cuda", 0)
    t = lambda a: torch.from_numpy(a).to(d, dt)
    return pos, [t(rng.normal(size=3 * N)) for _ in range(4)]


def _call(fn, s, F, T, gU, gW, outs, dtype_code):
    from doublyperiodicstokes_amd import _native as Nt
    vp = lambda t: ctypes.c_void_p(t.data_ptr())
    gF, gT, gX = outs
    rc = fn(s._h, vp(F), vp(T), vp(gU), vp(gW), vp(gF), vp(gT), vp(gX), dtype_code, None)
    Nt.check(rc)


def _rel(a, b):
    a, b = a.double().cpu().numpy(), b.double().cpu().numpy()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_noalloc_adjoint_matches_allocating_adjoint(precision):
    import torch
    from doublyperiodicstokes_amd import _native as Nt
    L = Nt.load()
    dt = torch.float64 if precision == "double" else torch.float32
    code = Nt.DPS_F64 if precision == "double" else Nt.DPS_F32
    tol = 1e-12 if precision == "double" else 1e-5
    # first a larger bottom-wall solver, then a smaller slit solver
    for geom, Lb, Lz, n, nz, N, zlo, zhi in (("bottom", 32.0, 12.0, 64, 48, 96, 0.5, 9.0),
                                              ("slit", 16.0, 10.0, 32, 25, 24, 0.3, 9.7)):
        print("no-alloc adjoint: %s %s N=%d grid %dx%dx%d" % (precision, geom, N, n, n, nz), flush=True)
        s = _solver(geom, Lb, Lz, n, nz, N, precision)
        pos, (F, T, gU, gW) = _inputs(torch, N, Lb, zlo, zhi, dt, n + nz)
        s.setPositions(pos)
        outs_na = [torch.empty(3 * N, dtype=dt, device="cuda") for _ in range(3)]
        outs = [torch.empty(3 * N, dtype=dt, device="cuda") for _ in range(3)]
        # one forward Mdot first: the first use of each code object and of the
        # FFT plans loads them, which is set-up, not this call's doing
        torch.cuda.synchronize()
        freew = torch.cuda.mem_get_info()[0]
        s.mdot_device(F, T)
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        print("  free memory change of the first forward Mdot: %d B" % (free0 - freew), flush=True)
        _call(L.dps_mdot_adjoint_noalloc, s, F, T, gU, gW, outs_na, code)
        torch.cuda.synchronize()
        free1 = torch.cuda.mem_get_info()[0]
        _call(L.dps_mdot_adjoint, s, F, T, gU, gW, outs, code)
        torch.cuda.synchronize()
        free2 = torch.cuda.mem_get_info()[0]
        errs = [_rel(a, b) for a, b in zip(outs_na, outs)]
        print("  rel-L2 dL/dF %.3e dL/dT %.3e dL/dx %.3e; free memory change: no-alloc %d B, "
              "allocating %d B" % (errs[0], errs[1], errs[2], free1 - free0, free2 - free1), flush=True)
        assert max(errs) <= tol
        assert free1 == free0
        if geom == "bottom":
            assert free2 < free1   # the allocating entry point's first call allocates
        # a second no-allocation call on the same solver reproduces the first
        again = [torch.empty(3 * N, dtype=dt, device="cuda") for _ in range(3)]
        _call(L.dps_mdot_adjoint_noalloc, s, F, T, gU, gW, again, code)
        torch.cuda.synchronize()
        assert max(_rel(a, b) for a, b in zip(again, outs)) <= tol
        s.clear()


def test_noalloc_adjoint_refuses_solvers_beyond_its_scratch():
    import torch
    from doublyperiodicstokes_amd import _native as Nt
    L = Nt.load()
    N = 16
    # second velocity grid 3 * 384^2 * 96 doubles = 340 MB > 256 MiB
    s = _solver("bottom", 192.0, 12.0, 384, 96, N, "double")
    pos, (F, T, gU, gW) = _inputs(torch, N, 192.0, 0.5, 9.0, torch.float64, 5)
    s.setPositions(pos)
    outs = [torch.empty(3 * N, dtype=torch.float64, device="cuda") for _ in range(3)]
    vp = lambda t: ctypes.c_void_p(t.data_ptr())
    rc = L.dps_mdot_adjoint_noalloc(s._h, vp(F), vp(T), vp(gU), vp(gW), vp(outs[0]), vp(outs[1]),
                                    vp(outs[2]), Nt.DPS_F64, None)
    msg = L.dps_last_error().decode()
    print("refusal: rc %d, %s" % (rc, msg), flush=True)
    assert rc == Nt.DPS_ERR_ARG and "scratch" in msg
    s.clear()
