"""The batched form (dps_batch_*): independent instances of different boxes,
grids, particle counts and geometries captured into one HIP graph.  Every
replay -- repeated, with eager work and synchronisations between, after a
position update, after autotuning -- matches the instances' own eager Mdot
(FP64 rel-L2 <= 1e-12; mixed <= 1e-6), and one instance matches the oracle
(<= 1e-11)."""
import ctypes
import json

import numpy as np
import pytest

from oracle.dp_oracle import OracleDPStokes
from test_gpu_parity import rel_l2, say

pytestmark = pytest.mark.gpu

ETA = 1 / (6 * np.pi)
# geometry, L, Lz, n, nz, N, torque
INSTANCES = [
    ("bottom", 8.0, 6.3, 16, 17, 20, False),
    ("bottom", 12.0, 9.0, 24, 25, 60, False),
    ("slit", 10.0, 8.0, 20, 21, 40, True),
    ("nowall", 16.0, 10.0, 32, 33, 80, False),
    ("bottom", 16.0, 12.0, 32, 41, 120, False),
    ("periodic", 8.0, 8.0, 16, 16, 30, False),
]


def _solver(inst, precision):
    from doublyperiodicstokes_amd import DPStokes, StokesParameters
    geom, L, Lz, n, nz, N, torque = inst
    par = StokesParameters(viscosity=ETA, Lx=L, Ly=L, zmin=0.0, zmax=Lz, w=6, w_d=6 if torque else 0,
                           beta=1.785 * 6, beta_d=1.785 * 6, nx=n, ny=n, nz=nz, mode=geom, precision=precision)
    s = DPStokes()
    s.initialize(par, N)
    return s


def _positions(inst, rng):
    geom, L, Lz, n, nz, N, torque = inst
    zlo = 1.6 if geom == "nowall" else 0.0
    zhi = Lz - 1.6 if geom in ("nowall", "bottom") else Lz
    return np.stack([rng.uniform(0, L, N), rng.uniform(0, L, N), rng.uniform(zlo, zhi, N)], 1).ravel()


@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_batch_replays_match_eager(precision):
    import torch
    from doublyperiodicstokes_amd import DPStokesBatch
    say("batch of %d instances, %s" % (len(INSTANCES), precision))
    dev = torch.device("cuda", 0)
    dt = torch.float64 if precision == "double" else torch.float32
    tol = 1e-12 if precision == "double" else 1e-6
    rng = np.random.RandomState(8)
    solvers = [_solver(i, precision) for i in INSTANCES]
    pos = [_positions(i, rng) for i in INSTANCES]
    F = [torch.tensor(rng.normal(size=3 * i[5]), dtype=dt, device=dev) for i in INSTANCES]
    T = [torch.tensor(rng.normal(size=3 * i[5]), dtype=dt, device=dev) if i[6] else None for i in INSTANCES]
    for s, p in zip(solvers, pos):
        s.setPositions(p)

    def eager():
        out = [s.mdot_device(f, t) for s, f, t in zip(solvers, F, T)]
        torch.cuda.synchronize()
        return [(u.double().cpu().numpy(), None if w is None else w.double().cpu().numpy()) for u, w in out]

    def check(U, W, ref, what):
        for k, ((u_ref, w_ref), u, w) in enumerate(zip(ref, U, W)):
            e = rel_l2(u.double().cpu().numpy(), u_ref)
            ew = 0.0 if w_ref is None else rel_l2(w.double().cpu().numpy(), w_ref)
            assert e <= tol and ew <= tol, (what, k, e, ew)

    ref = eager()
    b = DPStokesBatch(solvers, nstreams=3)
    for rep in range(3):
        U, W = b.mdot_device(F, T)
        torch.cuda.synchronize()
        check(U, W, ref, "replay %d" % rep)
        junk = torch.randn(1 << 16, device=dev).sum()   # eager work between replays
        torch.cuda.synchronize()
        del junk
    # a position update of one instance is seen by the next replay
    pos[0] = _positions(INSTANCES[0], rng)
    solvers[0].setPositions(pos[0])
    ref = eager()
    U, W = b.mdot_device(F, T)
    torch.cuda.synchronize()
    check(U, W, ref, "after position update")
    # autotuned spreading chunks, then a fresh capture
    chunks = b.tune(F, T)
    say("  tuned z chunks %s" % (chunks,))
    assert all(c in (4, 8, 12, 16, 24, 32, 48) for c in chunks)
    ref = eager()
    U, W = b.mdot_device(F, T)
    torch.cuda.synchronize()
    check(U, W, ref, "after tuning")
    if precision == "double":
        geom, L, Lz, n, nz, N, torque = INSTANCES[4]
        o = OracleDPStokes(geom, L, L, 0.0, Lz, n, n, nz, ETA, 6, 1.785 * 6)
        o.set_positions(pos[4])
        v, _ = o.mdot(F[4].cpu().numpy())
        e = rel_l2(U[4].cpu().numpy(), v)
        say("  instance 4 vs oracle rel-L2 %.3e" % e)
        assert e <= 1e-11
    b.clear()
    for s in solvers:
        s.clear()


def test_batch_tune_cache(tmp_path):
    """A second tune with the same cache file applies the stored chunks
    (dps_batch_set_chunks, no timing); replays then still match eager.  A
    chunk that cannot fit is refused and changes nothing."""
    import torch
    from doublyperiodicstokes_amd import DPStokesBatch
    from doublyperiodicstokes_amd import _native as N
    say("batch tuning cache")
    dev = torch.device("cuda", 0)
    rng = np.random.RandomState(9)
    insts = INSTANCES[:3]
    solvers = [_solver(i, "mixed") for i in insts]
    for s, i in zip(solvers, insts):
        s.setPositions(_positions(i, rng))
    F = [torch.tensor(rng.normal(size=3 * i[5]), dtype=torch.float32, device=dev) for i in insts]
    T = [torch.tensor(rng.normal(size=3 * i[5]), dtype=torch.float32, device=dev) if i[6] else None
         for i in insts]
    path = str(tmp_path / "tune.json")
    b = DPStokesBatch(solvers, nstreams=2)
    first = b.tune(F, T, cache=path)
    with open(path) as f:
        stored = json.load(f)["chunks"]
    assert sorted(stored.values()) == sorted(first)
    second = b.tune(F, T, cache=path)
    say("  tuned %s, from cache %s" % (first, second))
    assert second == first
    ref = [s.mdot_device(f, t) for s, f, t in zip(solvers, F, T)]
    ref = [(u.double().cpu().numpy(), None if w is None else w.double().cpu().numpy()) for u, w in ref]
    U, W = b.mdot_device(F, T)
    torch.cuda.synchronize()
    for k, ((u_ref, w_ref), u, w) in enumerate(zip(ref, U, W)):
        assert rel_l2(u.double().cpu().numpy(), u_ref) <= 1e-6, k
        if w_ref is not None:
            assert rel_l2(w.double().cpu().numpy(), w_ref) <= 1e-6, k
    bad = (ctypes.c_int * 3)(first[0], 4096, first[2])
    with pytest.raises(ValueError):
        N.check(N.load().dps_batch_set_chunks(b._b, bad, 1, None))
    b.clear()
    for s in solvers:
        s.clear()
