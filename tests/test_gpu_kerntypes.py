"""FCMJoint('hip') for every kernel type of the auto-configuration
(GridAndKernelConfig.py:48-58, 184-208) against the FP64 oracle configured
the same way.

Types with torques: 1 (ES 5/5), 3 (6/5), 4 (6/4), 5 (5/4); without: 1 (w 5),
2 (w 4).  Odd widths take the half-cell rule of the support start
(ParticleList.cpp:493-497), and unequal monopole/dipole widths take the
unpaired spreading and interpolation entries.  The auto-configuration itself
is pinned against the reference's function in test_gridconfig_golden.py.
FP64 <= 1e-11, mixed <= 1e-5 (rel-L2 of U and W).
"""
import numpy as np
import pytest

from doublyperiodicstokes_amd.gridconfig import configure_grid_and_kernels_xy, configure_grid_and_kernels_z
from oracle import dp_oracle as o
from test_gpu_parity import rel_l2, say

pytestmark = pytest.mark.gpu

ETA = 1 / (6 * np.pi)
CASES = [(kt, True, dom) for kt in (1, 3, 4, 5) for dom in ("DPBW", "DPSC")] + \
        [(kt, False, dom) for kt in (1, 2) for dom in ("DPBW", "DPSC", "DP")]


def _oracle(kt, ht, dom, L, zmin, zmax):
    Lx, Ly, hx, hy, nx, ny, w, wd, cb, cbd, b, bd = configure_grid_and_kernels_xy(
        L, L, np.ones(1), np.array([kt]), 0, ht, False)
    Lz, hz, nz, a = configure_grid_and_kernels_z(zmin, zmax, hx, w, wd, dom, 1.5, False)
    kw = dict(alpha=1 / (2 * cb[0]))
    if ht:
        kw.update(w_d=int(wd[0]), betaw_d=float(bd[0] * wd[0]), alpha_d=1 / (2 * cbd[0]), has_torque=True)
    geom = {"DP": "nowall", "DPBW": "bottom", "DPSC": "slit"}[dom]
    return o.OracleDPStokes(geom, Lx, Ly, a, a + Lz, nx, ny, nz, ETA, int(w[0]), float(b[0] * w[0]), **kw), \
        (nx, ny, nz, int(w[0]), int(wd[0]))


@pytest.mark.parametrize("kt,ht,dom", CASES, ids=["kt%d_%s_%s" % (k, "FT" if t else "F", d) for k, t, d in CASES])
def test_kernel_type_matches_oracle(kt, ht, dom):
    from doublyperiodicstokes_amd import FCMJoint
    L, zmin, zmax, N = 24.0, 0.0, 8.0, 12
    rng = np.random.RandomState(kt + 10 * ht)
    zlo, zhi = {"DP": (3.0, 5.0), "DPBW": (0.3, 7.5), "DPSC": (0.3, 7.7)}[dom]
    pos = np.stack([rng.uniform(0, L, N), rng.uniform(0, L, N), rng.uniform(zlo, zhi, N)], 1).ravel()
    F = rng.normal(size=3 * N)
    T = rng.normal(size=3 * N) if ht else None
    ref, shape = _oracle(kt, ht, dom, L, zmin, zmax)
    say("kernType %d torque=%s %s grid %dx%dx%d w=%d w_d=%d" % ((kt, ht, dom) + shape))
    ref.set_positions(pos)
    U_ref, W_ref = ref.mdot(F, T)
    for prec, tol in (("double", 1e-11), ("mixed", 1e-5)):
        s = FCMJoint("hip", precision=prec)
        s.Initialize(numberParticles=N, hydrodynamicRadius=1.0, viscosity=ETA, domType=dom,
                     has_torque=ht, xmax=L, ymax=L, zmin=zmin, zmax=zmax, kernType=kt)
        s.SetPositions(pos)
        U, W = s.Mdot(F, T) if ht else s.Mdot(F)
        eu = rel_l2(U, U_ref)
        ew = rel_l2(W, W_ref) if ht else 0.0
        say("  %s rel-L2 U %.3e W %.3e" % (prec, eu, ew))
        assert eu <= tol and ew <= tol
        s.Clean()
