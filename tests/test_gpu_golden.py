"""The HIP path, through the joint interface FCMJoint('hip'), against the
reference's own stored mobilities (tests/golden, see tests/golden/make_golden.py).

Fixtures: 135 x 135 grids (kernType 0, torques), the doubled 270 x 270 grids
(ref=True, 12-point supports, Nz 121 / 109), and translation-only runs on
90 x 90 (w = 4) and 120 x 120 (w = 6) grids.  Every grid but the 135 one has
Nyquist modes, handled here by r2c/c2r transforms whose Nyquist rows equal the
reference's c2c real part.

FP64 path: every normalised column at every stored configuration, <= 1e-10
of the column's max-norm.  Mixed path (FP32 grids): <= 1e-5 on a subset.
"""
import numpy as np
import pytest

import _mobility as mob

pytestmark = pytest.mark.gpu

try:
    import torch  # noqa: F401  (first import can be slow on a fresh box)
except Exception:
    torch = None

FIXTURES = [("mobility_bw", "self"), ("mobility_sc", "self"), ("mobility_bw", "pair"),
            ("mobility_sc", "pair"), ("mobility_bw_ref2x", "self"), ("mobility_sc_ref", "self"),
            ("mobility_sc_ref", "pair"), ("mobility_bw_w4", "self"), ("mobility_sc_w4", "self"),
            ("mobility_bw_w4", "pair"), ("mobility_sc_w4", "pair")]


def joint(name, kind, precision):
    from doublyperiodicstokes_amd import FCMJoint
    c = mob.config(name, kind)
    s = FCMJoint("hip", precision=precision)
    s.Initialize(numberParticles=c["nP"], hydrodynamicRadius=c["radius"], viscosity=c["eta"],
                 domType=c["domType"], has_torque=c["has_torque"], xmax=c["L"], ymax=c["L"],
                 zmin=c["zmin"], zmax=c["zmax"], kernType=c["kernType"], optInd=0, ref=c["ref"])
    assert (s.par.nx, s.par.ny, s.par.nz) == (c["nx"], c["ny"], c["nz"])
    return s, c


def cases(name, kind, stride=1):
    g = mob.load(name)
    if kind == "self":
        return [(mob.self_position(name, i), g["self_M"][i])
                for i in range(0, len(g["self_heights"]), stride)]
    # row 0 of the pair heights is unused by pair_mobility.py:56
    return [(mob.pair_position(name, iS, iH), g["pair_M"][iS, iH])
            for iS in range(3) for iH in range(1, len(g["pair_heights"]), stride)]


def check(name, kind, precision, stride, tol):
    s, c = joint(name, kind, precision)
    nP, ht = c["nP"], c["has_torque"]
    worst = 0.0
    for pos, ref in cases(name, kind, stride):
        s.SetPositions(pos)
        cols = mob.columns(s.Mdot, nP, range(ref.shape[1]), c["eta"], ht)
        for j, col in cols.items():
            err = np.abs(col - ref[:, j]).max() / np.abs(ref[:, j]).max()
            worst = max(worst, err)
    s.Clean()
    print("%s %s %s grid %dx%dx%d worst column error %.3e" %
          (name, kind, precision, c["nx"], c["ny"], c["nz"], worst), flush=True)
    assert worst <= tol


@pytest.mark.parametrize("name,kind", FIXTURES)
def test_mobility_fp64(name, kind):
    check(name, kind, "double", 1, 1e-10)


@pytest.mark.parametrize("name,kind", FIXTURES)
def test_mobility_mixed(name, kind):
    check(name, kind, "mixed", 6, 1e-5)
