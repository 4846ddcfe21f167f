"""The HIP path, through the joint interface FCMJoint('hip'), against the
reference's own stored mobilities (tests/golden, see tests/golden/make_golden.py).

FP64 path: every normalised column at all stored heights, <= 1e-10 of the
column's max-norm.  Mixed path (FP32 grids): <= 1e-5.
"""
import numpy as np
import pytest

import _mobility as mob

pytestmark = pytest.mark.gpu

try:
    import torch  # noqa: F401  (first import can be slow on a fresh box)
except Exception:
    torch = None


def joint(tag, nP, precision):
    from doublyperiodicstokes_amd import FCMJoint
    c = mob.config(tag, nP)
    s = FCMJoint("hip", precision=precision)
    s.Initialize(numberParticles=nP, hydrodynamicRadius=c["radius"], viscosity=c["eta"],
                 domType=c["domType"], has_torque=True, xmax=c["L"], ymax=c["L"], zmin=c["zmin"],
                 zmax=c["zmax"], kernType=0, optInd=0, ref=False)
    return s, c


def check(tag, nP, precision, cases, tol):
    g = mob.load(tag)
    s, c = joint(tag, nP, precision)
    worst = 0.0
    for pos, ref in cases(g):
        s.SetPositions(pos)
        cols = mob.columns(s.Mdot, nP, range(6 * nP), c["eta"])
        for j, col in cols.items():
            err = np.abs(col - ref[:, j]).max() / np.abs(ref[:, j]).max()
            worst = max(worst, err)
    s.Clean()
    print("%s %s nP=%d worst column error %.3e" % (tag, precision, nP, worst), flush=True)
    assert worst <= tol


@pytest.mark.parametrize("tag", ["bw", "sc"])
def test_self_mobility_fp64(tag):
    check(tag, 1, "double",
          lambda g: [(mob.self_position(tag, i), g["self_M"][i]) for i in range(len(g["self_heights"]))],
          1e-10)


@pytest.mark.parametrize("tag", ["bw", "sc"])
def test_pair_mobility_fp64(tag):
    check(tag, 2, "double",
          lambda g: [(mob.pair_position(tag, iS, iH), g["pair_M"][iS, iH])
                     for iS in range(3) for iH in (1, 7, 12, 25)], 1e-10)


@pytest.mark.parametrize("tag", ["bw", "sc"])
def test_self_mobility_mixed(tag):
    check(tag, 1, "mixed",
          lambda g: [(mob.self_position(tag, i), g["self_M"][i]) for i in range(0, 60, 6)], 1e-5)
