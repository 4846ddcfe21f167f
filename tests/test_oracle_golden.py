"""The FP64 oracle against the reference's own stored mobilities.

tests/golden/mobility_{bw,sc}.npz are the reference's figure data
(figures/Figures/data/{self,pair}_mobility_{bw,sc}.mat, produced by
figures/self_mobility.py and figures/pair_mobility.py through
FCMJoint('cpu')).  The oracle, configured as FCMJoint.Initialize configures
those runs, must reproduce them to rounding: every stage of the restatement
(kernels, supports, images, transforms, BVP, corrections, k = 0, curls)
enters these numbers.  Tolerance: max-norm error of each normalised column
relative to the column's max-norm, <= 1e-11 (observed ~2e-14).
"""
import numpy as np
import pytest

import _mobility as mob

TOL = 1e-11


@pytest.mark.parametrize("tag", ["bw", "sc"])
def test_self_mobility_matches_reference_data(tag):
    g = mob.load(tag)
    o = mob.oracle(tag, 1)
    eta = mob.config(tag, 1)["eta"]
    # one particle below a radius from the wall, one in the bulk
    for iH, cols in ((2, (0, 2, 3, 5)), (33, (0, 4))):
        o.set_positions(mob.self_position(tag, iH))
        got = mob.columns(o.mdot, 1, cols, eta)
        for j, col in got.items():
            ref = g["self_M"][iH][:, j]
            err = np.abs(col - ref).max() / np.abs(ref).max()
            assert err <= TOL, (tag, iH, j, err)


@pytest.mark.parametrize("tag", ["bw", "sc"])
def test_pair_mobility_matches_reference_data(tag):
    g = mob.load(tag)
    o = mob.oracle(tag, 2)
    eta = mob.config(tag, 2)["eta"]
    iS, iH = 1, 7     # separation 4, height 0.78 (below one radius)
    o.set_positions(mob.pair_position(tag, iS, iH))
    got = mob.columns(o.mdot, 2, (0, 5, 6, 11), eta)
    for j, col in got.items():
        ref = g["pair_M"][iS, iH][:, j]
        err = np.abs(col - ref).max() / np.abs(ref).max()
        assert err <= TOL, (tag, j, err)
