"""The FP64 oracle against the reference's own stored mobilities.

tests/golden/mobility_*.npz are the reference's figure data
(figures/Figures/data/{self,pair}_mobility_*.mat, produced by
figures/self_mobility.py and figures/pair_mobility.py through
FCMJoint('cpu'); see tests/golden/make_golden.py).  The oracle, configured as
FCMJoint.Initialize configures each run, must reproduce them to rounding:
every stage of the restatement (kernels, supports, images, transforms, BVP,
corrections, k = 0, curls) enters these numbers.

Grids: 135 x 135 (kernType 0 with torques; odd, no Nyquist modes);
270 x 270 (ref=True, the doubled grid with the coarse grid's kernels: 12-point
supports); 90 x 90 (kernType 2, w = 4, translation only) and 120 x 120 (w = 6,
translation only).  The even grids carry Nyquist modes, so they pin the
reference's Nyquist conventions (Dx, Dy zeroed there, Solvers.py:298-304; the
real part of the backward c2c transform kept, FCM.py:406-407).  The self runs
put the particle at x = y = L/2, a grid node on the 90 and 270 grids, so the
offsets at the support edge sit on the kernel's square-root singularity and
pin the rounding of the offset (a fused multiply-add in the reference build).

Tolerance: max-norm error relative to the max-norm of the stored column (or of
a random combination of columns, one Mdot), <= 1e-11 (observed 5e-15..2e-14).
"""
import numpy as np
import pytest

import _mobility as mob

TOL = 1e-11


def _combo(name, kind, o, pos, Mref, seed):
    c = mob.config(name, kind)
    o.set_positions(pos)
    v = np.random.RandomState(seed).normal(size=Mref.shape[1])
    return mob.combination_error(o.mdot, c["nP"], c["eta"], c["has_torque"], Mref, v)


@pytest.mark.parametrize("tag", ["bw", "sc"])
def test_self_mobility_matches_reference_data(tag):
    """kernType 0 with torques, 135 x 135 grid: one particle below a radius
    from the wall, one in the bulk."""
    name = "mobility_" + tag
    g = mob.load(name)
    o = mob.oracle(name, "self")
    for iH in (2, 33):
        err = _combo(name, "self", o, mob.self_position(name, iH), g["self_M"][iH], iH)
        assert err <= TOL, (tag, iH, err)


@pytest.mark.parametrize("tag", ["bw", "sc"])
def test_pair_mobility_matches_reference_data(tag):
    name = "mobility_" + tag
    g = mob.load(name)
    o = mob.oracle(name, "pair")
    # separation 4, height 0.78 (below one radius), one column at a time
    iS, iH = 1, 7
    o.set_positions(mob.pair_position(name, iS, iH))
    got = mob.columns(o.mdot, 2, (0, 11), mob.config(name, "pair")["eta"])
    for j, col in got.items():
        ref = g["pair_M"][iS, iH][:, j]
        err = np.abs(col - ref).max() / np.abs(ref).max()
        assert err <= TOL, (tag, j, err)


@pytest.mark.parametrize("tag", ["bw", "sc"])
def test_translation_only_w4_self_mobility(tag):
    """kernType 2 (w = 4), 90 x 90 grid, particle on a grid node."""
    name = "mobility_%s_w4" % tag
    g = mob.load(name)
    assert mob.config(name, "self")["nx"] == 90 and mob.config(name, "self")["w"] == 4
    o = mob.oracle(name, "self")
    for iH in (2, 10, 40):
        err = _combo(name, "self", o, mob.self_position(name, iH), g["self_M"][iH], iH)
        assert err <= TOL, (tag, iH, err)


@pytest.mark.parametrize("tag", ["bw", "sc"])
def test_translation_only_w6_pair_mobility(tag):
    """kernType 0 without torques (pair_mobility.py:39), 120 x 120 grid."""
    name = "mobility_%s_w4" % tag
    g = mob.load(name)
    assert mob.config(name, "pair")["nx"] == 120 and mob.config(name, "pair")["w"] == 6
    o = mob.oracle(name, "pair")
    for iS, iH in ((0, 3), (2, 20)):
        err = _combo(name, "pair", o, mob.pair_position(name, iS, iH), g["pair_M"][iS, iH], iS + iH)
        assert err <= TOL, (tag, iS, iH, err)


@pytest.mark.parametrize("name,kind", [("mobility_bw_ref2x", "self"), ("mobility_sc_ref", "pair")])
def test_doubled_grid_mobility(name, kind):
    """ref=True: 270 x 270 grid, Nz 121 (bottom wall) / 109 (slit), with the
    kernels of the 135 x 135 grid (12-point supports).  One Mdot each."""
    g = mob.load(name)
    c = mob.config(name, kind)
    assert (c["nx"], c["ny"]) == (270, 270) and c["nz"] in (121, 109)
    o = mob.oracle(name, kind)
    if kind == "self":
        err = _combo(name, kind, o, mob.self_position(name, 2), g["self_M"][2], 2)
    else:
        err = _combo(name, kind, o, mob.pair_position(name, 1, 7), g["pair_M"][1, 7], 17)
    assert err <= TOL, (name, err)
