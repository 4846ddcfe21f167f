"""The auto-configuration restatement (doublyperiodicstokes_amd/gridconfig.py)
against vectors of the reference's own GridAndKernelConfig.py
(tests/golden/gridconfig.npz, written by tests/golden/make_gridconfig.py).

2304 cases: kernel types 0-5 (translation-rotation 0, 1, 3, 4, 5 and
translation-only 0, 1, 2), three radii, three boxes, optInd 0, 1, 3 and -1,
ref False/True, the four domain types.  Integers (grid sizes, kernel widths)
must be equal, floats equal to 1e-15 relative, and the cases the reference
rejects must be rejected.
"""
import numpy as np
import pytest

from doublyperiodicstokes_amd.gridconfig import (configure_grid_and_kernels_xy,
                                                 configure_grid_and_kernels_z)
from _mobility import HERE

import os

G = dict(np.load(os.path.join(HERE, "golden", "gridconfig.npz"), allow_pickle=False))


def _run(row):
    Lx, Ly, rad, kt, ht, opt, ref, dom, zmin, zmax = row
    dom = str(G["dom_codes"][int(dom)])
    r = configure_grid_and_kernels_xy(Lx, Ly, np.array([rad]), np.array([int(kt)]), int(opt), bool(ht),
                                      bool(ref))
    Lx2, Ly2, hx, hy, Nx, Ny, wm, wd, cbm, cbd, bm, bd = r
    Lz, hz, Nz, a = configure_grid_and_kernels_z(zmin, zmax, hx, wm, wd, dom, 1.5, bool(ref))
    first = lambda v: float(np.ravel(v)[0]) if np.size(v) else 0.0
    f = [Lx2, Ly2, hx, hy, first(cbm), first(cbd), first(bm), first(bd), Lz, hz, a, first(wm)]
    i = [int(Nx), int(Ny), int(first(wm)), int(first(wd)), int(Nz)]
    return np.array(f), np.array(i)


@pytest.mark.parametrize("kt", [0, 1, 2, 3, 4, 5])
def test_matches_reference_vectors(kt):
    rows = np.nonzero(G["inputs"][:, 3] == kt)[0]
    assert rows.size > 0
    n_ok = 0
    for r in rows:
        if G["ok"][r] != 1:
            with pytest.raises((ValueError, IndexError)):
                _run(G["inputs"][r])
            continue
        f, i = _run(G["inputs"][r])
        assert (i == G["out_i"][r]).all(), (G["inputs"][r], i, G["out_i"][r])
        want = G["out_f"][r]
        assert np.allclose(f, want, rtol=1e-15, atol=0), (G["inputs"][r], f, want)
        n_ok += 1
    assert n_ok > 0
