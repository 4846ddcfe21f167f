"""Mobility-matrix invariants of the HIP path against the reference's own
stored distributions (tests/golden/mobility_asym_posdef.npz, written by
figures/pair_mobility_matrix_asym.py through FCMJoint('cpu')).

The reference draws 99 unseeded pair configurations (x, y ~ U[0, L),
z ~ U[0, zmax)), builds the normalised mobility matrix column by column
(figures/mobility_matrix.py) and stores ||M - M^T|| / ||M|| and the smallest
eigenvalue of M.  Here the same distribution is drawn with a fixed seed and
M is built through FCMJoint('hip') in FP64.  Checks:
  - every M is positive definite (smallest real eigenvalue > 0);
  - the largest asymmetry is at most twice the reference's largest;
  - the asymmetry and eigenvalue samples are consistent with the
    reference's samples (two-sample Kolmogorov-Smirnov, p > 1e-3): the
    discrete operator, not only its size, is the reference's.
"""
import numpy as np
import pytest

import _mobility as mob

pytestmark = pytest.mark.gpu

try:
    import torch  # noqa: F401
except Exception:
    torch = None


@pytest.mark.parametrize("tag,width", [("bw", "w6"), ("sc", "w6"), ("bw", "w4"), ("sc", "w4")])
def test_asymmetry_and_positive_definiteness(tag, width):
    from scipy.stats import ks_2samp
    from doublyperiodicstokes_amd import FCMJoint
    ref = mob.load("mobility_asym_posdef")
    L, zmax, eta = float(ref["L"]), float(ref["zmax"]), float(ref["eta"])
    has_torque = width == "w6"
    s = FCMJoint("hip", precision="double")
    s.Initialize(numberParticles=2, hydrodynamicRadius=1.0, viscosity=eta,
                 domType={"bw": "DPBW", "sc": "DPSC"}[tag], has_torque=has_torque, xmax=L,
                 ymax=L, zmin=0.0, zmax=zmax, kernType=0, optInd=0, ref=False)
    rng = np.random.RandomState(7)
    asym, mineig = [], []
    for _ in range(99):
        pos = np.array([rng.random_sample() * L, rng.random_sample() * L, rng.random_sample() * zmax,
                        rng.random_sample() * L, rng.random_sample() * L, rng.random_sample() * zmax])
        s.SetPositions(pos)
        cols = mob.columns(s.Mdot, 2, range(12 if has_torque else 6), eta, has_torque)
        M = np.stack([cols[j] for j in sorted(cols)], axis=1)
        asym.append(np.linalg.norm(M - M.T) / np.linalg.norm(M))
        mineig.append(np.sort(np.linalg.eigvals(M).real)[0])
    s.Clean()
    asym, mineig = np.array(asym), np.array(mineig)
    ra, rp = ref["asym_%s_%s" % (tag, width)], ref["posdef_%s_%s" % (tag, width)]
    pa, pp = ks_2samp(asym, ra).pvalue, ks_2samp(mineig, rp).pvalue
    print("%s %s asym median %.3e max %.3e (ref %.3e %.3e, KS p %.3f); min eig %.3e "
          "(ref %.3e, KS p %.3f)" % (tag, width, np.median(asym), asym.max(), np.median(ra),
                                     ra.max(), pa, mineig.min(), rp.min(), pp), flush=True)
    assert mineig.min() > 0
    assert asym.max() <= 2 * ra.max()
    assert pa > 1e-3 and pp > 1e-3
