"""Write the golden mobility fixtures from the reference's stored figure data.

Reads (MATLAB v5 data files, loaded with scipy.io.loadmat -- data only) from
figures/Figures/data/, all written through FCMJoint('cpu') with L = 76.8,
zmin = 0, zmax = 19.2, eta = 1/(4 sqrt(pi)), radius 1, optInd 0:

  mobility_{bw,sc}.npz        self_mobility_{bw,sc}.mat, pair_mobility_{bw,sc}.mat
      kernType 0 with torques, ref False: 135 x 135 grid (Nz 61 / 55)
  mobility_bw_ref2x.npz       self_mobility_bw_ref2x.mat
  mobility_sc_ref.npz         self_mobility_sc_ref.mat, pair_mobility_sc_ref.mat
      kernType 0 with torques, ref True: 270 x 270 grid (Nz 121 / 109)
  mobility_{bw,sc}_w4.npz     self_mobility_{bw,sc}_w4.mat, pair_mobility_{bw,sc}_w4.mat
      translation only (has_torque False).  The self runs use kernType 2
      (w = 4, 90 x 90 grid, Nz 43 / 37); the pair runs pass kernType=0 to
      Initialize whatever the script's kernType (pair_mobility.py:39), so
      they are w = 6 without torque (120 x 120 grid, Nz 57 / 49).

Self runs (figures/self_mobility.py:8-80): one particle at (L/2, L/2, h) for
60 heights, matrix 6x6 (3x3 translation only).  Pair runs
(figures/pair_mobility.py:7-95): separations 3, 4, 8 along x, 26 heights
(row 0 unused), matrix 12x12 (6x6).  Matrices are normalised as
figures/mobility_matrix.py:19-30 does.  Every grid above is even, so all of
them but the 135 x 135 runs carry Nyquist modes.

  mobility_asym_posdef.npz    mobility_asymm_posdef_{bw,sc}_{w6,w4}.mat
      figures/pair_mobility_matrix_asym.py: 99 random (unseeded) pair
      configurations, x, y ~ U[0, L), z ~ U[0, zmax); ||M - M^T|| / ||M|| and
      the smallest eigenvalue of the normalised M.  Initialize gets
      kernType=0 whatever the script's kernType (line 23): 'w6' has torques
      (12 x 12), 'w4' has none (6 x 6).  Entry 0 of each array is unused.

Usage (needs the reference tree; the fixtures are committed):
    python tests/golden/make_golden.py /root/reference
"""
import os
import sys

import numpy as np
import scipy.io as sio

HERE = os.path.dirname(os.path.abspath(__file__))
COMMON = dict(L=76.8, zmin=0.0, zmax=19.2, eta=1 / (4 * np.sqrt(np.pi)), radius=1.0)

# fixture -> (self .mat or None, pair .mat or None, Initialize arguments of each)
FIXTURES = {
    "mobility_bw": ("self_mobility_bw", "pair_mobility_bw",
                    dict(kernType=0, has_torque=True, ref=False),
                    dict(kernType=0, has_torque=True, ref=False)),
    "mobility_sc": ("self_mobility_sc", "pair_mobility_sc",
                    dict(kernType=0, has_torque=True, ref=False),
                    dict(kernType=0, has_torque=True, ref=False)),
    "mobility_bw_ref2x": ("self_mobility_bw_ref2x", None,
                          dict(kernType=0, has_torque=True, ref=True), None),
    "mobility_sc_ref": ("self_mobility_sc_ref", "pair_mobility_sc_ref",
                        dict(kernType=0, has_torque=True, ref=True),
                        dict(kernType=0, has_torque=True, ref=True)),
    "mobility_bw_w4": ("self_mobility_bw_w4", "pair_mobility_bw_w4",
                       dict(kernType=2, has_torque=False, ref=False),
                       dict(kernType=0, has_torque=False, ref=False)),
    "mobility_sc_w4": ("self_mobility_sc_w4", "pair_mobility_sc_w4",
                       dict(kernType=2, has_torque=False, ref=False),
                       dict(kernType=0, has_torque=False, ref=False)),
}


def main(ref_root):
    data = os.path.join(ref_root, "figures", "Figures", "data")
    for name, (sname, pname, sargs, pargs) in FIXTURES.items():
        out = dict(COMMON)
        if sname:
            s = sio.loadmat(os.path.join(data, sname + ".mat"))
            out.update(self_heights=np.asarray(s["heights"], dtype=np.float64).ravel(),
                       self_M=np.asarray(s["M"], dtype=np.float64),
                       self_kernType=sargs["kernType"], self_has_torque=sargs["has_torque"],
                       self_ref=sargs["ref"])
        if pname:
            p = sio.loadmat(os.path.join(data, pname + ".mat"))
            out.update(pair_heights=np.asarray(p["heights"], dtype=np.float64).ravel(),
                       pair_M=np.asarray(p["M"], dtype=np.float64),
                       pair_seps=np.array([3.0, 4.0, 8.0]),
                       pair_kernType=pargs["kernType"], pair_has_torque=pargs["has_torque"],
                       pair_ref=pargs["ref"])
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **out)
        print("wrote", path, os.path.getsize(path), "bytes")
    out = dict(COMMON)
    for tag in ("bw", "sc"):
        for w in ("w6", "w4"):
            m = sio.loadmat(os.path.join(data, "mobility_asymm_posdef_%s_%s.mat" % (tag, w)))
            out["asym_%s_%s" % (tag, w)] = np.asarray(m["Masym"], dtype=np.float64).ravel()[1:]
            out["posdef_%s_%s" % (tag, w)] = np.asarray(m["Posdef"], dtype=np.float64).ravel()[1:]
    path = os.path.join(HERE, "mobility_asym_posdef.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
