"""Write the golden mobility fixtures from the reference's stored figure data.

Reads (MATLAB v5 data files, loaded with scipy.io.loadmat -- data only):
  figures/Figures/data/self_mobility_bw.mat, self_mobility_sc.mat
      written by figures/self_mobility.py: 60 heights, 6x6 normalised
      self-mobility of one particle at (L/2, L/2, h)
  figures/Figures/data/pair_mobility_bw.mat, pair_mobility_sc.mat
      written by figures/pair_mobility.py:29-60: separations 3, 4, 8 along x,
      26 heights (row 0 unused), 12x12 normalised pair mobility
Both scripts use FCMJoint('cpu') with L = 76.8, zmin = 0, zmax = 19.2,
eta = 1/(4 sqrt(pi)), radius 1, kernType 0 (torques on), optInd 0, ref False.
The matrices are normalised as figures/mobility_matrix.py:19-25 does.

Usage (needs the reference tree; the fixtures are committed):
    python tests/golden/make_golden.py /root/reference
"""
import os
import sys

import numpy as np
import scipy.io as sio

HERE = os.path.dirname(os.path.abspath(__file__))


def main(ref_root):
    data = os.path.join(ref_root, "figures", "Figures", "data")
    for tag in ("bw", "sc"):
        s = sio.loadmat(os.path.join(data, "self_mobility_%s.mat" % tag))
        p = sio.loadmat(os.path.join(data, "pair_mobility_%s.mat" % tag))
        out = os.path.join(HERE, "mobility_%s.npz" % tag)
        np.savez_compressed(out,
                            self_heights=np.asarray(s["heights"], dtype=np.float64).ravel(),
                            self_M=np.asarray(s["M"], dtype=np.float64),
                            pair_heights=np.asarray(p["heights"], dtype=np.float64).ravel(),
                            pair_M=np.asarray(p["M"], dtype=np.float64),
                            pair_seps=np.array([3.0, 4.0, 8.0]),
                            L=76.8, zmin=0.0, zmax=19.2, eta=1 / (4 * np.sqrt(np.pi)), radius=1.0)
        print("wrote", out, os.path.getsize(out), "bytes")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
