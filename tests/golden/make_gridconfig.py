"""Write tests/golden/gridconfig.npz: input/output vectors of the reference's
grid and kernel auto-configuration (source/cpu/python/GridAndKernelConfig.py:
configure_grid_and_kernels_xy, lines 4-361, and configure_grid_and_kernels_z,
lines 363-428), run from the reference tree.

No stored output of the reference covers kernel types 1, 3, 4 and 5 or most
optInd choices, so these vectors come from the reference function itself.
It is imported from /root/reference; numpy 2 removed the alias np.int (the
builtin int) that the reference's function uses (lines 74, 124-125, 151,
218, 296-297, 328, 359), so it is restored for the call.  The function's
printed tables go to /dev/null.

Each case is (Lx, Ly, radius, kernType, has_torque, optInd, ref, domType,
zmin, zmax); a case the reference rejects is stored with ok = -1
(ValueError) or -2 (IndexError: an optInd past the candidate grids).

Usage (needs the reference tree; the fixture is committed):
    python tests/golden/make_gridconfig.py /root/reference
"""
import contextlib
import importlib.util
import io
import itertools
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def load_reference(ref_root):
    path = os.path.join(ref_root, "source", "cpu", "python", "GridAndKernelConfig.py")
    spec = importlib.util.spec_from_file_location("reference_gridconfig", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def cases():
    boxes = [(76.8, 76.8), (50.0, 64.0), (128.0, 96.0)]
    radii = [0.7, 1.0, 1.3]
    doms = [("DP", 2.0, 14.0), ("DPBW", 0.0, 19.2), ("DPSC", 0.0, 10.0), ("TP", 0.0, 12.0)]
    out = []
    for (Lx, Ly), rad, opt, ref in itertools.product(boxes, radii, (0, 1, 3, -1), (False, True)):
        for kt in (0, 1, 3, 4, 5):
            for dom, zmin, zmax in doms:
                out.append((Lx, Ly, rad, kt, True, opt, ref, dom, zmin, zmax))
        for kt in (0, 1, 2):
            for dom, zmin, zmax in doms:
                out.append((Lx, Ly, rad, kt, False, opt, ref, dom, zmin, zmax))
    return out


def main(ref_root):
    G = load_reference(ref_root)
    if not hasattr(np, "int"):
        np.int = int
    rows_in, rows_f, rows_i, ok = [], [], [], []
    doms = {"DP": 0, "DPBW": 1, "DPSC": 2, "TP": 3}
    for c in cases():
        Lx, Ly, rad, kt, ht, opt, ref, dom, zmin, zmax = c
        rows_in.append([Lx, Ly, rad, kt, int(ht), opt, int(ref), doms[dom], zmin, zmax])
        try:
            with contextlib.redirect_stdout(io.StringIO()):
                r = G.configure_grid_and_kernels_xy(Lx, Ly, np.array([rad]), np.array([kt]), opt, ht, ref)
                Lx2, Ly2, hx, hy, Nx, Ny, wm, wd, cbm, cbd, bm, bd = r
                Lz, hz, Nz, a = G.configure_grid_and_kernels_z(zmin, zmax, hx, wm, wd, dom, 1.5, ref)
        except (ValueError, IndexError) as e:
            ok.append(-1 if isinstance(e, ValueError) else -2)
            rows_f.append([np.nan] * 12)
            rows_i.append([0] * 5)
            continue
        first = lambda v: float(np.ravel(v)[0]) if np.size(v) else 0.0
        ok.append(1)
        rows_f.append([Lx2, Ly2, hx, hy, first(cbm), first(cbd), first(bm), first(bd), Lz, hz, a,
                       first(wm)])
        rows_i.append([int(Nx), int(Ny), int(first(wm)), int(first(wd)), int(Nz)])
    path = os.path.join(HERE, "gridconfig.npz")
    np.savez_compressed(path, inputs=np.array(rows_in, dtype=np.float64),
                        out_f=np.array(rows_f, dtype=np.float64), out_i=np.array(rows_i, dtype=np.int64),
                        ok=np.array(ok, dtype=np.int64),
                        in_cols=np.array("Lx Ly radius kernType has_torque optInd ref domType zmin zmax".split()),
                        f_cols=np.array("Lx Ly hx hy cbeta_m cbeta_d beta_m beta_d Lz hz z0 w_m".split()),
                        i_cols=np.array("Nx Ny w_m w_d Nz".split()),
                        dom_codes=np.array(["DP", "DPBW", "DPSC", "TP"]))
    print("wrote", path, os.path.getsize(path), "bytes;", sum(k == 1 for k in ok), "of", len(ok), "cases accepted")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
