"""The planar transforms with the y pass in k_fft_y (every Ny the kernel
handles, powers of two and 3 * 2^k, odd and even Nx so that the flattened
(slice, kx) column tiles straddle slices at different offsets) and the x pass
in k_fft_x (every length it handles), against the FP64 oracle, through Mdot
with torques.  Bounds as everywhere: rel-L2 <= 1e-11 FP64, <= 1e-5 mixed."""
import numpy as np
import pytest

from test_gpu_parity import TOL64, TOLMIX, gpu_for, make_case, oracle_for, rel_l2, say

pytestmark = pytest.mark.gpu

# name, geometry, Lx, Ly, Lz, nx, ny, nz, N, torque
CASES = [
    ("fy16", "nowall", 8.0, 8.0, 6.3, 16, 16, 13, 20, True),
    ("fy64_oddx", "slit", 13.5, 32.0, 6.3, 27, 64, 13, 30, True),
    ("fy128", "bottom", 12.0, 64.0, 6.0, 24, 128, 11, 30, True),
    ("fy256_oddx", "slit", 7.5, 128.0, 5.0, 15, 256, 9, 30, True),
    ("fy512", "nowall", 8.0, 256.0, 6.0, 16, 512, 11, 30, True),
    ("fy1024_oddx", "bottom", 6.5, 512.0, 6.0, 13, 1024, 9, 30, True),
    # power-of-two Nx as well: the x pass in k_fft_x (Nx/2 = 4x4 ... 32x32)
    ("fx32", "slit", 16.0, 16.0, 6.3, 32, 32, 13, 20, True),
    ("fx64", "slit", 32.0, 16.0, 6.3, 64, 32, 13, 30, True),
    ("fx512", "bottom", 256.0, 8.0, 6.0, 512, 16, 11, 30, True),
    ("fx2048", "nowall", 1024.0, 8.0, 6.0, 2048, 16, 9, 30, True),
    # sizes 3 * 2^k: radix-3 register transforms in either pass
    ("fy48", "slit", 6.5, 24.0, 6.3, 13, 48, 13, 20, True),
    ("fxy96", "bottom", 48.0, 48.0, 6.0, 96, 96, 11, 30, True),
    ("fxy192", "slit", 96.0, 96.0, 6.3, 192, 192, 13, 40, True),
    ("fx384", "nowall", 192.0, 8.0, 6.0, 384, 16, 9, 30, True),
    ("fy384", "bottom", 8.0, 192.0, 6.0, 16, 384, 9, 30, True),
    ("fx768", "slit", 384.0, 8.0, 6.3, 768, 16, 13, 30, True),
    ("fy768", "bottom", 8.0, 384.0, 6.0, 16, 768, 9, 30, True),
    ("fx1536", "nowall", 768.0, 8.0, 6.0, 1536, 16, 9, 30, True),
]


@pytest.mark.parametrize("precision", ["double", "mixed"])
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_fft_y_pass_mdot(case, precision):
    say("case %s %s y-pass FFT mdot" % (case[0], precision))
    pos, F, T = make_case(case, seed=7)
    ref = oracle_for(case)
    ref.set_positions(pos)
    v_ref, w_ref = ref.mdot(F, T)
    s = gpu_for(case, precision)
    s.setPositions(pos)
    v, w = s.Mdot(F, T)
    tol = TOL64 if precision == "double" else TOLMIX
    ev, ew = rel_l2(v, v_ref), rel_l2(w, w_ref)
    say("  rel-L2 U = %.3e  W = %.3e" % (ev, ew))
    assert ev <= tol and ew <= tol
    s.clear()
