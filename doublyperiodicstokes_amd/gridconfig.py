"""Grid and ES-kernel auto-configuration for the joint interface.

Restates ``source/cpu/python/GridAndKernelConfig.py`` of the reference:
``configure_grid_and_kernels_xy`` (lines 4-361), ``configure_grid_and_kernels_z``
(lines 363-428) and the FFT-friendly size search (lines 430-459).  It maps a
hydrodynamic radius, a kernel type and a box to a grid spacing, grid sizes,
kernel widths and ES ``beta`` values; ``FCMJoint.Initialize`` calls it.

The four polynomial fits below (c(beta) = R_h/(w h) for monopole/dipole
kernels and their inverses) are data of the reference
(``GridAndKernelConfig.py:472-526``) and must be identical for the same grid
to be chosen.  The reference's error tables (``errm*``/``errd*``) only feed
the printed error column; candidate grids are rejected on the same beta*w
range test, so they are not needed here.
"""
import numpy as np

# c(beta*w) = R_h / (w h) fits and inverses (GridAndKernelConfig.py:472-526)
CBETAM = np.array([
    1.377360225705533e-14, -2.780166372816629e-12, 2.448353284693753e-10,
    -1.237430413656322e-08, 3.979148502148348e-07, -8.561766184042201e-06,
    0.0001272679847510219, -0.001354403465648405, 0.01097749452611932,
    -0.07481288697153615, 0.5542048130466118])
CBETAD = np.array([
    -4.547131634839802e-15, 9.269149081395723e-13, -8.100791156357088e-11,
    3.923458768903104e-09, -1.120891440829999e-07, 1.758862781905333e-06,
    -6.91617984068661e-06, -0.0002952983095968742, 0.006854380595771486,
    -0.07755569487992697, 0.6862730778816093])
CBETAM_INV = np.array([
    4131643418.193291, -10471683395.26777, 11833009228.6429, -7851132955.882548,
    3388121732.651829, -994285251.2185925, 201183449.7086889, -27776767.88241613,
    2515647.646492857, -136305.2970161326, 3445.959503226691])
CBETAD_INV = np.array([
    39129605.95992675, -141095001.9380851, 226461386.9399433, -213225975.7007014,
    130589579.4920038, -54470021.75408537, 15723274.21737542, -3119398.842468358,
    411336.473729197, -33239.91568304499, 1311.988715157221])

# (w_m, beta_m, h_opt divisor) without torque; (w_m, w_d, beta_m, beta_d, h_opt divisor) with
# torque, per kernel type (GridAndKernelConfig.py:48-58 and 184-208)
_NO_TORQUE = {0: (6, 1.714, 1.5539), 1: (5, 1.886, 1.3437), 2: (4, 1.785, 1.2047)}
_TORQUE = {0: (6, 6, 1.3267, 2.2160, 1.7305), 1: (5, 5, 1.3055, 2.2320, 1.5598),
           3: (6, 5, 1.3261, 1.7584, 1.7309), 4: (6, 4, 1.5728, 1.5115, 1.6121),
           5: (5, 4, 1.3536, 1.7147, 1.5382)}


def _is_valid(n):
    while n % 2 == 0:
        n /= 2
    while n % 3 == 0:
        n /= 3
    while n % 5 == 0:
        n /= 5
    while n % 7 == 0:
        n /= 7
    return n == 1


def _find_best(N):
    for i in range(0, N):
        if _is_valid(N + i):
            return N + i
        if _is_valid(N - i):
            return N - i
    return 0


def fft_friendly_sizes(N, sep, count):
    """Up to ``count`` distinct 2^a3^b5^c7^d sizes near N..N+sep (lines 449-459)."""
    Ns = []
    c = 0
    for i in range(N, N + sep):
        n = _find_best(i)
        if n not in Ns:
            if c == count:
                break
            Ns.append(n)
            c += 1
    return np.array(Ns, dtype=np.int64)


def configure_grid_and_kernels_xy(Lx, Ly, pRs, kernTypes, optInd, has_torque=False, ref=False,
                                  verbose=False):
    """x-y grid and kernels (GridAndKernelConfig.py:4-361).

    Returns ``Lx, Ly, hx, hy, Nx, Ny, wm, wd, cbeta_m, cbeta_d, betafP, betadP``
    with the reference's meaning (beta per unit width, so beta*w is the ES
    exponent).  Raises ``ValueError`` for unsupported kernel types or an
    out-of-range candidate, as the reference does.
    """
    pRs = np.asarray(pRs, dtype=np.float64)
    kernTypes = np.asarray(kernTypes, dtype=int)
    nP = pRs.size
    minPr = np.min(pRs)
    wm = np.zeros(nP, dtype=np.ushort)
    betafP = np.zeros(nP)
    h = np.zeros(nP)
    if not has_torque:
        for iP in range(nP):
            if kernTypes[iP] not in _NO_TORQUE:
                raise ValueError('translation-only kernel type not supported')
            wm[iP], betafP[iP], div = _NO_TORQUE[kernTypes[iP]]
            hopt = minPr / div
            h[iP] = pRs[iP] / (wm[iP] * np.polyval(CBETAM, betafP[iP] * wm[iP]))
        h = np.min(h)
        Nx_all = fft_friendly_sizes(int(np.floor(Lx / h)), 100, 10)
        hx_all = Lx / Nx_all
        betaf_test = np.zeros((nP, len(Nx_all)))
        ok = np.ones((nP, len(Nx_all)), dtype=bool)
        betaf_opt = np.zeros(nP)
        for iP in range(nP):
            betaf_opt[iP] = np.polyval(CBETAM_INV, pRs[iP] / (wm[iP] * hopt)) / wm[iP]
            if betaf_opt[iP] * wm[iP] < 4 or betaf_opt[iP] * wm[iP] > 18:
                raise ValueError('Could not find optimal (h,beta) in range for particle #%d' % iP)
            for j in range(len(hx_all)):
                betaf_test[iP, j] = np.polyval(CBETAM_INV, pRs[iP] / (wm[iP] * hx_all[j])) / wm[iP]
                b = betaf_test[iP, j] * wm[iP]
                ok[iP, j] = not (b < 4 or b > 18)
        Nxopt = int(np.round(Lx / hopt))
        Nyopt = int(np.round(Ly / hopt))
        if optInd >= 0:
            if not ok[:, optInd].all():
                raise ValueError('Out of range interpolation')
            betafP = betaf_test[:, optInd].copy()
            Nx = int(Nx_all[optInd])
            hx = hy = hx_all[optInd]
            Ny = Ly / hy
            if np.round(Ny) != Ny:
                Ny = int(np.ceil(Ny))
                Ly = hy * Ny
        else:
            betafP = betaf_opt
            Lx, Ly, Nx, Ny = hopt * Nxopt, hopt * Nyopt, Nxopt, Nyopt
            hx = hy = hopt
        cbeta_m = np.array([pRs[iP] / (wm[iP] * hx) for iP in range(nP)])
        cbeta_d = 0
        wd = np.array([0])
        betadP = np.array([0])
        if ref:
            hx /= 2
            Nx *= 2
            hy /= 2
            Ny *= 2
    else:
        wd = np.zeros(nP, dtype=np.ushort)
        betadP = np.zeros(nP)
        for iP in range(nP):
            if kernTypes[iP] not in _TORQUE:
                raise ValueError('translation-rotation kernel type not supported')
            wm[iP], wd[iP], betafP[iP], betadP[iP], div = _TORQUE[kernTypes[iP]]
            hopt = minPr / div
            h[iP] = pRs[iP] / (wm[iP] * np.polyval(CBETAM, betafP[iP] * wm[iP]))
        h = np.min(h)
        Nx_all = fft_friendly_sizes(int(np.floor(Lx / h)), 20, 10)
        hx_all = Lx / Nx_all
        betaf_test = np.zeros((nP, len(Nx_all)))
        betad_test = np.zeros((nP, len(Nx_all)))
        okm = np.ones((nP, len(Nx_all)), dtype=bool)
        okd = np.ones((nP, len(Nx_all)), dtype=bool)
        betaf_opt = np.zeros(nP)
        betad_opt = np.zeros(nP)
        for iP in range(nP):
            betaf_opt[iP] = np.polyval(CBETAM_INV, pRs[iP] / (wm[iP] * hopt)) / wm[iP]
            betad_opt[iP] = np.polyval(CBETAD_INV, pRs[iP] / (wd[iP] * hopt)) / wd[iP]
            if (betaf_opt[iP] * wm[iP] < 4 or betaf_opt[iP] * wm[iP] > 18) or \
               (betad_opt[iP] * wd[iP] < 4 or betad_opt[iP] * wd[iP] > 18):
                raise ValueError('Could not find optimal (h,beta) in range for particle #%d' % iP)
            for j in range(len(hx_all)):
                betaf_test[iP, j] = np.polyval(CBETAM_INV, pRs[iP] / (wm[iP] * hx_all[j])) / wm[iP]
                betad_test[iP, j] = np.polyval(CBETAD_INV, pRs[iP] / (wd[iP] * hx_all[j])) / wd[iP]
                bm = betaf_test[iP, j] * wm[iP]
                bd = betad_test[iP, j] * wd[iP]
                okm[iP, j] = not (bm < 4 or bm > 18)
                okd[iP, j] = not (bd < 4 or bd > 18)
        Nxopt = int(np.round(Lx / hopt))
        Nyopt = int(np.round(Ly / hopt))
        if optInd >= 0:
            if not okm[:, optInd].all() or not okd[:, optInd].all():
                raise ValueError('Out of range interpolation')
            betafP = betaf_test[:, optInd].copy()
            betadP = betad_test[:, optInd].copy()
            Nx = int(Nx_all[optInd])
            hx = hy = hx_all[optInd]
            Ny = Ly / hy
            if np.round(Ny) != Ny:
                Ny = int(np.ceil(Ny))
                Ly = hy * Ny
        else:
            betafP = betaf_opt
            betadP = betad_opt
            Lx, Ly, Nx, Ny = hopt * Nxopt, hopt * Nyopt, Nxopt, Nyopt
            hx = hy = hopt
        cbeta_m = np.array([pRs[iP] / (wm[iP] * hx) for iP in range(nP)])
        cbeta_d = np.array([pRs[iP] / (wd[iP] * hx) for iP in range(nP)])
        if ref:
            hx /= 2
            Nx *= 2
            hy /= 2
            Ny *= 2
    if verbose:
        print('grid: Nx=%d Ny=%d h=%.16g Lx=%.16g Ly=%.16g' % (Nx, Ny, hx, Lx, Ly))
    return Lx, Ly, hx, hy, int(Nx), int(Ny), wm, wd, cbeta_m, cbeta_d, betafP, betadP


def configure_grid_and_kernels_z(minZ, maxZ, h, wm, wd, domType, fac=1.5, ref=False):
    """z grid (GridAndKernelConfig.py:363-428).  Returns ``Lz, hz, Nz, a``."""
    if domType not in ('TP', 'DP', 'DPBW', 'DPSC'):
        raise ValueError('domType is not supported')
    hz = h * (domType == 'TP')
    if domType in ('TP', 'DPSC'):
        msep_down = msep_up = 0
    elif domType == 'DPBW':
        msep_down = 0
        msep_up = fac * np.max(wm) * h / 2
    else:
        msep_down = msep_up = fac * np.max(wm) * h / 2
    if ref:
        msep_down *= 2.0
        msep_up *= 2.0
    a = minZ - msep_down
    b = maxZ + msep_up
    Lz = (b - a)
    if domType == 'TP':
        Nz = Lz / hz
        if np.round(Nz) != Nz:
            Nz = int(np.ceil(Nz))
            Lz = hz * Nz
    else:
        H = Lz / 2
        Nz = int(np.ceil(1 * np.pi / (np.arccos(-h / H) - np.pi / 2)))
        Nz = int(fft_friendly_sizes(Nz, 100, 1)[0]) + 1
    return Lz, hz, int(Nz), a
