"""The solver's device arithmetic, compiled for the host, against the oracle.

tests/native/core_check.cpp runs the exact code of dps_core.h (stencils,
per-mode sequential BVP), dps_seg.h (the lane-group BVP of the fused
spectral kernel, with one host thread per lane) and dps_tables.h (z grids,
Chebyshev tables) on the CPU, so these checks need no GPU.
"""
import ctypes

import numpy as np
import pytest
from numpy.polynomial import chebyshev as C

from oracle import dp_oracle as o

np.seterr(all="ignore")
GEOM = {"nowall": 1, "bottom": 2, "slit": 3}


def P(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def D(x):
    return ctypes.c_double(x)


def mode_inputs(Nx, Ny, Nz, Lx, Ly, seed):
    rng = np.random.RandomState(seed)
    ix = np.tile(np.arange(Nx), Ny)
    iy = np.repeat(np.arange(Ny), Nx)
    kx = 2 * np.pi * np.where(ix <= Nx // 2, ix, ix - Nx) / Lx
    ky = 2 * np.pi * np.where(iy <= Ny // 2, iy, iy - Ny) / Ly
    nyqx = ((Nx % 2 == 0) & (ix == Nx // 2)).astype(np.int32)
    nyqy = ((Ny % 2 == 0) & (iy == Ny // 2)).astype(np.int32)
    M = Nx * Ny
    Cf = (rng.normal(size=(M, Nz, 3)) + 1j * rng.normal(size=(M, Nz, 3))) * np.exp(-0.3 * np.arange(Nz))[None, :, None]
    fin = np.ascontiguousarray(np.stack([Cf.real, Cf.imag], -1).transpose(0, 2, 1, 3))   # [M][3][Nz][2]
    return kx, ky, nyqx, nyqy, Cf, fin


def to_complex(u):
    return (u[..., 0] + 1j * u[..., 1]).transpose(0, 2, 1)   # [M][Nz][3]


@pytest.mark.parametrize("geometry", ["nowall", "bottom", "slit"])
def test_sequential_bvp_matches_oracle(corecheck, geometry):
    """bvp_mode/bvp_k0 (one wavenumber per lane) against the oracle's
    vectorised per-mode chain, wall corrections included: the kernel returns
    correction coefficients of the analytic profiles, which are summed here in
    value space and transformed like the oracle's correction."""
    Nx, Ny, Nz, Lx, Ly, Lz, eta = 16, 12, 17, 8.0, 6.0, 5.0, 0.8
    H = Lz / 2
    ops = o.DPOperators(Nx, Ny, Nz, Lx, Ly, H)
    kx, ky, nyqx, nyqy, Cf, fin = mode_inputs(Nx, Ny, Nz, Lx, Ly, 3)
    U, Pp = o.dp_solve_no_wall(ops, Cf, eta)
    z = o.clencurt(Nz, 0, Lz)[0]
    if geometry == "bottom":
        o.bottom_wall_correction(ops, U, Pp, z, eta)
        o.k0_wall_velocity(ops, Cf[0], U, eta, "bottom")
    elif geometry == "slit":
        o.slit_channel_correction(ops, U, Pp, z, Lz, eta)
        o.k0_wall_velocity(ops, Cf[0].real + 1j * Cf[0].real, U, eta, "slit")
    M = Nx * Ny
    uout = np.zeros_like(fin)
    corr = np.zeros((M, 3, 4, 2))
    corecheck.cc_bvp(Nz, D(H), D(Lz), D(eta), GEOM[geometry], M, P(kx), P(ky), P(nyqx), P(nyqy),
                     P(fin), P(uout), P(corr))
    Uo = to_complex(uout)
    cc = corr[..., 0] + 1j * corr[..., 1]
    prof = np.zeros((M, Nz, 3), dtype=np.complex128)
    for m in range(1, M):
        K = np.hypot(kx[m], ky[m])
        e1, e2 = np.exp(-K * z), np.exp(K * (z - Lz))
        basis = [e1, e2, z * e1, z * e2] if geometry == "slit" else [e1, z * e1]
        for c in range(3):
            for b, f in enumerate(basis):
                prof[m, :, c] += cc[m, c, b] * f
    tot = Uo.copy()
    tot[1:] += o._cheb_of_profiles(prof)[1:]
    assert np.abs(tot - U).max() <= 1e-12 * np.abs(U).max()


@pytest.mark.parametrize("geometry", ["nowall", "bottom", "slit"])
def test_nyquist_mirror_is_a_sign_flip_of_the_wall_correction(corecheck, geometry):
    """The Hermitian average of a Nyquist-touching mode with its mirror
    (solved on the conjugated forcing with the mirror's grid wavenumbers)
    needs only one BVP solve: the mirror's velocity is the conjugate of the
    mode's own, and its conjugated wall correction is the correction of the
    mode's forcing at wavenumbers (-kx_mirror, -ky_mirror), whose BVP part is
    again the mode's own.  Checked with the sequential solver."""
    Nx, Ny, Nz, Lx, Ly, Lz, eta = 8, 6, 17, 8.0, 6.0, 5.0, 0.8
    kx, ky, nyqx, nyqy, Cf, fin = mode_inputs(Nx, Ny, Nz, Lx, Ly, 11)
    ix = np.tile(np.arange(Nx), Ny)
    iy = np.repeat(np.arange(Ny), Nx)
    sel = np.where(((nyqx == 1) | (nyqy == 1)) & ~((ix == 0) & (iy == 0)))[0]
    ixm, iym = (Nx - ix[sel]) % Nx, (Ny - iy[sel]) % Ny
    kxm = 2 * np.pi * np.where(ixm <= Nx // 2, ixm, ixm - Nx) / Lx
    kym = 2 * np.pi * np.where(iym <= Ny // 2, iym, iym - Ny) / Ly
    nxm = ((Nx % 2 == 0) & (ixm == Nx // 2)).astype(np.int32)
    nym = ((Ny % 2 == 0) & (iym == Ny // 2)).astype(np.int32)
    M = len(sel)
    f = np.ascontiguousarray(fin[sel])
    fc = f.copy()
    fc[..., 1] *= -1

    def solve(ff, a, b, na, nb):
        u, c = np.zeros_like(ff), np.zeros((M, 3, 4, 2))
        corecheck.cc_bvp(Nz, D(Lz / 2), D(Lz), D(eta), GEOM[geometry], M,
                         P(np.ascontiguousarray(a, dtype=np.float64)), P(np.ascontiguousarray(b, dtype=np.float64)),
                         P(np.ascontiguousarray(na, dtype=np.int32)), P(np.ascontiguousarray(nb, dtype=np.int32)),
                         P(ff), P(u), P(c))
        return u[..., 0] + 1j * u[..., 1], c[..., 0] + 1j * c[..., 1]

    uA, cA = solve(f, kx[sel], ky[sel], nyqx[sel], nyqy[sel])
    uM, cM = solve(fc, kxm, kym, nxm, nym)
    uB, cB = solve(f, -kxm, -kym, nxm, nym)
    su, sc = np.abs(uA).max(), max(np.abs(cA).max(), 1e-300)
    assert np.abs(np.conj(uM) - uA).max() <= 1e-13 * su
    assert np.abs(uB - uA).max() <= 1e-13 * su
    assert np.abs(np.conj(cM) - cB).max() <= 1e-13 * sc
    if geometry != "nowall":
        # the corrections themselves do differ (else the test pins nothing)
        assert np.abs(cB - cA).max() > 1e-3 * sc


SHAPES = [(17, 16, 2, 3, 0), (33, 16, 4, 1, 1), (61, 16, 4, 3, 0), (96, 16, 6, 1, 1), (96, 16, 6, 1, 0),
          (128, 16, 8, 1, 1), (150, 32, 8, 1, 1), (13, 4, 4, 3, 1)]


@pytest.mark.parametrize("geometry", ["nowall", "bottom", "slit"])
@pytest.mark.parametrize("shape", SHAPES, ids=["nz%d_lpm%d_e%d_nrv%d_stage%d" % s for s in SHAPES])
def test_lane_group_bvp_matches_sequential(corecheck, geometry, shape):
    """seg_bvp_mode/seg_bvp_k0 (LPM lanes of E rows, affine scans across
    lanes) against the sequential solver, padding rows included; stage = the
    in-place variant that stages the x/y right-hand sides in the forcing."""
    Nz, lpm, e, nrv, stage = shape
    Nx, Ny, Lx, Ly, Lz, eta = 4, 3, 8.0, 6.0, 5.0, 0.8
    kx, ky, nyqx, nyqy, Cf, fin = mode_inputs(Nx, Ny, Nz, Lx, Ly, Nz)
    M = Nx * Ny
    u1, u2 = np.zeros_like(fin), np.zeros_like(fin)
    c1, c2 = np.zeros((M, 3, 4, 2)), np.zeros((M, 3, 4, 2))
    args = (Nz, D(Lz / 2), D(Lz), D(eta), GEOM[geometry], M, P(kx), P(ky), P(nyqx), P(nyqy), P(fin))
    corecheck.cc_bvp(*args, P(u1), P(c1))
    assert corecheck.cc_bvp_seg(*args, P(u2), P(c2), lpm, e, nrv, stage) == 0
    assert np.abs(u2 - u1).max() <= 1e-13 * np.abs(u1).max()
    assert np.abs(c2 - c1).max() <= 1e-13 * max(np.abs(c1).max(), 1e-300)


@pytest.mark.parametrize("N", [9, 24, 61])
def test_cheb_tables_match_oracle(corecheck, N):
    H = 1.9
    b = {k: np.zeros(N) for k in ("sd0", "sm2", "sp2", "fm1", "fp1")}
    scal = np.zeros(3)
    rows = {k: np.zeros(N + 2) for k in ("R1", "R2", "L1", "L2")}
    corecheck.cc_cheb_tables(N, D(H), *(P(b[k]) for k in ("sd0", "sm2", "sp2", "fm1", "fp1")), P(scal),
                             *(P(rows[k]) for k in ("R1", "R2", "L1", "L2")))
    S, F = o.second_integral_matrix(N, H), o.first_integral_matrix(N, H)
    i = np.arange(N)
    np.testing.assert_allclose(b["sd0"], S[i, i], rtol=1e-14, atol=0)
    np.testing.assert_allclose(b["sm2"][2:], S[i[2:], i[2:] - 2], rtol=1e-14, atol=0)
    np.testing.assert_allclose(b["sp2"], S[i, i + 2], rtol=1e-14, atol=0)
    np.testing.assert_allclose(b["fm1"][1:], F[i[1:], i[1:] - 1], rtol=1e-14, atol=0)
    np.testing.assert_allclose(b["fp1"], F[i, i + 1], rtol=1e-14, atol=0)
    assert scal[0] == S[0, N] and scal[1] == S[1, N + 1] and scal[2] == F[0, N + 1]
    for k, ref in zip(("R1", "R2", "L1", "L2"), o.no_wall_bc_rows(N)):
        np.testing.assert_allclose(rows[k], ref, rtol=1e-14, atol=1e-15)


@pytest.mark.parametrize("N", [5, 8, 33, 96])
def test_folded_cheb_transform_is_chebyshev_interpolation(corecheck, N):
    """Forward = inverse of evaluation at the nodes t_j = cos(pi j/(N-1));
    inverse = evaluation."""
    Tf, Ti = np.zeros((N, N)), np.zeros((N, N))
    corecheck.cc_cheb_transform(N, P(Tf), P(Ti))
    t = np.cos(np.pi * np.arange(N) / (N - 1))
    V = C.chebvander(t, N - 1)          # V[j][k] = T_k(t_j)
    np.testing.assert_allclose(Ti, V, atol=1e-13)
    np.testing.assert_allclose(Tf @ V, np.eye(N), atol=1e-12)


def test_kernel_constants_match_oracle(corecheck):
    corecheck.cc_es_norm.restype = ctypes.c_double
    corecheck.cc_dipole_thresh.restype = ctypes.c_double
    L = o.lib()
    for a, bw in ((1.5, 10.71), (2.0, 8.0), (0.9, 13.0)):
        assert corecheck.cc_es_norm(D(a), D(bw)) == pytest.approx(L.orc_es_norm(a, bw), rel=1e-15)
        assert corecheck.cc_dipole_thresh(D(a), D(bw)) == pytest.approx(L.orc_dipole_thresh(a, bw), rel=1e-14)


@pytest.mark.parametrize("walls", [(0, 0), (0, 1), (1, 1)])
def test_extended_z_grid_and_fold_tables(corecheck, walls):
    """Ghost nodes mirror interior nodes across the ends (ParticleList.cpp:419-451);
    fold factors: -1 (velocity-like, mirror_inv) / +1 (mirror) at walls, 0 at
    open ends, and (1 + s) on the wall node itself."""
    Nz, Lz, th = 25, 6.0, 1.3
    zG, zW = o.clencurt(Nz, 0, Lz)
    n = 3 * Nz
    z, w, fa, fb = (np.zeros(n) for _ in range(4))
    kt = np.zeros(n, dtype=np.int32)
    info = np.zeros(3, dtype=np.int32)
    nzeff = corecheck.cc_zext(P(zG), P(zW), Nz, D(th), walls[0], walls[1], P(z), P(w), P(kt), P(fa), P(fb), P(info))
    up, down = info[0], info[1]
    n2 = Nz + 2 * (Nz - 1)
    zo, wo = np.zeros(n2), np.zeros(n2)
    iu, idn = ctypes.c_int(), ctypes.c_int()
    nzo = o.lib().orc_zext(o._p(zG), o._p(zW), Nz, th, o._p(zo), o._p(wo), ctypes.byref(iu), ctypes.byref(idn))
    assert nzeff == nzo and up == iu.value and down == idn.value
    np.testing.assert_array_equal(z[:nzeff], zo[:nzeff])
    np.testing.assert_array_equal(w[:nzeff], wo[:nzeff])
    for e in range(nzeff):
        k = kt[e]
        if e < up:
            assert z[e] == pytest.approx(2 * Lz - zG[k]) and fa[e] == (-1.0 if walls[0] else 0.0)
        elif e >= up + Nz:
            assert z[e] == pytest.approx(-zG[k]) and fa[e] == (-1.0 if walls[1] else 0.0)
        else:
            assert k == e - up
    if walls[1]:
        assert fa[up + Nz - 1] == 0.0 and fb[up + Nz - 1] == 2.0


def test_stencils_match_oracle_locate(corecheck):
    """Supports and 1-D weights of stencil_support/stencil_weights_* against
    the oracle's locate (ParticleList.cpp:453-524)."""
    s = o.OracleDPStokes("bottom", 8.0, 7.0, 0.0, 6.0, 16, 14, 21, 0.8, 6, 10.71)
    rng = np.random.RandomState(5)
    N = 40
    pos = np.stack([rng.uniform(0, 8, N), rng.uniform(0, 7, N), rng.uniform(0, 6, N)], 1).ravel()
    s.set_positions(pos)
    loc, k = s._loc, s.mono
    wx, wy, zs = loc["wx"], loc["wy"], loc["zstride"]
    sup = np.zeros(4 * N, dtype=np.int32)
    arr = lambda m: np.zeros(N * m)
    Wx, Dx, Wy, Dy = arr(wx), arr(wx), arr(wy), arr(wy)
    wzmax = zs
    Wz, Dz, Wzi, Dzi = arr(wzmax), arr(wzmax), arr(wzmax), arr(wzmax)
    corecheck.cc_stencil(N, P(s.xs), D(s.hx), D(s.hy), D(k.betaw), D(k.alpha), D(k.thresh), D(1 / k.norm ** 3),
                         wx, wy, P(loc["zext"]), P(loc["zext_w"]), loc["Nzeff"], wzmax, P(sup), P(Wx), P(Dx),
                         P(Wy), P(Dy), P(Wz), P(Dz), P(Wzi), P(Dzi))
    sup = sup.reshape(N, 4)
    phi = lambda r: np.exp(k.betaw * (np.sqrt(1 - r * r / k.alpha ** 2) - 1))
    even = 1 - wx % 2
    np.testing.assert_array_equal(sup[:, 0], loc["xclose"] - wx // 2 + even)
    np.testing.assert_array_equal(sup[:, 1], loc["yclose"] - wy // 2 + even)
    np.testing.assert_array_equal(sup[:, 2], loc["indl"])
    np.testing.assert_array_equal(sup[:, 3], loc["indr"] - loc["indl"] + 1)
    np.testing.assert_allclose(Wx, phi(loc["xunwrap"]) / k.norm ** 3, rtol=1e-14)
    np.testing.assert_allclose(Wy, phi(loc["yunwrap"]), rtol=1e-14)
    for p in range(N):
        nz = sup[p, 3]
        zu = loc["zunwrap"][p * zs:p * zs + nz]
        np.testing.assert_allclose(Wz[p * wzmax:p * wzmax + nz], phi(zu), rtol=1e-14)
        np.testing.assert_allclose(Wzi[p * wzmax:p * wzmax + nz], phi(zu) * loc["pt_wts"][p * zs:p * zs + nz],
                                   rtol=1e-14)


@pytest.mark.parametrize("R", [2, 4, 8, 16, 32])
def test_register_fft_matches_numpy(corecheck, R):
    """The y-pass kernel's register DFT (dps_fft.h) against numpy's FFT, both
    directions (inverse unnormalised)."""
    import ctypes
    rng = np.random.RandomState(R)
    x = rng.normal(size=R) + 1j * rng.normal(size=R)
    for d, want in ((-1, np.fft.fft(x)), (1, R * np.fft.ifft(x))):
        re = np.ascontiguousarray(x.real)
        im = np.ascontiguousarray(x.imag)
        p = ctypes.POINTER(ctypes.c_double)
        assert corecheck.cc_fft_reg(R, d, re.ctypes.data_as(p), im.ctypes.data_as(p)) == 0
        assert np.abs(re + 1j * im - want).max() <= 1e-14 * np.abs(want).max()
