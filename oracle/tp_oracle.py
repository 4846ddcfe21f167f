"""FP64 CPU restatement of the reference's triply periodic (TP) Mdot.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the HIP path's
'periodic' mode; the product never imports it.

Follows:
  - FCM.py:149-441 (Mdot sequence, TP branches; the z grid is uniform and
    periodic, hz = hx, FCM.py:182-186 with GridAndKernelConfig.py:393-417);
  - ParticleList.cpp:285-393 (locateOnGridUnifZ: support selection by
    xclose = floor(x/h) (+1 for odd w past the half cell), points
    xclose + j - w/2 + even, offsets clamped to +thresh);
  - SpreadInterp.cpp:11-300 (spreadUnifZ / interpUnifZ: periodic in x, y, z;
    interpolation weight hx hy hz);
  - ESKernels.h:8-46 (3-D kernel exp(beta (sum sqrt(1 - x_i^2/alpha^2) - 3))
    and its derivative kernels), normalised by the 1-D integral cubed;
  - Solvers.py:197-255 and TPTools.cpp:6-46 (Fourier solve: p = -(D.f)/k^2,
    u = (f - D p)/(eta k^2), D = i k with the unpaired Nyquist entries zeroed,
    k^2 not zeroed, the k = 0 mode dropped);
  - Transform.cpp (TP: unnormalised forward 3-D DFT, backward divided by
    Nx Ny Nz, real part kept).

Reference quirk not reproduced: locateOnGridUnifZ subtracts the unit cell's
minZ from y instead of z (ParticleList.cpp:297); positions here are shifted by
zmin in z, which is what the reference intends and what it computes when
zmin = 0.
"""
import ctypes

import numpy as np

from oracle.dp_oracle import KernelSpec, _p, lib


def _wrap_offsets(xs, h, N, w, thresh):
    """Support grid indices (wrapped) and clamped offsets for one direction
    (ParticleList.cpp:339-368)."""
    even = 1 - (w % 2)
    close = np.floor(xs / h).astype(np.int64)   # (int)(x/h) for x >= 0
    if w % 2:
        close += ((xs / h - close) > 0.5).astype(np.int64)
    j = np.arange(w)
    idx = close[:, None] + j[None, :] - w // 2 + even
    # index * h - x rounded once, as the reference build computes it (fused
    # multiply-add; see orc_locate in oracle/spread_interp.c), then clamped
    xs = np.ascontiguousarray(xs, dtype=np.float64)
    close = np.ascontiguousarray(close, dtype=np.int64)
    off = np.zeros((xs.shape[0], w))
    lib().orc_fma_offsets(xs.shape[0], w, _p(xs),
                          close.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)), h, thresh, _p(off))
    return np.mod(idx, N), off


def _kern3(kind, ox, oy, oz, betaw, alpha):
    """3-D ES kernel (kind 0) or its x/y/z derivative (1/2/3) on the support
    tensor grid (ESKernels.h:8-46)."""
    a2 = alpha * alpha
    sx, sy, sz = (np.sqrt(1 - o * o / a2) for o in (ox, oy, oz))
    e = np.exp(betaw * (sx[:, None, None, :] + sy[:, None, :, None] + sz[:, :, None, None] - 3))
    if kind == 0:
        return e
    o, s = [(ox, sx), (oy, sy), (oz, sz)][kind - 1]
    d = -betaw * o / (a2 * s)
    shape = [(1, 1, 1, -1), (1, 1, -1, 1), (1, -1, 1, 1)][kind - 1]
    return e * d.reshape((d.shape[0],) + shape[1:])


class OracleTPStokes:
    """FP64 CPU Mdot, triply periodic box [0,Lx) x [0,Ly) x [zmin, zmin+Lz)."""

    def __init__(self, Lx, Ly, zmin, zmax, nx, ny, nz, eta, w, betaw, alpha=None,
                 w_d=None, betaw_d=None, alpha_d=None, has_torque=False, reg_kernel=False):
        self.Lx, self.Ly, self.zmin = float(Lx), float(Ly), float(zmin)
        self.Lz = float(zmax) - float(zmin)
        self.Nx, self.Ny, self.Nz = int(nx), int(ny), int(nz)
        self.hx, self.hy, self.hz = self.Lx / self.Nx, self.Ly / self.Ny, self.Lz / self.Nz
        self.eta = float(eta)
        if alpha is None or alpha < 0:
            alpha = w * self.hx / 2
        self.mono = KernelSpec(w, betaw, alpha)
        self.has_torque = has_torque
        if has_torque:
            if alpha_d is None or alpha_d < 0:
                alpha_d = w_d * self.hx / 2
            # useRegKernel (FCM.py:313-319): plain ES kernel of the dipole
            # width, support alpha_d
            self.dip = KernelSpec(w_d, betaw_d, alpha_d, dipole=not reg_kernel)
        self.reg_kernel = bool(reg_kernel) and has_torque

    def set_positions(self, positions):
        xs = np.array(positions, dtype=np.float64).reshape(-1, 3).copy()
        xs[:, 2] -= self.zmin
        # periodic wrap as ParticleList.cpp:561-580 (x, y; z likewise here)
        for d, L in enumerate((self.Lx, self.Ly, self.Lz)):
            v = xs[:, d]
            lo = v < 0
            v[lo] += (1 - np.trunc(v[lo] / L)) * L
            hi = v > L
            v[hi] -= np.trunc(v[hi] / L) * L
        self.xs = xs
        self.N = xs.shape[0]
        self._loc = self._locate(self.mono)
        self._tloc = self._locate(self.dip) if self.has_torque else None

    def _locate(self, ks):
        w = [int(round(2 * ks.thresh / h)) for h in (self.hx, self.hy, self.hz)]
        out = []
        for d, (h, N) in enumerate(((self.hx, self.Nx), (self.hy, self.Ny), (self.hz, self.Nz))):
            out.append(_wrap_offsets(self.xs[:, d], h, N, w[d], ks.thresh))
        return out

    def _weights(self, kind, loc, ks):
        (ix, ox), (iy, oy), (iz, oz) = loc
        return _kern3(kind, ox, oy, oz, ks.betaw, ks.alpha) / ks.norm ** 3

    def _flat_index(self, loc):
        (ix, _), (iy, _), (iz, _) = loc
        # grid (Ny, Nx, Nz) flattened; support tensor axes (z, y, x)
        return (iy[:, None, :, None] * self.Nx + ix[:, None, None, :]) * self.Nz + iz[:, :, None, None]

    def _spread(self, kind, loc, ks, f):
        W = self._weights(kind, loc, ks)
        g = self._flat_index(loc)
        G = np.zeros((self.Ny * self.Nx * self.Nz, 3))
        for d in range(3):
            np.add.at(G[:, d], g.ravel(), (W * f[:, d, None, None, None]).ravel())
        return G.reshape(self.Ny, self.Nx, self.Nz, 3)

    def _interp(self, kind, loc, ks, uG):
        W = self._weights(kind, loc, ks) * (self.hx * self.hy * self.hz)
        g = self._flat_index(loc)
        U = uG.reshape(-1, 3)
        return np.stack([np.sum(U[g, d] * W, axis=(1, 2, 3)) for d in range(3)], 1)

    def spread_forces(self, forces, torques=None):
        """Force grid plus the torque curl; with the regular torque kernel the
        torques are spread as a vector field into self.tgrid instead and their
        curl is taken in solve()."""
        F = self._spread(0, self._loc, self.mono, np.reshape(forces, (-1, 3)))
        self.tgrid = None
        if self.has_torque and torques is not None and np.size(torques) > 0:
            T = np.reshape(torques, (-1, 3))
            if self.reg_kernel:
                self.tgrid = self._spread(0, self._tloc, self.dip, T)
                return F
            dxT = self._spread(1, self._tloc, self.dip, T)
            dyT = self._spread(2, self._tloc, self.dip, T)
            dzT = self._spread(3, self._tloc, self.dip, T)
            F[..., 0] += 0.5 * (dyT[..., 2] - dzT[..., 1])
            F[..., 1] += 0.5 * (dzT[..., 0] - dxT[..., 2])
            F[..., 2] += 0.5 * (dxT[..., 1] - dyT[..., 0])
        return F

    def _operators(self):
        def kv(N, L):
            return 2 * np.pi * np.concatenate((np.arange(0, np.floor(N / 2) + 1),
                                               np.arange(-np.ceil(N / 2) + 1, 0))) / L
        ky, kx, kz = np.meshgrid(kv(self.Ny, self.Ly), kv(self.Nx, self.Lx), kv(self.Nz, self.Lz),
                                 indexing='ij')
        Ksq = kx ** 2 + ky ** 2 + kz ** 2
        Dx, Dy, Dz = 1j * kx, 1j * ky, 1j * kz
        if self.Nx % 2 == 0:
            Dx[:, self.Nx // 2, :] = 0
        if self.Ny % 2 == 0:
            Dy[self.Ny // 2, :, :] = 0
        if self.Nz % 2 == 0:
            Dz[:, :, self.Nz // 2] = 0
        return Ksq, Dx, Dy, Dz

    @staticmethod
    def curl_hat(C, Dx, Dy, Dz):
        """Half the Fourier curl of C (..., 3) (FCM.py:376-395, Solvers.py:137-157)."""
        out = np.empty_like(C)
        out[..., 0] = 0.5 * (Dy * C[..., 2] - Dz * C[..., 1])
        out[..., 1] = 0.5 * (Dz * C[..., 0] - Dx * C[..., 2])
        out[..., 2] = 0.5 * (Dx * C[..., 1] - Dy * C[..., 0])
        return out

    def solve(self, F, tgrid=None):
        """Fourier solve; with a torque grid (regular torque kernel) half its
        curl is added to the forcing and half the curl of the velocity is
        transformed back into self.utG."""
        Fh = np.fft.fftn(F, axes=(0, 1, 2))
        Ksq, Dx, Dy, Dz = self._operators()
        if tgrid is not None:
            Fh = Fh + self.curl_hat(np.fft.fftn(tgrid, axes=(0, 1, 2)), Dx, Dy, Dz)
        Ksq = Ksq.copy()
        Ksq[0, 0, 0] = 1.0
        f, g, h = Fh[..., 0], Fh[..., 1], Fh[..., 2]
        p = -(Dx * f + Dy * g + Dz * h) / Ksq
        val = 1.0 / (self.eta * Ksq)
        U = np.stack([val * (f - Dx * p), val * (g - Dy * p), val * (h - Dz * p)], -1)
        U[0, 0, 0, :] = 0
        self.utG = None
        if tgrid is not None:
            self.utG = np.real(np.fft.ifftn(self.curl_hat(U, Dx, Dy, Dz), axes=(0, 1, 2)))
        return np.real(np.fft.ifftn(U, axes=(0, 1, 2)))

    def interp_force_kernel(self, kind, uG):
        """Interpolate uG with the force kernel (kind 0) or its x/y/z
        derivative (kind 1/2/3) as a function of the particle position: zero
        where an offset was clamped (the weight is then the constant
        phi(thresh))."""
        if kind == 0:
            return self._interp(0, self._loc, self.mono, uG).ravel()
        (ix, ox), (iy, oy), (iz, oz) = self._loc
        o = (ox, oy, oz)[kind - 1]
        keep = (np.abs(o) < self.mono.thresh).astype(np.float64)
        W = _kern3(kind, ox, oy, oz, self.mono.betaw, self.mono.alpha) / self.mono.norm ** 3
        shape = [(1, 1, 1, -1), (1, 1, -1, 1), (1, -1, 1, 1)][kind - 1]
        W = W * keep.reshape((keep.shape[0],) + shape[1:]) * (self.hx * self.hy * self.hz)
        with np.errstate(invalid="ignore"):
            W = np.nan_to_num(W, nan=0.0, posinf=0.0, neginf=0.0)
        g = self._flat_index(self._loc)
        U = uG.reshape(-1, 3)
        return np.stack([np.sum(U[g, d] * W, axis=(1, 2, 3)) for d in range(3)], 1).ravel()

    def interp_torque_kernel_d2(self, k, b, uG):
        """Interpolate uG with d2/(do_k do_b) of the torque kernel as a
        function of the particle position (zero when o_k was clamped):
        phi (d_k d_b + [k == b] d'_k), d_i = -beta o_i / (alpha^2 s_i),
        d'_i = -beta / (alpha^2 s_i^3), s_i = sqrt(1 - o_i^2 / alpha^2)."""
        ks = self.dip
        (ix, ox), (iy, oy), (iz, oz) = self._tloc
        offs = (ox, oy, oz)
        a2 = ks.alpha * ks.alpha
        shapes = [(1, 1, 1, -1), (1, 1, -1, 1), (1, -1, 1, 1)]
        W = _kern3(0, ox, oy, oz, ks.betaw, ks.alpha) / ks.norm ** 3 * (self.hx * self.hy * self.hz)

        def axis(i, v):
            return v.reshape((v.shape[0],) + shapes[i][1:])

        with np.errstate(invalid="ignore", divide="ignore"):
            s_k = np.sqrt(1 - offs[k] ** 2 / a2)
            s_b = np.sqrt(1 - offs[b] ** 2 / a2)
            d_k = -ks.betaw * offs[k] / (a2 * s_k)
            d_b = -ks.betaw * offs[b] / (a2 * s_b)
            if k == b:
                W = W * axis(k, d_k * d_k - ks.betaw / (a2 * s_k ** 3))
            else:
                W = W * axis(k, d_k) * axis(b, d_b)
            # a clamped o_k is constant (no derivative along k); a clamped
            # o_b keeps its value d_b(thresh) in the forward weight
            W = W * axis(k, (np.abs(offs[k]) < ks.thresh).astype(float))
            W = np.nan_to_num(W, nan=0.0, posinf=0.0, neginf=0.0)
        g = self._flat_index(self._tloc)
        U = uG.reshape(-1, 3)
        return np.stack([np.sum(U[g, d] * W, axis=(1, 2, 3)) for d in range(3)], 1).ravel()

    def mdot(self, forces, torques=None):
        forces = np.asarray(forces, dtype=np.float64)
        tp = self.has_torque and torques is not None and np.size(torques) > 0
        F = self.spread_forces(forces, torques if tp else None)
        uG = self.solve(F, self.tgrid)
        self.uG = uG
        v = self._interp(0, self._loc, self.mono, uG).ravel()
        if not tp:
            return v, None
        if self.reg_kernel:
            return v, self._interp(0, self._tloc, self.dip, self.utG).ravel()
        dxv = self._interp(1, self._tloc, self.dip, uG)
        dyv = self._interp(2, self._tloc, self.dip, uG)
        dzv = self._interp(3, self._tloc, self.dip, uG)
        om = np.empty((self.N, 3))
        om[:, 0] = -0.5 * (dyv[:, 2] - dzv[:, 1])
        om[:, 1] = -0.5 * (dzv[:, 0] - dxv[:, 2])
        om[:, 2] = -0.5 * (dxv[:, 1] - dyv[:, 0])
        return v, om.ravel()
