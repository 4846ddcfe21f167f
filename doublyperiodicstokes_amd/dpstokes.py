"""The GPU mobility solver behind the joint interface.

``StokesParameters`` and ``DPStokes`` mirror the reference's pybind module
``uammd`` (``source/gpu/python_wrapper/uammd_python.cpp:59-129``) so that a
caller of ``uammd.DPStokes`` can switch to this module:

    par = StokesParameters(viscosity=..., Lx=..., Ly=..., zmin=..., zmax=...,
                           w=6, w_d=6, beta=10.71, beta_d=10.71,
                           nx=256, ny=256, nz=96, mode='slit')
    solver = DPStokes()
    solver.initialize(par, numberParticles)
    solver.setPositions(positions)                       # [x0,y0,z0,x1,...]
    solver.Mdot(forces=F, torques=T, velocities=MF, angularVelocities=MT)
    solver.clear()

Differences from the reference module, all deliberate:

* ``precision`` ('double' or 'mixed') is chosen per solver instead of at
  compile time; ``getPrecision()`` reports the default ('double', or
  'single' when ``DPSTOKES_PRECISION=mixed``).
* ``alpha``/``alpha_d`` are physical support half-widths and default to
  ``w*h/2`` -- the CPU reference's value (``ParticleList.cpp:229``).
* The torque kernel is the derivative ES kernel of the CPU reference
  (``w_d``, ``beta_d``, ``alpha_d``); ``w_d = 0`` disables torques.
* Output arrays of the wrong dtype raise ``TypeError`` instead of being
  silently ignored.
* Besides numpy arrays, every method also accepts device-resident torch
  tensors; ``mdot_device`` returns device tensors without host copies.

All arithmetic happens in the HIP library; if it is not built this module
raises on use.
"""
import ctypes
import json
import os
import tempfile

import numpy as np

from . import _native as N

_MODES = {"nowall": N.DPS_MODE_NOWALL, "bottom": N.DPS_MODE_BOTTOM, "slit": N.DPS_MODE_SLIT,
          "periodic": N.DPS_MODE_PERIODIC}
_SOLVERS = {"auto": N.DPS_SOLVER_AUTO, "lane": N.DPS_SOLVER_LANE}


def getPrecision():
    """Default grid precision of new solvers: 'double' or 'single' (mixed)."""
    return "single" if os.environ.get("DPSTOKES_PRECISION", "double") == "mixed" else "double"


class StokesParameters:
    """Parameters of a doubly (or triply) periodic Stokes solve (uammd.StokesParameters).

    ``beta``/``beta_d`` are the ES exponents beta*w; ``<= 0`` selects 1.785*w.
    ``mode`` is 'nowall', 'bottom' or 'slit'.  ``solver`` picks the spectral
    solve: 'auto' (fused lane-group kernel when nz <= 512) or 'lane' (one
    wavenumber per lane, separate z-transform kernels).  ``useRegKernel``
    selects the CPU reference's regular-kernel torque path (FCM.py:149,
    372-396, 427-439): torques spread with the plain ES kernel of the dipole
    width and curls taken spectrally (it runs the 'lane' kernels).
    """

    def __init__(self, viscosity=1.0, Lx=0.0, Ly=0.0, zmin=0.0, zmax=0.0, w=1, w_d=1,
                 alpha=-1.0, alpha_d=-1.0, beta=-1.0, beta_d=-1.0, nx=-1, ny=-1, nz=-1,
                 mode="none", precision=None, solver="auto", useRegKernel=False):
        self.viscosity = viscosity
        self.Lx, self.Ly = Lx, Ly
        self.zmin, self.zmax = zmin, zmax
        self.w, self.w_d = w, w_d
        self.alpha, self.alpha_d = alpha, alpha_d
        self.beta, self.beta_d = beta, beta_d
        self.nx, self.ny, self.nz = nx, ny, nz
        self.mode = mode
        self.precision = precision if precision is not None else (
            "mixed" if getPrecision() == "single" else "double")
        self.solver = solver
        self.useRegKernel = bool(useRegKernel)

    def __str__(self):
        return ("viscosity: %g\nLx: %g\nLy: %g\nzmin: %g\nzmax: %g\nw: %g\nw_d: %g\nalpha: %g\n"
                "alpha_d: %g\nbeta: %g\nbeta_d: %g\nnx: %d\nny: %d\nnz: %d\nmode: %s\nprecision: %s\n"
                % (self.viscosity, self.Lx, self.Ly, self.zmin, self.zmax, self.w, self.w_d,
                   self.alpha, self.alpha_d, self.beta, self.beta_d, self.nx, self.ny, self.nz,
                   self.mode, self.precision))

    def to_c(self):
        if self.mode not in _MODES:
            raise ValueError("mode must be one of 'nowall', 'bottom', 'slit', 'periodic' (got %r)"
                             % (self.mode,))
        if self.precision not in ("double", "mixed"):
            raise ValueError("precision must be 'double' or 'mixed'")
        if self.solver not in _SOLVERS:
            raise ValueError("solver must be 'auto' or 'lane'")
        p = N.DpsParams()
        p.viscosity = float(self.viscosity)
        p.Lx, p.Ly = float(self.Lx), float(self.Ly)
        p.zmin, p.zmax = float(self.zmin), float(self.zmax)
        p.w, p.w_d = int(self.w), int(self.w_d)
        p.alpha, p.alpha_d = float(self.alpha), float(self.alpha_d)
        p.beta, p.beta_d = float(self.beta), float(self.beta_d)
        p.nx, p.ny, p.nz = int(self.nx), int(self.ny), int(self.nz)
        p.solver = _SOLVERS[self.solver]
        p.reg_kernel = 1 if self.useRegKernel else 0
        p.mode = _MODES[self.mode]
        p.precision = N.DPS_PREC_MIXED if self.precision == "mixed" else N.DPS_PREC_DOUBLE
        return p


def _torch():
    import torch
    return torch


def _stream():
    torch = _torch()
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _dtype_code(dt):
    torch = _torch()
    if dt in (np.float32, torch.float32):
        return N.DPS_F32
    if dt in (np.float64, torch.float64):
        return N.DPS_F64
    raise TypeError("arrays must be float32 or float64 (got %s)" % (dt,))


class DPStokes:
    """Doubly periodic Stokes mobility on one MI355X (uammd.DPStokes)."""

    def __init__(self):
        self._h = None
        self.N = 0
        self.par = None

    # -- lifetime -------------------------------------------------------------
    def initialize(self, par, numberParticles):
        """Create (or re-create) the solver for ``numberParticles`` particles."""
        L = N.load()
        torch = _torch()
        if not torch.cuda.is_available():
            raise RuntimeError("DPStokes needs a HIP device")
        self.clear()
        h = ctypes.c_void_p()
        cpar = par.to_c()
        N.check(L.dps_create(ctypes.byref(cpar), int(numberParticles), ctypes.byref(h)))
        self._h = h
        self.N = int(numberParticles)
        self.par = par
        self.has_torque = int(par.w_d) > 0
        self.dtype = torch.float32 if par.precision == "mixed" else torch.float64
        self.device = torch.device("cuda", torch.cuda.current_device())

    def clear(self):
        """Release device memory; the solver must be initialised again to be used."""
        if self._h is not None:
            N.load().dps_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.clear()
        except Exception:
            pass

    def _need(self):
        if self._h is None:
            raise RuntimeError("DPStokes is not initialized. Call Initialize first")

    # -- data movement --------------------------------------------------------
    def _to_device(self, a, name, size):
        torch = _torch()
        if isinstance(a, torch.Tensor):
            t = a.detach()
            if t.device.type != "cuda":
                t = t.to(self.device)
        else:
            arr = np.asarray(a)
            if arr.dtype not in (np.float32, np.float64):
                arr = arr.astype(np.float64)
            t = torch.from_numpy(np.ascontiguousarray(arr)).to(self.device)
        t = t.reshape(-1).contiguous()
        if t.numel() != size:
            raise ValueError("%s must hold %d values (got %d)" % (name, size, t.numel()))
        if t.dtype not in (torch.float32, torch.float64):
            t = t.to(torch.float64)
        return t

    def setPositions(self, positions):
        """Positions [x0,y0,z0,...]; x, y wrap periodically, z must be in [zmin, zmax]."""
        self._need()
        p = self._to_device(positions, "positions", 3 * self.N)
        N.check(N.load().dps_set_positions(self._h, ctypes.c_void_p(p.data_ptr()),
                                           _dtype_code(p.dtype), _stream()))

    def update_positions_device(self, positions, velocities=None, dt=0.0, check=True):
        """One position update kept on the device (ParticleList::update,
        ParticleList.cpp:550-753): ``positions += dt * velocities`` in place on a
        device tensor, x and y rewrapped into the unit cell, then relocated as
        by setPositions.  ``velocities`` must have the positions' dtype.  With
        ``check`` (the default) the stream is synchronised and a particle
        that left the z extent raises ValueError, as the reference exits;
        ``check=False`` leaves the host unsynchronised (step loops, graph
        capture) and ``check_positions()`` reports later."""
        self._need()
        torch = _torch()
        if not (isinstance(positions, torch.Tensor) and positions.is_cuda and positions.is_contiguous()):
            raise TypeError("positions must be a contiguous device tensor")
        if positions.numel() != 3 * self.N or positions.dtype not in (torch.float32, torch.float64):
            raise ValueError("positions must hold 3N float32 or float64 values")
        v = None
        if velocities is not None:
            if not (isinstance(velocities, torch.Tensor) and velocities.is_cuda and velocities.is_contiguous()
                    and velocities.dtype == positions.dtype and velocities.numel() == 3 * self.N):
                raise TypeError("velocities must be a contiguous device tensor like positions")
            v = ctypes.c_void_p(velocities.data_ptr())
        N.check(N.load().dps_update_positions(self._h, ctypes.c_void_p(positions.data_ptr()), v,
                                              float(dt), _dtype_code(positions.dtype), _stream()))
        if check:
            self.check_positions()

    def check_positions(self):
        """Synchronise and raise ValueError if a position update since the last
        check moved a particle out of the z extent (dps_check_positions)."""
        self._need()
        N.check(N.load().dps_check_positions(self._h, _stream()))

    def mdot_device(self, forces, torques=None, out_dtype=None):
        """Mdot on device tensors; returns (velocities, angular velocities or None)."""
        self._need()
        torch = _torch()
        F = self._to_device(forces, "forces", 3 * self.N)
        dt = out_dtype if out_dtype is not None else F.dtype
        F = F.to(dt)
        T = None
        if torques is not None and self.has_torque:
            T = self._to_device(torques, "torques", 3 * self.N)
            if T.numel() == 0:
                T = None
            else:
                T = T.to(dt)
        U = torch.empty(3 * self.N, dtype=dt, device=self.device)
        W = torch.empty(3 * self.N, dtype=dt, device=self.device) if T is not None else None
        N.check(N.load().dps_mdot(self._h, ctypes.c_void_p(F.data_ptr()),
                                  ctypes.c_void_p(T.data_ptr()) if T is not None else None,
                                  ctypes.c_void_p(U.data_ptr()),
                                  ctypes.c_void_p(W.data_ptr()) if W is not None else None,
                                  _dtype_code(dt), _stream()))
        return U, W

    def mdot_adjoint_device(self, forces, grad_velocities, torques=None, grad_angular=None, positions=True):
        """Adjoint of Mdot at the current positions (dps_mdot_adjoint): returns
        (dL/dF, dL/dT or None, dL/dx or None) for L = <gU, U> + <gW, W>.
        dL/dT comes when torques or grad_angular are given; position
        gradients with torques need the derivative torque kernels (not
        useRegKernel)."""
        self._need()
        torch = _torch()
        gU = self._to_device(grad_velocities, "grad_velocities", 3 * self.N)
        dt = gU.dtype
        F = self._to_device(forces, "forces", 3 * self.N).to(dt)
        T = gW = None
        if torques is not None and self.has_torque:
            T = self._to_device(torques, "torques", 3 * self.N).to(dt)
        if grad_angular is not None and self.has_torque:
            gW = self._to_device(grad_angular, "grad_angular", 3 * self.N).to(dt)
        if not positions and T is not None and gW is None:
            gW = torch.zeros(3 * self.N, dtype=dt, device=self.device)
        gF = torch.empty(3 * self.N, dtype=dt, device=self.device)
        gT = torch.empty(3 * self.N, dtype=dt, device=self.device) if (gW is not None or T is not None) else None
        gX = torch.empty(3 * self.N, dtype=dt, device=self.device) if positions else None
        vp = ctypes.c_void_p
        ptr = lambda t: vp(t.data_ptr()) if t is not None else None
        N.check(N.load().dps_mdot_adjoint(self._h, ptr(F), ptr(T), ptr(gU), ptr(gW), ptr(gF), ptr(gT), ptr(gX),
                                          _dtype_code(dt), _stream()))
        return gF, gT, gX

    def Mdot(self, forces, torques=None, velocities=None, angularVelocities=None):
        """Velocities (and angular velocities when torques are given).

        As in the reference, ``velocities``/``angularVelocities`` are output
        arrays written in place; if they are omitted new numpy arrays are
        returned.  Torques of size < 3N are treated as absent.
        """
        self._need()
        tq = None
        if torques is not None and np.size(torques) >= 3 * self.N and self.N > 0:
            tq = torques
        out_np = np.float64 if self.par.precision == "double" else np.float32
        if isinstance(velocities, np.ndarray):
            out_np = velocities.dtype.type
        torch = _torch()
        dt = torch.float32 if out_np == np.float32 else torch.float64
        U, W = self.mdot_device(forces, tq, out_dtype=dt)
        Uh = U.cpu().numpy()
        Wh = W.cpu().numpy() if W is not None else np.zeros(0, dtype=out_np)
        if velocities is not None:
            if not isinstance(velocities, np.ndarray) or velocities.dtype != Uh.dtype:
                raise TypeError("velocities must be a numpy array of dtype %s" % Uh.dtype)
            velocities.reshape(-1)[:] = Uh
        else:
            velocities = Uh
        if W is not None:
            if angularVelocities is not None:
                if not isinstance(angularVelocities, np.ndarray) or angularVelocities.dtype != Wh.dtype:
                    raise TypeError("angularVelocities must be a numpy array of dtype %s" % Wh.dtype)
                angularVelocities.reshape(-1)[:] = Wh
            else:
                angularVelocities = Wh
        elif angularVelocities is None:
            angularVelocities = Wh
        return velocities, angularVelocities

    # -- stages (parity testing, adjoint building blocks) --------------------
    def info(self):
        self._need()
        out = (ctypes.c_int * 8)()
        N.check(N.load().dps_get_info(self._h, out))
        keys = ("nx", "ny", "nz", "N", "nzeff", "nzeff_d", "wzmax", "wzmax_d")
        return dict(zip(keys, list(out)))

    def stage_spread(self, forces, torques=None):
        """Spread forces (+ torque curl); returns the grid (Ny, Nx, Nz, 3) FP64 on device."""
        self._need()
        torch = _torch()
        i = self.info()
        F = self._to_device(forces, "forces", 3 * self.N)
        T = None
        if torques is not None and self.has_torque:
            T = self._to_device(torques, "torques", 3 * self.N).to(F.dtype)
        G = torch.empty((i["ny"], i["nx"], i["nz"], 3), dtype=torch.float64, device=self.device)
        N.check(N.load().dps_stage_spread(self._h, ctypes.c_void_p(F.data_ptr()),
                                          ctypes.c_void_p(T.data_ptr()) if T is not None else None,
                                          _dtype_code(F.dtype), ctypes.c_void_p(G.data_ptr()), _stream()))
        return G

    def stage_solve(self, grid):
        """Velocity grid from a force grid, both (Ny, Nx, Nz, 3) FP64."""
        self._need()
        torch = _torch()
        G = grid.to(self.device, torch.float64).contiguous()
        out = torch.empty_like(G)
        N.check(N.load().dps_stage_solve(self._h, ctypes.c_void_p(G.data_ptr()),
                                         ctypes.c_void_p(out.data_ptr()), _stream()))
        return out

    def stage_interp(self, grid, with_angular=False, dtype=None):
        """Interpolate a (Ny, Nx, Nz, 3) FP64 velocity grid to the particles."""
        self._need()
        torch = _torch()
        G = grid.to(self.device, torch.float64).contiguous()
        dt = dtype or torch.float64
        U = torch.empty(3 * self.N, dtype=dt, device=self.device)
        W = torch.empty(3 * self.N, dtype=dt, device=self.device) if with_angular and self.has_torque else None
        N.check(N.load().dps_stage_interp(self._h, ctypes.c_void_p(G.data_ptr()), ctypes.c_void_p(U.data_ptr()),
                                          ctypes.c_void_p(W.data_ptr()) if W is not None else None,
                                          _dtype_code(dt), _stream()))
        return U, W

    def set_profiling(self, enable=True):
        self._need()
        N.check(N.load().dps_set_profiling(self._h, int(bool(enable))))

    def timings(self):
        """Per-stage microseconds of the last profiled Mdot (HIP events)."""
        self._need()
        out = (ctypes.c_double * 8)()
        N.check(N.load().dps_get_timings(self._h, out))
        keys = ("locate", "spread", "fft_fwd", "z_fwd", "bvp", "z_inv", "fft_inv", "interp")
        return dict(zip(keys, list(out)))


def tune_key(solver, torques, dtype):
    """Cache key of one batch instance: device architecture, every solver
    parameter, particle count, whether torques are spread, input dtype."""
    torch = _torch()
    try:
        props = torch.cuda.get_device_properties(solver.device)
        arch = getattr(props, "gcnArchName", "") or props.name
    except Exception:
        arch = "unknown"
    par = {k: v for k, v in sorted(vars(solver.par).items())}
    return json.dumps({"arch": arch, "par": par, "N": int(solver.N), "torques": bool(torques),
                       "dtype": str(dtype)}, sort_keys=True, default=str)


def load_tune_cache(path):
    """Stored tuning choices {key: chunk}; empty when the file is absent."""
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        data = json.load(f)
    if not isinstance(data, dict) or data.get("version") != 1 or not isinstance(data.get("chunks"), dict):
        raise ValueError("%s is not a tuning cache of this library" % path)
    return dict(data["chunks"])


def save_tune_cache(path, chunks):
    """Write {key: chunk} atomically (a temporary file renamed over ``path``)."""
    d = os.path.dirname(os.path.abspath(path))
    fd, tmp = tempfile.mkstemp(prefix=".tune-", dir=d)
    try:
        with os.fdopen(fd, "w") as f:
            json.dump({"version": 1, "chunks": {k: int(v) for k, v in chunks.items()}}, f, indent=1,
                      sort_keys=True)
        os.replace(tmp, path)
    except BaseException:
        if os.path.exists(tmp):
            os.unlink(tmp)
        raise


class DPStokesBatch:
    """Many independent solvers run as one (dps_batch_*): every instance's
    Mdot captured into one HIP graph, forked over ``nstreams`` streams, and
    replayed while the same input tensors are passed.  The solvers stay
    owned by the caller and must outlive the batch."""

    def __init__(self, solvers, nstreams=4):
        for s in solvers:
            s._need()
        self.solvers = list(solvers)
        arr = (ctypes.c_void_p * len(self.solvers))(*[s._h.value for s in self.solvers])
        b = ctypes.c_void_p()
        N.check(N.load().dps_batch_create(arr, len(self.solvers), int(nstreams), ctypes.byref(b)))
        self._b = b
        self._out = {}

    def clear(self):
        if getattr(self, "_b", None) is not None and self._b.value:
            N.load().dps_batch_destroy(self._b)
        self._b = None

    def __del__(self):
        try:
            self.clear()
        except Exception:
            pass

    def _inputs(self, forces, torques):
        if len(forces) != len(self.solvers):
            raise ValueError("need one force array per instance")
        F = [s._to_device(f, "forces", 3 * s.N) for s, f in zip(self.solvers, forces)]
        dt = F[0].dtype
        F = [f.to(dt) for f in F]
        T = [None] * len(F)
        if torques is not None:
            if len(torques) != len(self.solvers):
                raise ValueError("need one torque array (or None) per instance")
            T = [s._to_device(t, "torques", 3 * s.N).to(dt) if (t is not None and s.has_torque) else None
                 for s, t in zip(self.solvers, torques)]
        return F, T, dt

    @staticmethod
    def _ptrs(ts):
        return (ctypes.c_void_p * len(ts))(*[t.data_ptr() if t is not None else None for t in ts])

    def tune(self, forces, torques=None, cache=None):
        """Autotune the spreading z chunk of every instance; returns the chunks.

        With ``cache`` (path of a JSON file), instances whose configuration
        the file holds take their stored chunk without timing, provided every
        instance of the batch is found there; otherwise all are timed and the
        file is updated with the choices.  A choice is measured on the
        positions set when it was tuned.  A file that is not a tuning cache
        of this library raises ValueError and is left as it is."""
        F, T, dt = self._inputs(forces, torques)
        keys = [tune_key(s, t is not None, dt) for s, t in zip(self.solvers, T)]
        stored = load_tune_cache(cache) if cache else {}
        if cache and all(k in stored for k in keys):
            chunks = (ctypes.c_int * len(F))(*[int(stored[k]) for k in keys])
            N.check(N.load().dps_batch_set_chunks(self._b, chunks, int(any(t is not None for t in T)), _stream()))
            return list(chunks)
        out = (ctypes.c_int * len(F))()
        N.check(N.load().dps_batch_tune(self._b, self._ptrs(F), self._ptrs(T), _dtype_code(dt), _stream(), out))
        if cache:
            stored.update(zip(keys, list(out)))
            save_tune_cache(cache, stored)
        return list(out)

    def mdot_device(self, forces, torques=None):
        """Velocities (and angular velocities where torques are given) of every
        instance, as lists of device tensors owned by the batch (overwritten by
        the next call)."""
        torch = _torch()
        F, T, dt = self._inputs(forces, torques)
        key = (dt, tuple(t is not None for t in T))
        if key not in self._out:
            dev = self.solvers[0].device
            self._out[key] = ([torch.empty(3 * s.N, dtype=dt, device=dev) for s in self.solvers],
                              [torch.empty(3 * s.N, dtype=dt, device=dev) if t is not None else None
                               for s, t in zip(self.solvers, T)])
        U, W = self._out[key]
        N.check(N.load().dps_batch_mdot(self._b, self._ptrs(F), self._ptrs(T), self._ptrs(U), self._ptrs(W),
                                        _dtype_code(dt), _stream()))
        return U, W
