"""Joint interface of the reference (``python_interface/common_interface_wrapper.py:21-133``)
with the MI355X solver as its backend.

``FCMJoint('hip')`` (alias ``'gpu'``) keeps the reference signature:

    solver = FCMJoint('hip')
    solver.Initialize(numberParticles, hydrodynamicRadius, viscosity, domType,
                      has_torque, xmax, ymax, zmin, zmax, xmin=0, ymin=0,
                      kernType=0, optInd=0, ref=False)
    solver.SetPositions(positions)          # [x0,y0,z0,...], FP64 numpy
    MF, MT = solver.Mdot(forces, torques)   # MT is empty without torques
    solver.Clean()

Grid and kernels come from the same auto-configuration as the reference CPU
backend (``gridconfig``), and the kernel support is the CPU's physical
``alpha = R_h / (2 c(beta))`` (``FCM.py:296-310``).  The z domain is the
configured one, ``[a, a + Lz]`` with ``a = zmin - msep_down``, as the CPU path
uses (``FCM.py:191-232``).  Triply periodic ``'TP'`` uses a uniform periodic z
grid with ``hz = hx`` and ``Lz`` rounded up to whole cells
(``GridAndKernelConfig.py:393-417``).  The reference's import-time guard on
``FCM_CPUCONFIG_LAUNCHED`` belongs to its CPU thread pinning and is not
needed here.
"""
import numpy as np

from .dpstokes import DPStokes, StokesParameters
from .gridconfig import configure_grid_and_kernels_xy, configure_grid_and_kernels_z

_DOM = {"DP": "nowall", "DPBW": "bottom", "DPSC": "slit", "TP": "periodic"}


class FCMJoint:
    """Common interface for the doubly and triply periodic FCM solvers."""

    def __init__(self, device="hip", precision="double"):
        if device not in ("hip", "gpu"):
            raise ValueError("this backend provides device 'hip' (alias 'gpu')")
        self.device = device
        self.precision = np.float64
        self.grid_precision = precision
        self.created = False
        self.solver = None

    def Initialize(self, numberParticles, hydrodynamicRadius, viscosity, domType, has_torque,
                   xmax, ymax, zmin, zmax, xmin=0, ymin=0, kernType=0, optInd=0, ref=False,
                   verbose=False, useRegKernel=False):
        """Configure grid and kernels and create the solver (re-callable).
        ``useRegKernel`` is the keyword python_interface/dpstokes_common.py:19-31
        passes: torques through the plain ES kernel and spectral curls
        (FCM.py:149, 372-396, 427-439) instead of the derivative kernels."""
        self.Clean()
        if domType not in _DOM:
            raise ValueError("Domain type is not supported")
        self.has_torque = bool(has_torque)
        self.viscosity = viscosity
        self.radP = hydrodynamicRadius * np.ones(numberParticles)
        self.kernTypes = kernType * np.ones(numberParticles, dtype=int)
        self.domType = domType
        Lx, Ly, hx, hy, nx, ny, w, w_d, cbeta, cbeta_d, beta, beta_d = configure_grid_and_kernels_xy(
            xmax - xmin, ymax - ymin, self.radP, self.kernTypes, optInd, self.has_torque, ref,
            verbose=verbose)
        Lz, hz, nz, a = configure_grid_and_kernels_z(zmin, zmax, hx, w, w_d, domType, fac=1.5, ref=ref)
        self.xmin, self.ymin = xmin, ymin
        self.zmin, self.zmax = zmin, zmax
        self.Lx, self.Ly = Lx, Ly
        par = StokesParameters(
            viscosity=viscosity, Lx=Lx, Ly=Ly, zmin=a, zmax=a + Lz,
            w=int(w[0]), w_d=int(w_d[0]) if self.has_torque else 0,
            beta=float(beta[0] * w[0]),
            beta_d=float(beta_d[0] * w_d[0]) if self.has_torque else -1.0,
            alpha=float(self.radP[0] / (2 * cbeta[0])),
            alpha_d=float(self.radP[0] / (2 * cbeta_d[0])) if self.has_torque else -1.0,
            nx=nx, ny=ny, nz=nz, mode=_DOM[domType], precision=self.grid_precision,
            useRegKernel=useRegKernel)
        self.par = par
        self.solver = DPStokes()
        self.solver.initialize(par, numberParticles)
        self.created = True

    def Clean(self):
        """Release all memory allocated by the module."""
        if self.created and self.solver is not None:
            self.solver.clear()
        self.created = False

    def SetPositions(self, positions):
        """Set the positions used by the next Mdot calls."""
        if not self.created:
            raise RuntimeError("DPStokes is not initialized. Call Initialize first")
        p = np.array(positions, dtype=np.float64).reshape(-1).copy()
        if self.xmin:
            p[0::3] -= self.xmin
        if self.ymin:
            p[1::3] -= self.ymin
        self.solver.setPositions(p)

    def Mdot(self, forces, torques=np.empty(0)):
        """Linear (and, with torques, angular) velocities of the particles."""
        if not self.has_torque:
            torques = np.empty(0)
        MF = np.zeros(np.size(forces), self.precision)
        MT = np.zeros(np.size(torques), self.precision)
        tq = torques if np.size(torques) > 0 else None
        self.solver.Mdot(forces=np.asarray(forces, dtype=np.float64),
                         torques=None if tq is None else np.asarray(tq, dtype=np.float64),
                         velocities=MF, angularVelocities=MT if tq is not None else None)
        return MF, MT
