"""Doubly periodic Stokes mobility (Mdot) on AMD MI355X.

Drop-in backend for the joint interface of stochasticHydroTools/
DoublyPeriodicStokes: ``FCMJoint('hip')`` and the ``uammd``-compatible
``StokesParameters``/``DPStokes`` pair, backed by hand-written gfx950 HIP
kernels in ``lib/libdpstokes_hip.so``.
"""
from .dpstokes import DPStokes, DPStokesBatch, StokesParameters, getPrecision
from .fcm_joint import FCMJoint
from .gridconfig import configure_grid_and_kernels_xy, configure_grid_and_kernels_z

__all__ = ["DPStokes", "DPStokesBatch", "StokesParameters", "getPrecision", "FCMJoint",
           "configure_grid_and_kernels_xy", "configure_grid_and_kernels_z"]
