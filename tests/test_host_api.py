"""Host-side logic of the drop-in layer (no device calls)."""
import numpy as np
import pytest

from doublyperiodicstokes_amd import DPStokes, FCMJoint, StokesParameters, getPrecision
from doublyperiodicstokes_amd.gridconfig import (configure_grid_and_kernels_xy,
                                                 configure_grid_and_kernels_z, fft_friendly_sizes)


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def test_grid_configuration_of_the_stored_reference_runs():
    """FCMJoint.Initialize's choice for figures/self_mobility.py (L = 76.8,
    zmax = 19.2, radius 1, kernType 0, torques): the grid under which the
    oracle reproduces the reference's stored mobilities to rounding."""
    Lx, Ly, hx, hy, nx, ny, w, wd, cb, cbd, b, bd = configure_grid_and_kernels_xy(
        76.8, 76.8, np.array([1.0]), np.array([0]), 0, True, False)
    assert (nx, ny, int(w[0]), int(wd[0])) == (135, 135, 6, 6)
    assert hx == pytest.approx(76.8 / 135)
    assert Lx == pytest.approx(76.8)
    Lz, hz, nz, a = configure_grid_and_kernels_z(0.0, 19.2, hx, w, wd, "DPBW", 1.5, False)
    assert nz == 61 and a == 0.0 and Lz == pytest.approx(19.2 + 1.5 * 6 * hx / 2)
    Lz, hz, nz, a = configure_grid_and_kernels_z(0.0, 19.2, hx, w, wd, "DPSC", 1.5, False)
    assert nz == 55 and a == 0.0 and Lz == pytest.approx(19.2)


def test_grid_configuration_rejects_unknown_domains():
    with pytest.raises(ValueError):
        configure_grid_and_kernels_z(0.0, 10.0, 0.5, np.array([6]), np.array([6]), "XX", 1.5, False)


def test_fft_friendly_sizes_are_smooth():
    for n in fft_friendly_sizes(127, 10, 4):
        m = n
        for p in (2, 3, 5, 7):
            while m % p == 0:
                m //= p
        assert m == 1 and n >= 127


def test_stokes_parameters_validation():
    p = StokesParameters(viscosity=1.0, Lx=8, Ly=8, zmin=0, zmax=4, w=6, w_d=6, nx=32, ny=32, nz=17,
                         mode="slit", precision="double")
    c = p.to_c()
    assert (c.nx, c.ny, c.nz, c.mode, c.solver, c.reg_kernel) == (32, 32, 17, 3, 0, 0)
    p.useRegKernel = True
    assert p.to_c().reg_kernel == 1
    assert "mode: slit" in str(p)
    with pytest.raises(ValueError):
        StokesParameters(mode="none").to_c()              # uammd's unknown-mode value
    with pytest.raises(ValueError):
        StokesParameters(mode="slit", precision="half").to_c()
    with pytest.raises(ValueError):
        StokesParameters(mode="slit", solver="fast").to_c()


def test_precision_strings():
    assert getPrecision() in ("single", "double")


def test_joint_interface_argument_errors():
    with pytest.raises(ValueError):
        FCMJoint("cpu")
    s = FCMJoint("hip")
    with pytest.raises(ValueError):
        s.Initialize(4, 1.0, 1.0, "XYZ", False, 20.0, 20.0, 0.0, 10.0)
    with pytest.raises(RuntimeError):
        s.SetPositions(np.zeros(12))


def test_uninitialised_solver_raises():
    s = DPStokes()
    with pytest.raises(RuntimeError):
        s.setPositions(np.zeros(3))
    with pytest.raises(RuntimeError):
        s.Mdot(np.zeros(3))
    s.clear()   # harmless before initialize (the reference crashes here)


@pytest.mark.skipif(gpu_available(), reason="checks the behaviour without a HIP device")
def test_no_device_fails_loudly():
    s = DPStokes()
    p = StokesParameters(viscosity=1.0, Lx=8, Ly=8, zmin=0, zmax=4, w=6, w_d=0, nx=32, ny=32, nz=17,
                         mode="bottom", precision="double")
    with pytest.raises(RuntimeError):
        s.initialize(p, 4)


def test_triply_periodic_configuration():
    """'TP' (uammd mode 'periodic'): uniform z with hz = hx and Lz rounded up
    to whole cells (GridAndKernelConfig.py:393-417)."""
    from doublyperiodicstokes_amd import _native as N
    from doublyperiodicstokes_amd.gridconfig import configure_grid_and_kernels_z
    p = StokesParameters(mode="periodic", Lx=16.0, Ly=16.0, zmin=0.0, zmax=10.0, w=6, nx=32, ny=32, nz=20).to_c()
    assert p.mode == N.DPS_MODE_PERIODIC == 4
    Lz, hz, Nz, a = configure_grid_and_kernels_z(0.0, 10.3, 0.5, np.array([6]), np.array([6]), "TP")
    assert hz == 0.5 and Nz == 21 and Lz == 10.5 and a == 0.0
