"""Position updates kept on the device (dps_update_positions, the counterpart
of ParticleList::update, ParticleList.cpp:550-753): a few explicit Euler steps
x <- x + dt * M F on the GPU, positions never leaving the device, against the
same steps taken with the FP64 oracle on the host.  Bound: the FP64 path's
rel-L2 <= 1e-11 per step compounds over the steps, so the positions are held
to 1e-10 of the box size."""
import numpy as np
import pytest

from oracle import dp_oracle as o
from test_gpu_parity import CASES, gpu_for, make_case, oracle_for, say

pytestmark = pytest.mark.gpu


def wrap_like_reference(x, L):
    """ParticleList.cpp:561-580: images counted by truncation (x = L stays L)."""
    x = np.array(x, dtype=np.float64)
    lo = x < 0
    x[lo] += (1 - np.trunc(x[lo] / L)) * L
    hi = x > L
    x[hi] -= np.trunc(x[hi] / L) * L
    return x


@pytest.mark.parametrize("case", [CASES[1], CASES[2]], ids=[CASES[1][0], CASES[2][0]])
def test_euler_steps_on_device_match_oracle(case):
    import torch
    name, geom, Lx, Ly, Lz, nx, ny, nz, N, torque = case
    say("case %s: 4 Euler steps on the device" % name)
    pos, F, T = make_case(case, seed=7)
    # keep clear of the walls for the steps taken
    pos[2::3] = 0.25 * Lz + 0.5 * pos[2::3] / Lz * Lz
    dt = 0.05
    ref = oracle_for(case)
    s = gpu_for(case, "double")
    p_dev = torch.tensor(pos, dtype=torch.float64, device="cuda")
    Fd = torch.tensor(F, dtype=torch.float64, device="cuda")
    Td = torch.tensor(T, dtype=torch.float64, device="cuda") if T is not None else None
    s.setPositions(p_dev)
    p_ref = pos.copy()
    for step in range(4):
        U, W = s.mdot_device(Fd, Td)
        s.update_positions_device(p_dev, U, dt)
        ref.set_positions(p_ref)
        v, w = ref.mdot(F, T)
        p_ref = p_ref + dt * v
        p_ref[0::3] = wrap_like_reference(p_ref[0::3], Lx)
        p_ref[1::3] = wrap_like_reference(p_ref[1::3], Ly)
    got = p_dev.cpu().numpy()
    err = np.abs(got - p_ref).max() / max(Lx, Ly, Lz)
    say("  max |x_gpu - x_oracle| / L = %.3e" % err)
    assert err <= 1e-10
    # one more Mdot at the final positions agrees too
    U, W = s.mdot_device(Fd, Td)
    ref.set_positions(p_ref)
    v, w = ref.mdot(F, T)
    e = float(np.linalg.norm(U.cpu().numpy() - v) / np.linalg.norm(v))
    say("  Mdot at the final positions rel-L2 = %.3e" % e)
    assert e <= 1e-10
    s.clear()


def test_update_wraps_in_place_and_reports_z_range():
    import torch
    case = CASES[2]
    name, geom, Lx, Ly, Lz, nx, ny, nz, N, torque = case
    say("case %s: wrap and range" % name)
    pos, F, T = make_case(case, seed=8)
    s = gpu_for(case, "double")
    pos[0:3] = [-0.25, Ly + 3.0, 0.5 * Lz]
    pos[3:6] = [2 * Lx + 1.0, -2.5 * Ly, 0.5 * Lz]
    pos[6:9] = [Lx, 0.0, 0.5 * Lz]
    p_dev = torch.tensor(pos, dtype=torch.float64, device="cuda")
    s.update_positions_device(p_dev)          # no move: wrap and locate only
    got = p_dev.cpu().numpy()
    want = pos.copy()
    want[0::3] = wrap_like_reference(pos[0::3], Lx)
    want[1::3] = wrap_like_reference(pos[1::3], Ly)
    assert np.array_equal(got, want)
    # a step that carries particle 0 above the slit raises, as the reference exits
    vel = torch.zeros(3 * N, dtype=torch.float64, device="cuda")
    vel[2] = 1.0
    with pytest.raises(ValueError):
        s.update_positions_device(p_dev, vel, 0.75 * Lz)
    s.clear()


def test_step_loop_without_host_sync_and_as_a_graph():
    """dps_update_positions leaves the host unsynchronised, so a time step
    (Mdot, then the update) can be captured in a HIP graph.  Five Euler steps
    run eagerly with check=False; the same step captured once with
    torch.cuda.graph and replayed five times, with eager work and a
    synchronisation between replays, gives each step's positions (FP64;
    spreading adds with float atomics, so to 1e-12 of the box).  An update
    that leaves the z extent is reported by the next check only."""
    import torch
    case = CASES[1]
    name, geom, Lx, Ly, Lz, nx, ny, nz, N, torque = case
    say("case %s: step loop without host sync, and as a graph" % name)
    pos, F, T = make_case(case, seed=9)
    pos[2::3] = 0.25 * Lz + 0.5 * pos[2::3]
    dt = 0.05
    Fd = torch.tensor(F, dtype=torch.float64, device="cuda")
    Td = torch.tensor(T, dtype=torch.float64, device="cuda") if T is not None else None
    s = gpu_for(case, "double")
    p_e = torch.tensor(pos, dtype=torch.float64, device="cuda")
    s.setPositions(p_e)
    traj = []
    for _ in range(5):
        U, _W = s.mdot_device(Fd, Td)
        s.update_positions_device(p_e, U, dt, check=False)
        traj.append(p_e.clone())
    s.check_positions()
    g_s = gpu_for(case, "double")
    p_g = torch.tensor(pos, dtype=torch.float64, device="cuda")
    g_s.setPositions(p_g)
    g_s.mdot_device(Fd, Td)     # loads code objects and FFT plans outside the capture
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        Ug, _Wg = g_s.mdot_device(Fd, Td)
        g_s.update_positions_device(p_g, Ug, dt, check=False)
    worst = 0.0
    for k in range(5):
        graph.replay()
        torch.cuda.synchronize()
        _ = (Fd * 2.0).sum().item()          # eager work between replays
        g_s.check_positions()
        e = float((p_g - traj[k]).abs().max()) / max(Lx, Ly, Lz)
        worst = max(worst, e)
    say("  graph replays vs eager steps: max |dx| / L = %.3e" % worst)
    assert worst <= 1e-12
    # out of range: not reported by the update, reported by the check
    vel = torch.zeros(3 * N, dtype=torch.float64, device="cuda")
    vel[2] = 1.0
    s.update_positions_device(p_e, vel, 2.0 * Lz, check=False)
    with pytest.raises(ValueError):
        s.check_positions()
    s.clear()
    g_s.clear()
