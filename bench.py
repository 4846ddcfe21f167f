#!/usr/bin/env python3
"""Benchmark: doubly periodic Stokes Mdot on one MI355X.

Workload (default): slit channel, walls at z=0 and z=32, N=16384 particles,
Lx=Ly=128, x,y uniform, z uniform in [1.5, 30.5], grid 256x256x96, ES kernels
w=w_d=6 with beta*w=10.71 for forces and torques, forces and torques iid
N(0,1) (numpy RandomState(0), draw order x, y, z, F, T).  A step is one Mdot
(forces + torques -> velocities + angular velocities) with inputs resident
in HBM, mixed precision (FP32 grids/FFT/z-transform, FP64 BVP).

Prints one JSON line.  Besides throughput it reports the relative L2
difference of the mixed path against the FP64 GPU path on the same inputs,
per-stage times from HIP events on the solver's stream, the roofline line of
the dominant kernel, and an FP64 CPU baseline (the oracle restatement of the
reference CPU algorithm) timed on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
# MI355X FP32 peak (dense matrix cores; the vector ALUs reach it only with
# packed FP32, which this build does not emit: half of it without)
FP32_PEAK_TFLOPS = 157.3
HERE = os.path.dirname(os.path.abspath(__file__))
# kernel that runs each stage (the "bvp" stage is the fused spectral kernel)
STAGE_KERNEL = {"spread": "k_spread", "fft_fwd": "k_fft_x_fwd + k_fft_y", "z_fwd": "k_spectral",
                "bvp": "k_spectral", "z_inv": "k_spectral", "fft_inv": "k_fft_y + k_fft_x_inv", "interp": "k_interp"}
# unfused spectral chain (one wavenumber per lane)
STAGE_KERNEL_LANE = dict(STAGE_KERNEL, z_fwd="k_zfwd", bvp="k_bvp", z_inv="k_zinv")

WORKLOADS = {
    # name: (geometry, N, Lx, Ly, zmax, zlo, zhi, nx, ny, nz, torque)
    "slit16k": ("slit", 16384, 128.0, 128.0, 32.0, 1.5, 30.5, 256, 256, 96, True),
    "bottom2k": ("bottom", 2048, 64.0, 64.0, 24.0, 1.5, 20.0, 128, 128, 64, False),
    # differentiation workload: the top of the z grid is 24 + 1.5 w h / 2
    "bottom8k": ("bottom", 8192, 96.0, 96.0, 26.25, 1.5, 24.0, 192, 192, 80, True),
    # sedimented monolayer: z = 1.1 + |N(0, 0.3)|, h = 1 (alpha = 3), Lz = 8
    "monolayer256k": ("bottom", 262144, 1024.0, 1024.0, 8.0, None, None, 1024, 1024, 48, True),
}


def make_inputs(wl, seed=0):
    geom, N, Lx, Ly, zmax, zlo, zhi, nx, ny, nz, torque = wl
    rng = np.random.RandomState(seed)
    x = rng.uniform(0, Lx, N)
    y = rng.uniform(0, Ly, N)
    if zlo is None:
        # monolayer: z = 1.1 + |N(0, 0.3)|; forces (N(0,.1), N(0,.1), -1 + N(0,.1))
        z = 1.1 + np.abs(rng.normal(0, 0.3, N))
        F = rng.normal(0, 0.1, 3 * N)
        F[2::3] -= 1.0
    else:
        z = rng.uniform(zlo, zhi, N)
        F = rng.normal(0, 1, 3 * N)
    T = rng.normal(0, 1, 3 * N) if torque else None
    return np.stack([x, y, z], 1).ravel(), F, T


def stage_bytes(wl, s, reg=False):
    """Algorithmic HBM bytes of each stage (s = bytes per grid real).  The
    regular torque kernel doubles the grids and transforms, adds a curl to the
    z-forward stage (read 2 spectra, write 1) and a z-forward, a curl and a
    second inverse to the z-inverse stage."""
    geom, N, Lx, Ly, zmax, zlo, zhi, nx, ny, nz, torque = wl
    ng = 3 * nx * ny * nz
    nm = 3 * nz * ny * (nx // 2 + 1)
    npart = N * (6 if torque else 3)
    g = 2 if reg else 1
    spec = nm * 2 * s
    return {
        "spread": g * ng * s + npart * s,
        "fft_fwd": g * (ng * s + spec),
        "z_fwd": g * 2 * spec + (3 * spec if reg else 0),
        "bvp": 2 * spec,
        "z_inv": 2 * spec + (3 * 2 * spec + 2 * spec if reg else 0),
        "fft_inv": g * (spec + ng * s),
        "interp": g * ng * s + npart * s,
    }


def z_transform_flop(wl):
    """FLOPs of the forward and inverse Chebyshev z transforms of one Mdot:
    per mode, component and real part, Nz coefficients each a sum over the
    ceil(Nz/2) folded node pairs (forward), and Nz node values from the
    ceil(Nz/2) even/odd coefficient pairs (inverse); one FMA = 2 FLOPs."""
    geom, N, Lx, Ly, zmax, zlo, zhi, nx, ny, nz, torque = wl
    nm = ny * (nx // 2 + 1)
    nh = (nz + 1) // 2
    return 2 * (2 * nz * nh * 2 * 3 * nm)


def gpu_solver(wl, precision, reg=False, solver="auto"):
    from doublyperiodicstokes_amd import DPStokes, StokesParameters
    geom, N, Lx, Ly, zmax, zlo, zhi, nx, ny, nz, torque = wl
    par = StokesParameters(viscosity=1 / (6 * np.pi), Lx=Lx, Ly=Ly, zmin=0.0, zmax=zmax,
                           w=6, w_d=6 if torque else 0, beta=1.785 * 6, beta_d=1.785 * 6,
                           nx=nx, ny=ny, nz=nz, mode=geom, precision=precision, useRegKernel=reg,
                           solver=solver)
    s = DPStokes()
    s.initialize(par, N)
    return s


def time_steps(torch, fn, steps, warmup):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1000.0 / steps


def measured_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE, each scaled by the counter's reading for a copy
    kernel of known size), or None when no summary covers this kernel."""
    path = os.path.join(HERE, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        d = json.load(f)
    e = d.get(workload, {}).get(kernel)
    if not e:
        return None, None
    return e["bytes_per_launch"], e["source"]


def copy_probe(torch, mb):
    """A plain device copy of known size (reads and writes `mb` MiB), run so
    that a profiler can calibrate its byte counters against it."""
    n = mb * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    del src, dst


def cpu_baseline(wl, pos, F, T, budget_s, reg=False):
    """FP64 oracle Mdot timed on the host, as the reference's timing harness
    does it (source/cpu/examples/mobility_bench.py:312-442): one untimed
    warm-up call, then timed calls until `budget_s` seconds or 20 calls, with
    the harness's phases (SI force, SI torque, FCT+iFCT, Solve).  Returns the
    rate, the call count, the timed seconds, the thread count, the mean
    seconds per phase and the oracle's (U, W) of the last call."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle import dp_oracle as o
    geom, N, Lx, Ly, zmax, zlo, zhi, nx, ny, nz, torque = wl
    np.seterr(all="ignore")
    ref = o.OracleDPStokes(geom, Lx, Ly, 0.0, zmax, nx, ny, nz, 1 / (6 * np.pi), 6, 1.785 * 6,
                           has_torque=torque, w_d=6 if torque else None,
                           betaw_d=1.785 * 6 if torque else None, reg_kernel=reg)
    ref.set_positions(pos)
    call = (lambda: ref.mdot(F, T) + ({},)) if reg else (lambda: ref.mdot_phases(F, T))
    call()
    n = 0
    phases = {}
    t0 = time.perf_counter()
    while True:
        U, W, ph = call()
        for k, v in ph.items():
            phases[k] = phases.get(k, 0.0) + v
        n += 1
        el = time.perf_counter() - t0
        if el > budget_s or n >= 20:
            break
    return n / el, n, el, o.num_threads(), {k: v / n for k, v in phases.items()}, U, W


def batch_workload(n_inst=64, seed=0):
    """64 independent bottom-wall suspensions: N log-uniform in [256, 32768],
    Lx = Ly for area fraction 0.3 (a = 1), box height log-uniform in [12, 48],
    h = 0.5 (w = 6, alpha = 1.5); Nx = Ny the smallest 2^k or 3*2^k with
    Nx h >= L (then L = Nx h); the z grid from the reference's bottom-wall rule
    for the height; x, y uniform, z uniform in [1.5, 0.8 height]; forces and
    torques iid N(0, 1).  numpy RandomState(seed), draw order: all N, all
    heights, then per instance x, y, z, F, T."""
    from doublyperiodicstokes_amd.gridconfig import configure_grid_and_kernels_z
    rng = np.random.RandomState(seed)
    Ns = np.round(np.exp(rng.uniform(np.log(256), np.log(32768), n_inst))).astype(int)
    heights = np.exp(rng.uniform(np.log(12), np.log(48), n_inst))
    h = 0.5
    out = []
    for N, H in zip(Ns, heights):
        L = np.sqrt(N * np.pi / 0.3)
        n = min(c for c in [2 ** k * m for k in range(4, 12) for m in (2, 3)] if c * h >= L)
        L = n * h
        Lz, hz, nz, a = configure_grid_and_kernels_z(0.0, H, h, np.array([6]), np.array([6]), "DPBW")
        x = rng.uniform(0, L, N)
        y = rng.uniform(0, L, N)
        z = rng.uniform(1.5, 0.8 * H, N)
        F = rng.normal(size=3 * N)
        T = rng.normal(size=3 * N)
        out.append(dict(N=int(N), L=float(L), zmax=float(a + Lz), zmin=float(a), n=int(n), nz=int(nz),
                        pos=np.stack([x, y, z], 1).ravel(), F=F, T=T))
    return out


def bench_batch(args):
    """Batched Mdot of the 64-instance workload: one captured graph per step."""
    import torch
    from doublyperiodicstokes_amd import DPStokes, DPStokesBatch, StokesParameters
    insts = batch_workload()
    dev = torch.device("cuda", 0)
    solvers, F, T = [], [], []
    t0 = time.perf_counter()
    for d in insts:
        par = StokesParameters(viscosity=1 / (6 * np.pi), Lx=d["L"], Ly=d["L"], zmin=d["zmin"], zmax=d["zmax"],
                               w=6, w_d=6, beta=1.785 * 6, beta_d=1.785 * 6, nx=d["n"], ny=d["n"], nz=d["nz"],
                               mode="bottom", precision="mixed")
        s = DPStokes()
        s.initialize(par, d["N"])
        s.setPositions(d["pos"])
        solvers.append(s)
        F.append(torch.from_numpy(d["F"].astype(np.float32)).to(dev))
        T.append(torch.from_numpy(d["T"].astype(np.float32)).to(dev))
    setup_s = time.perf_counter() - t0
    b = DPStokesBatch(solvers, nstreams=4)
    chunks = b.tune(F, T)
    ms_graph = time_steps(torch, lambda: b.mdot_device(F, T), args.steps, args.warmup)

    def eager():
        for s, f, t in zip(solvers, F, T):
            s.mdot_device(f, t)
    ms_eager = time_steps(torch, eager, max(3, args.steps // 2), 1)
    ntot = sum(d["N"] for d in insts)
    out = {
        "metric": "batched Mdot/s: 64 independent bottom-wall instances (forces + torques) per step",
        "value": 1000.0 / ms_graph,
        "unit": "batches/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_graph,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 grids + f64 BVP",
        "data": "synthetic",
        "config": {"workload": "64 bottom-wall instances, N log-uniform [256, 32768] (total %d), height "
                               "log-uniform [12, 48], h = 0.5, w = w_d = 6, mixed" % ntot},
        "particles_per_s": ntot * 1000.0 / ms_graph,
        "eager_ms_per_step": ms_eager,
        "graph_speedup_vs_eager": ms_eager / ms_graph,
        "grids": sorted(set((d["n"], d["nz"]) for d in insts)),
        "tuned_z_chunks": chunks,
        "setup_s": setup_s,
    }
    print(json.dumps(out), flush=True)


def bench_adjoint(args, wl):
    """Forward Mdot and its adjoint (gradients with respect to positions,
    forces and torques of L = <g, U> + <h, W>, g, h iid N(0, 1)) on one
    workload, mixed precision, with the FP64 path's gradients for comparison."""
    import torch
    geom, N, Lx, Ly, zmax, zlo, zhi, nx, ny, nz, torque = wl
    pos, F, T = make_inputs(wl)
    rng = np.random.RandomState(1)
    gU = rng.normal(size=3 * N)
    gW = rng.normal(size=3 * N) if torque else None
    dev = torch.device("cuda", 0)
    res = {}
    for prec, dt in (("mixed", torch.float32), ("double", torch.float64)):
        s = gpu_solver(wl, prec)
        s.setPositions(pos)
        t = lambda a: torch.from_numpy(a).to(dev, dt) if a is not None else None
        Fd, Td, gUd, gWd = t(F), t(T), t(gU), t(gW)
        steps = args.steps if prec == "mixed" else max(3, args.steps // 5)
        ms_f = time_steps(torch, lambda: s.mdot_device(Fd, Td), steps, args.warmup)
        ms_a = time_steps(torch, lambda: s.mdot_adjoint_device(Fd, gUd, torques=Td, grad_angular=gWd),
                          steps, max(1, args.warmup // 2))
        gF, gT, gX = s.mdot_adjoint_device(Fd, gUd, torques=Td, grad_angular=gWd)
        torch.cuda.synchronize()
        res[prec] = (ms_f, ms_a, gX.double().cpu().numpy(), gF.double().cpu().numpy())
        del s
    ms_f, ms_a = res["mixed"][0], res["mixed"][1]
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
    out = {
        "metric": "adjoint Mdot/s (forward + gradients w.r.t. positions, forces%s)" % (" and torques" if torque else ""),
        "value": 1000.0 / (ms_f + ms_a),
        "unit": "forward+adjoint/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_f + ms_a,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 grids + f64 BVP",
        "data": "synthetic",
        "config": {"workload": "%s: %s N=%d L=%gx%g zmax=%g grid %dx%dx%d w=w_d=6 betaw=10.71 %s"
                   % (args.workload, geom, N, Lx, Ly, zmax, nx, ny, nz, "F+T" if torque else "F"),
                   "precision": "mixed"},
        "forward_ms": ms_f,
        "adjoint_ms": ms_a,
        "fp64_forward_ms": res["double"][0],
        "fp64_adjoint_ms": res["double"][1],
        "rel_l2_dLdx_mixed_vs_fp64": rel(res["mixed"][2], res["double"][2]),
        "rel_l2_dLdF_mixed_vs_fp64": rel(res["mixed"][3], res["double"][3]),
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="slit16k", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-budget", type=float, default=15.0,
                    help="seconds of timed CPU-baseline Mdot work after one warm-up call (0 disables)")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--reg-kernel", action="store_true",
                    help="torques through the regular kernel with spectral curls (useRegKernel)")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="also run a plain copy of this many MiB (profiler byte-counter calibration)")
    ap.add_argument("--batch", action="store_true",
                    help="the 64-instance batched workload instead of the single slit instance")
    ap.add_argument("--solver", default="auto", choices=["auto", "lane"],
                    help="spectral solve: fused lane-group kernel (auto) or MFMA z-GEMMs + one-lane BVP (lane)")
    ap.add_argument("--adjoint", action="store_true",
                    help="time the forward Mdot and its adjoint (position, force and torque gradients)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.batch:
        bench_batch(args)
        return
    if args.adjoint:
        bench_adjoint(args, WORKLOADS[args.workload])
        return
    import torch
    wl = WORKLOADS[args.workload]
    geom, N, Lx, Ly, zmax, zlo, zhi, nx, ny, nz, torque = wl
    pos, F, T = make_inputs(wl)
    dev = torch.device("cuda", 0)
    Fd = torch.from_numpy(F.astype(np.float32)).to(dev)
    Td = torch.from_numpy(T.astype(np.float32)).to(dev) if torque else None

    reg = args.reg_kernel and torque
    mixed = gpu_solver(wl, "mixed", reg, args.solver)
    mixed.setPositions(pos)
    ms = time_steps(torch, lambda: mixed.mdot_device(Fd, Td), args.steps, args.warmup)
    U32, W32 = mixed.mdot_device(Fd, Td)
    torch.cuda.synchronize()

    # per-stage times from HIP events recorded on the solver's stream
    mixed.set_profiling(True)
    acc = None
    nprof = 10
    for _ in range(nprof):
        mixed.mdot_device(Fd, Td)
        t = mixed.timings()
        acc = t if acc is None else {k: acc[k] + t[k] for k in acc}
    mixed.set_profiling(False)
    stage_us = {k: v / nprof for k, v in acc.items() if k != "locate"}
    kernels = STAGE_KERNEL
    if stage_us["z_fwd"] == 0 and stage_us["z_inv"] == 0:
        # fused spectral kernel: the z transforms run inside the "bvp" launch
        del stage_us["z_fwd"], stage_us["z_inv"]
    else:
        kernels = STAGE_KERNEL_LANE
    sb = stage_bytes(wl, 4, reg)
    stage_gbs = {k: sb[k] / (stage_us[k] * 1e-6) / 1e9 for k in stage_us}
    dom = max(stage_us, key=lambda k: stage_us[k])
    traffic, traffic_src = (None, None) if reg else measured_traffic(args.workload, kernels[dom])

    out = {
        "metric": "Mdot solves/sec, N=16k slit, mixed vs FP64 at rel-L2<=1e-5; spread HBM GB/s",
        "value": 1000.0 / ms,
        "unit": "Mdot/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 grids + f64 BVP",
        "data": "synthetic",
        "config": {"workload": "%s: %s N=%d L=%gx%g zmax=%g grid %dx%dx%d w=w_d=6 betaw=10.71 %s%s%s"
                   % (args.workload, geom, N, Lx, Ly, zmax, nx, ny, nz, "F+T" if torque else "F",
                      " regular torque kernel" if reg else "",
                      " lane solver" if args.solver == "lane" else ""),
                   "precision": "mixed"},
        "stage_us": stage_us,
        "stage_GBps": stage_gbs,
        "spread_GBps": stage_gbs["spread"],
        "roofline": {"bound": "hbm", "kernel": kernels[dom], "stage": dom, "achieved": stage_gbs[dom],
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": stage_gbs[dom] / HBM_PEAK_GBS,
                     "traffic": traffic, "traffic_source": traffic_src, "algorithmic_bytes": sb[dom],
                     "launch_us": stage_us[dom]},
    }
    if kernels["bvp"] == "k_spectral" and not reg:
        # the Chebyshev z transforms inside the fused solve, forward and
        # inverse over folded node pairs: v_mfma 16x16x4 FP32 products for
        # 16-lane groups (17 <= Nz <= 32 or 49 <= Nz <= 128), FP32 VALU sums
        # for the other shapes
        zf = z_transform_flop(wl)
        nz = wl[9]
        unit = "MFMA FP32" if (16 < nz <= 32 or 48 < nz <= 128) else "VALU FP32"
        out["z_transform"] = {"kernel": "k_spectral", "unit": unit, "flop_per_launch": zf,
                              "tflops_over_kernel_time": zf / (stage_us["bvp"] * 1e-6) / 1e12,
                              "peak_tflops": FP32_PEAK_TFLOPS}
    del mixed
    U64 = W64 = None
    if not args.no_fp64:
        f64 = gpu_solver(wl, "double", reg, args.solver)
        f64.setPositions(pos)
        Fd64 = torch.from_numpy(F).to(dev)
        Td64 = torch.from_numpy(T).to(dev) if torque else None
        ms64 = time_steps(torch, lambda: f64.mdot_device(Fd64, Td64), max(5, args.steps // 5), 3)
        U64, W64 = f64.mdot_device(Fd64, Td64)
        torch.cuda.synchronize()
        u64 = U64.cpu().numpy()
        out["fp64_ms_per_step"] = ms64
        out["rel_l2_U_mixed_vs_fp64"] = float(np.linalg.norm(U32.double().cpu().numpy() - u64) / np.linalg.norm(u64))
        if torque:
            w64 = W64.cpu().numpy()
            out["rel_l2_W_mixed_vs_fp64"] = float(np.linalg.norm(W32.double().cpu().numpy() - w64) / np.linalg.norm(w64))
        del f64
    if args.copy_probe_mb > 0:
        copy_probe(torch, args.copy_probe_mb)
    if args.cpu_budget > 0:
        rate, n, el, thr, phases, U_ref, W_ref = cpu_baseline(wl, pos, F, T, args.cpu_budget, reg)
        out["cpu_baseline"] = {"value": rate, "unit": "Mdot/s", "cores": thr, "kind": "port",
                               "sample": "1 warm-up + %d timed full-size FP64 oracle Mdot calls on the same "
                                         "inputs (%.1f s)" % (n, el),
                               "phase_ms": {k: 1000.0 * v for k, v in phases.items()}}
        out["speedup_vs_cpu"] = out["value"] / rate
        if "fp64_ms_per_step" in out:
            out["fp64_speedup_vs_cpu"] = (1000.0 / out["fp64_ms_per_step"]) / rate
            out["fp64_over_mixed_time"] = out["fp64_ms_per_step"] / out["ms_per_step"]
        # parity of the timed GPU paths against the oracle on the same inputs
        rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
        out["rel_l2_U_mixed_vs_oracle"] = rel(U32.double().cpu().numpy(), U_ref)
        if torque:
            out["rel_l2_W_mixed_vs_oracle"] = rel(W32.double().cpu().numpy(), W_ref)
        if U64 is not None:
            out["rel_l2_U_fp64_vs_oracle"] = rel(U64.cpu().numpy(), U_ref)
            if torque:
                out["rel_l2_W_fp64_vs_oracle"] = rel(W64.cpu().numpy(), W_ref)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
