"""Benchmark: pressure-Poisson solve to relative residual 1e-8 at 128^3 on one MI355X.

Workload `ppe_128_duct`: 128^3 cells, x periodic (uniform), y and z Roberts
beta=1.05 walls (the rectangular-duct mesh), homogeneous Neumann walls, right-
hand side (1/dt) div(U*) of a seeded 1e-3 SplitMix64 face-noise field, x0 = 0.
One step = one PPE solve (x reset, modify_RHS, PCG loop, BCs on x), inputs
resident in HBM; `value` is the learned + mixed solver's time (the metric's
method; --no-learned: the mixed solver's), beside the plain FP64 Jacobi-PCG
and the mixed solver without the network.  Prints one JSON line.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "PPE time-to-1e-8 per timestep (learned+mixed vs plain FP64 PCG); HBM GB/s"
HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0  # dense BF16 MFMA
N = 128
BETA = 1.05
TOL = 1e-8
DT = 1e-3


def make_inputs(M):
    hx = M.nodes_uniform(0.0, 1.0, N)
    hy = M.nodes_roberts_both(0.0, 1.0, N, BETA)
    m = M.mesh(hx, hy, hy)
    types = ["Periodic", "Periodic", "Neumann", "Neumann", "Neumann", "Neumann"]
    # U*: seeded face noise, then b = (1/dt) div(U*) on the GPU
    us = []
    for d in range(3):
        f = M.SF(m, 1 << d)
        n = f.size
        a = splitmix(n, 3 + d, 1e-3).reshape(f.shape, order="F")
        us.append(f.upload(a))
    b = M.SF(m, "CC")
    M.div(b, *us)
    M.synchronize()
    bh = b.download() * (1.0 / DT)
    b.upload(bh)
    x = M.SF(m, "CC").set_BCs(types, None, all_Neumann=True)
    return m, x, b, types, (hx, hy), bh


def splitmix(n, seed, amp):
    """SplitMix64 counter form -> uniform [-amp, amp)."""
    q = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + q * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return amp * (2.0 * u - 1.0)


def load_traffic(name):
    """HBM bytes per launch of a kernel, from its committed rocprofv3 PMC summary
    (the line `hbm_bytes_per_launch: N`), or None."""
    p = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(p):
        return None
    with open(p) as fh:
        for line in fh:
            if line.startswith("hbm_bytes_per_launch:"):
                return int(float(line.split(":", 1)[1]))
    return None


def host_cpu():
    """(cores this process may run on, nproc, CPU model name)."""
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return cores, os.cpu_count() or cores, model


def cpu_baseline(hs, bh, gpu_iters):
    """The oracle's FP64 Jacobi-PCG (the reference's operator sequence in C,
    OpenMP) on the same mesh and right-hand side, solved to the same 1e-8 on
    every core this process may use; one core beside it on a bounded sample
    (10 iterations, extrapolated to the solve's iteration count)."""
    from oracle import oracle as O

    hx, hy = hs
    g = O.Grid(hx, hy, hy)
    bcs = O.BCs.make(["Periodic", "Periodic", "Neumann", "Neumann", "Neumann", "Neumann"])
    b = np.asfortranarray(bh)
    cores, nproc, model = host_cpu()
    O.set_threads(cores)
    x = np.zeros(g.shape(0), order="F")
    t0 = time.perf_counter()
    it, _, _, _ = O.pcg_lap(g, 0, bcs, True, x, b, 20000, TOL, 1e-300)
    full_ms = 1e3 * (time.perf_counter() - t0)
    O.set_threads(1)
    ts = {}
    for k in (2, 10):
        x1 = np.zeros(g.shape(0), order="F")
        t0 = time.perf_counter()
        O.pcg_lap(g, 0, bcs, True, x1, b, k, 0.0, 0.0)
        ts[k] = time.perf_counter() - t0
    per_iter = (ts[10] - ts[2]) / 8.0
    one_ms = 1e3 * (max(ts[2] - 2 * per_iter, 0.0) + per_iter * it)
    return {"value": round(full_ms, 3), "unit": "ms", "cores": cores, "kind": "port",
            "sample": f"the oracle's FP64 Jacobi-PCG solve of the same 128^3 PPE to 1e-8 ({it} iterations; the GPU's "
                      f"FP64 solve took {gpu_iters}), OpenMP on the {cores} cores this process may use",
            "iterations": it, "gpu_fp64_iterations": gpu_iters, "nproc": nproc, "cpu_model": model,
            "omp_threads": cores,
            "single_core": {"value": round(one_ms, 3), "unit": "ms", "cores": 1,
                            "sample": f"2 and 10 iterations timed ({ts[2]:.2f} s, {ts[10]:.2f} s), extrapolated to "
                                      f"{it} iterations; {1e3 * per_iter:.1f} ms per iteration"}}


def load_rocprof_avg(name, kernel_prefix, grid=None):
    """Average duration (us) of a kernel in a committed rocprofv3 summary, or
    None.  Summary lines read `kernel: <name> grid: <g> calls: <c> avg_us: <t>`
    (the --rocprof-summary output); every line whose name starts with
    kernel_prefix (a string or a tuple of them) and whose grid (work-items) is
    `grid`, when given, counts, weighted by its calls."""
    p = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(p):
        return None
    tot = n = 0.0
    with open(p) as fh:
        for line in fh:
            if not line.startswith("kernel:") or "avg_us:" not in line:
                continue
            kname = line[len("kernel:"):].split(" grid:")[0].strip()
            kname = kname[5:] if kname.startswith("void ") else kname  # the return type and namespace
            kname = kname[7:] if kname.startswith("moons::") else kname
            if not kname.startswith(tuple(kernel_prefix) if isinstance(kernel_prefix, (list, tuple)) else kernel_prefix):
                continue
            f = line.split()
            g = f[f.index("grid:") + 1] if "grid:" in f else None
            if grid is not None and g != str(grid):
                continue
            c = float(f[f.index("calls:") + 1])
            tot += c * float(f[f.index("avg_us:") + 1])
            n += c
    return round(tot / n, 3) if n else None


def rocprof_summary(d, top=60):
    """Per (kernel, grid size) dispatch count and duration statistics of the
    rocprofv3 kernel-trace CSVs under directory d, longest total first, and
    per (kernel, grid size, counter) statistics of its counter-collection CSVs."""
    import csv
    import glob

    files = sorted(glob.glob(os.path.join(d, "**", "*kernel_trace.csv"), recursive=True))
    groups = {}
    for fpath in files:
        with open(fpath) as fh:
            for r in csv.DictReader(fh):
                if "Grid_Size_X" in r:
                    grid = int(r["Grid_Size_X"]) * int(r["Grid_Size_Y"]) * int(r["Grid_Size_Z"])
                else:
                    grid = int(r.get("Grid_Size", 0))
                dur = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-3
                groups.setdefault((r["Kernel_Name"], grid), []).append(dur)
    rows = sorted(groups.items(), key=lambda kv: -sum(kv[1]))
    total = sum(sum(v) for v in groups.values())
    print(f"trace files: {len(files)}; {sum(len(v) for v in groups.values())} dispatches, "
          f"{total / 1e3:.3f} ms of kernel time")
    for (kname, grid), v in rows[:top]:
        print(f"kernel: {kname[:120]} grid: {grid} calls: {len(v)} avg_us: {sum(v) / len(v):.3f} "
              f"min_us: {min(v):.3f} max_us: {max(v):.3f} total_ms: {sum(v) / 1e3:.3f}")
    # PMC runs (--pmc): per (kernel, grid, counter) mean, min and max over dispatches
    cfiles = sorted(glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True))
    counters = {}
    for fpath in cfiles:
        with open(fpath) as fh:
            for r in csv.DictReader(fh):
                key = (r["Kernel_Name"], int(r.get("Grid_Size", 0)), r["Counter_Name"])
                counters.setdefault(key, []).append(float(r["Counter_Value"]))
    if cfiles:
        print(f"counter files: {len(cfiles)}")
    for (kname, grid, cname), v in sorted(counters.items()):
        print(f"kernel: {kname[:120]} grid: {grid} counter: {cname} dispatches: {len(v)} "
              f"mean: {sum(v) / len(v):.1f} min: {min(v):.1f} max: {max(v):.1f}")


def run_ensemble(M, args):
    """64 independent MHD lid-driven cavities (sizes from {32,48,64,96}^3, Re in
    [100, 1000], Ha in [0, 50], seeded 1e-3 noise, dt 5e-4, the reference's
    5-iteration solver settings); one step = one time step of every member,
    each member's step a captured HIP graph, replayed over 4 streams."""
    from moons_amd.ensemble import Ensemble

    t0 = time.perf_counter()
    ens = Ensemble(n=64, sizes=(32, 48, 64, 96), seed=5, dt=5e-4, n_streams=4)
    setup_s = time.perf_counter() - t0
    ens.step(args.warmup)
    t1 = time.perf_counter()
    ens.step(args.steps)
    ms = 1e3 * (time.perf_counter() - t1) / args.steps
    sizes = {}
    for mb in ens.members:
        sizes[mb.N] = sizes.get(mb.N, 0) + 1
    # per-member error against the plain FP64 path: the same members stepped
    # eagerly (no graph, one stream) for the same number of steps
    plain = Ensemble(n=64, sizes=(32, 48, 64, 96), seed=5, dt=5e-4, n_streams=1, capture=False)
    for mb in plain.members:
        mb.st.step(args.warmup + args.steps)
    M.synchronize()

    def rel(a, b):
        den = float(np.linalg.norm(b))
        return float(np.linalg.norm(a - b)) / (den if den > 0 else 1.0)

    errs = []
    for a, b in zip(ens.members, plain.members):
        e = max([rel(a.U[d].download(), b.U[d].download()) for d in range(3)] +
                [rel(a.B[d].download(), b.B[d].download()) for d in range(3)] +
                [rel(a.p.download(), b.p.download())])
        errs.append(e)
    per_member = [{"N": mb.N, "Re": round(mb.Re, 1), "Ha": round(mb.Ha, 2), "rel_l2": e}
                  for mb, e in zip(ens.members, errs)]
    out = {"metric": "ensemble MHD time steps (member-steps) per second", "value": round(64 * 1e3 / ms, 2),
           "unit": "member-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": "synthetic",
           "config": {"workload": "ensemble_64_mhd_cavity", "members": 64, "sizes": {str(k): v for k, v in sorted(sizes.items())},
                      "cells": ens.cells, "Re": [100, 1000], "Ha": [0, 50], "dt": 5e-4, "solver_iterations": 5,
                      "execution": "per-member HIP graph, 4 streams"},
           "cell_steps_per_s": round(ens.cells * 1e3 / ms, 1), "setup_s": round(setup_s, 1),
           "per_member_error_vs_plain_fp64": {"max_rel_l2": max(errs), "fields": "U, B, p",
                                              "plain": "the same members stepped eagerly, no graph",
                                              "members": per_member}}
    print(json.dumps(out), flush=True)


def run_mhd(M, args):
    """MHD lid-driven cavity time steps at N^3, eager on the library stream
    (no graph): the 64^3 MHD configuration's parameters (Re 400, Ha 10, Rem 1
    low-Rem induction, B0 = (0, 1, 0), dt 5e-4, 1e-3 seeded velocity noise,
    insulating B walls, backward-Euler momentum diffusion); the PPE solved by
    the mixed solver to 1e-8, the momentum, induction and phi solves with 5
    iterations.  One step = one time step."""
    from moons_amd.ensemble import PSEUDO_VACUUM, noise_velocity

    N = args.mhd
    m = M.mesh(*[M.nodes_uniform(0.0, 1.0, N)] * 3)
    U = noise_velocity(m, 2)
    p = M.SF(m, "CC").fill(0.0).set_BCs(["Neumann"] * 6, None, all_Neumann=True)
    B = [M.SF(m, 1 << d).fill(0.0).set_BCs(*PSEUDO_VACUUM[d]) for d in range(3)]
    phi = M.SF(m, "CC").fill(0.0).set_BCs(["Dirichlet"] * 6)
    isp = (5, 1e-6, 1e-13, 1)
    st = M.MHD_step(m, U, p, B=B, phi=phi, Re=400.0, Ha=10.0, Rem=1.0, dt=5e-4, alpha_U=1.0, B0=(0.0, 1.0, 0.0),
                    isp_U=isp, isp_p=(10000, TOL, 1e-13, 1), isp_B=isp, isp_phi=isp, prec_p="MG_FP32")
    st.step(args.warmup)
    M.synchronize()
    t0 = time.perf_counter()
    info = st.step(args.steps)
    M.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / args.steps
    out = {"metric": "MHD time steps per second", "value": round(1e3 / ms, 2), "unit": "steps/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"mhd_cavity_{N}", "cells": f"{N}^3", "Re": 400, "Ha": 10, "Rem": 1, "dt": 5e-4,
                      "ppe": "mixed FP32 V-cycle in FP64 FPCG to 1e-8", "other_solves": "5 iterations",
                      "execution": "eager, one stream"},
           "last_step_iterations": info["iterations"]}
    print(json.dumps(out), flush=True)


def run_mhd_md(M, args):
    """The MHD cavity of run_mhd with the induction on its own mesh: the N^3
    fluid mesh extended by 11 uniform cells at every wall (ext_uniform_IO,
    the reference's mesh bank extension), sigma = 1 on the fluid and a
    one-cell wall and 0.1 beyond, B = 0 and phi = 0 Dirichlet at the far
    boundary.  Each step embeds the fluid's edge velocity, solves B and phi
    on the (N + 22)^3 mesh and extracts J x B.  One step = one time step."""
    from moons_amd.ensemble import noise_velocity

    N = args.mhd_md
    hn = M.nodes_uniform(0.0, 1.0, N)
    m = M.mesh(hn, hn, hn)
    hi = M.nodes_ext_uniform(hn, 11)
    hw = M.nodes_ext_uniform(hn, 1)
    mi, mw = M.mesh(hi, hi, hi), M.mesh(hw, hw, hw)
    md, mdw = M.mesh_domain(m, mi), M.mesh_domain(mw, mi)
    U = noise_velocity(m, 2)
    p = M.SF(m, "CC").fill(0.0).set_BCs(["Neumann"] * 6, None, all_Neumann=True)
    B = [M.SF(mi, 1 << d).fill(0.0).set_BCs(["Dirichlet"] * 6) for d in range(3)]
    phi = M.SF(mi, "CC").fill(0.0).set_BCs(["Dirichlet"] * 6)
    sig = [M.SF(mi, dl) for dl in (6, 5, 3)]  # edge x, y, z
    M.set_sigma_inv(sig, mdw, 0.1)
    isp = (5, 1e-6, 1e-13, 1)
    st = M.MHD_step(m, U, p, B=B, phi=phi, Re=400.0, Ha=10.0, Rem=1.0, dt=5e-4, alpha_U=1.0, B0=(0.0, 1.0, 0.0),
                    isp_U=isp, isp_p=(10000, TOL, 1e-13, 1), isp_B=isp, isp_phi=isp, prec_p="MG_FP32", m_ind=mi,
                    domain=md, sigma_inv=sig)
    st.step(args.warmup)
    M.synchronize()
    t0 = time.perf_counter()
    info = st.step(args.steps)
    M.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / args.steps
    Ni = N + 22
    out = {"metric": "MHD time steps per second", "value": round(1e3 / ms, 2), "unit": "steps/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"mhd_cavity_{N}_induction_{Ni}", "cells": f"{N}^3 fluid, {Ni}^3 induction",
                      "Re": 400, "Ha": 10, "Rem": 1, "dt": 5e-4, "sigma": "1 on fluid + 1-cell wall, 0.1 beyond",
                      "ppe": "mixed FP32 V-cycle in FP64 FPCG to 1e-8", "other_solves": "5 iterations",
                      "execution": "eager, one stream"},
           "last_step_iterations": info["iterations"]}
    print(json.dumps(out), flush=True)


def run_duct(M, args):
    """The duct MHD flow at N^3 (x periodic, Roberts beta 1.05 walls in y and
    z, Re 1000, Ha 100, thin conducting walls c_w 0.01, mean pressure gradient,
    finite Rem, dt 1e-3): the PPE by the mixed solver to 1e-8, phi to 1e-12,
    U and B with the reference's 5-iteration settings.  One step = one time
    step, eager on the library stream."""
    from moons_amd.duct import Duct

    N = args.duct
    # The fields of this set-up grow without bound after a few steps (max|B|
    # 3-8x a step from the fourth, NaN by the 18th at 128^3): the reference's
    # thin-wall ghost rule for cell-located components, reproduced here, is
    # (2c+1)/(2c-1) times the interior value on min walls (c = c_w/dh, 1.34 at
    # c_w 0.01 on this mesh), an anti-diffusive wall.  So the duct is timed
    # over the first 7 steps: 2 untimed, then at most 5.
    warmup, steps = min(args.warmup, 2), min(args.steps, 5)
    # phi cleaning by the mixed multigrid too (the reference's Jacobi beside it)
    D = Duct(n=(N, N, N), prec_p="MG_FP32", prec_phi="MG_FP32")
    D.step(warmup)
    M.synchronize()
    t0 = time.perf_counter()
    info = D.step(steps)
    M.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / steps
    ms_j, info_j, agree = None, {"iterations": None}, None
    if not args.duct_no_jacobi:
        DJ = Duct(n=(N, N, N), prec_p="MG_FP32", prec_phi="Lap_SF")
        DJ.step(warmup)
        M.synchronize()
        t0 = time.perf_counter()
        info_j = DJ.step(steps)
        M.synchronize()
        ms_j = 1e3 * (time.perf_counter() - t0) / steps
        # the two steps' fields agree (both solves to their tolerances)
        fd, fj = D.fields(), DJ.fields()
        agree = {f"{k}{d}": float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
                 for k in ("U", "B", "J") for d, (a, b) in enumerate(zip(fd[k], fj[k]))}
    out = {"metric": "duct MHD time steps per second", "value": round(1e3 / ms, 2), "unit": "steps/s", "n_gpus": 1,
           "steps": steps, "warmup": warmup, "ms_per_step": round(ms, 3), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"duct_mhd_{N}", "cells": f"{N}^3", "Re": 1000, "Ha": 100, "c_w": 0.01,
                      "mesh": "x periodic uniform [-1,1]; y,z Roberts beta=1.05 on [-1,1]", "dt": 1e-3,
                      "ppe": "mixed FP32 V-cycle in FP64 FPCG to 1e-8",
                      "phi": "mixed FP32 V-cycle in FP64 FPCG to 1e-12", "U, B": "5 iterations",
                      "execution": "eager, one stream"},
           "last_step_iterations": info["iterations"],
           "phi_jacobi": {"ms_per_step": None if ms_j is None else round(ms_j, 3),
                          "last_step_iterations": info_j["iterations"],
                          "phi": "the reference's Jacobi-PCG to 1e-12"},
           "fields_vs_phi_jacobi_rel_l2": agree}
    print(json.dumps(out), flush=True)


def run_fft(M, args):
    """The cosine-transform direct solver (the reference's FFT pressure path,
    one independent 2D Neumann problem per plane normal to x) at N^3 on a
    uniform grid, inputs resident in HBM; one step = one solve (pack, DCT-II
    along y and z as FP64 GEMMs, mode scaling, inverse, unpack, BCs)."""
    N = args.fft
    hs = [M.nodes_uniform(0.0, 1.0, N)] * 3
    m = M.mesh(*hs)
    f = M.SF(m, "CC").upload(np.asfortranarray(splitmix((N + 2) ** 3, 21, 1.0).reshape((N + 2,) * 3, order="F")))
    u = M.SF(m, "CC").fill(0.0).set_BCs(["Neumann"] * 6, None, all_Neumann=True)
    S = M.FFT_solver_SF(m, 1)
    for _ in range(args.warmup):
        S.solve(u, f)
    M.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        S.solve(u, f)
    M.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / args.steps
    flop = 2 * 2 * N ** 3 * (N + N)  # forward and inverse, along y and z: 2 n^3 N each
    tfs = flop / (ms * 1e-3) / 1e12
    out = {"metric": "cosine-transform PPE solve time", "value": round(ms, 4), "unit": "ms", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": False,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"fft_ppe_{N}", "cells": f"{N}^3", "direction": 1, "mesh": "uniform",
                      "transforms": "DCT-II along y and z as FP64 GEMMs (rocBLAS)"},
           "roofline": {"bound": "mfma", "achieved": round(tfs, 2), "peak": 78.6, "unit": "TFLOP/s",
                        "frac": round(tfs / 78.6, 4), "traffic": None,
                        "flop_per_solve": flop, "note": "whole solve (GEMMs plus pack, scale, unpack, BCs)"}}
    print(json.dumps(out), flush=True)


def run_energy(M, args):
    """The energy equation at N^3 on a uniform grid, inputs resident in HBM:
    one step = sources (advection by a seeded velocity, scale -1) and method 3
    (AB2, backward-Euler I + c div(k grad) with k = 1 on every face, Pe 100,
    dt 5e-4) solved by the nrg_diffusion Jacobi-PCG to 1e-8.  Also times the
    nrg_diffusion matvec kernel (read p and the three face arrays of k, write
    q: 40 B/cell) against the HBM peak."""
    from moons_amd.ensemble import noise_velocity

    N = args.energy
    hs = [M.nodes_uniform(0.0, 1.0, N)] * 3
    m = M.mesh(*hs)
    U = noise_velocity(m, 2)
    T_bcs = (["Dirichlet", "Dirichlet", "Neumann", "Neumann", "Neumann", "Neumann"], [1.0, 0.0, 0, 0, 0, 0])
    x = np.linspace(0.0, 1.0, N + 2)
    T0 = np.asfortranarray(np.broadcast_to((1.0 - x)[:, None, None], (N + 2,) * 3))
    T = M.SF(m, "CC").upload(T0).set_BCs(*T_bcs)
    k = [M.SF(m, 1 << d).fill(1.0) for d in range(3)]
    st = M.Energy_step(m, T, U, k, dt=5e-4, method=3, Pe=100.0, advection=(True, -1.0), diffusion=(False, 0.0),
                       isp_T=(1000, TOL, 1e-300, 1))
    st.step(args.warmup)
    M.synchronize()
    t0 = time.perf_counter()
    info = st.step(args.steps)
    M.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / args.steps
    # the nrg_diffusion PCG's kernels, 50 timed launches each after 4 iterations
    xk = M.SF(m, "CC").upload(T0).set_BCs(*T_bcs)
    P = M.PCG_solver_SF(m, xk, M.iter_solver_params(4, 0.0, 0.0, 1), op="nrg_diffusion")
    P.set_k(k)
    P.update_MFP(5e-4 * (-1.0 / 100.0))
    P.solve(xk, M.SF(m, "CC").upload(T0))
    kms = P.time_kernels(50)
    cells = N ** 3
    per = {"nrg_matvec": 40, "pcg_update": 48, "pcg_pupdate": 24}
    kern = {n: {"us": round(1e3 * t, 2), "GBs": round(per[n] * cells / (t * 1e-3) / 1e9, 1)}
            for n, t in zip(per, kms)}
    gbs = per["nrg_matvec"] * cells / (kms[0] * 1e-3) / 1e9
    out = {"metric": "energy time steps per second", "value": round(1e3 / ms, 2), "unit": "steps/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"energy_{N}", "cells": f"{N}^3", "method": 3, "Pe": 100, "dt": 5e-4,
                      "k": "1 on every face", "T_solve": "nrg_diffusion Jacobi-PCG to 1e-8",
                      "sources": "advection (scale -1) by a seeded velocity"},
           "last_step_iterations": info["iterations"], "kernels": kern,
           "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": 8000.0, "unit": "GB/s",
                        "frac": round(gbs / 8000.0, 4), "traffic": load_traffic("energy_step_128.md") if N == 128 else None,
                        "kernel": "nrg_matvec",
                        "bytes_per_cell": 40}}
    # the same step with every source of add_all_energy_sources on (advection,
    # diffusion into L, KE diffusion, viscous dissipation, Joule heating of a
    # seeded J with sigmaInv 1, volumetric heating) and the 3-stage RK with
    # method 5: three sources-and-solve stages a step
    J = [M.SF(m, dl).upload(np.asfortranarray(splitmix(int(np.prod(m.shape(dl))), 40 + q, 0.1)
                                              .reshape(m.shape(dl), order="F")))
         for q, dl in enumerate((6, 5, 3))]
    sig = M.SF(m, "CC").fill(1.0)
    T2 = M.SF(m, "CC").upload(T0).set_BCs(*T_bcs)
    st2 = M.Energy_step(m, T2, U, k, dt=5e-4, method=5, Pe=100.0, advection=(True, -1.0), diffusion=(True, 0.01),
                        isp_T=(1000, TOL, 1e-300, 1), KE_diffusion=(True, 0.01), viscous_dissipation=(True, 0.01),
                        joule_heating=(True, 0.01), volumetric_heating=(True, 0.01), rk_stages=3, J=J, sigma_inv=sig)
    st2.step(args.warmup)
    M.synchronize()
    t0 = time.perf_counter()
    info2 = st2.step(args.steps)
    M.synchronize()
    out["all_sources_rk3"] = {"ms_per_step": round(1e3 * (time.perf_counter() - t0) / args.steps, 3),
                              "last_stage_iterations": info2["iterations"],
                              "config": "method 5 with the 3-stage RK (three stages a step); every source on"}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--ensemble", action="store_true",
                    help="batched-ensemble benchmark instead: 64 MHD cavities of sizes {32,48,64,96}^3 advanced by "
                         "HIP-graph replays on 4 streams")
    ap.add_argument("--mhd", type=int, default=0, metavar="N",
                    help="MHD time-step benchmark instead: one lid-driven MHD cavity at N^3, eager")
    ap.add_argument("--mhd-md", type=int, default=0, metavar="N",
                    help="the MHD benchmark with the induction on the N^3 fluid mesh extended by 11 cells a wall")
    ap.add_argument("--duct", type=int, default=0, metavar="N",
                    help="duct MHD time-step benchmark instead (thin conducting walls, Ha 100), at N^3")
    ap.add_argument("--fft", type=int, default=0, metavar="N",
                    help="cosine-transform direct PPE solver benchmark instead, at N^3")
    ap.add_argument("--duct-no-jacobi", action="store_true",
                    help="with --duct, skip the comparison run with the reference's Jacobi phi cleaning")
    ap.add_argument("--energy", type=int, default=0, metavar="N",
                    help="energy-equation step benchmark instead (method 3, k = 1), at N^3")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip training and timing the learned component (U-net initial guess in the V-cycle)")
    ap.add_argument("--kernels-only", action="store_true",
                    help="4 PCG iterations, then 50 timed launches of each iteration kernel "
                         "(no early-exited launches; for profiler runs)")
    ap.add_argument("--rocprof-summary", metavar="DIR",
                    help="print the per-kernel summary of the rocprofv3 kernel-trace CSVs under DIR and exit")
    args = ap.parse_args()
    if args.rocprof_summary:
        rocprof_summary(args.rocprof_summary)
        return
    if args.gpus != 1:
        print(f"bench.py: this library runs on one GPU only (asked for --gpus {args.gpus}); "
              "sharding the solve over several GPUs is not implemented.", file=sys.stderr)
        sys.exit(2)

    import moons_amd as M

    M.init(0)
    if args.ensemble:
        run_ensemble(M, args)
        return
    if args.mhd:
        run_mhd(M, args)
        return
    if args.mhd_md:
        run_mhd_md(M, args)
        return
    if args.duct:
        run_duct(M, args)
        return
    if args.fft:
        run_fft(M, args)
        return
    if args.energy:
        run_energy(M, args)
        return
    m, x, b, types, hs, bh = make_inputs(M)
    P = M.PCG_solver_SF(m, x, M.iter_solver_params(20000, TOL, 1e-300, 1))
    if args.kernels_only:
        P.set_ISP(M.iter_solver_params(4, 0.0, 0.0, 1))
        x.fill(0.0)
        P.solve(x, b)
        kms = P.time_kernels(50)
        names = ["pcg_matvec", "pcg_update", "pcg_pupdate"]
        per = {"pcg_matvec": 16, "pcg_update": 48, "pcg_pupdate": 24}
        print(json.dumps({n: {"us": round(1e3 * t, 2), "GBs": round(per[n] * N ** 3 / (t * 1e-3) / 1e9, 1)}
                          for n, t in zip(names, kms)}), flush=True)
        # the multigrid kernels of the mixed solver, 50 launches each
        xk = M.SF(m, "CC").set_BCs(types, None, all_Neumann=True)
        Pk = M.PCG_solver_SF(m, xk, M.iter_solver_params(2, 0.0, 0.0, 1), prec="MG_FP32")
        Pk.solve(xk, b)
        mk = Pk.time_MG(50)
        names = ["line_half_128", "restrict_128_to_64", "prolong_64_to_128", "tail", "vcycle", "fused_sweep_128",
                 "line_half_64", "fused_sweep_64", "half_sweep_32", "fused_sweep_32"]
        print(json.dumps({n + "_us": round(1e3 * v, 2) for n, v in zip(names, mk)}), flush=True)
        # calibration for the PMC byte counters: 20 device-to-device copies of
        # one 130^3 FP64 field (17,576,000 bytes read and written per copy)
        c = M.SF(m, "CC")
        for _ in range(20):
            c.assign(b)
        M.synchronize()
        return

    def timed(solver, xf):
        def step():
            xf.fill(0.0)
            return solver.solve(xf, b)

        for _ in range(args.warmup):
            step()
        M.synchronize()
        infos = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            infos.append(step())
        M.synchronize()
        t1 = time.perf_counter()
        return 1e3 * (t1 - t0) / args.steps, infos[-1]

    ms_step, info = timed(P, x)
    iters = info.iterations
    rel = info.res_norm / info.res_norm0
    # mixed precision: FP32 multigrid V-cycle inside the FP64 flexible PCG
    xm = M.SF(m, "CC").set_BCs(types, None, all_Neumann=True)
    Pm = M.PCG_solver_SF(m, xm, M.iter_solver_params(2000, TOL, 1e-300, 1), prec="MG_FP32")
    mg_ms, mg_info = timed(Pm, xm)
    mg_levels, mg_coarsest = Pm.mg_info()
    # the same solver with the 128^3 duct configuration's stated multigrid
    # (4 levels, 2 pre- and 2 post-smoothing red-black sweeps)
    xc = M.SF(m, "CC").set_BCs(types, None, all_Neumann=True)
    Pc = M.PCG_solver_SF(m, xc, M.iter_solver_params(2000, TOL, 1e-300, 1), prec="MG_FP32")
    Pc.set_MG(4, 2, 2, 40)
    mg4_ms, mg4_info = timed(Pc, xc)
    # learned + mixed, the metric's method: the U-net (trained here, on the
    # GPU, on pairs from random right-hand sides on the 32^3 level of this
    # mesh and from cavity runs with Re in [100, 1000], Ha in [0, 100]) as the
    # initial guess of the 32^3 level inside the FP32 V-cycle
    learned_out = None
    Pl = xl = None
    if not args.no_learned:
        t_l0 = time.perf_counter()
        from moons_amd import learned

        def coarsen(hn):
            phys = np.asarray(hn)[1:-1]
            c = phys[::2]
            return np.concatenate([[c[0] - (c[1] - c[0])], c, [c[-1] + (c[-1] - c[-2])]])

        lx, ly = np.asarray(hs[0]), np.asarray(hs[1])
        for _ in range(2):
            lx, ly = coarsen(lx), coarsen(ly)
        pairs = learned.random_pairs([lx, ly, ly], types, 128, seed=2)
        pairs += learned.run_pairs(32, 4, 12, seed=3, Ha_max=100.0)
        net = M.UNet(seed=0)
        hist = net.train([p[0] for p in pairs], [p[1] for p in pairs], epochs=30, batch=4, lr=1e-3, seed=0)
        t_train = time.perf_counter() - t_l0
        # the network in every V-cycle, and in the first V-cycle of each solve only
        xe = M.SF(m, "CC").set_BCs(types, None, all_Neumann=True)
        Pe = M.PCG_solver_SF(m, xe, M.iter_solver_params(2000, TOL, 1e-300, 1), prec="MG_FP32")
        Pe.set_MG_UNet(net, 2, cycles=0)
        e_ms, e_info = timed(Pe, xe)
        xl = M.SF(m, "CC").set_BCs(types, None, all_Neumann=True)
        Pl = M.PCG_solver_SF(m, xl, M.iter_solver_params(2000, TOL, 1e-300, 1), prec="MG_FP32")
        Pl.set_MG_UNet(net, 2, cycles=1)
        l_ms, l_info = timed(Pl, xl)
        learned_out = {"ms": round(l_ms, 4), "iterations": l_info.iterations,
                       "final_rel_residual": l_info.res_norm / l_info.res_norm0,
                       "placement": "U-net initial guess of multigrid level 2 (32^3) in the first FP32 V-cycle of "
                                    "each solve",
                       "every_vcycle": {"ms": round(e_ms, 4), "iterations": e_info.iterations},
                       "training": {"pairs": len(pairs), "grid": "32^3", "epochs": 30, "seconds": round(t_train, 1),
                                    "sources": "random right-hand sides on the 32^3 level of this mesh; MHD cavity "
                                               "runs with Re in [100, 1000], Ha in [0, 100]",
                                    "loss_first": round(float(hist[0]), 4), "loss_last": round(float(hist[-1]), 4)},
                       "gain_over_mixed": {"ms_without_network": round(mg_ms, 4),
                                           "iterations_without_network": mg_info.iterations,
                                           "speedup": round(mg_ms / l_ms, 3), "pays": bool(l_ms < mg_ms)}}
        # the ceiling of any coarse-level initial guess: the V-cycle's part from
        # level l down cycled K times per visit makes that level's correction
        # nearly exact, what a perfect network placed there would give
        ceiling = {}
        xg = M.SF(m, "CC").set_BCs(types, None, all_Neumann=True)
        Pg = M.PCG_solver_SF(m, xg, M.iter_solver_params(2000, TOL, 1e-300, 1), prec="MG_FP32")
        for lv, K in ((2, 4), (2, 16), (1, 4), (1, 16)):
            Pg.set_MG_gamma(lv, K)
            g_ms, g_info = timed(Pg, xg)
            ceiling[f"level {lv} ({N >> lv}^3), {K} cycles"] = {"ms": round(g_ms, 4), "iterations": g_info.iterations}
        Pg.set_MG_gamma(-1, 1)
        learned_out["coarse_guess_ceiling"] = ceiling
        # the network's dominant kernel: the 16 -> 16 convolution at 32^3
        # (k_conv_tile<1, 1>), 2 x 27 x 16 x 16 flops a voxel, timed in situ
        # over the every-V-cycle solves
        with M.ktimer("unet_conv16") as ku:
            for _ in range(max(1, min(args.steps, 5))):
                xe.fill(0.0)
                Pe.solve(xe, b)
        M.synchronize()
        conv_flop = 2 * 27 * 16 * 16 * 32 ** 3
        conv_tf = conv_flop / (ku.avg_ms * 1e-3) / 1e12 if ku.launches else 0.0
        conv_rp = load_rocprof_avg("ppe_128_bench_trace.txt", "k_conv_tile<1, 1>")
        learned_out["roofline"] = {
            "bound": "mfma", "kernel": "U-net 16->16 3x3x3 convolution at 32^3 (k_conv_tile<1, 1>, BF16 MFMA 16x16x32)",
            "achieved": round(conv_tf, 2), "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(conv_tf / BF16_PEAK_TFLOPS, 5), "flops": conv_flop,
            "in_situ_us": round(1e3 * ku.avg_ms, 3), "in_situ_launches": ku.launches, "rocprof_avg_us": conv_rp,
            "frac_rocprof": round(conv_flop / (conv_rp * 1e-6) / 1e12 / BF16_PEAK_TFLOPS, 5) if conv_rp else None}
        # the same kernel where it is not latency-bound: one application of
        # the network at 128^3 (the solver applies it at 32^3, about a launch's
        # latency of work)
        x128 = np.zeros((4, N + 2, N + 2, N + 2), dtype=np.float32)
        x128[:, 1:-1, 1:-1, 1:-1] = np.random.default_rng(7).standard_normal((4, N, N, N)).astype(np.float32)
        net.forward(x128)
        with M.ktimer("unet_conv16") as k128:
            for _ in range(3):
                net.forward(x128)
        M.synchronize()
        fl128 = 2 * 27 * 16 * 16 * N ** 3
        tf128 = fl128 / (k128.avg_ms * 1e-3) / 1e12 if k128.launches else 0.0
        learned_out["roofline"]["at_128"] = {"achieved": round(tf128, 1), "frac": round(tf128 / BF16_PEAK_TFLOPS, 4),
                                             "in_situ_us": round(1e3 * k128.avg_ms, 2), "flops": fl128}
        del x128

    # every solution against the FP64 Jacobi-PCG converged to 1e-11 (constant removed)
    xr = M.SF(m, "CC").fill(0.0).set_BCs(types, None, all_Neumann=True)
    Pr = M.PCG_solver_SF(m, xr, M.iter_solver_params(20000, 1e-11, 1e-300, 1))
    conv_info = Pr.solve(xr, b)

    def centered(f):
        a = f.download()[1:-1, 1:-1, 1:-1]
        return a - a.mean()

    xc = centered(xr)

    def vs_converged(f):
        return float(np.linalg.norm(centered(f) - xc) / np.linalg.norm(xc))

    agreement = {"reference": f"FP64 Jacobi-PCG to 1e-11 ({conv_info.iterations} iterations)",
                 "plain_fp64": vs_converged(x), "mixed": vs_converged(xm)}
    if Pl is not None:
        agreement["learned_mixed"] = vs_converged(xl)
    xa, xb = x.download()[1:-1, 1:-1, 1:-1], xm.download()[1:-1, 1:-1, 1:-1]
    xa, xb = xa - xa.mean(), xb - xb.mean()
    mg_vs_fp64 = float(np.linalg.norm(xb - xa) / np.linalg.norm(xa))

    # in-situ kernel durations: HIP events on the library stream around every
    # launch of the dominant kernel, over one more pass of the measured
    # method's timed solves (that pass is not in `value`)
    Pmeas, xmeas = (Pl, xl) if Pl is not None else (Pm, xm)
    nk = max(1, min(args.steps, 5))
    with M.ktimer("mg_line_fine") as kt:
        for _ in range(nk):
            xmeas.fill(0.0)
            Pmeas.solve(xmeas, b)
    M.synchronize()
    with M.ktimer("pcg_update") as kf:
        x.fill(0.0)
        P.solve(x, b)
    M.synchronize()

    ncell = N ** 3
    # FP64 PCG kernels, 50 back-to-back launches each (the dispatch gap included)
    kms = P.time_kernels(50)
    alg = {"pcg_matvec": 16 * ncell, "pcg_update": 48 * ncell, "pcg_pupdate": 24 * ncell}
    names = list(alg)
    upd_gbs = alg["pcg_update"] / (kf.avg_ms * 1e-3) / 1e9
    iter_us = 1e3 * sum(kms)
    # k_pcg_update handles 8 cells a work-item; the learned leg's 32^3 solves use it too
    upd_rp = load_rocprof_avg("ppe_128_bench_trace.txt", "k_pcg_update", grid=ncell // 8)
    fp64_roofline = {"bound": "hbm", "kernel": "pcg_update", "achieved": round(upd_gbs, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(upd_gbs / HBM_PEAK_GBS, 4),
                     "traffic": load_traffic("pcg_update_128_traffic.txt"),
                     "in_situ_us": round(1e3 * kf.avg_ms, 3), "in_situ_launches": kf.launches,
                     "rocprof_avg_us": upd_rp,
                     "frac_rocprof": round(alg["pcg_update"] / (upd_rp * 1e-6) / 1e9 / HBM_PEAK_GBS, 4) if upd_rp else None,
                     "back_to_back_us": {n: round(1e3 * t, 2) for n, t in zip(names, kms)},
                     "iteration_GBs": round(88 * ncell / (iter_us * 1e-6) / 1e9, 1)}
    # the mixed solver's dominant kernel: the finest level's zebra line
    # half-sweep (y- and z-lines, 4 (nu1 + nu2) launches per V-cycle), 6 B per
    # cell of the level: the other colour's x (2 B), this colour's f (2 B) and
    # x written (2 B), FP32
    mk = Pm.time_MG(50)
    sm_bytes = 6 * ncell
    sm_gbs = sm_bytes / (kt.avg_ms * 1e-3) / 1e9
    line_rp = load_rocprof_avg("ppe_128_bench_trace.txt", ("k_mg_line<1, 128,", "k_mg_line<2, 128,"))
    roofline = {"bound": "hbm", "kernel": "mg_line (128^3 zebra line half-sweep, y and z)",
                "achieved": round(sm_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(sm_gbs / HBM_PEAK_GBS, 4), "traffic": load_traffic("mg_line_128_traffic.txt"),
                "algorithmic_bytes": sm_bytes, "in_situ_us": round(1e3 * kt.avg_ms, 3),
                "in_situ_launches": kt.launches, "rocprof_avg_us": line_rp,
                "frac_rocprof": round(sm_bytes / (line_rp * 1e-6) / 1e9 / HBM_PEAK_GBS, 4) if line_rp else None,
                "back_to_back_us": {"mg_line_128": round(1e3 * mk[0], 2), "mg_restrict_128_to_64": round(1e3 * mk[1], 2),
                                    "mg_prolong_64_to_128": round(1e3 * mk[2], 2), "mg_tail": round(1e3 * mk[3], 2),
                                    "mg_line_64": round(1e3 * mk[6], 2), "mg_fused_sweep_32": round(1e3 * mk[9], 2)},
                "vcycle_us": round(1e3 * mk[4], 2),
                "mixed_iteration_us": round(1e3 * mg_ms / max(mg_info.iterations, 1), 1)}

    # PCIe-inclusive rate of the host-buffer boundary (upload b, solve, download x)
    t2 = time.perf_counter()
    b.upload(bh)
    x.fill(0.0)
    P.solve(x, b)
    x.download()
    pcie_ms = 1e3 * (time.perf_counter() - t2)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(hs, bh, iters)

    if learned_out:
        method, best, best_iters = "learned_mixed_unet_mg_fp32_in_fp64_fpcg", learned_out["ms"], learned_out["iterations"]
    else:
        method, best, best_iters = "mixed_mg_fp32_in_fp64_fpcg", mg_ms, mg_info.iterations
    out = {
        "metric": METRIC,
        "value": round(best, 4),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(best, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "ppe_128_duct", "cells": "128^3", "mesh": "x periodic uniform; y,z Roberts beta=1.05",
                   "bcs": "homogeneous Neumann walls, all_Neumann", "rhs": "(1/dt) div(U*), U* = 1e-3 SplitMix64 face noise",
                   "tol_rel": TOL, "method": method, "iterations": best_iters},
        "speedup_vs_plain_fp64": round(ms_step / best, 3),
        "plain_fp64_ms": round(ms_step, 4),
        "plain_fp64_iterations": iters,
        "plain_fp64_final_rel_residual": rel,
        "mixed_mg_ms": round(mg_ms, 4),
        "mixed_mg_iterations": mg_info.iterations,
        "mixed_mg_final_rel_residual": mg_info.res_norm / mg_info.res_norm0,
        "mixed_mg_levels": mg_levels,
        "mixed_mg_coarsest_cells": list(mg_coarsest),
        "mixed_vs_fp64_solution_rel_l2": mg_vs_fp64,
        "solution_vs_converged_fp64_rel_l2": agreement,
        "mixed_mg_4levels_2x2_ms": round(mg4_ms, 4),
        "mixed_mg_4levels_2x2_iterations": mg4_info.iterations,
        "speedup_mixed_vs_plain": round(ms_step / mg_ms, 3),
        "learned_mixed": learned_out,
        "pcie_inclusive_ms": round(pcie_ms, 3),
        "roofline": roofline,
        "plain_fp64_roofline": fp64_roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
