"""Benchmark of the second-order FV residual path on one MI355X.

Workload: 3-D Euler flow past a cylinder, O-grid hexes 256 (circumferential) x 128 (radial,
geometric ratio 1.05 from r=0.5 to r=10) x 4 (span, dz = 2*pi*0.5/256) = 131,072 cells,
M = 0.3, second order with the reference limiter, cells renumbered by reverse Cuthill-McKee.
A step is one solver iteration (boundary conditions, LSQ gradient + limiter, flux + local
time-step update, residual norm) with the state resident in HBM.

Prints one JSON line: cell-updates/s of the binary64 solver (value), the roofline of its
dominant kernel measured live with HIP events on the library's stream, the CPU oracle on the
host cores, the mixed-precision throughput and the time to a 1e-3 residual drop.

  python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu] [--no-converge]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

P_INF, T_INF = 1.0, 1.0 / 1.4
MACH = (0.3, 0.0, 0.0)
HBM_PEAK_GBPS = 8000.0
BCS = [("wall", "slip"), ("farfield", "farfield"), ("zmin", "slip"), ("zmax", "slip")]
WORKLOAD = dict(n_theta=256, n_r=128, n_z=4, r0=0.5, r1=10.0, ratio=1.05, dz=2 * math.pi * 0.5 / 256)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def elements(gr):
    w = WORKLOAD
    return gr.Elements.ogrid(w["n_theta"], w["n_r"], w["n_z"], w["r0"], w["r1"], w["ratio"], w["dz"])


def make_solver(gr, m, dtype, cfl):
    sv = gr.Solver(m, dtype=dtype, cfl=cfl)
    for name, kind in BCS:
        sv.set_bc(name, gr.GR_BC_FARFIELD if kind == "farfield" else gr.GR_BC_SLIP, MACH, P_INF, T_INF)
    sv.init_freestream(MACH, P_INF, T_INF)
    return sv


def iteration_bytes(s, su, sg=8):
    """Algorithmic bytes of one iteration in the reference's phase accounting (each array
    element counted once per phase: BC, gradient, limiter, ghost copy, flux+update)."""
    N, G, F = s.n_cells, s.n_ghosts, s.n_faces
    c = (s.n_cell_slots - G) / N
    return su * (75 * N + 85 * G) + sg * (16 * N + 9 * G + 10 * F) + 4 * ((5 * c + 3) * N + 2 * G)


def timed_steps(sv, steps):
    t0 = time.perf_counter()
    res = sv.iterate(steps)          # returns after the stream is synchronised
    t1 = time.perf_counter()
    assert res.iterations == steps, res.iterations
    return t1 - t0


def cpu_baseline(gr, e, budget_s=15.0):
    """The FP64 oracle (the reference algorithm restated in C, Jacobi form) on the host cores."""
    import oracle as orc
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    om = orc.OracleMesh(gr.ElementArrays.of(e))
    s = orc.EulerState.zeros(om.n_cells + om.n_ghosts)
    orc.euler_init(om, s, MACH, P_INF, T_INF)
    bcs = [(om.region(n), orc.BC_FARFIELD if k == "farfield" else orc.BC_SLIP, MACH, P_INF, T_INF) for n, k in BCS]
    orc.iterate(om, s, bcs, 0.5, 1, threads=threads)   # warm-up
    iters, t = 0, 0.0
    while t < budget_s and iters < 200:
        t0 = time.perf_counter()
        orc.iterate(om, s, bcs, 0.5, 2, threads=threads)
        t += time.perf_counter() - t0
        iters += 2
    return dict(value=om.n_cells * iters / t, unit="cell-updates/s", cores=threads, kind="port",
                sample=f"{iters} iterations of the same 131,072-cell O-grid case, binary64, "
                       f"{threads} OpenMP threads, {t:.1f} s")


CGRID_BCS = [("airfoil", "slip"), ("farfield", "farfield"), ("zmin", "slip"), ("zmax", "slip")]


def cgrid_machs(n, seed=2024):
    """Mach ~ U[0.5, 0.85], angle of attack ~ U[-2, 6] degrees (seeded)."""
    rng = np.random.default_rng(seed)
    M = rng.uniform(0.5, 0.85, n)
    a = np.deg2rad(rng.uniform(-2.0, 6.0, n))
    return np.stack([M * np.cos(a), M * np.sin(a), np.zeros(n)], axis=1)


def make_batch(gr, m, dtype, cfl, machs):
    b = gr.Batch(m, len(machs), dtype=dtype, cfl=cfl)
    for name, kind in CGRID_BCS:
        b.set_bc(name, gr.GR_BC_FARFIELD if kind == "farfield" else gr.GR_BC_SLIP)
    b.set_freestream(machs, P_INF, T_INF)
    return b


def batched_leg(gr, args):
    """64 instances on the NACA0012 C-grid (384x96x1 = 36,864 hexes) per launch."""
    log("bench: batched C-grid leg")
    m = gr.Mesh.build(gr.Elements.cgrid_naca())
    m.renumber_rcm()
    m.compute_metrics()
    s = m.sizes()
    machs = cgrid_machs(64)
    res = {"workload": "NACA0012 C-grid 384x96x1 (36,864 hexes) x 64 instances, Mach U[0.5,0.85], "
                       "AoA U[-2,6] deg", "instances": 64, "cells": s.n_cells}
    steps = max(20, args.steps // 4)
    for name, dtype in (("f64", gr.GR_F64), ("mixed", gr.GR_F32)):
        b = make_batch(gr, m, dtype, args.cfl, machs)
        b.iterate(args.warmup)
        t0 = time.perf_counter()
        b.iterate(steps)
        dt = time.perf_counter() - t0
        prof = b.profile(20)
        kb = b.kernel_bytes()
        res[name] = {"value": s.n_cells * 64 * steps / dt, "unit": "cell-updates/s", "ms_per_step": dt / steps * 1e3,
                     "kernels": {k: {"ms": prof[k], "bytes": kb[k], "GBps": kb[k] / (prof[k] * 1e-3) / 1e9}
                                 for k in prof}}
    if not args.no_converge:
        b = make_batch(gr, m, gr.GR_F64, args.cfl, machs)
        t0 = time.perf_counter()
        its, conv, ran = b.iterate(20000, tol=1e-3)
        wall = time.perf_counter() - t0
        per_iter = wall / max(ran, 1)
        amort = its * per_iter / 64.0
        res["time_to_1e-3_f64_cold"] = {"batch_seconds": wall, "batch_iterations": int(ran),
                                        "converged": int(conv.sum()),
                                        "iterations_p50": float(np.percentile(its, 50)),
                                        "iterations_p99": float(np.percentile(its, 99)),
                                        "seconds_per_instance_p50": float(np.percentile(amort, 50)),
                                        "seconds_per_instance_p99": float(np.percentile(amort, 99)),
                                        "normalisation": "residual of iteration 0 per instance"}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--cfl", type=float, default=0.5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-converge", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--profile-iters", type=int, default=100)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the path would shard by mesh partition with a halo exchange per iteration)", file=sys.stderr)
        sys.exit(2)

    import graphite_amd as gr
    if gr.device_count() < 1:
        raise SystemExit("bench.py: no HIP device visible")

    log("bench: building O-grid mesh")
    e = elements(gr)
    m = gr.Mesh.build(e)
    m.renumber_rcm()
    m.compute_metrics()
    s = m.sizes()

    # binary64 solver: timed region
    sv = make_solver(gr, m, gr.GR_F64, args.cfl)
    sv.iterate(args.warmup)
    log(f"bench: timing {args.steps} steps (binary64)")
    dt = timed_steps(sv, args.steps)
    value = s.n_cells * args.steps / dt

    # live kernel durations (HIP events on the library's stream) right after the timed region
    prof = sv.profile(args.profile_iters)
    kb = sv.kernel_bytes()
    dom = max(("grad_lim", "flux_update"), key=lambda k: prof[k])
    achieved = kb[dom] / (prof[dom] * 1e-3) / 1e9
    per_kernel = {k: dict(ms=prof[k], bytes=kb[k], GBps=kb[k] / (prof[k] * 1e-3) / 1e9 if prof[k] > 0 else None)
                  for k in prof}

    # mixed precision (binary32 state/gradient/limiter/flux, binary64 metrics and norms)
    sv32 = make_solver(gr, m, gr.GR_F32, args.cfl)
    sv32.iterate(args.warmup)
    dt32 = timed_steps(sv32, args.steps)
    prof32 = sv32.profile(args.profile_iters)
    kb32 = sv32.kernel_bytes()

    out = {
        "metric": "cell-updates/s + %HBM peak; time-to-1e-3 residual, warm-start vs plain",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "cylinder O-grid 256x128x4 hexes (131,072 cells), 3-D Euler M=0.3, "
                               "second order + reference limiter, plain FP64 solver",
                   "cells": s.n_cells, "ghosts": s.n_ghosts, "faces": s.n_faces, "cfl": args.cfl},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBPS, "traffic": None,
                     "algorithmic_bytes_per_launch": kb[dom], "avg_launch_ms": prof[dom]},
        "kernels": per_kernel,
        "iteration_bytes_reference_phases": iteration_bytes(s, 8),
        "iteration_GBps_reference_phases": value * iteration_bytes(s, 8) / s.n_cells / 1e9,
        "mixed_precision": {"value": s.n_cells * args.steps / dt32, "ms_per_step": dt32 / args.steps * 1e3,
                            "flux_update_GBps": kb32["flux_update"] / (prof32["flux_update"] * 1e-3) / 1e9,
                            "grad_lim_GBps": kb32["grad_lim"] / (prof32["grad_lim"] * 1e-3) / 1e9,
                            "iteration_GBps_reference_phases":
                                s.n_cells * args.steps / dt32 * iteration_bytes(s, 4) / s.n_cells / 1e9},
    }

    # achievable HBM rate of a plain copy kernel (2 x 1 GiB per launch, far beyond the caches)
    out["stream_copy_GBps"] = gr.stream_copy_gbps(1 << 30, 20)

    if not args.no_batch:
        out["batched"] = batched_leg(gr, args)

    if not args.no_converge:
        log("bench: time to a 1e-3 residual drop (binary64, cold start)")
        for name, dtype in (("f64", gr.GR_F64), ("mixed", gr.GR_F32)):
            svc = make_solver(gr, m, dtype, args.cfl)
            t0 = time.perf_counter()
            r = svc.iterate(20000, tol=1e-3)
            t1 = time.perf_counter()
            out[f"time_to_1e-3_{name}"] = {"seconds": t1 - t0, "iterations": r.iterations,
                                            "converged": bool(r.converged), "ratio": r.ratio,
                                            "normalisation": "residual of iteration 0 (reference semantics)"}
    if not args.no_cpu:
        log("bench: CPU oracle baseline")
        out["cpu_baseline"] = cpu_baseline(gr, e)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
