"""Benchmark of the second-order FV residual path on one MI355X.

Workload of `value`: the transonic swept wing -- NACA0012 section, 30 deg leading-edge sweep,
taper 0.56, semispan 1.5 root chords; 20 prism layers (growth 1.2) on the wall and tetrahedra
out to a half sphere of 25 chords (1,069,112 cells), lengths scaled by 0.01; M 0.80, AoA 2.5
deg; second order with the reference limiter, mixed precision (binary32 state, gradient,
limiter and flux; binary64 metrics, least-squares inverse and residual norms); cells
renumbered along a Hilbert curve of their centres.  A step is one solver iteration (boundary conditions, LSQ
gradient + limiter, flux + local time-step update, residual norm) with the state resident in
HBM.

Prints one JSON line: cell-updates/s of that solver (value), the roofline of its dominant
kernel measured live with HIP events on the library's stream, the wing in binary64 and its
solve to a 1e-3 residual drop with lift and drag, the CPU oracle on the wing (all host cores
and one core), and the other cases: the cylinder O-grid in both precisions with its mixed-
precision error at the 1e-3 point, the batched C-grid (64 flows per launch), the learned warm
start against cold batched and cold single-flow solves, and the incompressible 64^3 cavity
solved to its continuity criterion.

  python bench.py [--steps K] [--warmup W] [--gpus 1] [--only-headline] [--only-nsi] [--only-mlp] [--no-cpu]
                  [--no-converge]
                  [--no-ogrid] [--no-batch] [--no-warm] [--no-nsi]
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


def to_json(obj) -> str:
    """One JSON line; numpy scalars become Python numbers."""
    return json.dumps(obj, default=lambda o: o.item() if hasattr(o, "item") else str(o))


def elements(gr):
    w = WORKLOAD
    return gr.Elements.ogrid(w["n_theta"], w["n_r"], w["n_z"], w["r0"], w["r1"], w["ratio"], w["dz"])


def make_solver(gr, m, dtype, cfl, tune=True):
    """Solver of the O-grid case; with tune, the tile size is chosen by the autotuner."""
    sv = gr.Solver(m, dtype=dtype, cfl=cfl)
    for name, kind in BCS:
        sv.set_bc(name, gr.GR_BC_FARFIELD if kind == "farfield" else gr.GR_BC_SLIP, MACH, P_INF, T_INF)
    sv.init_freestream(MACH, P_INF, T_INF)
    sv.tile_cells = sv.autotune(20) if tune else 51
    return sv


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` measured with rocprofv3 (FETCH_SIZE, WRITE_SIZE, calibrated
    against a copy kernel), from the summaries under profiles/; None if none records it."""
    d = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles")
    for name in sorted(os.listdir(d)) if os.path.isdir(d) else []:
        with open(os.path.join(d, name)) as fh:
            for line in fh:
                f = line.split()
                if len(f) == 3 and f[0] == "traffic_bytes_per_launch" and f[1] == kernel:
                    return float(f[2])
    return None


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


def batched_leg(gr, args):
    """64 instances on the NACA0012 C-grid (384x96x1 = 36,864 hexes) per launch."""
    from graphite_amd import warmstart as ws
    log("bench: batched C-grid leg")
    m = gr.Mesh.build(gr.Elements.cgrid_naca())
    m.renumber_hilbert()
    m.compute_metrics()
    s = m.sizes()
    machs = ws.sample_instances(64, seed=2024)
    res = {"workload": "NACA0012 C-grid 384x96x1 (36,864 hexes) x 64 instances, Mach U[0.5,0.85], "
                       "AoA U[-2,6] deg", "instances": 64, "cells": s.n_cells}
    steps = max(20, args.steps // 4)
    for name, dtype in (("f64", gr.GR_F64), ("mixed", gr.GR_F32)):
        b = ws.make_batch(m, machs, dtype, args.cfl)
        schedule = b.autotune(10)
        b.iterate(args.warmup)
        t0 = time.perf_counter()
        b.iterate(steps)
        dt = time.perf_counter() - t0
        prof = b.profile(20)
        kb = b.kernel_bytes()
        res[name] = {"value": s.n_cells * 64 * steps / dt, "unit": "cell-updates/s", "ms_per_step": dt / steps * 1e3,
                     "schedule": schedule, "kernels": kernel_table(prof, kb, f"batch_{name}")}
    return res


def warm_start_leg(gr, args):
    """Learned warm start on the C-grid: 512 training flows generated by the batched solver,
    BF16-MFMA training, then 64 new flows to 1e-3 of their freestream residual: warm-started
    batched vs cold batched vs cold single-flow solves; lift/drag after equal further work."""
    from graphite_amd import warmstart as ws
    log("bench: warm-start leg (training set)")
    m = gr.Mesh.build(gr.Elements.cgrid_naca())
    m.renumber_hilbert()
    m.compute_metrics()
    wd = ws.wall_distance(m)
    N = m.n_cells
    t0 = time.perf_counter()
    n_flows = args.warm_flows
    x, t, origin, _ = ws.generate_training_set(m, ws.sample_instances(n_flows, seed=1), tol=1e-5, max_iter=6000,
                                               dtype=gr.GR_F32)
    t_data = time.perf_counter() - t0
    # oversample rows of cells within half a chord of the wall (x3)
    rows, per = [], 64
    for b0 in range(0, n_flows, per):
        nb = min(per, n_flows - b0)
        base = b0 * N
        cells = np.repeat(np.arange(N), nb)
        r = base + np.arange(N * nb)
        near = r[wd[cells] < 0.5]
        rows += [r, near, near]
    idx = np.concatenate(rows)
    log("bench: warm-start leg (training)")
    net, norm, losses, t_train = ws.train(x[idx], t[idx], origin, steps=args.warm_steps, batch=32768, lr=2e-3,
                                          lr_final=2e-5)
    test = ws.sample_instances(64, seed=99)
    log("bench: warm-start leg (cold and warm solves)")
    cold = ws.make_batch(m, test, tune=True)
    fs = cold.residual_norms()
    cs = ws.solve(cold, 1e-3, fs, 20000)
    warm = ws.make_batch(m, test, tune=True)
    t0 = time.perf_counter()
    _, used = warm.warm_start(net, wd, norm)
    t_pred = time.perf_counter() - t0
    wsol = ws.solve(warm, 1e-3, fs, 20000)
    t_warm = t_pred + wsol.seconds
    # cold single-flow solves (the plain solver, one flow at a time)
    plain = []
    for k in range(64):
        sv = gr.Solver(m, gr.GR_F64, cfl=args.cfl)
        for name, kind in ws.CGRID_BCS:
            sv.set_bc(name, kind, test[k], ws.P_INF, ws.T_INF)
        sv.init_freestream(test[k], ws.P_INF, ws.T_INF)
        t1 = time.perf_counter()
        r = sv.iterate(20000, tol=1e-3)
        plain.append((time.perf_counter() - t1, r.iterations, bool(r.converged)))
    t_plain = sum(p[0] for p in plain)
    s_ref = 0.1                                     # chord 1 x extrusion depth 0.1

    def coeffs(b):
        F = b.force("airfoil")
        return np.array([gr.lift_drag(F[k], test[k], 1.4, ws.P_INF, s_ref) for k in range(64)])
    c3, w3 = coeffs(cold), coeffs(warm)
    log(f"bench: warm-start leg (converged lift/drag: {args.warm_settle} + {args.warm_window} further iterations)")
    agree = ws.force_agreement(cold, warm, test, "airfoil", s_ref, args.warm_settle, args.warm_window, args.warm_sample,
                               log=log)
    rel = lambda a, b: (np.abs(a - b).max(axis=0) / np.abs(b[:, 0]).max()).tolist()
    it_w, it_c = wsol.iterations, cs.iterations
    forces = warm_forces(gr, args, m, net, norm, test, s_ref) if args.warm_forces else None
    return {
        "time_to_settled_forces": forces,
        "workload": "NACA0012 C-grid 384x96x1, 64 new flows (Mach U[0.5,0.85], AoA U[-2,6] deg), binary64, "
                    "to 1e-3 of each flow's freestream residual",
        "training": {"flows": n_flows, "data_seconds": t_data, "rows": int(len(idx)), "steps": args.warm_steps,
                     "batch": 32768,
                     "train_seconds": t_train, "loss_first": float(losses[0]), "loss_last": float(losses[-1]),
                     "network": "4 layers 6-256-256-256-5, tanh, bf16 MFMA (v_mfma_f32_16x16x32_bf16), fp32 accumulation"},
        "predictions_kept": int(used.sum()),
        "prediction_seconds": t_pred,
        "seconds": {"warm_batched": t_warm, "cold_batched": cs.seconds, "cold_single_flows": t_plain},
        "speedup_vs_cold_single_flows": t_plain / t_warm,
        "speedup_vs_cold_batched": cs.seconds / t_warm,
        "iterations": {"warm_p50": float(np.percentile(it_w, 50)), "warm_p99": float(np.percentile(it_w, 99)),
                       "cold_p50": float(np.percentile(it_c, 50)), "cold_p99": float(np.percentile(it_c, 99)),
                       "warm_batch": wsol.batch_iterations, "cold_batch": cs.batch_iterations},
        "seconds_per_flow": {"cold_single_p50": float(np.percentile([p[0] for p in plain], 50)),
                             "cold_single_p99": float(np.percentile([p[0] for p in plain], 99)),
                             "warm_batched_amortised": t_warm / 64, "cold_batched_amortised": cs.seconds / 64},
        "converged": {"warm": int(wsol.converged.sum()), "cold": int(cs.converged.sum()),
                      "single": int(sum(p[2] for p in plain))},
        "lift_drag": {"s_ref": s_ref,
                      "rel_diff_at_1e-3_warm_vs_cold": rel(w3, c3),
                      "rel_diff_at_1e-3_scale": "largest |C_L| of the sweep",
                      "converged_window_means": agree},
        "normalisation": "residual of the freestream state of each flow, for warm and cold starts alike",
    }


def warm_forces(gr, args, m, net, norm, test, s_ref):
    """Time to settled forces: cold (freestream) and warm-started batches of the 64 test flows
    run args.force_horizon iterations with lift and drag sampled every 100; the converged value
    of each flow is the cold run's mean over its last args.force_window iterations (the scheme
    ends in a limit cycle); for each flow and start, the iterations until the trailing-window
    mean of both coefficients is within 1e-4 (and 1e-3) of it and stays there."""
    from graphite_amd import warmstart as ws
    every, H, W = 100, args.force_horizon, args.force_window
    dtype = gr.GR_F32 if args.force_mixed else gr.GR_F64
    res = {"horizon_iterations": H, "window_iterations": W, "sample_every": every,
           "dtype": "mixed" if args.force_mixed else "f64",
           "criterion": "trailing-window mean of C_L and C_D within tol of the cold run's mean over its last window, "
                        "relative to max(|C|, 0.05 for lift, 0.005 for drag), from then on"}
    if args.force_train_iters > 0:
        # a network trained on settled solutions: every training flow run force_train_iters
        # iterations (no residual stop), then the same training as the leg's network
        log(f"bench: settled training set ({args.warm_flows} flows x {args.force_train_iters} iterations)")
        t0 = time.perf_counter()
        x, t, origin, _ = ws.generate_training_set(m, ws.sample_instances(args.warm_flows, seed=1), tol=1e-30,
                                                   max_iter=args.force_train_iters, dtype=gr.GR_F32, log=log)
        t_data = time.perf_counter() - t0
        net, norm, losses, t_train = ws.train(x, t, origin, steps=args.warm_steps, batch=32768, lr=2e-3,
                                              lr_final=2e-5)
        res["settled_training"] = {"flows": args.warm_flows, "iterations": args.force_train_iters,
                                   "data_seconds": t_data, "train_seconds": t_train,
                                   "loss_first": float(losses[0]), "loss_last": float(losses[-1])}
        del x, t
    cold = ws.make_batch(m, test, dtype, tune=True)
    warm = ws.make_batch(m, test, dtype, tune=True)
    t0 = time.perf_counter()
    warm.warm_start(net, ws.wall_distance(m), norm)
    t_pred = time.perf_counter() - t0
    sc, secs_c, rc = ws.force_history(cold, test, "airfoil", s_ref, H, every, log=log, tag=" cold")
    sw, secs_w, rw = ws.force_history(warm, test, "airfoil", s_ref, H, every, log=log, tag=" warm")
    ref = sc[-(W // every):].mean(0)
    per_iter_c, per_iter_w = secs_c / H, secs_w / H
    for tol in (1e-3, 1e-4):
        tc = ws.time_to_forces(sc, ref, every, W // 4, tol)
        tw = ws.time_to_forces(sw, ref, every, W // 4, tol)
        both = (tc > 0) & (tw > 0)
        key = f"tol_{tol:.0e}"
        res[key] = {"flows_met_cold": int((tc > 0).sum()), "flows_met_warm": int((tw > 0).sum()),
                    "iterations_cold_p50_p99": [float(np.percentile(tc[tc > 0], q)) if (tc > 0).any() else None for q in (50, 99)],
                    "iterations_warm_p50_p99": [float(np.percentile(tw[tw > 0], q)) if (tw > 0).any() else None for q in (50, 99)],
                    "speedup_batch_wall": (float(tc[both].max() * per_iter_c) / float(tw[both].max() * per_iter_w + t_pred))
                    if both.any() else None,
                    "speedup_per_flow_p50": float(np.median(tc[both] / tw[both])) if both.any() else None}
        log(f"bench: time to forces within {tol:.0e}: {res[key]}")
    # settled apart or still drifting: the warm-cold difference of the final-window means against
    # each run's own change between its last two windows (relative to max(|C|, floor), the larger
    # of lift and drag)
    wn = W // every
    if sc.shape[0] >= 2 * wn:
        scale = np.maximum(np.abs(ref), np.array([0.05, 0.005]))[None]
        mw, mc_prev, mw_prev = sw[-wn:].mean(0), sc[-2 * wn:-wn].mean(0), sw[-2 * wn:-wn].mean(0)
        d = (np.abs(mw - ref) / scale[0]).max(1)
        drift = np.maximum((np.abs(ref - mc_prev) / scale[0]).max(1), (np.abs(mw - mw_prev) / scale[0]).max(1))
        pct = lambda v: [float(np.median(v)), float(v.max())]
        res["final_windows"] = {
            "warm_vs_cold_rel_diff_p50_max": pct(d), "own_drift_p50_max": pct(drift),
            "flows_apart_by_more_than_1e-3": int((d > 1e-3).sum()),
            "flows_apart_by_3x_their_drift_and_1e-3": int(((d > 1e-3) & (d > 3 * drift)).sum()),
            "definition": "d = |warm - cold| of the last-window means, drift = larger change of either run's mean "
                          "between its last two windows; both relative to max(|C|, 0.05 lift / 0.005 drag)"}
        log(f"bench: final windows: {res['final_windows']}")
    res["restarted"] = {"cold": int(rc.sum()), "warm": int(rw.sum())}
    res["batch_seconds_per_iteration"] = {"cold": per_iter_c, "warm": per_iter_w}
    return res


def mlp_leg(gr, args):
    """The warm-start network's BF16-MFMA kernels on their own: training steps (batch 32,768 rows
    of a resident synthetic set of 2^21 rows) and inference of 2^21 rows, FLOP rates of the
    products the MFMAs execute (padded shapes; the first layer's forward product over 32 input
    columns, hidden 256 x 3, out 32: 2 x (32*256 + 2*256*256 + 256*32) = 294,912 FLOP per row
    forward; a training step adds the input-gradient products of layers 4..2,
    2 x (32*256 + 2*256*256), and the weight-gradient products of all four layers,
    2 x (64*256 + 2*256*256 + 256*32): 884,736 FLOP per row)."""
    log("bench: MLP leg")
    rng = np.random.default_rng(0)
    rows, batch = 1 << 21, 32768
    x = rng.standard_normal((rows, 6)).astype(np.float32)
    t = np.tanh(x[:, :5] * 0.5 + 0.1 * x[:, 1:6]).astype(np.float32)
    net = gr.Mlp(6, 5, seed=0)
    net.set_data(x, t)
    net.train(batch, 20, 1e-3, 1e-3)                      # warm-up
    steps = args.mlp_steps
    t0 = time.perf_counter()
    net.train(batch, steps, 1e-3, 1e-4)                   # returns after the losses are copied back
    dt = time.perf_counter() - t0
    fwd_row = 2 * (32 * 256 + 2 * 256 * 256 + 256 * 32)
    train_row = fwd_row + 2 * (32 * 256 + 2 * 256 * 256) + 2 * (64 * 256 + 2 * 256 * 256 + 256 * 32)
    t1 = time.perf_counter()
    net.forward(x)
    dt_inf = time.perf_counter() - t1
    peak = 2.5e15
    return {"workload": "MLP 6-256-256-256-5 (padded 64-256-256-256-32), bf16 MFMA with fp32 accumulation",
            "train": {"batch": batch, "steps": steps, "ms_per_step": dt / steps * 1e3,
                      "TFLOPs": train_row * batch * steps / dt / 1e12,
                      "frac_of_bf16_dense_peak": train_row * batch * steps / dt / peak},
            "inference": {"rows": rows, "seconds": dt_inf, "TFLOPs": fwd_row * rows / dt_inf / 1e12,
                          "frac_of_bf16_dense_peak": fwd_row * rows / dt_inf / peak,
                          "note": "includes the host-to-device copy of the inputs and the copy back"}}


def nsi_leg(gr, args):
    """Lid-driven cavity 64^3 (262,144 hexes), Re 400, lid velocity 1: steps and seconds to a
    continuity ratio of 1e-4 (continuity residual over the largest of the first five steps), in
    mixed precision (binary64 velocity, pressure and face fluxes; binary32 gradients, increments
    and CG vectors) and in binary64."""
    log("bench: incompressible cavity leg")
    n = 64
    m = gr.Mesh.build(gr.Elements.box(n, n, n))
    m.renumber_rcm()
    m.compute_metrics()
    res = {"workload": "lid-driven cavity 64^3 hexes (262,144 cells), Re 400, lid velocity 1, dt 0.005, "
                       "Jacobi-PCG to 1e-4 of |b| per step, to a continuity ratio of 1e-4",
           "cells": m.n_cells,
           "ratio_definition": "sum over cells of |net outflow of the predicted face fluxes| of a step over the "
                               "largest of the first five steps"}
    for name, dtype in (("mixed", gr.GR_F32), ("f64", gr.GR_F64)):
        ns = gr.Nsi(m, nu=1.0 / 400.0, dt=0.005, dtype=dtype, cg_max=500, cg_tol=1e-4)
        ns.set_wall("ymax", (1.0, 0.0, 0.0))
        t0 = time.perf_counter()
        steps, ratio = ns.iterate(args.nsi_steps, 1e-4)
        dt = time.perf_counter() - t0
        _, its = ns.history(steps)
        prof, cgb = ns.profile(20), ns.cg_bytes()
        kernels = {k: {"ms": v, "bytes": cgb.get(k), "GBps": cgb[k] / (v * 1e-3) / 1e9 if k in cgb else None,
                       "frac_of_peak": cgb[k] / (v * 1e-3) / 1e9 / HBM_PEAK_GBPS if k in cgb else None}
                   for k, v in prof.items()}
        res[name] = {"steps": steps, "seconds": dt, "ratio": ratio, "converged": bool(ratio < 1e-4),
                     "kernels": kernels,
                     "steps_per_second": steps / dt, "cell_updates_per_second": m.n_cells * steps / dt,
                     "cg_iterations_mean": float(np.mean(its)), "cg_iterations_total": int(np.sum(its)),
                     "us_per_cg_iteration": dt / max(1, int(np.sum(its))) * 1e6}
        log(f"bench: cavity {name}: {steps} steps to ratio {ratio:.2e} in {dt:.2f} s")
    return res


def cpu_model():
    """Model name of the host CPU (from /proc/cpuinfo; None where it is not readable)."""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def wing_cpu_baseline(gr, e, mach, budget_s=12.0):
    """The FP64 oracle (the reference algorithm restated in C, Jacobi form) on the swept wing,
    timed on all host cores and on one core: a bounded sample of iterations each."""
    import oracle as orc
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    log("bench: CPU oracle baseline (building the oracle's mesh of the wing)")
    om = orc.OracleMesh(gr.ElementArrays.of(e))
    s = orc.EulerState.zeros(om.n_cells + om.n_ghosts)
    orc.euler_init(om, s, mach, P_INF, T_INF)
    bcs = [(om.region(n), orc.BC_FARFIELD if k == "farfield" else orc.BC_SLIP, mach, P_INF, T_INF)
           for n, k in WING_BCS]
    res = {}
    for tag, th, cap in (("all", threads, 40), ("one", 1, 6)):
        orc.iterate(om, s, bcs, 0.5, 1, threads=th)   # warm-up
        iters, t = 0, 0.0
        while t < budget_s and iters < cap:
            t0 = time.perf_counter()
            orc.iterate(om, s, bcs, 0.5, 1, threads=th)
            t += time.perf_counter() - t0
            iters += 1
        res[tag] = (om.n_cells * iters / t, iters, t, th)
        log(f"bench: CPU oracle {th} thread(s): {iters} iterations in {t:.1f} s")
    v, iters, t, th = res["all"]
    v1, iters1, t1, _ = res["one"]
    return dict(value=v, unit="cell-updates/s", cores=th, kind="port",
                sample=f"{iters} iterations of the same 1.07 M-cell swept-wing case from freestream, binary64, "
                       f"{th} OpenMP threads, {t:.1f} s",
                single_core={"value": v1, "unit": "cell-updates/s", "cores": 1,
                             "sample": f"{iters1} iterations, 1 thread, {t1:.1f} s"},
                cpu_model=cpu_model(), host_cpus=os.cpu_count())


WING_SCALE = 0.01
WING_MACH, WING_AOA = 0.8, 2.5
WING_BCS = (("wall", "slip"), ("symmetry", "slip"), ("farfield", "farfield"))


def wing_mach():
    a = math.radians(WING_AOA)
    return (WING_MACH * math.cos(a), WING_MACH * math.sin(a), 0.0)


def wing_solver(gr, m, dtype, cfl, tune_iters=10):
    mach = wing_mach()
    sv = gr.Solver(m, dtype=dtype, cfl=cfl)
    for name, kind in WING_BCS:
        sv.set_bc(name, gr.GR_BC_FARFIELD if kind == "farfield" else gr.GR_BC_SLIP, mach, P_INF, T_INF)
    sv.init_freestream(mach, P_INF, T_INF)
    sv.tile_cells = sv.autotune(tune_iters)
    return sv


def shard_leg(gr, e, m, c2o, steps, parts=8):
    """The wing cut into `parts` contiguous ranges of its Hilbert order, as `parts` GPUs would own
    it: each range's cells with its boundary faces and a halo region for the faces to the other
    ranges (ghosts without a boundary condition, read from the state array where an exchange
    would write them).  Each range timed alone in mixed precision; the slowest sets the
    iteration time of a partitioned run (before any exchange)."""
    from graphite_amd.partition import shard_elements
    log(f"bench: swept wing in {parts} shards")
    n = m.n_cells
    bounds = [n * k // parts for k in range(parts + 1)]
    mach = wing_mach()
    rows = []
    for k in range(parts):
        sm = gr.Mesh.build(shard_elements(e, m, c2o, bounds[k], bounds[k + 1]))
        sm.renumber_hilbert()
        sm.compute_metrics()
        ss = sm.sizes()
        halo = int(np.diff(sm.connectivity()["ghost_cell_starts"])[sm.region_id("halo")]) \
            if "halo" in sm.region_names() else 0
        sv = gr.Solver(sm, dtype=gr.GR_F32, cfl=0.5)
        for name, kind in WING_BCS:
            if name in sm.region_names():
                sv.set_bc(name, gr.GR_BC_FARFIELD if kind == "farfield" else gr.GR_BC_SLIP, mach, P_INF, T_INF)
        sv.init_freestream(mach, P_INF, T_INF)
        tile = sv.autotune(10)
        sv.iterate(11)
        dt = timed_steps(sv, steps)
        prof = sv.profile(20)
        rows.append({"cells": ss.n_cells, "halo_faces": halo, "ghosts": ss.n_ghosts, "tile_cells": tile,
                     "ms_per_step": dt / steps * 1e3, "grad_lim_ms": prof["grad_lim"],
                     "flux_update_ms": prof["flux_update"]})
        del sv, sm
    slowest = max(r["ms_per_step"] for r in rows)
    return {"parts": parts, "partition": "contiguous ranges of the Hilbert-curve cell order, one halo layer",
            "shards": rows, "slowest_ms_per_step": slowest}


def kernel_table(prof, kb, traffic_tag=None, lb=None):
    """Per kernel: mean launch (ms), algorithmic bytes and rate, measured HBM traffic; with lb
    (the tile solver's layout bytes: its per-slot records and padded state records, each element
    once) also those and the measured traffic over them, the share of data fetched twice."""
    out = {}
    for k in prof:
        tr = measured_traffic(f"{traffic_tag}:{k}") if traffic_tag else None
        out[k] = {"ms": prof[k], "bytes": kb[k],
                  "GBps": kb[k] / (prof[k] * 1e-3) / 1e9 if prof[k] > 0 else None,
                  "frac_of_peak": kb[k] / (prof[k] * 1e-3) / 1e9 / HBM_PEAK_GBPS if prof[k] > 0 else None,
                  "traffic": tr}
        if lb is not None:
            out[k]["layout_bytes"] = lb[k]
            out[k]["traffic_over_layout"] = tr / lb[k] if tr and lb[k] > 0 else None
    return out


def field_errors(a, b, n):
    """Mixed precision (a) against binary64 (b) per field over the first n cells: max |a-b| over
    max |b| (max-rel) and ||a-b|| / ||b|| (L2)."""
    out = {}
    for name, x, y in zip(("rho", "rhou", "rhoe"), a, b):
        x, y = np.asarray(x)[:n], np.asarray(y)[:n]
        out[name] = {"max_rel": float(np.abs(x - y).max() / np.abs(y).max()),
                     "l2_rel": float(np.linalg.norm(x - y) / np.linalg.norm(y))}
    return out


def ogrid_leg(gr, args):
    """Cylinder O-grid (C2): binary64 and mixed-precision throughput, time to a 1e-3 drop and the
    mixed-precision per-field error at that point."""
    log("bench: O-grid leg")
    m = gr.Mesh.build(elements(gr))
    m.renumber_hilbert()
    m.compute_metrics()
    s = m.sizes()
    res = {"workload": "cylinder O-grid 256x128x4 hexes (131,072 cells), M 0.3, second order + reference limiter",
           "cells": s.n_cells, "ghosts": s.n_ghosts, "faces": s.n_faces}
    for name, dtype in (("f64", gr.GR_F64), ("mixed", gr.GR_F32)):
        sv = make_solver(gr, m, dtype, args.cfl)
        sv.iterate(args.warmup + 1)
        dt = timed_steps(sv, args.steps)
        prof = sv.profile(args.profile_iters)
        kb = sv.kernel_bytes()
        su = 8 if dtype == gr.GR_F64 else 4
        res[name] = {"value": s.n_cells * args.steps / dt, "unit": "cell-updates/s", "ms_per_step": dt / args.steps * 1e3,
                     "tile_cells": sv.tile_cells, "kernels": kernel_table(prof, kb, f"ogrid_{name}", sv.layout_bytes()),
                     "iteration_bytes_reference_phases": iteration_bytes(s, su),
                     "iteration_GBps_reference_phases": s.n_cells * args.steps / dt * iteration_bytes(s, su) / s.n_cells / 1e9}
    if not args.no_converge:
        states = {}
        for name, dtype in (("f64", gr.GR_F64), ("mixed", gr.GR_F32)):
            svc = make_solver(gr, m, dtype, args.cfl)
            t0 = time.perf_counter()
            r = svc.iterate(20000, tol=1e-3)
            t1 = time.perf_counter()
            res[f"time_to_1e-3_{name}"] = {"seconds": t1 - t0, "iterations": r.iterations, "converged": bool(r.converged),
                                           "ratio": r.ratio, "normalisation": "residual of iteration 0 (reference semantics)"}
            states[name] = svc.get_state()
        res["mixed_precision_error_at_1e-3"] = {
            "fields": field_errors(states["mixed"], states["f64"], s.n_cells),
            "definition": "each solve run to its own 1e-3 drop; max_rel = max|mixed-f64| / max|f64|, "
                          "l2_rel = ||mixed-f64|| / ||f64|| over the interior cells"}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--cfl", type=float, default=0.5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-converge", action="store_true")
    ap.add_argument("--no-ogrid", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--profile-iters", type=int, default=100)
    ap.add_argument("--no-warm", action="store_true")
    ap.add_argument("--no-nsi", action="store_true")
    ap.add_argument("--no-shards", action="store_true", help="skip timing the wing's 8 shards")
    ap.add_argument("--only-headline", action="store_true", help="the swept-wing headline alone (profiling)")
    ap.add_argument("--only-nsi", action="store_true", help="the cavity leg alone (profiling)")
    ap.add_argument("--only-mlp", action="store_true", help="the warm-start network's kernels alone (profiling)")
    ap.add_argument("--mlp-steps", type=int, default=200)
    ap.add_argument("--wing-iters", type=int, default=6000)
    ap.add_argument("--nsi-steps", type=int, default=12000, help="step limit of the cavity solve to its criterion")
    ap.add_argument("--warm-flows", type=int, default=512)
    ap.add_argument("--warm-steps", type=int, default=8000)
    ap.add_argument("--warm-settle", type=int, default=20000,
                    help="further iterations of both warm and cold flows before their forces are compared")
    ap.add_argument("--warm-window", type=int, default=20000, help="iterations the compared forces are averaged over")
    ap.add_argument("--warm-sample", type=int, default=500, help="iterations between force samples in the window")
    ap.add_argument("--warm-forces", action="store_true",
                    help="also measure the iterations to settled lift/drag, warm vs cold (long)")
    ap.add_argument("--force-horizon", type=int, default=140000)
    ap.add_argument("--force-window", type=int, default=40000)
    ap.add_argument("--force-train-iters", type=int, default=0,
                    help="with --warm-forces: train another network on flows run this many iterations (settled)")
    ap.add_argument("--force-mixed", action="store_true", help="with --warm-forces: batches in mixed precision")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the path would shard by mesh partition with a halo exchange per iteration)", file=sys.stderr)
        sys.exit(2)

    import graphite_amd as gr
    if gr.device_count() < 1:
        raise SystemExit("bench.py: no HIP device visible")
    if args.only_nsi:
        print(to_json({"nsi_cavity": nsi_leg(gr, args)}), flush=True)
        return
    if args.only_mlp:
        print(to_json({"mlp": mlp_leg(gr, args)}), flush=True)
        return

    log("bench: building the swept-wing mesh")
    e = gr.Elements.wing(scale=WING_SCALE)
    m = gr.Mesh.build(e)
    c2o, _ = m.renumber_hilbert()
    m.compute_metrics()
    s = m.sizes()
    mach = wing_mach()

    # headline: mixed-precision solver on the wing, both graph parities captured in the warm-up
    sv = wing_solver(gr, m, gr.GR_F32, args.cfl)
    sv.iterate(max(args.warmup, 1))
    sv.iterate(1)
    log(f"bench: timing {args.steps} steps (swept wing, mixed precision)")
    dt = timed_steps(sv, args.steps)
    value = s.n_cells * args.steps / dt

    # live kernel durations (HIP events on the library's stream) right after the timed region
    prof = sv.profile(args.profile_iters)
    kb = sv.kernel_bytes()
    dom = max(("grad_lim", "flux_update"), key=lambda k: prof[k])
    achieved = kb[dom] / (prof[dom] * 1e-3) / 1e9

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
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "transonic swept wing: NACA0012 section, 30 deg sweep, taper 0.56, semispan 1.5 "
                               "chords, 20 prism layers (growth 1.2) + tetrahedra in a half sphere of 25 chords, "
                               "lengths x0.01, cells in Hilbert-curve order; M 0.80, AoA 2.5 deg, second order + reference limiter; mixed "
                               "precision (binary32 state/gradient/limiter/flux, binary64 metrics, LSQ inverse "
                               "and residual norms)",
                   "cells": s.n_cells, "ghosts": s.n_ghosts, "faces": s.n_faces, "cfl": args.cfl,
                   "tile_cells": sv.tile_cells, "step": "one solver iteration (BCs, LSQ gradient + limiter, "
                                                        "flux + local time-step update, residual norm)"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBPS, "traffic": measured_traffic(f"wing_mixed:{dom}"),
                     "algorithmic_bytes_per_launch": kb[dom], "avg_launch_ms": prof[dom]},
        "kernels": kernel_table(prof, kb, "wing_mixed", sv.layout_bytes()),
        "iteration_bytes_reference_phases": iteration_bytes(s, 4),
        "iteration_GBps_reference_phases": value * iteration_bytes(s, 4) / s.n_cells / 1e9,
    }
    if not args.only_headline:
        wing = {}
        sv64 = wing_solver(gr, m, gr.GR_F64, args.cfl)
        sv64.iterate(max(args.warmup, 1))
        sv64.iterate(1)
        steps64 = max(20, args.steps // 2)
        dt64 = timed_steps(sv64, steps64)
        wing["f64"] = {"value": s.n_cells * steps64 / dt64, "unit": "cell-updates/s", "ms_per_step": dt64 / steps64 * 1e3,
                       "tile_cells": sv64.tile_cells,
                       "kernels": kernel_table(sv64.profile(20), sv64.kernel_bytes(), "wing_f64",
                                               sv64.layout_bytes())}
        del sv64
        if not args.no_converge:
            log("bench: swept wing to a 1e-3 residual drop (mixed precision, cold start)")
            svc = wing_solver(gr, m, gr.GR_F32, args.cfl)
            svc.iterate(1)   # captures both graph parities; the solve below restarts from freestream
            svc.init_freestream(mach, P_INF, T_INF)
            t0 = time.perf_counter()
            r = svc.iterate(args.wing_iters, tol=1e-3)
            t1 = time.perf_counter()
            s_ref = 1.5 * 0.5 * (1.0 + 0.56) * WING_SCALE * WING_SCALE
            cl, cd = gr.lift_drag(svc.force("wall"), mach, P_INF / T_INF, P_INF, s_ref)
            wing["solve_mixed"] = {"iterations": r.iterations, "seconds": t1 - t0, "converged_to_1e-3": bool(r.converged),
                                   "residual_drop_reached": r.ratio,
                                   "normalisation": "residual of iteration 0 (reference semantics)",
                                   "lift_drag": {"C_L": cl, "C_D": cd, "s_ref": s_ref}}
            del svc
        if not args.no_shards:
            sh = shard_leg(gr, e, m, c2o, max(20, args.steps // 4))
            full = dt / args.steps * 1e3
            sh["whole_mesh_ms_per_step"] = full
            sh["kernel_speedup_at_8"] = full / sh["slowest_ms_per_step"]
            wing["shards_1_of_8"] = sh
        out["swept_wing"] = wing
        out["stream_copy_GBps"] = gr.stream_copy_gbps(1 << 30, 20)
        if not args.no_ogrid:
            out["ogrid"] = ogrid_leg(gr, args)
        if not args.no_batch:
            out["batched"] = batched_leg(gr, args)
        if not args.no_warm:
            out["warm_start"] = warm_start_leg(gr, args)
            out["mlp"] = mlp_leg(gr, args)
        if not args.no_nsi:
            out["nsi_cavity"] = nsi_leg(gr, args)
    if not args.no_cpu:
        out["cpu_baseline"] = wing_cpu_baseline(gr, e, mach)
    else:
        out["cpu_baseline"] = None
    print(to_json(out), flush=True)


if __name__ == "__main__":
    main()
