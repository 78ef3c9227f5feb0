"""Learned warm start of the batched Euler solver on a NACA0012 C-grid.

A 4-layer, 256-wide MLP (bfloat16 matrix-core kernels, binary32 accumulation) maps
(cell centre, wall distance, Mach, angle of attack) to the converged conservative state.  It is
trained on solutions the batched solver itself generates; at run time the library predicts
every instance's field on the device, the plain solver judges each prediction by its residual
(falling back to freestream where it is not better) and then finishes every instance.

Convergence is measured against each instance's *freestream* residual for both cold and warm
starts, so the warm start is not credited with a smaller starting residual.
"""
from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np

from . import GR_BC_FARFIELD, GR_BC_SLIP, GR_F64, Batch, Mesh, Mlp, normalise, warm_norm

P_INF = 1.0
T_INF = 1.0 / 1.4          # rho_inf = 1.4, c_inf = 1
RHO_INF = P_INF / T_INF
CGRID_BCS = (("airfoil", GR_BC_SLIP), ("farfield", GR_BC_FARFIELD), ("zmin", GR_BC_SLIP), ("zmax", GR_BC_SLIP))
LANES = 64


def sample_instances(n: int, seed: int) -> np.ndarray:
    """Far-field Mach vectors: Mach ~ U[0.5, 0.85], angle of attack ~ U[-2, 6] degrees."""
    rng = np.random.default_rng(seed)
    M = rng.uniform(0.5, 0.85, n)
    a = np.deg2rad(rng.uniform(-2.0, 6.0, n))
    return np.stack([M * np.cos(a), M * np.sin(a), np.zeros(n)], axis=1)


def make_batch(mesh: Mesh, machs, dtype: int = GR_F64, cfl: float = 0.5, tune: bool = False) -> Batch:
    """A batch of the C-grid flows at freestream; with tune, its launch schedule is chosen by
    the autotuner (results do not depend on it)."""
    b = Batch(mesh, len(machs), dtype=dtype, cfl=cfl)
    for name, kind in CGRID_BCS:
        b.set_bc(name, kind)
    b.set_freestream(machs, P_INF, T_INF)
    if tune:
        b.autotune(10)
    return b


def wall_distance(mesh: Mesh) -> np.ndarray:
    return mesh.wall_distance(["airfoil"])


def wall_origin(mesh: Mesh, region: str = "airfoil") -> np.ndarray:
    """Area-weighted centroid of a wall region's faces (the feature origin)."""
    mt, c = mesh.metrics(), mesh.connectivity()
    r = mesh.region_id(region)
    N = mesh.n_cells
    faces = [c["cell_faces"][c["cell_faces_starts"][N + g]]
             for g in range(c["ghost_cell_starts"][r], c["ghost_cell_starts"][r + 1])]
    a = mt["face_areas"][faces]
    return (mt["face_centers"][faces] * a[:, None]).sum(0) / a.sum()


@dataclass
class SolveStats:
    iterations: np.ndarray     # per instance
    converged: np.ndarray      # per instance
    batch_iterations: int
    seconds: float


def solve(b: Batch, tol: float, reference, max_iter: int, active=None) -> SolveStats:
    """Iterate to `tol` of `reference`.  An instance that diverges restarts from freestream
    (the plain solver is the fallback) and its iterations before the restart are counted."""
    t0 = time.perf_counter()
    its, conv, div, ran = b.iterate_masked(max_iter, tol=tol, reference=reference, active=active)
    if div.any():
        b.reset_freestream(div)
        its2, conv2, div2, ran2 = b.iterate_masked(max_iter, tol=tol, reference=reference, active=div)
        its = np.where(div, its + its2, its)
        conv = np.where(div, conv2, conv)
        ran += ran2
    return SolveStats(np.asarray(its), np.asarray(conv, bool), int(ran), time.perf_counter() - t0)


def generate_training_set(mesh: Mesh, machs, tol: float, max_iter: int, dtype: int = GR_F64, log=None):
    """Solve the instances in batches of 64 to `tol` (freestream-normalised residual) and
    return the raw rows (x [n*N, 6], t [n*N, 5]) and per-batch solve statistics."""
    wd = wall_distance(mesh)
    origin = wall_origin(mesh)
    xs, ts, stats = [], [], []
    for s in range(0, len(machs), LANES):
        b = make_batch(mesh, machs[s:s + LANES], dtype)
        fs = b.residual_norms()
        st = solve(b, tol, fs, max_iter)
        x, t = b.warm_rows(wd, origin)
        xs.append(x)
        ts.append(t)
        stats.append(st)
        if log:
            log(f"training set: instances {s}..{s + len(machs[s:s + LANES]) - 1} solved in {st.batch_iterations} "
                f"iterations, {st.seconds:.2f} s, converged {int(st.converged.sum())}")
        del b
    return np.concatenate(xs), np.concatenate(ts), origin, stats


def train(x, t, origin, steps: int, batch: int = 32768, lr: float = 2e-3, lr_final: float = 2e-5, seed: int = 0):
    """Train the warm-start network on raw rows; returns (network, normalisation, losses, seconds)."""
    norm = warm_norm(x, t, origin)
    xn, tn = normalise(x, t, norm)
    net = Mlp(6, 5, seed=seed)
    net.set_data(xn, tn)
    t0 = time.perf_counter()
    losses = net.train(batch, steps, lr, lr_final, seed)
    return net, norm, losses, time.perf_counter() - t0


def cold_and_warm(mesh: Mesh, net: Mlp, norm, machs, tol: float, max_iter: int, dtype: int = GR_F64):
    """Solve the instances from freestream and from the network's prediction (batched), both to
    `tol` of the freestream residual.  Returns (cold stats, warm stats, warm-start seconds,
    kept flags, (cold batch, warm batch), freestream residuals)."""
    wd = wall_distance(mesh)
    cold = make_batch(mesh, machs, dtype)
    fs = cold.residual_norms()
    cs = solve(cold, tol, fs, max_iter)
    warm = make_batch(mesh, machs, dtype)
    t0 = time.perf_counter()
    fs_w, used = warm.warm_start(net, wd, norm)
    t_pred = time.perf_counter() - t0
    ws = solve(warm, tol, fs_w, max_iter)
    ws.seconds += t_pred
    return cs, ws, t_pred, used, (cold, warm), fs


def force_agreement(cold: Batch, warm: Batch, machs, region: str, s_ref: float, settle_iter: int,
                    window_iter: int = 0, sample_every: int = 500, log=None) -> dict:
    """Lift and drag of warm- and cold-started solves of the same flows once both have run the
    plain solver for settle_iter further iterations.  The reference's scheme ends in a limit
    cycle on the C-grid (squared residual 1e-6 to 1e-4, lift oscillating by about 0.3%
    peak to peak), so a converged force is the mean over a window of window_iter iterations,
    sampled every sample_every; with window_iter 0 the forces at the end are compared.  A flow
    that diverges restarts from freestream (the plain solver is the fallback) and is counted.
    Lift differences are relative to each flow's mean |C_L| (flows with |C_L| of at least
    0.05), drag differences to each flow's mean C_D.  log(msg), if given, reports progress.
    Returns a dict of the statistics."""
    from . import lift_drag
    n = len(machs)

    def coeffs(b):
        F = b.force(region)
        return np.array([lift_drag(F[k], machs[k], RHO_INF, P_INF, s_ref) for k in range(n)])

    restarted = {"cold": np.zeros(n, bool), "warm": np.zeros(n, bool)}

    def run(tag, b, iters, chunk, samples=None):
        done = 0
        while done < iters:
            step = min(chunk, iters - done)
            div = b.iterate_masked(step)[2]
            if div.any():
                restarted[tag] |= div
                b.reset_freestream(div)
            done += step
            if samples is not None:
                samples.append(coeffs(b))
            if log and done % 20000 == 0:
                log(f"force agreement: {tag} +{done}")

    for tag, b in (("cold", cold), ("warm", warm)):
        run(tag, b, settle_iter, 10000)
    sc, sw = [], []
    if window_iter > 0:
        run("cold", cold, window_iter, sample_every, sc)
        run("warm", warm, window_iter, sample_every, sw)
    else:
        sc, sw = [coeffs(cold)], [coeffs(warm)]
    sc, sw = np.array(sc), np.array(sw)
    mc, mw = sc.mean(0), sw.mean(0)
    cl = np.abs(mc[:, 0])
    lifting = cl >= 0.05
    d_cl = np.abs(mw[:, 0] - mc[:, 0])[lifting] / cl[lifting]
    d_cd = np.abs(mw[:, 1] - mc[:, 1]) / np.abs(mc[:, 1])
    amp = (sc.max(0) - sc.min(0))[:, 0][lifting] / cl[lifting]
    return {"settle_iterations": settle_iter, "window_iterations": window_iter, "samples": len(sc),
            "flows": n, "lifting_flows": int(lifting.sum()),
            "restarted_cold": int(restarted["cold"].sum()), "restarted_warm": int(restarted["warm"].sum()),
            "mean_cl_rel_diff_p50_max": [float(np.median(d_cl)), float(d_cl.max())],
            "mean_cd_rel_diff_p50_max": [float(np.median(d_cd)), float(d_cd.max())],
            "cold_cl_cycle_amplitude_p50_max": [float(np.median(amp)), float(amp.max())]}


def force_history(b: Batch, machs, region: str, s_ref: float, iters: int, every: int, log=None, tag: str = ""):
    """Run the batch `iters` further iterations (frozen flows stay frozen: iterate_masked with no
    tolerance runs every flow), sampling lift and drag every `every` iterations.  A flow that
    diverges restarts from freestream.  Returns (samples [n_samples, n, 2], seconds of solver
    time, restarted flags)."""
    from . import lift_drag
    n = len(machs)
    out, secs = [], 0.0
    restarted = np.zeros(n, bool)
    done = 0
    while done < iters:
        step = min(every, iters - done)
        t0 = time.perf_counter()
        div = b.iterate_masked(step)[2]
        secs += time.perf_counter() - t0
        if div.any():
            restarted |= div
            b.reset_freestream(div)
        done += step
        F = b.force(region)
        out.append([lift_drag(F[k], machs[k], RHO_INF, P_INF, s_ref) for k in range(n)])
        if log and done % 20000 == 0:
            log(f"force history{tag}: +{done}")
    return np.array(out), secs, restarted


def time_to_forces(samples: np.ndarray, ref: np.ndarray, every: int, window: int, tol: float,
                   floor=(0.05, 0.005)) -> np.ndarray:
    """Iterations until a flow's lift and drag, averaged over the trailing `window` iterations,
    are within `tol` of ref [n, 2] relative to max(|ref|, floor) (the floors keep the test
    meaningful for flows with almost no lift) and stay so to the end of the samples; -1 where
    that never happens.  samples [n_samples, n, 2] every `every` iterations from the start."""
    w = max(1, window // every)
    ns, n, _ = samples.shape
    csum = np.cumsum(np.concatenate([np.zeros((1, n, 2)), samples]), axis=0)
    means = (csum[w:] - csum[:-w]) / w                    # mean over samples j-w+1..j, j >= w-1
    scale = np.maximum(np.abs(ref), np.asarray(floor, float)[None, :])
    rel = np.abs(means - ref[None]) / scale[None]
    ok = (rel <= tol).all(axis=2)                          # [ns - w + 1, n]
    out = np.full(n, -1, int)
    for k in range(n):
        bad = np.nonzero(~ok[:, k])[0]
        first = 0 if len(bad) == 0 else bad[-1] + 1       # from here on every window agrees
        if first < ok.shape[0]:
            out[k] = (first + w) * every
    return out
