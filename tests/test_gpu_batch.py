"""Batched solver (instance-inner layout, one wavefront per cell): each instance must evolve
exactly as a single-instance solve with the same far field (binary64, bit for bit), and the
single-instance solver is itself pinned to the FP64 oracle in test_gpu_parity.py."""
import numpy as np
import pytest

from conftest import normwise
from flowcases import P_INF, T_INF, perturbed_state

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def box(gr):
    e = gr.Elements.box(8, 8, 8, tets=True, jitter=0.1, seed=5)
    m = gr.Mesh.build(e)
    m.renumber_rcm()
    m.compute_metrics()
    return m


def machs(n, seed=0):
    rng = np.random.default_rng(seed)
    M = rng.uniform(0.5, 0.85, n)
    a = np.deg2rad(rng.uniform(-2, 6, n))
    return np.stack([M * np.cos(a), M * np.sin(a), np.zeros(n)], axis=1)


BOX_BCS = [("xmin", 1), ("xmax", 1), ("ymin", 0), ("ymax", 0), ("zmin", 1), ("zmax", 0)]


def single(gr, m, mach, n_iter, dtype, state=None, tol=0.0, ref=0.0, corrected=False):
    sv = gr.Solver(m, dtype=dtype, cfl=0.5, corrected=corrected)
    for name, kind in BOX_BCS:
        sv.set_bc(name, kind, mach, P_INF, T_INF)
    if state is None:
        sv.init_freestream(mach, P_INF, T_INF)
    else:
        sv.set_state(*state)
    res = sv.iterate(n_iter, tol=tol, reference=ref)
    return sv, res


@pytest.mark.parametrize("corrected", [False, True])
def test_batch_instances_equal_single_solves(gr, box, corrected):
    m = box
    M = machs(5, seed=1)
    b = gr.Batch(m, 5, gr.GR_F64, cfl=0.5, corrected=corrected)
    for name, kind in BOX_BCS:
        b.set_bc(name, kind)
    b.set_freestream(M, P_INF, T_INF)
    # perturb two instances away from their freestream
    cc = m.metrics()["cell_centers"]
    states = {1: perturbed_state(cc, mach=M[1]), 3: perturbed_state(cc, mach=M[3], amp=0.05)}
    for k, st in states.items():
        b.set_state(k, *st)
    its, conv, ran = b.iterate(15)
    assert list(its) == [15] * 5 and ran == 15
    for k in range(5):
        sv, res = single(gr, m, M[k], 15, gr.GR_F64, state=states.get(k), corrected=corrected)
        for x, y in zip(b.get_state(k), sv.get_state()):
            assert np.array_equal(x, y), k
        # states are bitwise equal; the residual sums add the same cell terms in another order
        assert np.allclose(b.history(k, 15), sv.history(15), rtol=1e-13, atol=0), k


def test_batch_per_instance_convergence_freezes(gr, box):
    m = box
    M = machs(3, seed=2)
    cc = m.metrics()["cell_centers"]
    b = gr.Batch(m, 3, gr.GR_F64, cfl=0.5)
    for name, kind in BOX_BCS:
        b.set_bc(name, kind)
    b.set_freestream(M, P_INF, T_INF)
    amps = [0.02, 0.1, 0.05]
    states = [perturbed_state(cc, mach=M[k], amp=amps[k]) for k in range(3)]
    for k in range(3):
        b.set_state(k, *states[k])
    tol = 0.05
    its, conv, ran = b.iterate(300, tol=tol)
    for k in range(3):
        sv, res = single(gr, m, M[k], 300, gr.GR_F64, state=states[k], tol=tol)
        assert its[k] == res.iterations and conv[k] == bool(res.converged), k
        for x, y in zip(b.get_state(k), sv.get_state()):
            assert np.array_equal(x[:m.n_cells], y[:m.n_cells]), k
    assert ran == its.max()


def test_batch_mixed_precision_close_to_fp64(gr, box):
    m = box
    M = machs(4, seed=3)
    out = {}
    for dt in (gr.GR_F64, gr.GR_F32):
        b = gr.Batch(m, 4, dt, cfl=0.5)
        for name, kind in BOX_BCS:
            b.set_bc(name, kind)
        b.set_freestream(M, P_INF, T_INF)
        b.iterate(40)
        out[dt] = [b.get_state(k) for k in range(4)]
    for k in range(4):
        for x, y in zip(out[gr.GR_F32][k], out[gr.GR_F64][k]):
            # measured on MI355X: 1.1e-4 (binary32 rounding amplified by the limiter switches)
            assert normwise(x, y) < 3e-4


def test_batch_cgrid_64_instances(gr):
    print("case: C5 NACA0012 C-grid 384x96x1, 64 instances", flush=True)
    e = gr.Elements.cgrid_naca()
    m = gr.Mesh.build(e)
    m.renumber_rcm()
    m.compute_metrics()
    M = machs(64, seed=2024)
    b = gr.Batch(m, 64, gr.GR_F64, cfl=0.5)
    bcs = [("airfoil", gr.GR_BC_SLIP), ("farfield", gr.GR_BC_FARFIELD), ("zmin", gr.GR_BC_SLIP),
           ("zmax", gr.GR_BC_SLIP)]
    for name, kind in bcs:
        b.set_bc(name, kind)
    b.set_freestream(M, P_INF, T_INF)
    its, conv, ran = b.iterate(6)
    assert ran == 6
    for k in (0, 37, 63):
        sv = gr.Solver(m, dtype=gr.GR_F64, cfl=0.5)
        for name, kind in bcs:
            sv.set_bc(name, kind, M[k], P_INF, T_INF)
        sv.init_freestream(M[k], P_INF, T_INF)
        sv.iterate(6)
        for x, y in zip(b.get_state(k), sv.get_state()):
            assert np.array_equal(x, y), k
    r = b.residual_norms()
    assert np.all(np.isfinite(r)) and r.shape == (64,)


@pytest.mark.parametrize("corrected", [False, True])
def test_batch_binary32_equals_single_binary32(gr, box, corrected):
    """The batched kernels and the single-flow tile kernels evaluate the same binary32
    arithmetic (gradient rows premultiplied by their limiters in both): bitwise equal states,
    with the reference's transposed momentum reconstruction and with the corrected one."""
    m = box
    M = machs(3, seed=11)
    b = gr.Batch(m, 3, gr.GR_F32, cfl=0.5, corrected=corrected)
    for name, kind in BOX_BCS:
        b.set_bc(name, kind)
    b.set_freestream(M, P_INF, T_INF)
    b.iterate(12)
    for k in range(3):
        sv, res = single(gr, m, M[k], 12, gr.GR_F32, corrected=corrected)
        for x, y in zip(b.get_state(k), sv.get_state()):
            assert np.array_equal(x[:m.n_cells], y[:m.n_cells]), k


def test_batch_profile_leaves_downloads_unchanged(gr, box):
    """The batch's timing runs write its second state buffer: iterate -> profile -> get_state
    must return bitwise what iterate -> get_state returns, ghosts included."""
    m = box
    M = machs(3, seed=4)

    def run(profile):
        b = gr.Batch(m, 3, gr.GR_F64, cfl=0.5)
        for name, kind in BOX_BCS:
            b.set_bc(name, kind)
        b.set_freestream(M, P_INF, T_INF)
        b.set_state(1, *perturbed_state(m.metrics()["cell_centers"], mach=M[1]))
        b.iterate(11)
        if profile:
            b.profile(3)
        return [b.get_state(k) for k in range(3)]

    for sa, sb in zip(run(False), run(True)):
        for x, y in zip(sa, sb):
            assert np.array_equal(x, y)


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_autotuned_and_every_schedule_equal_single_solves(gr, box, dtype):
    """The launch schedule (cells per block 2/4/8, flux occupancy cap, flux instance groups 1/2/4,
    graph length) changes no cell's arithmetic: after autotuning (which leaves the state
    unchanged) and under every fixed schedule, each instance equals its single-flow solve bit for
    bit."""
    print(f"case: batch schedules vs single solves, {dtype}", flush=True)
    m = box
    dt = gr.GR_F64 if dtype == "f64" else gr.GR_F32
    M = machs(4, seed=21)
    ref = [single(gr, m, M[k], 12, dt)[0].get_state() for k in range(4)]

    def batch():
        b = gr.Batch(m, 4, dt, cfl=0.5)
        for name, kind in BOX_BCS:
            b.set_bc(name, kind)
        b.set_freestream(M, P_INF, T_INF)
        return b

    def check(b, tag):
        its, _, ran = b.iterate(12)
        assert list(its) == [12] * 4, tag
        for k in range(4):
            for x, y in zip(b.get_state(k), ref[k]):
                assert np.array_equal(x[:m.n_cells], y[:m.n_cells]), (tag, k)

    b = batch()
    before = [b.get_state(k) for k in range(4)]
    chosen = b.autotune(5)
    print("autotuned schedule:", chosen, flush=True)
    assert chosen == b.schedule
    for k in range(4):                      # autotuning left the state as it was
        for x, y in zip(b.get_state(k), before[k]):
            assert np.array_equal(x, y)
    check(b, "autotuned")
    for cpb in (2, 4, 8):
        for cap in (False, True):
            for groups in ((1, 2, 4) if dtype == "f64" else (1, 2)):
                b = batch()
                b.schedule = {"cells_per_block": cpb, "flux_waves_cap": cap, "graph_iters": 4,
                              "instance_groups": groups}
                check(b, (cpb, cap, groups))
