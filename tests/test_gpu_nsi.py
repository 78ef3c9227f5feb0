"""Incompressible pressure-correction solver on the device against the FP64 oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NU = 1.0 / 400.0


def rel(a, b):
    d = np.linalg.norm(np.asarray(a) - np.asarray(b))
    n = np.linalg.norm(np.asarray(b))
    return d / n if n > 0 else d


def pair(gr, orc, e, walls, dtype=None, dt=0.02, cg_max=40, cg_tol=0.0):
    m = gr.Mesh.build(e)
    m.compute_metrics()
    om = orc.OracleMesh(gr.ElementArrays.of(e))
    ns = gr.Nsi(m, NU, dt, gr.GR_F64 if dtype is None else dtype, cg_max=cg_max, cg_tol=cg_tol)
    for name, u in walls:
        ns.set_wall(name, u)
    ost = orc.NsiState(om)
    ow = [(om.region(name), u) for name, u in walls]
    return m, om, ns, ost, ow


@pytest.mark.parametrize("tets", [False, True])
def test_steps_match_oracle(gr, orc, tets):
    e = gr.Elements.box(6, 5, 4, tets=tets, jitter=0.1 if tets else 0.0, seed=3)
    walls = [("ymax", (1.0, 0.0, 0.0)), ("xmin", (0.0, 0.2, 0.0))]
    m, om, ns, ost, ow = pair(gr, orc, e, walls)
    conts = []
    for _ in range(8):
        c, _ = orc.nsi_step(om, ost, ow, NU, 0.02, 40, 0.0)
        conts.append(c)
    c_dev, its = ns.step(8)
    assert list(its) == [40] * 8
    u, p, F = ns.get_state()
    N = om.n_cells
    print("normwise u, p, F:", rel(u[:N], ost.u[:N]), rel(p[:N], ost.p[:N]), rel(F, ost.F))
    assert rel(u[:N], ost.u[:N]) < 1e-10
    assert rel(p[:N], ost.p[:N]) < 1e-10
    assert rel(F, ost.F) < 1e-10
    assert np.allclose(c_dev, conts, rtol=1e-9, atol=0)


def test_first_step_without_correction_is_exact(gr, orc):
    e = gr.Elements.box(5, 5, 5)
    m, om, ns, ost, ow = pair(gr, orc, e, [("ymax", (1.0, 0.0, 0.0))], cg_max=0)
    orc.nsi_step(om, ost, ow, NU, 0.02, 0, 0.0)
    ns.step(1)
    u, p, F = ns.get_state()
    assert np.array_equal(u[:om.n_cells], ost.u[:om.n_cells])
    assert np.array_equal(F, ost.F)


def test_projection_removes_divergence(gr, orc):
    e = gr.Elements.box(6, 6, 6, tets=True, jitter=0.1, seed=5)
    m, om, ns, ost, ow = pair(gr, orc, e, [("ymax", (1.0, 0.0, 0.0))], cg_max=800, cg_tol=1e-12)
    c, its = ns.step(3)
    assert its.max() < 800
    u, p, F = ns.get_state()
    cfs, cf, cc = om.cell_faces_starts, om.cell_faces, om.cell_connects
    worst = max(abs(sum(F[cf[q]] if i < cc[q] else -F[cf[q]] for q in range(cfs[i], cfs[i + 1])))
                for i in range(om.n_cells))
    assert worst < 1e-9 * np.abs(F).max()


def test_mixed_precision_close_to_fp64(gr, orc):
    e = gr.Elements.box(8, 8, 8)
    m = gr.Mesh.build(e)
    m.compute_metrics()
    out = {}
    for dt in (gr.GR_F64, gr.GR_F32):
        ns = gr.Nsi(m, NU, 0.02, dt, cg_max=100, cg_tol=1e-6)
        ns.set_wall("ymax", (1.0, 0.0, 0.0))
        ns.step(30)
        out[dt] = ns.get_state()
    N = m.n_cells
    eu = rel(out[gr.GR_F32][0][:N], out[gr.GR_F64][0][:N])
    ep = rel(out[gr.GR_F32][1][:N], out[gr.GR_F64][1][:N])
    print("binary32 bulk vs binary64 after 30 steps: u", eu, "p", ep)
    assert eu < 1e-4 and ep < 1e-3


def test_cavity_develops_and_matches_oracle(gr, orc):
    """12^3 lid-driven cavity, Re 400, 150 steps: device binary64 against the oracle, and the
    primary vortex's centre-line profile (u_x < 0 near the middle, u_x -> 1 under the lid)."""
    n = 12
    e = gr.Elements.box(n, n, n)
    m, om, ns, ost, ow = pair(gr, orc, e, [("ymax", (1.0, 0.0, 0.0))], dt=0.02, cg_max=300, cg_tol=1e-10)
    for _ in range(150):
        orc.nsi_step(om, ost, ow, NU, 0.02, 300, 1e-10)
    ns.step(150)
    u, p, F = ns.get_state()
    N = om.n_cells
    print("normwise after 150 steps u, p:", rel(u[:N], ost.u[:N]), rel(p[:N], ost.p[:N]))
    assert rel(u[:N], ost.u[:N]) < 1e-8
    assert rel(p[:N], ost.p[:N]) < 1e-7
    ux = u[:N, 0].reshape(n, n, n)[n // 2, :, n // 2]      # [k, j, i]: x = z = 1/2 column
    assert ux.min() < -0.05 and ux[-1] > 0.3


def test_profile_leaves_the_solve_unchanged(gr):
    """gr_nsi_profile times the step's parts on the current state and restores it: a profile
    between two runs of steps changes nothing, bit for bit."""
    print("case: NSI profile between steps, 8^3 cavity, mixed", flush=True)
    e = gr.Elements.box(8, 8, 8)
    m = gr.Mesh.build(e)
    m.compute_metrics()
    out = []
    for with_profile in (False, True):
        ns = gr.Nsi(m, NU, 0.01, gr.GR_F32, cg_max=100, cg_tol=1e-6)
        ns.set_wall("ymax", (1.0, 0.0, 0.0))
        ns.step(10)
        if with_profile:
            prof = ns.profile(5)
            assert all(v > 0 for v in prof.values()), prof
        c, its = ns.step(5)
        out.append((ns.get_state(), c, its))
    for x, y in zip(out[0][0], out[1][0]):
        assert np.array_equal(x, y)
    assert np.array_equal(out[0][1], out[1][1]) and np.array_equal(out[0][2], out[1][2])
