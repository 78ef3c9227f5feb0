"""Wall force (lift/drag) and wall distance on the device."""
import numpy as np
import pytest

from flowcases import P_INF, T_INF, ogrid_elements, perturbed_state

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cyl(gr):
    m = gr.Mesh.build(gr.Elements.ogrid(64, 16, 2, 0.5, 10.0, 1.1, 0.05))
    m.compute_metrics()
    return m


def test_uniform_pressure_gives_no_force(gr, cyl):
    sv = gr.Solver(cyl, gr.GR_F64)
    sv.init_freestream((0.3, 0.0, 0.0), P_INF, T_INF)
    F = sv.force("wall")
    assert np.abs(F).max() < 1e-13


def test_force_matches_host_sum(gr, cyl):
    m = cyl
    mt, c = m.metrics(), m.connectivity()
    rho, rhou, rhoe = perturbed_state(mt["cell_centers"], mach=(0.3, 0.1, 0.0), amp=0.2)
    sv = gr.Solver(m, gr.GR_F64)
    sv.set_state(rho, rhou, rhoe)
    F = sv.force("wall")
    r = m.region_id("wall")
    N = m.n_cells
    exp = np.zeros(3)
    for g in range(c["ghost_cell_starts"][r], c["ghost_cell_starts"][r + 1]):
        q = c["cell_faces_starts"][N + g]
        i, f = c["cell_connects"][q], c["cell_faces"][q]
        p = 0.4 * (rhoe[i] - 0.5 * rhou[i] @ rhou[i] / rho[i])
        n = mt["face_normals"][f]
        n = n * np.sign(n @ (mt["face_centers"][f] - mt["cell_centers"][i]))
        exp += p * n * mt["face_areas"][f]
    assert np.allclose(F, exp, rtol=1e-12, atol=1e-14)
    cl, cd = gr.lift_drag(F, (0.3, 0.0, 0.0), P_INF / T_INF, P_INF, 1.0)
    assert np.isfinite(cl) and np.isfinite(cd)


def test_batch_force_equals_single(gr, cyl):
    m = cyl
    M = np.array([[0.3, 0.0, 0.0], [0.5, 0.05, 0.0]])
    b = gr.Batch(m, 2, gr.GR_F64)
    for nm, k in (("wall", 0), ("farfield", 1), ("zmin", 0), ("zmax", 0)):
        b.set_bc(nm, k)
    b.set_freestream(M, P_INF, T_INF)
    b.iterate(20)
    Fb = b.force("wall")
    for k in range(2):
        sv = gr.Solver(m, gr.GR_F64)
        for nm, kind in (("wall", 0), ("farfield", 1), ("zmin", 0), ("zmax", 0)):
            sv.set_bc(nm, kind, M[k], P_INF, T_INF)
        sv.init_freestream(M[k], P_INF, T_INF)
        sv.iterate(20)
        assert np.array_equal(Fb[k], sv.force("wall"))


def test_wall_distance_brute_force(gr):
    m = gr.Mesh.build(gr.Elements.box(6, 5, 4, tets=True, jitter=0.1, seed=4))
    m.compute_metrics()
    d = m.wall_distance(["xmin", "ymax"])
    mt, c = m.metrics(), m.connectivity()
    N = m.n_cells
    wf = []
    for nm in ("xmin", "ymax"):
        r = m.region_id(nm)
        for g in range(c["ghost_cell_starts"][r], c["ghost_cell_starts"][r + 1]):
            wf.append(c["cell_faces"][c["cell_faces_starts"][N + g]])
    wc = mt["face_centers"][wf]
    x = mt["cell_centers"][:N]
    exp = np.sqrt(((x[:, None, :] - wc[None, :, :]) ** 2).sum(-1)).min(1)
    assert np.allclose(d, exp, rtol=1e-14, atol=1e-15)
    with pytest.raises(gr.GraphiteError):
        m.wall_distance(["nope"])
