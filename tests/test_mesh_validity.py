"""Meshes the scheme cannot run on are rejected after the metrics, by the oracle (CPU) and by
the product's device check (GPU): non-positive volumes, faces that do not separate their two
cell centres (n.d = 0, where the reference's binary64 sign() returns 0 and silently drops the
face from the flux, include/datatypes.cuh:209-211; or a tangled cell), singular least-squares
matrices and degenerate faces.  The reference runs such meshes without a message
(src/mesh.cu:102-143)."""
import numpy as np
import pytest

TET = 10
TRI = 5


def _two_tets(apex_b):
    """Tets (0,1,2,3) and (0,1,2,4) sharing the face (0,1,2) in the plane z = 0; node 3 above."""
    nodes = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0.3, 0.3, 1.0], apex_b], float)
    tris = [(0, 1, 3), (1, 2, 3), (0, 2, 3), (0, 1, 4), (1, 2, 4), (0, 2, 4)]
    return nodes, [0, 1, 2, 3, 0, 1, 2, 4], tris


def _one_tet(nodes):
    nodes = np.asarray(nodes, float)
    return nodes, [0, 1, 2, 3], [(0, 1, 2), (0, 1, 3), (1, 2, 3), (0, 2, 3)]


def _elements(gr, nodes, enodes, tris):
    n_el = len(enodes) // 4
    return gr.Elements.create(nodes, [TET] * n_el, np.arange(0, 4 * n_el + 1, 4), enodes,
                              [("wall", [TRI] * len(tris), np.arange(0, 3 * len(tris) + 1, 3),
                                np.array(tris).reshape(-1))])


CASES = {
    "valid": _two_tets([0.3, 0.3, -1.0]),
    # apex of the second tet on the first tet's side of the shared face: the two cell centres
    # lie on one side of it (tangled mesh)
    "tangled": _two_tets([0.2, 0.2, 0.5]),
    # four coplanar nodes: zero volume, n.d = 0 on every face, singular LSQ matrix
    "flat": _one_tet([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0.3, 0.3, 0.0]]),
    # three collinear nodes: a face of zero area (normal 0/0)
    "degenerate_face": _one_tet([[0, 0, 0], [1, 0, 0], [2, 0, 0], [0, 1, 1]]),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_oracle_rejects_invalid_meshes(gr, orc, case):
    e = _elements(gr, *CASES[case])
    om = orc.OracleMesh(gr.ElementArrays.of(e), validate=False)
    if case == "valid":
        assert om.invalid == (0, 0, 0, 0)
        orc.OracleMesh(gr.ElementArrays.of(e))
        return
    expect = {"tangled": lambda c: c[1] == 2 and c[0] == 0 and c[3] == 0,
              "flat": lambda c: c[0] == 1 and c[1] == 4 and c[2] == 1,
              "degenerate_face": lambda c: c[3] >= 1}[case]
    assert expect(om.invalid), om.invalid
    with pytest.raises(ValueError):
        orc.OracleMesh(gr.ElementArrays.of(e))


@pytest.mark.gpu
@pytest.mark.parametrize("case", sorted(CASES))
def test_device_metrics_reject_invalid_meshes(gr, case):
    print("case: mesh validity", case, flush=True)
    e = _elements(gr, *CASES[case])
    m = gr.Mesh.build(e)
    if case == "valid":
        m.compute_metrics()
        assert m.metrics()["cell_volumes"][:2].min() > 0
        return
    with pytest.raises(gr.GraphiteError) as ei:
        m.compute_metrics()
    assert ei.value.code == gr.GR_ERR_TOPOLOGY
    print(str(ei.value))
    # nothing may run on the rejected metrics
    with pytest.raises(gr.GraphiteError):
        gr.Solver(m, gr.GR_F64)


@pytest.mark.gpu
def test_device_check_accepts_the_benchmark_meshes(gr):
    """The strict separation rule holds on every generated case (checked on the CPU with the
    oracle's metrics too: smallest |n.(x_f - x_c)| / V^(1/3) is 0.0027 on the C-grid)."""
    for e in (gr.Elements.box(20, 20, 20, tets=True, jitter=0.1, seed=2024),
              gr.Elements.ogrid(256, 128, 4, 0.5, 10.0, 1.05, 2 * np.pi * 0.5 / 256),
              gr.Elements.cgrid_naca()):
        m = gr.Mesh.build(e)
        m.renumber_rcm()
        m.compute_metrics()
