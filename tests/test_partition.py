"""One GPU's share of a mesh (graphite_amd.partition): the range's cells, its boundary faces under
their regions and the cut faces as a halo region.  Connectivity only (host), no GPU."""
import numpy as np
import pytest

from graphite_amd.partition import shard_elements


def _mesh(gr, tets):
    e = gr.Elements.box(6, 5, 4, tets=tets, jitter=0.1 if tets else 0.0, seed=3)
    m = gr.Mesh.build(e)
    c2o, _ = m.renumber_hilbert()
    return e, m, c2o


def _ghosts_by_region(m):
    con = m.connectivity()
    g = con["ghost_cell_starts"]
    return dict(zip(m.region_names(), np.diff(g).tolist()))


@pytest.mark.parametrize("tets", [False, True])
def test_whole_range_rebuilds_the_mesh(gr, tets):
    e, m, c2o = _mesh(gr, tets)
    n = m.n_cells
    sub = gr.Mesh.build(shard_elements(e, m, c2o, 0, n))
    a, b = sub.sizes(), m.sizes()
    assert (a.n_cells, a.n_faces, a.n_ghosts) == (b.n_cells, b.n_faces, b.n_ghosts)
    assert _ghosts_by_region(sub) == _ghosts_by_region(m)
    assert "halo" not in sub.region_names()


@pytest.mark.parametrize("tets", [False, True])
def test_shards_cover_the_mesh_and_cut_faces_become_halo(gr, tets):
    e, m, c2o = _mesh(gr, tets)
    n = m.n_cells
    con = m.connectivity()
    cfs, cc = con["cell_faces_starts"].astype(np.int64), con["cell_connects"].astype(np.int64)
    nodes, kinds, starts, enodes, _ = e.arrays()
    bounds = np.linspace(0, n, 5).astype(int)
    owner = np.searchsorted(bounds, np.arange(n), side="right") - 1
    # interior faces whose cells lie in different shards, each seen from both sides
    cell_of_slot = np.repeat(np.arange(n), np.diff(cfs[:n + 1]))
    nb = cc[:cfs[n]]
    interior = nb < n
    cut_slots = interior & (owner[cell_of_slot] != owner[np.minimum(nb, n - 1)])
    total_cells, boundary = 0, {}
    for s in range(4):
        lo, hi = bounds[s], bounds[s + 1]
        sh = shard_elements(e, m, c2o, lo, hi)
        sub = gr.Mesh.build(sh)
        assert sub.n_cells == hi - lo
        total_cells += sub.n_cells
        g = _ghosts_by_region(sub)
        mine = cut_slots & (owner[cell_of_slot] == s)
        assert g.get("halo", 0) == int(mine.sum())
        for k, v in g.items():
            if k != "halo":
                boundary[k] = boundary.get(k, 0) + v
        # the shard's cells are the range's elements, in the range's order
        _, k2, s2, n2, _ = sh.arrays()
        for c in (0, (hi - lo) // 2, hi - lo - 1):
            o = int(c2o[lo + c])
            assert k2[c] == kinds[o]
            assert np.array_equal(n2[s2[c]:s2[c + 1]], enodes[starts[o]:starts[o + 1]])
    assert total_cells == n
    assert boundary == _ghosts_by_region(m)


def test_range_checked(gr):
    e, m, c2o = _mesh(gr, False)
    with pytest.raises(ValueError):
        shard_elements(e, m, c2o, 5, 5)
    with pytest.raises(ValueError):
        shard_elements(e, m, c2o, 0, m.n_cells + 1)
