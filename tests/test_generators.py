"""Synthetic mesh generators: closed-form counts, positive volumes, watertightness and the
areas/volumes the geometry fixes (metrics from the FP64 oracle)."""
import numpy as np
import pytest

WING_SMALL = dict(n_around=32, n_prism=4, n_tet=5, n_span=6, n_cap=4, radius=12.0)


def wing_counts(ni, npr, nt, ns, nc, n_taper):
    """Below the tip: 2 prisms per wall-layer hex and 6 tetrahedra per outer hex.  Above it, up to
    the transition row k = ns + n_taper: 2 prisms per wall-layer hex, a pyramid and 4
    tetrahedra at j = npr, 2 prisms per outer hex; on the transition row the wall-layer hexes
    are a pyramid and 4 tetrahedra each; above it 2 prisms per hex."""
    below = ns * ni * (2 * npr + 6 * nt)
    taper = n_taper * ni * (2 * npr + 5 + 2 * (nt - 1))
    transition = ni * (5 * npr + 5 + 2 * (nt - 1))
    above = (nc - n_taper - 1) * ni * 2 * (npr + nt)
    return below + taper + transition + above


def taper_rows(kinds, ni, npr):
    """Rows over which the tip tapers, from the pyramid count ni * (n_taper + 1 + npr)."""
    n_pyr = int((np.asarray(kinds) == 14).sum())
    assert n_pyr % ni == 0
    return n_pyr // ni - 1 - npr


def test_wing_counts_and_markers(gr):
    e = gr.Elements.wing(**WING_SMALL)
    m = gr.Mesh.build(e)
    s = m.sizes()
    x, kinds, starts, nodes, markers = e.arrays()
    nta = taper_rows(kinds, 32, 4)
    assert 1 <= nta <= 2
    assert s.n_cells == wing_counts(32, 4, 5, 6, 4, nta)
    assert m.region_names() == ["wall", "farfield", "symmetry"]
    assert set(np.unique(kinds)) == {gr.GR_TET, gr.GR_WEDGE, gr.GR_PYRAMID}


def test_wing_default_size():
    import graphite_amd as gr
    gr.load()
    e = gr.Elements.wing()
    s = gr.Mesh.build(e).sizes()
    nta = taper_rows(e.arrays()[1], 88, 20)
    assert nta == 2
    assert s.n_cells == wing_counts(88, 20, 70, 20, 16, nta) == 1069112


def test_wing_cap_has_no_slivers(gr, orc):
    """The column of prisms above the tip reaches the dome in steps a fair fraction of its
    rings' spacing: its cells' largest/smallest face-area ratio stays below 1e4 (tall slivers
    there, ratios to 3e5, kept the transonic solve from converging)."""
    e = gr.Elements.wing()
    om = orc.OracleMesh(gr.ElementArrays.of(e))
    N = om.n_cells
    cfs, cf, fa = om.cell_faces_starts, om.cell_faces, om.face_areas
    cell = np.repeat(np.arange(N), np.diff(cfs[:N + 1]).astype(np.int64))
    a = fa[cf[:cfs[N]]]
    amax = np.zeros(N)
    np.maximum.at(amax, cell, a)
    amin = np.full(N, np.inf)
    np.minimum.at(amin, cell, a)
    ratio = amax / amin
    above = om.cell_centers[:N, 2] > 1.6     # above the tip (semispan 1.5)
    assert om.cell_volumes[:N].min() > 0
    assert ratio[above].max() < 1e4
    assert np.percentile(ratio[above], 99) < 2e3


def test_wing_geometry(gr, orc):
    R = WING_SMALL["radius"]
    e = gr.Elements.wing(**WING_SMALL)
    om = orc.OracleMesh(gr.ElementArrays.of(e))
    N = om.n_cells
    V = om.cell_volumes[:N]
    assert V.min() > 0
    # closure: the outward n A of a cell's faces sum to zero for planar faces (tetrahedra,
    # exactly); prism side faces between twisted layers (most of all in the ridge over the
    # tip) are non-planar, and the reference's face area (sum of fan-triangle areas) then
    # leaves a residue
    x, kinds, starts, nodes, _ = e.arrays()
    cfs, cf, cc = om.cell_faces_starts, om.cell_faces, om.cell_connects
    fn, fa, fc, xc = om.face_normals, om.face_areas, om.face_centers, om.cell_centers
    worst = {gr.GR_TET: 0.0, gr.GR_WEDGE: 0.0, gr.GR_PYRAMID: 0.0}
    for i in range(N):
        acc = np.zeros(3)
        for q in range(cfs[i], cfs[i + 1]):
            f = cf[q]
            n = fn[f] * np.sign(fn[f] @ (fc[f] - xc[i]))
            acc += n * fa[f]
        r = np.linalg.norm(acc) / fa[cf[cfs[i]:cfs[i + 1]]].sum()
        worst[int(kinds[i])] = max(worst[int(kinds[i])], r)
    assert worst[gr.GR_TET] < 1e-12
    assert worst[gr.GR_WEDGE] < 5e-2
    assert worst[gr.GR_PYRAMID] < 5e-2
    # symmetry plane: planar at z = 0, area of the inscribed 32-gon minus the root section
    r = om.region("symmetry")
    g0, g1 = om.ghost_cell_starts[r], om.ghost_cell_starts[r + 1]
    sym_faces = [cf[cfs[N + g]] for g in range(g0, g1)]
    assert np.abs(fc[sym_faces, 2]).max() < 1e-14
    # = regular 32-gon of radius R minus the root section polygon (nodes of the wall at z = 0,
    # in angular order around the section's centre; the section is convex)
    polygon = 0.5 * 32 * R * R * np.sin(2 * np.pi / 32)
    xyz = x.reshape(-1, 3)
    r_w = om.region("wall")
    wall_nodes = set()
    for g in range(om.ghost_cell_starts[r_w], om.ghost_cell_starts[r_w + 1]):
        f = cf[cfs[N + g]]
        wall_nodes.update(om.face_nodes[om.face_nodes_starts[f]:om.face_nodes_starts[f + 1]].tolist())
    sec = np.array([xyz[n] for n in sorted(wall_nodes) if abs(xyz[n, 2]) < 1e-14])
    assert len(sec) == 32
    ang = np.arctan2(sec[:, 1], sec[:, 0] - 0.5)
    sec = sec[np.argsort(ang)]
    root = 0.5 * abs(np.dot(sec[:, 0], np.roll(sec[:, 1], -1)) - np.dot(sec[:, 1], np.roll(sec[:, 0], -1)))
    assert 0.07 < root < 0.09                 # NACA 0012 section: about 0.685 t c^2 = 0.082
    assert fa[sym_faces].sum() == pytest.approx(polygon - root, rel=1e-13)
    # far field on the sphere
    r = om.region("farfield")
    g0, g1 = om.ghost_cell_starts[r], om.ghost_cell_starts[r + 1]
    ff_nodes = set()
    for g in range(g0, g1):
        f = cf[cfs[N + g]]
        ff_nodes.update(om.face_nodes[om.face_nodes_starts[f]:om.face_nodes_starts[f + 1]].tolist())
    pts = om.nodes.reshape(-1, 3)[sorted(ff_nodes)]
    rad = np.linalg.norm(pts - np.array([0.5, 0.0, 0.0]), axis=1)
    assert np.allclose(rad, R, rtol=1e-12)
    # total volume: inside the hemisphere, within its polyhedral approximation
    hemi = 2.0 / 3.0 * np.pi * R ** 3
    assert 0.95 * hemi < V.sum() < hemi


def test_wing_rejects_bad_parameters(gr):
    with pytest.raises(gr.GraphiteError):
        gr.Elements.wing(n_around=31)
    with pytest.raises(gr.GraphiteError):
        gr.Elements.wing(radius=3.0)
