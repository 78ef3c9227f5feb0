"""Connectivity (cuMeshFromSu2 numbering), .fvm/.su2 I/O, renumbering, colouring -- CPU only.

The product builder (sort-based, graphite_amd/csrc/connectivity.cpp), the oracle (literal
chained hash table, oracle/oracle_mesh.c) and the golden fixtures (pure-Python walk-through,
tests/golden/make_connectivity_golden.py) are three independent statements of
src/meshio.cu:713-1138; all must agree bit for bit.
"""
import glob
import json
import os

import numpy as np
import pytest

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "connectivity_*.json")))
KEYS = ["face_nodes_starts", "face_nodes", "cell_faces_starts", "cell_faces", "cell_connects",
        "ghost_cell_ids", "ghost_cell_starts"]


def elements_from_fixture(gr, c):
    markers = [(m["name"], m["kinds"], m["starts"], m["nodes"]) for m in c["markers"]]
    return gr.Elements.create(np.array(c["nodes"], float), c["kinds"], c["elem_starts"], c["elem_nodes"], markers)


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
def test_golden_fixture_product_and_oracle(gr, orc, path):
    c = json.load(open(path))
    e = elements_from_fixture(gr, c)
    got = gr.Mesh.build(e).connectivity()
    om = orc.OracleMesh(gr.ElementArrays.of(e))
    for k in KEYS:
        exp = np.array(c["expected"][k], np.uint32)
        assert np.array_equal(got[k], exp), f"product {k}"
        assert np.array_equal(getattr(om, k), exp), f"oracle {k}"


MESHES = {
    "hex_3x2x2": lambda gr: gr.Elements.box(3, 2, 2),
    "tets_jitter_4x3x2": lambda gr: gr.Elements.box(4, 3, 2, tets=True, jitter=0.1, seed=3),
    "c1_box_tets_20": lambda gr: gr.Elements.box(20, 20, 20, tets=True, jitter=0.1, seed=2024),
    "ogrid_32x16x2": lambda gr: gr.Elements.ogrid(32, 16, 2, 0.5, 10.0, 1.05, 0.05),
}


@pytest.mark.parametrize("name", list(MESHES))
def test_product_equals_oracle(gr, orc, name):
    e = MESHES[name](gr)
    got = gr.Mesh.build(e).connectivity()
    om = orc.OracleMesh(gr.ElementArrays.of(e))
    for k in KEYS:
        assert np.array_equal(got[k], getattr(om, k)), k


def check_invariants(c, n_cells, n_ghosts):
    cfs, cf, cc = c["cell_faces_starts"], c["cell_faces"], c["cell_connects"]
    nf = len(c["face_nodes_starts"]) - 1
    owners = [[] for _ in range(nf)]
    for i in range(n_cells + n_ghosts):
        for q in range(cfs[i], cfs[i + 1]):
            owners[cf[q]].append((i, cc[q]))
    for f, o in enumerate(owners):
        assert len(o) == 2, f"face {f} has {len(o)} owners"
        (a, na), (b, nb) = o
        assert na == b and nb == a, "cell_connects not symmetric"
    for g in range(n_ghosts):
        gid = n_cells + g
        assert cfs[gid + 1] - cfs[gid] == 1
        assert cc[cfs[gid]] < n_cells


def test_invariants_c1(gr):
    e = gr.Elements.box(20, 20, 20, tets=True, jitter=0.1, seed=2024)
    m = gr.Mesh.build(e)
    s = m.sizes()
    # closed-form counts of the 6-tet box (48k tets, 4800 boundary triangles)
    assert (s.n_cells, s.n_ghosts, s.n_faces, s.n_nodes) == (48000, 4800, 98400, 9261)
    check_invariants(m.connectivity(), s.n_cells, s.n_ghosts)


def test_counts_c2_ogrid(gr):
    e = gr.Elements.ogrid(256, 128, 4, 0.5, 10.0, 1.05, 2 * np.pi * 0.5 / 256)
    s = gr.Mesh.build(e).sizes()
    assert (s.n_cells, s.n_ghosts, s.n_faces, s.n_nodes) == (131072, 67584, 427008, 165120)


def test_unclosed_mesh_is_a_topology_error(gr):
    # a hex box without its markers: boundary faces have one cell (src/meshio.cu:1115-1128)
    x, k, s, n, _ = gr.Elements.box(2, 1, 1).arrays()
    e = gr.Elements.create(x, k, s, n, [])
    with pytest.raises(gr.GraphiteError) as ei:
        gr.Mesh.build(e)
    assert ei.value.code == 2


def test_marker_face_not_in_mesh_is_an_error(gr):
    x, k, s, n, m = gr.Elements.box(1, 1, 1).arrays()
    name, mk, ms, mn = m[0]
    mn = mn.copy()
    mn[0] = 7 if mn[0] != 7 else 6   # no longer a face of the hex
    bad = [(name, mk, ms, mn)] + m[1:]
    e = gr.Elements.create(x, k, s, n, bad)
    with pytest.raises(gr.GraphiteError):
        gr.Mesh.build(e)


def test_unknown_region_is_an_error(gr):
    m = gr.Mesh.build(gr.Elements.box(1, 1, 1))
    assert m.region_id("xmax") == 1
    with pytest.raises(gr.GraphiteError) as ei:
        m.region_id("nope")
    assert ei.value.code == 4


def test_bad_element_kind_rejected(gr):
    with pytest.raises(gr.GraphiteError):
        gr.Elements.create(np.zeros((4, 3)), [11], [0, 4], [0, 1, 2, 3])


@pytest.mark.parametrize("sb", [4, 8])
def test_fvm_round_trip(gr, tmp_path, sb):
    e = gr.Elements.box(3, 2, 2, tets=True, jitter=0.05, seed=5)
    m = gr.Mesh.build(e)
    p = str(tmp_path / "m.fvm")
    m.write_fvm(p, sb)
    m2 = gr.Mesh.read_fvm(p, sb)
    a, b = m.connectivity(), m2.connectivity()
    for k in KEYS:
        assert np.array_equal(a[k], b[k]), k
    if sb == 8:
        assert np.array_equal(a["nodes"], b["nodes"])
    else:
        assert np.array_equal(a["nodes"].astype(np.float32).astype(np.float64), b["nodes"])
    assert m2.region_names() == m.region_names()


def test_fvm_file_layout(gr, tmp_path):
    # header fields of the reference's .fvm writer (src/meshio.cu:1154-1198)
    m = gr.Mesh.build(gr.Elements.box(1, 1, 1))
    p = tmp_path / "one.fvm"
    m.write_fvm(str(p), 4)
    raw = p.read_bytes()
    u32 = np.frombuffer(raw, dtype="<u4")
    assert u32[0] == 8                             # nodes
    off = 1 + 8 * 3                                # 8 nodes x 3 float32
    assert u32[off] == 6 and u32[off + 1] == 24    # faces, face tags
    names = m.region_names()
    assert names == ["xmin", "xmax", "ymin", "ymax", "zmin", "zmax"]


def test_su2_round_trip_and_conversion(gr, tmp_path):
    e = gr.Elements.box(2, 2, 1, tets=True, jitter=0.05, seed=9)
    p = str(tmp_path / "m.su2")
    e.write_su2(p)
    e2 = gr.Elements.read_su2(p)
    a, b = e.arrays(), e2.arrays()
    for x, y in zip(a[:4], b[:4]):
        assert np.array_equal(x, y)
    assert [m[0] for m in a[4]] == [m[0] for m in b[4]]
    fvm = str(tmp_path / "m.fvm")
    gr.Mesh.from_su2(p, fvm, 8)
    c1 = gr.Mesh.read_fvm(fvm, 8).connectivity()
    c2 = gr.Mesh.build(e).connectivity()
    for k in KEYS:
        assert np.array_equal(c1[k], c2[k])


def test_su2_reader_rejects_garbage(gr, tmp_path):
    p = tmp_path / "bad.su2"
    p.write_text("NDIME= 3\nNELEM= 1\n10 0 1 2 3 0\nNPOIN= 2\n0 0 0 0\n")
    with pytest.raises(gr.GraphiteError):
        gr.Elements.read_su2(str(p))


def test_make_1d_matches_reference_layout(gr):
    # cuMeshMake1D, src/mesh.cu:540-634
    m = gr.Mesh.make_1d(4)
    c = m.connectivity()
    assert list(c["cell_faces_starts"]) == [0, 2, 4, 6, 8, 9, 10]
    assert list(c["cell_faces"]) == [0, 1, 1, 2, 2, 3, 3, 4, 0, 4]
    assert list(c["cell_connects"]) == [4, 1, 0, 2, 1, 3, 2, 5, 0, 3]
    assert list(c["ghost_cell_ids"]) == [4, 5] and list(c["ghost_cell_starts"]) == [0, 1, 2]


@pytest.mark.parametrize("method", ["rcm", "hilbert"])
def test_renumbering_is_an_exact_permutation(gr, method):
    e = gr.Elements.box(6, 5, 4, tets=True, jitter=0.1, seed=11)
    m0 = gr.Mesh.build(e)
    c0 = m0.connectivity()
    bw0 = m0.bandwidth()
    m1 = gr.Mesh.build(e)
    c2o, f2o = getattr(m1, "renumber_" + method)()
    c1 = m1.connectivity()
    s = m1.sizes()
    N, NG = s.n_cells, s.n_cells + s.n_ghosts
    assert sorted(c2o[:N]) == list(range(N)) and sorted(c2o[N:]) == list(range(N, NG))
    assert sorted(f2o) == list(range(s.n_faces))
    c_o2n = np.empty(NG, np.uint32)
    c_o2n[c2o] = np.arange(NG, dtype=np.uint32)
    f_o2n = np.empty(s.n_faces, np.uint32)
    f_o2n[f2o] = np.arange(s.n_faces, dtype=np.uint32)
    for new in range(NG):
        old = c2o[new]
        a0, a1 = c0["cell_faces_starts"][old], c0["cell_faces_starts"][old + 1]
        b0, b1 = c1["cell_faces_starts"][new], c1["cell_faces_starts"][new + 1]
        assert np.array_equal(f_o2n[c0["cell_faces"][a0:a1]], c1["cell_faces"][b0:b1])
        assert np.array_equal(c_o2n[c0["cell_connects"][a0:a1]], c1["cell_connects"][b0:b1])
    for new in range(s.n_faces):
        old = f2o[new]
        assert np.array_equal(
            c0["face_nodes"][c0["face_nodes_starts"][old]:c0["face_nodes_starts"][old + 1]],
            c1["face_nodes"][c1["face_nodes_starts"][new]:c1["face_nodes_starts"][new + 1]])
    if method == "rcm":
        assert m1.bandwidth() <= bw0
    # deterministic
    m2 = gr.Mesh.build(e)
    c2o2, f2o2 = getattr(m2, "renumber_" + method)()
    assert np.array_equal(c2o, c2o2) and np.array_equal(f2o, f2o2)


@pytest.mark.parametrize("n", [4, 8])
def test_hilbert_order_walks_face_neighbours(gr, n):
    """On an n^3 box of hexes (n a power of two) the cell centres fall on the curve's own
    2^k lattice, so the Hilbert order visits the cells as the order-k Hilbert curve does:
    every cell is followed by a face neighbour, and the walk starts in a corner.  A wrong
    rotation or reflection step of the key breaks the adjacency."""
    m = gr.Mesh.build(gr.Elements.box(n, n, n))
    c2o, _ = m.renumber_hilbert()
    c = m.connectivity()
    N = m.n_cells
    cfs, cc = c["cell_faces_starts"], c["cell_connects"]
    for i in range(N - 1):
        assert i + 1 in set(cc[cfs[i]:cfs[i + 1]].tolist()), i
    # original box numbering is x fastest: the first cell of the walk is a corner cell
    first = int(c2o[0])
    ijk = (first % n, (first // n) % n, first // (n * n))
    assert all(v in (0, n - 1) for v in ijk)


def test_face_colouring(gr):
    m = gr.Mesh.build(gr.Elements.box(5, 4, 3, tets=True, jitter=0.1, seed=1))
    col, n = m.face_colors()
    c = m.connectivity()
    s = m.sizes()
    cfs, cf = c["cell_faces_starts"], c["cell_faces"]
    for i in range(s.n_cells + s.n_ghosts):
        cols = col[cf[cfs[i]:cfs[i + 1]]]
        assert len(set(cols.tolist())) == len(cols), "two faces of one colour share a cell"
    assert n <= 2 * 4 - 1 + 1
