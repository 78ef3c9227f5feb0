"""VTU output (checked by parsing the XML against the mesh) and case-file validation; the
case driver CLI's usage error."""
import json
import os
import subprocess
import xml.etree.ElementTree as ET

import numpy as np
import pytest

HEX_FACES = [(0, 1, 2, 3), (4, 5, 6, 7), (0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (3, 0, 4, 7)]


def arrays_by_name(root):
    out = {}
    for da in root.iter("DataArray"):
        vals = da.text.split() if da.text else []
        out[da.get("Name") or "Points"] = (da, vals)
    return out


def test_vtu_polyhedra_match_the_mesh(gr, tmp_path):
    e = gr.Elements.box(3, 2, 2)
    m = gr.Mesh.build(e)
    x, kinds, starts, enodes, _ = e.arrays()
    s = m.sizes()
    n = s.n_cells + s.n_ghosts
    rng = np.random.default_rng(0)
    rho, rhou = rng.standard_normal(n), rng.standard_normal((n, 3))
    path = str(tmp_path / "out.vtu")
    gr.write_vtu(m, path, {"rho": rho, "rhou": rhou})
    root = ET.parse(path).getroot()
    piece = root.find("UnstructuredGrid/Piece")
    assert int(piece.get("NumberOfCells")) == s.n_cells and int(piece.get("NumberOfPoints")) == s.n_nodes
    A = arrays_by_name(root)
    assert np.array_equal(np.array(A["rho"][1], float), rho[:s.n_cells])          # 17 digits: exact
    assert A["rhou"][0].get("NumberOfComponents") == "3"
    assert np.array_equal(np.array(A["rhou"][1], float).reshape(-1, 3), rhou[:s.n_cells])
    pts = np.array(A["Points"][1], float).reshape(-1, 3)
    assert np.array_equal(pts, x.reshape(-1, 3))
    conn = np.array(A["connectivity"][1], int)
    offs = np.array(A["offsets"][1], int)
    assert set(A["types"][1]) == {"42"}
    faces = np.array(A["faces"][1], int)
    foffs = np.array(A["faceoffsets"][1], int)
    # cells are the hexes in element order; nodes = the element's nodes, sorted
    lo, flo = 0, 0
    for c in range(s.n_cells):
        el = enodes[starts[c]:starts[c + 1]]
        assert list(conn[lo:offs[c]]) == sorted(el)
        lo = offs[c]
        stream = faces[flo:foffs[c]]
        assert stream[0] == 6
        got, k = [], 1
        for _ in range(6):
            cnt = stream[k]
            got.append(frozenset(stream[k + 1:k + 1 + cnt]))
            k += 1 + cnt
        assert k == len(stream)
        assert set(got) == {frozenset(el[list(f)]) for f in HEX_FACES}
        flo = foffs[c]


def base_case(mesh_path):
    return {"mesh": mesh_path,
            "init": {"mach": {"x": 0.5, "y": 0.0, "z": 0.0}, "pressure": 1.0, "temperature": 0.7142857142857143},
            "bcs": {"xmin": {"type": "farfield", "mach": {"x": 0.5, "y": 0.0, "z": 0.0}, "pressure": 1.0,
                             "temperature": 0.7142857142857143},
                    "xmax": {"type": "slip"}},
            "controls": {"cfl": 0.5, "time": "steady", "tolerance": 1e-3, "max_iterations": 50},
            "save": {"file": "out.vtu"}}


@pytest.mark.parametrize("mutate,ok", [
    (lambda c: None, True),
    (lambda c: c.pop("mesh"), False),
    (lambda c: c["init"].pop("pressure"), False),
    (lambda c: c["bcs"]["xmin"].update(type="wall"), False),
    (lambda c: c["bcs"]["xmin"].pop("mach"), False),
    (lambda c: c["controls"].update(time="unsteady"), False),
    (lambda c: c["controls"].update(precision="f16"), False),
    (lambda c: c["controls"].update(precision="f32"), True),
])
def test_case_validation(gr, tmp_path, mutate, ok):
    case = base_case("mesh.su2")
    mutate(case)
    p = tmp_path / "case.json"
    p.write_text(json.dumps(case))
    if ok:
        gr.check_case(str(p))
    else:
        with pytest.raises(gr.GraphiteError) as ei:
            gr.check_case(str(p))
        assert ei.value.code == 3


def test_case_syntax_error(gr, tmp_path):
    p = tmp_path / "bad.json"
    p.write_text('{"mesh": "a.su2", "init": {')
    with pytest.raises(gr.GraphiteError) as ei:
        gr.check_case(str(p))
    assert ei.value.code == 3 and "json" in str(ei.value)


def test_cli_usage(gr):
    exe = os.path.join(os.path.dirname(gr.__file__), "graphite")
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 1 and "usage: graphite" in r.stdout
