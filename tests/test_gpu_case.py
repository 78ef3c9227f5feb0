"""The case driver end to end on the device: su2 -> fvm -> solve -> VTU, through the library
and through the graphite executable."""
import json
import os
import subprocess
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from flowcases import P_INF, T_INF

pytestmark = pytest.mark.gpu


def make_case(gr, tmp_path, tol=1e-2):
    e = gr.Elements.box(6, 5, 4, tets=True, jitter=0.1, seed=3)
    su2 = str(tmp_path / "box.su2")
    e.write_su2(su2)
    ff = {"type": "farfield", "mach": {"x": 0.4, "y": 0.1, "z": 0.0}, "pressure": P_INF, "temperature": T_INF}
    case = {"mesh": su2,
            "init": {"mach": {"x": 0.3, "y": 0.0, "z": 0.0}, "pressure": P_INF, "temperature": T_INF},
            "bcs": {"xmin": ff, "xmax": ff, "ymin": {"type": "slip"}, "ymax": {"type": "slip"}, "zmin": ff,
                    "zmax": ff},
            "controls": {"cfl": 0.5, "time": "steady", "tolerance": tol, "max_iterations": 5000},
            "save": {"file": str(tmp_path / "sol.vtu")}}
    p = tmp_path / "case.json"
    p.write_text(json.dumps(case))
    return str(p), case


def test_case_runs_and_writes_vtu(gr, tmp_path):
    path, case = make_case(gr, tmp_path)
    r = gr.run_case(path)
    assert r.converged and r.ratio < 1e-2
    root = ET.parse(case["save"]["file"]).getroot()
    rho = [da for da in root.iter("DataArray") if da.get("Name") == "rho"][0]
    vals = np.array(rho.text.split(), float)
    assert len(vals) == int(root.find("UnstructuredGrid/Piece").get("NumberOfCells"))
    assert np.isfinite(vals).all() and abs(vals.mean() - 1.4) < 0.2
    # the same case through the solver API gives the same iteration count
    m = gr.Mesh.read_fvm(case["mesh"][:-3] + "fvm")
    m.renumber_rcm()
    m.compute_metrics()
    sv = gr.Solver(m, gr.GR_F64, cfl=0.5)
    for name, b in case["bcs"].items():
        mach = b.get("mach", case["init"]["mach"])
        sv.set_bc(name, gr.GR_BC_FARFIELD if b["type"] == "farfield" else gr.GR_BC_SLIP,
                  (mach["x"], mach["y"], mach["z"]), b.get("pressure", P_INF), b.get("temperature", T_INF))
    sv.init_freestream((0.3, 0.0, 0.0), P_INF, T_INF)
    r2 = sv.iterate(5000, tol=1e-2)
    assert r2.iterations == r.iterations


def test_cli_runs_a_case(gr, tmp_path):
    path, case = make_case(gr, tmp_path)
    exe = os.path.join(os.path.dirname(gr.__file__), "graphite")
    r = subprocess.run([exe, path], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "converged" in r.stdout and os.path.exists(case["save"]["file"])
