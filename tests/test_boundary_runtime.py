"""The drop-in boundary at run time: a C++ program written against the reference's call surface
(include/graphite/graphite.hpp: cuMeshFromSu2, cuMeshFromFile, FvmFieldNew, FvmEuler*,
FvmFieldReduceSum), following PhysicsEuler's set-up and step protocol
(src/models/euler.cu:20-56, :99-175), is compiled, linked against libgraphite_hip.so and run;
its residual history and final state must match the FP64 oracle to 1e-12 (normwise)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, normwise

SRC = os.path.join(ROOT, "tests", "cpp", "physics_euler.cpp")
EXE = os.path.join(ROOT, "tests", "cpp", "physics_euler")
LIBDIR = os.path.join(ROOT, "graphite_amd")


def build_program(out=EXE):
    """Compile the program (g++ only, no GPU needed); the in-tree build makes it beforehand."""
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", f"-I{os.path.join(ROOT, 'include')}", SRC, "-o", out + ".tmp",
           f"-L{LIBDIR}", "-lgraphite_hip", f"-Wl,-rpath,{LIBDIR}"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    os.replace(out + ".tmp", out)
    return out


def program():
    return EXE if os.path.exists(EXE) else build_program()


def test_program_links_and_fails_loudly_without_a_device(gr, tmp_path):
    """Compiles and links against the shared library; with no device the first device call
    raises graphite_error and the program exits 2 with the library's message."""
    if gr.device_count() > 0:
        pytest.skip("a device is present (the GPU test runs the program)")
    exe = build_program(str(tmp_path / "physics_euler"))
    e = gr.Elements.box(3, 3, 3, tets=True)
    su2 = tmp_path / "box.su2"
    e.write_su2(str(su2))
    r = subprocess.run([exe, str(su2), str(tmp_path / "box.fvm"), "2", "0.5", "0.5", "0", "0",
                        str(tmp_path / "h.txt"), str(tmp_path / "s.bin"), "xmin:farfield"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 2, (r.returncode, r.stdout, r.stderr)
    assert "graphite error 5" in r.stderr


@pytest.mark.gpu
def test_reference_style_program_matches_oracle(gr, orc, tmp_path):
    print("case: PhysicsEuler-style C++ program on an 8^3 6-tet box, 30 iterations", flush=True)
    e = gr.Elements.box(8, 8, 8, tets=True, jitter=0.1, seed=11)
    su2 = str(tmp_path / "box.su2")
    e.write_su2(su2)
    mach = (0.5, 0.1, 0.0)
    names = ["xmin", "xmax", "ymin", "ymax", "zmin", "zmax"]
    kinds = ["farfield", "farfield", "slip", "farfield", "slip", "slip"]
    n_iter, cfl = 30, 0.5
    hist_path, state_path = str(tmp_path / "hist.txt"), str(tmp_path / "state.bin")
    r = subprocess.run([program(), su2, str(tmp_path / "box.fvm"), str(n_iter), str(cfl), *map(str, mach),
                        hist_path, state_path] + [f"{n}:{k}" for n, k in zip(names, kinds)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "binding_checks 1 1" in r.stdout
    hist = np.loadtxt(hist_path)
    state = np.fromfile(state_path, dtype=np.float64).reshape(-1, 5)

    om = orc.OracleMesh(gr.ElementArrays.of(gr.Elements.read_su2(su2)))
    assert state.shape[0] == om.n_cells + om.n_ghosts
    s = orc.EulerState.zeros(om.n_cells + om.n_ghosts)
    P, T = 1.0, 1.0 / 1.4
    orc.euler_init(om, s, mach, P, T)
    bcs = [(om.region(n), orc.BC_FARFIELD if k == "farfield" else orc.BC_SLIP, mach, P, T)
           for n, k in zip(names, kinds)]
    n, ohist = orc.iterate(om, s, bcs, cfl, n_iter, threads=8)
    assert n == n_iter
    assert normwise(hist[:, 0], ohist) <= 1e-12
    assert normwise(hist[:, 1], ohist / ohist[0]) <= 1e-12
    N = om.n_cells
    assert normwise(state[:N, 0], s.rho[:N]) <= 1e-12
    assert normwise(state[:N, 1:4], s.rhou[:N]) <= 1e-12
    assert normwise(state[:N, 4], s.rhoe[:N]) <= 1e-12
