"""The operator library as a device API for user kernels (include/graphite/fvm_ops.hpp; the
reference's fvmOp / fvmOpSetup, include/fvmops.cuh:30-77, and its operators :96-219).

tests/cpp/user_ops.hip is a user's code: it includes the header, gets the mesh and field views
through the C ABI (gr_mesh_get_view, gr_field_get_view) and launches kernels of its own.  One
composes two library operators (must equal the library's gr_field_operator results added on
the host, bit for bit); one is a new operator over for_each_face (closure and the divergence
of the position vector, closed forms)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

SRC = os.path.join(ROOT, "tests", "cpp", "user_ops.hip")
LIB = os.path.join(ROOT, "tests", "cpp", "libuser_ops.so")
HIPCC = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "bin", "hipcc")


def build_user_ops(out=LIB):
    """hipcc for gfx950 against the public headers, linked to libgraphite_hip.so (no GPU needed)."""
    cmd = [HIPCC, "-x", "hip", "-shared", "-fPIC", "-O3", "-std=c++17", "--offload-arch=gfx950",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", f"-I{os.path.join(ROOT, 'include')}",
           SRC, "-o", out + ".tmp", f"-L{os.path.join(ROOT, 'graphite_amd')}", "-lgraphite_hip",
           "-Wl,-rpath,$ORIGIN/../../graphite_amd"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    os.replace(out + ".tmp", out)
    return out


def user_lib(gr):
    gr.load()            # libgraphite_hip.so first: the user library binds to the same instance
    lib = C.CDLL(LIB if os.path.exists(LIB) else build_user_ops())
    lib.user_laplacian_plus_div.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_double)]
    lib.user_face_moments.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
    return lib


def test_header_exports_are_in_the_library(gr):
    """The views the header builds on are exported by the shared library (CPU check)."""
    lib = gr.load()
    for name in ("gr_mesh_get_view", "gr_field_get_view"):
        assert hasattr(lib, name)


def mesh_and_fields(gr, tets):
    e = gr.Elements.box(6, 5, 4, 1.0, 0.8, 0.6, tets=tets, jitter=0.12 if tets else 0.0, seed=21)
    m = gr.Mesh.build(e)
    m.compute_metrics()
    n = m.n_cells + m.sizes().n_ghosts
    rng = np.random.default_rng(5)
    p = gr.Field(m, 1, gr.GR_F64, "p")
    v = gr.Field(m, 3, gr.GR_F64, "v")
    p.upload(rng.uniform(0.5, 1.5, n))
    v.upload(rng.uniform(-1.0, 1.0, 3 * n))
    for f in (p, v):
        f.compute_gradient()
        f.compute_limiter()
        f.gradlim_bc()
    return m, p, v


@pytest.mark.gpu
@pytest.mark.parametrize("tets", [False, True], ids=["hex", "tet"])
def test_user_kernel_composing_operators_matches_library(gr, tets):
    print(f"case: user kernel laplacian(p) + div(v), tets={tets}", flush=True)
    lib = user_lib(gr)
    m, p, v = mesh_and_fields(gr, tets)
    out = np.zeros(m.n_cells)
    rc = lib.user_laplacian_plus_div(m._h, p._h, v._h, out.ctypes.data_as(C.POINTER(C.c_double)))
    assert rc == 0, gr.load().gr_last_error()
    want = p.operator(gr.GR_OP_LAPLACIAN_C) + v.operator(gr.GR_OP_DIV_C)
    assert np.array_equal(out, want), np.abs(out - want).max()


@pytest.mark.gpu
@pytest.mark.parametrize("tets", [False, True], ids=["hex", "tet"])
def test_user_operator_over_for_each_face(gr, tets):
    print(f"case: user operator over for_each_face, tets={tets}", flush=True)
    lib = user_lib(gr)
    m, _, _ = mesh_and_fields(gr, tets)
    out = np.zeros((m.n_cells, 4))
    rc = lib.user_face_moments(m._h, out.ctypes.data_as(C.POINTER(C.c_double)))
    assert rc == 0, gr.load().gr_last_error()
    # closed cells with planar faces: sum n A = 0 and sum (x_f - x_c).n A = 3 V
    # (terms are A/V, up to about 60 on these cells: rounding leaves ~1e-14; a wrong sign or
    # orientation leaves O(10))
    assert np.abs(out[:, :3]).max() < 1e-11
    assert np.abs(out[:, 3] - 3.0).max() < 1e-12
