"""gr_euler_residual_noalloc (the residual entry point that allocates nothing: its staging buffer
is a file-scope device array) against gr_euler_residual, through the C ABI: bitwise equal on two
calls in a row, the second on a smaller mesh, in both precisions; meshes above its stated
maximum are rejected."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

P_INF, T_INF = 1.0, 1.0 / 1.4
MAX_CELLS = 838860          # GR_RESIDUAL_NOALLOC_MAX_CELLS


def _state(gr, m, dtype, mach, seed):
    """Fields with a smooth perturbed state, boundary conditions applied and gradients/limiters
    prepared (the inputs gr_euler_residual reads)."""
    s = m.sizes()
    n = s.n_cells + s.n_ghosts
    rng = np.random.default_rng(seed)
    rho = 1.4 * (1.0 + 0.05 * rng.uniform(-1, 1, n))
    rhou = np.outer(rho, mach) * (1.0 + 0.05 * rng.uniform(-1, 1, (n, 1)))
    rhoe = P_INF / 0.4 * (1.0 + 0.05 * rng.uniform(-1, 1, n)) + 0.5 * (rhou ** 2).sum(axis=1) / rho
    f = [gr.Field(m, nc, dtype, name) for nc, name in ((1, "rho"), (3, "rhou"), (1, "rhoe"))]
    for x, v in zip(f, (rho, rhou, rhoe)):
        x.upload(v)
    for name in m.region_names():
        gr.euler_bc_farfield(m, *f, mach, P_INF, T_INF, name)
    gr.euler_step_prepare(m, *f)
    return f


def _residuals(gr, m, f):
    L = gr.load()
    N = m.n_cells
    a, b = np.zeros(5 * N), np.zeros(5 * N)
    args = [m._h] + [x._h for x in f]
    assert L.gr_euler_residual(*args, a.ctypes.data_as(C.POINTER(C.c_double))) == 0, L.gr_last_error()
    rc = L.gr_euler_residual_noalloc(*args, b.ctypes.data_as(C.POINTER(C.c_double)))
    assert rc == 0, L.gr_last_error()
    return a, b


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_noalloc_residual_equals_allocating_one(gr, dtype):
    dt = gr.GR_F64 if dtype == "f64" else gr.GR_F32
    mach = (0.5, 0.1, 0.0)
    meshes = []
    for n, seed in ((9, 1), (5, 2)):           # second call: a smaller mesh
        e = gr.Elements.box(n, n, n, tets=True, jitter=0.1, seed=seed)
        m = gr.Mesh.build(e)
        m.compute_metrics()
        meshes.append(m)
    for k, m in enumerate(meshes):
        print(f"case: noalloc residual call {k + 1}, {dtype}, {m.n_cells} cells", flush=True)
        f = _state(gr, m, dt, mach, seed=10 + k)
        a, b = _residuals(gr, m, f)
        assert np.abs(a).max() > 0
        assert np.array_equal(a, b), np.abs(a - b).max()
        # the state is left as it was (no update)
        a2, b2 = _residuals(gr, m, f)
        assert np.array_equal(a2, a) and np.array_equal(b2, b)


def test_noalloc_residual_rejects_meshes_above_its_maximum(gr):
    print("case: noalloc residual on a mesh above its maximum", flush=True)
    L = gr.load()
    m = gr.Mesh.make_1d(MAX_CELLS + 1)
    m.compute_metrics()
    f = [gr.Field(m, nc, gr.GR_F64, name) for nc, name in ((1, "rho"), (3, "rhou"), (1, "rhoe"))]
    out = np.zeros(5 * (MAX_CELLS + 1))
    rc = L.gr_euler_residual_noalloc(m._h, *[x._h for x in f], out.ctypes.data_as(C.POINTER(C.c_double)))
    assert rc == gr.GR_ERR_INVALID
    assert b"maximum" in L.gr_last_error()
